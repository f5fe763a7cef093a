"""Graphs of library calls (flutas_graph_*): the fused projection step
(fillps_rhs, FP64 solve, correc_update_bc, bounduvw, boundp) recorded
once and replayed must give the eager step's fields bit for bit on every
replay, with eager library work and synchronisations between the replays;
a capture that meets a synchronising call fails cleanly and leaves the
library usable."""
import numpy as np
import pytest

from cases import z_metrics

pytestmark = pytest.mark.gpu


def _setup(api, oracle, ng, seed):
    nx, ny, nz = ng
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, 1.4, 1.0)
    lengths = (2 * np.pi, np.pi, 1.0)
    dli = tuple(n / l for n, l in zip(ng, lengths))
    rng = np.random.default_rng(seed)
    sh = (nx + 2, ny + 2, nz + 2)
    host = {k: np.asfortranarray(rng.uniform(-1, 1, sh)) for k in ("u", "v", "w")}
    host["p"] = np.asfortranarray(rng.standard_normal(sh))
    return dzc, dzf, dzci, dzfi, dli, host


def _step(api, plan, ng, dli, dzc, dzf, dzfi_d, dzci_d, d, cbc):
    """The fused step of main__single_phase.f90:705-747 with the pressure
    buffer p and the increment buffer s (whose halo boundp sets at the
    end, so that the recording holds every kind of call)."""
    nx, ny, nz = ng
    dl = tuple(1.0 / x for x in dli)
    rx, ry, rz = plan.rhsb_device()
    dt = 1e-3
    api.fillps_rhs(nx, ny, nz, 1, 1, dli[0], dli[1], dli[2], dzfi_d, 1 / dt, 1.0,
                   d["u"], d["v"], d["w"], "ccc", cbc, rx, ry, rz, d["s"])
    plan.solve_async(d["s"])
    api.correc_update_bc(nx, ny, nz, 1, 1, dli[0], dli[1], dli[2], dzci_d, dt, 1.0,
                         d["s"], d["p"], d["p"], d["u"], d["v"], d["w"], cbc, (0.0,) * 6, dl,
                         dzc[0], dzc[nz])
    cbcvel = "PPPPDD" * 3
    api.bounduvw(cbcvel, ng, (0.0,) * 18, 1, 1, [0] * 6, dl, dzc, dzf, d["u"], d["v"], d["w"])
    api.boundp(cbc, ng, (0.0,) * 6, dl, dzc[0], dzc[nz], d["s"])


def test_step_graph_replays_match_eager(api, oracle):
    ng, cbc = (64, 48, 40), "PPPPNN"
    dzc, dzf, dzci, dzfi, dli, host = _setup(api, oracle, ng, 3)
    D = api.DeviceArray.from_host
    dzfi_d, dzci_d = D(dzfi), D(dzci)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc)
    d = {k: D(v) for k, v in host.items()}
    d["s"] = api.DeviceArray(host["p"].shape)
    d["s"].zero()
    print(f"eager step {ng} {cbc}", flush=True)
    _step(api, plan, ng, dli, dzc, dzf, dzfi_d, dzci_d, d, cbc)
    api.synchronize()
    ref = {k: d[k].to_host() for k in ("u", "v", "w", "p")}
    # record the same step on the same buffers
    for k, v in host.items():
        d[k].upload(v)
    with api.Graph() as g:
        _step(api, plan, ng, dli, dzc, dzf, dzfi_d, dzci_d, d, cbc)
    # recording ran nothing
    np.testing.assert_array_equal(d["p"].to_host(), host["p"])
    other = {k: D(v) for k, v in host.items()}
    other["s"] = api.DeviceArray(host["p"].shape)
    for rep in range(3):
        for k, v in host.items():
            d[k].upload(v)
        print(f"replay {rep}", flush=True)
        g.launch()
        api.synchronize()
        for k in ("u", "v", "w", "p"):
            np.testing.assert_array_equal(d[k].to_host(), ref[k], err_msg=f"replay {rep} {k}")
        # eager library work between the replays
        _step(api, plan, ng, dli, dzc, dzf, dzfi_d, dzci_d, other, cbc)
        api.synchronize()
    g.destroy()
    plan.destroy()


def test_capture_of_synchronous_call_fails_cleanly(api, oracle):
    ng, cbc = (32, 24, 16), "PPPPNN"
    dzc, dzf, dzci, dzfi, dli, host = _setup(api, oracle, ng, 4)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc)
    p = api.DeviceArray.from_host(host["p"])
    with pytest.raises(api.FlutasError):
        with api.Graph():
            plan.solve(p)  # synchronises: not recordable
    # the library still works after the failed capture
    q = api.DeviceArray.from_host(host["p"])
    plan.solve(q)
    assert np.all(np.isfinite(q.to_host()))
    plan.destroy()
