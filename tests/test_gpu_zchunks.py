"""z-chunked slab passes (flutas_psolve_set_zchunks): the planes of each
x-slab eliminated in chunks by separate workgroups and joined by a reduced
system per column.  Against the C FP64 oracle (src/solver_cpu.f90:20-223)
on the same seeded right-hand sides:

  * FP64, every chunk count 2..8 that the grid allows (3 and 5 included, so
    chunk boundaries fall at every plane-group multiple): <= 1e-12 relative
    max-norm, and <= 1e-13 against the whole-slab walk of the same plan;
  * non-periodic z with the pinned singular column (NN), DD, a face-centred
    z with a D top face (nz - 1 planes in the system, src/solver_cpu.f90:
    71-77), periodic z (cyclic closure), the 512-point blocked work field;
  * FP32 + FP64 refinement with chunks: relative residual <= 1e-10 and
    <= 1e-9 against the oracle;
  * the tuner times every count 1..8 the grid allows, picks one of them
    and leaves a correct solution.
"""
import numpy as np
import pytest

from cases import gauge_free, interior, relmax, z_metrics

pytestmark = pytest.mark.gpu

CASES = [
    # (ng, lengths, cbc, c_or_f, alpha)
    ((64, 32, 96), (2 * np.pi, np.pi, 1.0), "PPPPNN", "ccc", 1.5),
    ((32, 64, 80), (1.0, 1.0, 1.0), "NNNNNN", "ccc", 0.0),
    ((48, 32, 70), (1.0, 1.0, 1.0), "PPPPDD", "ccc", 1.2),
    ((32, 48, 64), (1.0, 1.0, 1.0), "DDNNND", "ccc", 1.0),
    ((32, 32, 66), (1.0, 1.0, 1.0), "PPPPDD", "ccf", 1.3),   # D top face: nz - 1 planes
    ((64, 32, 96), (1.0, 1.0, 2.0), "PPPPPP", "ccc", 0.0),   # cyclic z
    ((32, 64, 50), (1.0, 1.0, 1.0), "NNDDPP", "ccc", 0.0),   # cyclic, partial groups
    ((64, 512, 40), (1.0, 1.0, 1.0), "PPPPNN", "ccc", 1.5),  # blocked W (512-point y)
]


def _oracle(oracle, ng, lengths, cbc, c_or_f, alpha, seed):
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, ng[2], alpha, lengths[2])
    dli = tuple(n / l for n, l in zip(ng, lengths))
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(ng)
    singular = all(c in "PN" for c in cbc)
    if singular:
        w = dzf[1:ng[2] + 1][None, None, :]
        f = f - np.sum(f * w) / (np.sum(w) * ng[0] * ng[1])
    fh = np.zeros(tuple(n + 2 for n in ng), order="F")
    fh[1:-1, 1:-1, 1:-1] = f
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc, c_or_f=c_or_f)
    pref = fh.copy(order="F")
    s.solve(pref, pin_gauge=True)
    return fh, pref, dzci, dzfi, dli, singular


@pytest.mark.parametrize("ng,lengths,cbc,c_or_f,alpha", CASES)
def test_chunked_fp64_solve_matches_oracle(api, oracle, ng, lengths, cbc, c_or_f, alpha):
    fh, pref, dzci, dzfi, dli, singular = _oracle(oracle, ng, lengths, cbc, c_or_f, alpha,
                                                  sum(ng) + 7)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc, c_or_f=c_or_f)
    f = api.DeviceArray.from_host(fh)
    p = api.DeviceArray(fh.shape)
    assert plan.set_zchunks(1) == 1
    print(f"whole slabs {ng} {cbc} {c_or_f}", flush=True)
    plan.solve_into(f, p)
    whole = p.to_host()
    assert relmax(interior(whole), interior(pref)) <= 1e-12
    seen = set()
    for n in range(2, 9):
        got_n = plan.set_zchunks(n)
        assert 1 <= got_n <= n
        if got_n in seen or got_n == 1:
            continue
        seen.add(got_n)
        print(f"z chunks {got_n} {ng} {cbc} {c_or_f}", flush=True)
        p.zero()
        plan.solve_into(f, p)
        got = p.to_host()
        err = relmax(interior(got), interior(pref))
        assert err <= 1e-12, (got_n, err)
        assert relmax(interior(got), interior(whole)) <= 1e-13, got_n
        if singular:
            assert relmax(gauge_free(interior(got)), gauge_free(interior(pref))) <= 1e-12
        # halo untouched
        np.testing.assert_array_equal(got[0], 0.0)
    assert len(seen) >= 3
    plan.destroy()


@pytest.mark.parametrize("ng,lengths,cbc,c_or_f,alpha", [CASES[0], CASES[2], CASES[5]])
def test_chunked_refined_solve(api, oracle, ng, lengths, cbc, c_or_f, alpha):
    fh, pref, dzci, dzfi, dli, singular = _oracle(oracle, ng, lengths, cbc, c_or_f, alpha, 5)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc, c_or_f=c_or_f,
                           precision=api.FP32_REFINE, tol=1e-10)
    f = api.DeviceArray.from_host(fh)
    p = api.DeviceArray(fh.shape)
    for n in (1, 4):
        plan.set_zchunks(n)
        print(f"refined z chunks {plan.zchunks} {ng} {cbc}", flush=True)
        p.zero()
        plan.solve_into(f, p)
        it, rr = plan.stats()
        assert rr <= 1e-10, (n, it, rr)
        got = interior(p.to_host())
        ref = interior(pref)
        if singular:
            got, ref = gauge_free(got), gauge_free(ref)
        assert relmax(got, ref) <= 1e-9
    plan.destroy()


def test_tuner_keeps_a_timed_count(api, oracle):
    ng, lengths, cbc, c_or_f, alpha = CASES[0]
    fh, pref, dzci, dzfi, dli, _ = _oracle(oracle, ng, lengths, cbc, c_or_f, alpha, 9)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc)
    f = api.DeviceArray.from_host(fh)
    p = api.DeviceArray(fh.shape)
    print("tune", flush=True)
    chosen, ms = plan.tune(f, p, reps=3)
    assert chosen in ms and 1 in ms and all(v > 0 for v in ms.values())
    # every count the grid allows is timed, not only powers of two (96
    # planes in groups of 16: 3 chunks of 32 planes)
    assert 3 in ms and set(ms) <= set(range(1, 9))
    assert plan.zchunks == chosen
    assert relmax(interior(p.to_host()), interior(pref)) <= 1e-12
    plan.solve_into(f, p)
    assert relmax(interior(p.to_host()), interior(pref)) <= 1e-12
    with pytest.raises(api.FlutasError):
        plan.set_zchunks(9)
    plan.destroy()
