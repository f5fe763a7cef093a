"""Batched solves (flutas_batch_*, one persistent launch over a work queue
of every pass of every problem): every problem of an uneven, mixed-BC
batch must equal its own single-problem FP64 solve to 1e-13 relative
max-norm (the batched kernels run the same device bodies, with other
workgroup sizes and, after autotuning, other tile and plane-group sizes:
the compiler may contract the transform arithmetic differently), also on
a second replay and after new fields are bound to the captured graph; one
problem is also checked against the oracle.  The refined batch (FP32
transforms, FP64 residuals) must bring every problem to a relative
residual <= 1e-10, checked independently, and agree with the oracle to
1e-8 gauge-free."""
import numpy as np
import pytest

from cases import gauge_free, interior, relmax, z_metrics

PROBLEMS = [
    # (ng, lengths, cbc, alpha)
    ((64, 32, 40), (2 * np.pi, np.pi, 1.0), "PPPPNN", 1.5),
    ((32, 96, 18), (1.0, 1.0, 1.0), "NNNNNN", 0.0),
    ((128, 48, 24), (1.0, 2.0, 1.0), "PPPPDD", 1.2),
    ((256, 64, 10), (1.0, 1.0, 1.0), "PPPPPP", 0.0),
    ((48, 192, 34), (1.0, 1.0, 2.0), "NNPPPP", 0.0),
    ((96, 256, 4), (1.0, 1.0, 1.0), "DDNNNN", 0.0),
    ((192, 128, 70), (1.0, 1.0, 1.0), "PPPPNN", 1.8),
    # ND / DN pairs (DCT-IV / DST-IV tiles)
    ((64, 48, 20), (1.0, 1.0, 1.0), "NDPPNN", 1.3),
    ((32, 64, 16), (1.0, 1.0, 1.0), "PPDNDD", 0.0),
]


def _setup(oracle, seed, compatible=False):
    """The problems and seeded N(0,1) right-hand sides; compatible: the
    dz-weighted mean removed where the BCs make the operator singular (no
    D face), so that the residual can vanish."""
    problems, rhs = [], []
    rng = np.random.default_rng(seed)
    for ng, lengths, cbc, alpha in PROBLEMS:
        _, dzf, dzci, dzfi, _, _ = z_metrics(oracle, ng[2], alpha, lengths[2])
        dli = tuple(n / l for n, l in zip(ng, lengths))
        g = rng.standard_normal(ng)
        if compatible and "D" not in cbc:
            wz = dzf[1:ng[2] + 1][None, None, :]
            g -= np.sum(g * wz) / (np.sum(wz) * ng[0] * ng[1])
        f = np.zeros(tuple(n + 2 for n in ng), order="F")
        f[1:-1, 1:-1, 1:-1] = g
        problems.append((ng, dli, dzci, dzfi, cbc))
        rhs.append(f)
    return problems, rhs


def _single(api, problems, rhs):
    out = []
    for (ng, dli, dzci, dzfi, cbc), f in zip(problems, rhs):
        plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc)
        fd = api.DeviceArray.from_host(f)
        pd = api.DeviceArray(fd.shape)
        pd.zero()
        plan.solve_into(fd, pd)
        out.append(pd.to_host())
        plan.destroy()
    return out


@pytest.mark.gpu
def test_batch_matches_single_solves(api, oracle):
    problems, rhs = _setup(oracle, 11)
    ref = _single(api, problems, rhs)
    print(f"batch of {len(problems)} problems", flush=True)
    batch = api.BatchPlan(problems)
    fs = [api.DeviceArray.from_host(f) for f in rhs]
    ps = [api.DeviceArray(f.shape) for f in fs]
    for p in ps:
        p.zero()
    batch.set_fields(fs, ps)
    batch.solve()
    first = [p.to_host() for p in ps]
    for q, (p, r) in enumerate(zip(first, ref)):
        assert relmax(interior(p), interior(r)) <= 1e-13, f"problem {q}"
    # a second replay of the captured graph gives the same bits
    batch.solve()
    for q, (p, r) in enumerate(zip(ps, first)):
        np.testing.assert_array_equal(interior(p.to_host()), interior(r), err_msg=f"replay {q}")
    # the oracle on the first problem
    ng, dli, dzci, dzfi, cbc = problems[0]
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc)
    pr = rhs[0].copy(order="F")
    s.solve(pr)
    assert relmax(gauge_free(interior(ps[0].to_host())), gauge_free(interior(pr))) <= 1e-12
    batch.destroy()


@pytest.mark.gpu
def test_batch_tuned_and_rebound(api, oracle):
    problems, rhs = _setup(oracle, 12)
    batch = api.BatchPlan(problems)
    fs = [api.DeviceArray.from_host(f) for f in rhs]
    ps = [api.DeviceArray(f.shape) for f in fs]
    batch.set_fields(fs, ps)
    batch.tune(reps=2)
    th, gh = batch.tuning()
    assert set(th) <= {0, 1, 2} and set(gh) <= {0, 1, 2}
    wg, items, per_cu, lds = batch.launch_info()
    print(f"tuned: tiles {th}, groups {gh}, {wg} workgroups ({per_cu}/CU) over {items} items",
          flush=True)
    assert 1 <= per_cu and 1 <= wg <= items
    batch.solve()
    ref = _single(api, problems, rhs)
    for q, (p, r) in enumerate(zip(ps, ref)):
        assert relmax(interior(p.to_host()), interior(r)) <= 1e-13, f"tuned {q}"
    # new right-hand sides in new buffers, same captured graph
    _, rhs2 = _setup(oracle, 13)
    fs2 = [api.DeviceArray.from_host(f) for f in rhs2]
    ps2 = [api.DeviceArray(f.shape) for f in fs2]
    batch.set_fields(fs2, ps2)
    batch.solve()
    ref2 = _single(api, problems, rhs2)
    for q, (p, r) in enumerate(zip(ps2, ref2)):
        assert relmax(interior(p.to_host()), interior(r)) <= 1e-13, f"rebound {q}"
    batch.destroy()


@pytest.mark.gpu
def test_batch_rejects_unsupported(api, oracle):
    _, _, dzci, dzfi, _, _ = z_metrics(oracle, 16, 0.0, 1.0)
    with pytest.raises(api.FlutasError) as e:
        api.BatchPlan([((512, 64, 16), (1.0, 1.0, 1.0), dzci, dzfi, "PPPPNN")])
    assert e.value.code == api.FLUTAS_EUNSUPPORTED


def _refined_check(api, oracle, batch, problems, rhs, fs, ps, label):
    iters, rel = batch.stats()
    worst = 0.0
    for q, ((ng, dli, dzci, dzfi, cbc), f, fd, pd) in enumerate(zip(problems, rhs, fs, ps)):
        assert 1 <= iters[q] <= 4 and 0.0 <= rel[q] <= 1e-10, (label, q, iters[q], rel[q])
        # the residual measured apart from the batch, with a single FP64 plan
        plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc)
        check = plan.residual(fd, pd)
        plan.destroy()
        assert check <= 1e-10, (label, q, check)
        s = oracle.Solver(ng, dli, dzci, dzfi, cbc)
        ref = f.copy(order="F")
        s.solve(ref)
        err = relmax(gauge_free(interior(pd.to_host())), gauge_free(interior(ref)))
        worst = max(worst, err)
        assert err <= 1e-8, (label, q, err)
    print(f"{label}: iterations {iters}, worst oracle difference {worst:.2e}", flush=True)


@pytest.mark.gpu
def test_batch_refined_against_oracle(api, oracle):
    problems, rhs = _setup(oracle, 14, compatible=True)
    print(f"refined batch of {len(problems)} problems", flush=True)
    batch = api.BatchPlan(problems, precision=api.FP32_REFINE, tol=1e-10)
    fs = [api.DeviceArray.from_host(f) for f in rhs]
    ps = [api.DeviceArray(f.shape) for f in fs]
    batch.set_fields(fs, ps)
    batch.solve()
    _refined_check(api, oracle, batch, problems, rhs, fs, ps, "refined")
    # a replay gives the same bits; then tuned and rebound to new fields
    first = [p.to_host() for p in ps]
    batch.solve()
    for q, (p, r) in enumerate(zip(ps, first)):
        np.testing.assert_array_equal(p.to_host(), r, err_msg=f"refined replay {q}")
    batch.tune(reps=1)
    _, rhs2 = _setup(oracle, 15, compatible=True)
    fs2 = [api.DeviceArray.from_host(f) for f in rhs2]
    ps2 = [api.DeviceArray(f.shape) for f in fs2]
    batch.set_fields(fs2, ps2)
    batch.solve()
    _refined_check(api, oracle, batch, problems, rhs2, fs2, ps2, "refined, tuned, rebound")
    batch.destroy()


@pytest.mark.gpu
def test_batch_refined_nonfinite(api, oracle):
    """A NaN in one right-hand side is reported (ENONFINITE), not iterated."""
    problems, rhs = _setup(oracle, 16, compatible=True)
    bad = rhs[3].copy(order="F")
    bad[5, 6, 7] = np.nan
    batch = api.BatchPlan(problems, precision=api.FP32_REFINE, tol=1e-10)
    fs = [api.DeviceArray.from_host(f) for f in rhs[:3] + [bad] + rhs[4:]]
    ps = [api.DeviceArray(f.shape) for f in fs]
    batch.set_fields(fs, ps)
    with pytest.raises(api.FlutasError) as e:
        batch.solve()
    assert e.value.code == api.FLUTAS_ENONFINITE
    # the batch stays usable: clean right-hand sides solve
    fs[3].upload(rhs[3])
    batch.solve()
    _refined_check(api, oracle, batch, problems, rhs, fs, ps, "after NaN")
    batch.destroy()
