"""The timing entry points of a plan (flutas_psolve_profile,
flutas_psolve_time_stage): they report positive durations for every stage
of the four-pass FP64 pipeline, reject bad arguments and paths they do not
cover, and leave the plan solving as before (a solve after them gives the
same field bit for bit)."""
import numpy as np
import pytest

from cases import z_metrics

pytestmark = pytest.mark.gpu


def _plan(api, oracle, ng, cbc, alpha, precision=None):
    nx, ny, nz = ng
    _, _, dzci, dzfi, _, _ = z_metrics(oracle, nz, alpha, 1.0)
    dli = (nx / (2 * np.pi), ny / np.pi, float(nz))
    kw = {} if precision is None else {"precision": precision}
    return api.PoissonPlan(ng, dli, dzci, dzfi, cbc, **kw)


def test_time_stage_and_profile_leave_the_solve_unchanged(api, oracle):
    ng, cbc = (128, 64, 40), "PPPPNN"
    plan = _plan(api, oracle, ng, cbc, 1.5)
    rng = np.random.default_rng(11)
    sh = tuple(v + 2 for v in ng)
    f_h = np.asfortranarray(rng.standard_normal(sh))
    f = api.DeviceArray.from_host(f_h)
    p = api.DeviceArray(sh)
    plan.solve_into(f, p)
    ref = p.to_host()
    print(f"time_stage {ng} {cbc}", flush=True)
    for stage in range(4):
        ms = plan.time_stage(f, p, stage, reps=4)
        assert 0.0 < ms < 100.0, (stage, ms)
    prof = plan.profile(f, p, 3)
    assert len(prof) >= 4
    assert all(ms > 0.0 for _, ms in prof)
    # the input is untouched and a new solve gives the same field
    np.testing.assert_array_equal(f.to_host(), f_h)
    plan.solve_into(f, p)
    np.testing.assert_array_equal(p.to_host(), ref)
    for stage, reps in ((4, 1), (-1, 1), (0, 0), (0, 17)):
        with pytest.raises(api.FlutasError):
            plan.time_stage(f, p, stage, reps=reps)
    plan.destroy()


def test_time_stage_rejects_the_refined_path(api, oracle):
    ng, cbc = (64, 32, 24), "PPPPNN"
    plan = _plan(api, oracle, ng, cbc, 0.0, precision=api.FP32_REFINE)
    sh = tuple(v + 2 for v in ng)
    f = api.DeviceArray.from_host(np.asfortranarray(np.random.default_rng(2).standard_normal(sh)))
    p = api.DeviceArray(sh)
    with pytest.raises(api.FlutasError):
        plan.time_stage(f, p, 0)
    plan.destroy()
