"""FP32 fields (FluTAS built with _SINGLE_PRECISION, rp = FP32,
src/types.f90:6-12): the _f32 entry points against the FP64 oracle on the
same inputs rounded to FP32, to FP32 accuracy (stencils 1e-5, the solve
1e-4 of max|p| and a relative residual below 1e-5 of the FP64 operator),
on the fused four-pass pipeline and on the line pipeline."""
import ctypes as C

import numpy as np
import pytest

from cases import gauge_free, interior, relmax, z_metrics

pytestmark = pytest.mark.gpu

F = C.c_float


def _fp(api, a):
    return C.cast(C.c_void_p(a.ptr), C.POINTER(C.c_float))


def _r32(a):
    return np.asfortranarray(a.astype(np.float32))


@pytest.mark.parametrize("ng,cbc,alpha", [((32, 24, 16), "PPPPNN", 1.5),   # fused
                                          ((64, 32, 20), "NNDDPP", 0.0),   # fused, periodic z
                                          ((30, 20, 12), "DDNNNN", 1.2),   # line pipeline
                                          ((24, 16, 10), "NDPPDD", 0.0)])  # REDFT11 lines
def test_solve_f32_against_oracle(api, oracle, ng, cbc, alpha):
    nx, ny, nz = ng
    lengths = (2 * np.pi, np.pi, 1.0)
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, alpha, lengths[2])
    dli = tuple(n / l for n, l in zip(ng, lengths))
    rng = np.random.default_rng(sum(ng))
    g = rng.standard_normal(ng)
    if "D" not in cbc:
        wz = dzf[1:nz + 1][None, None, :]
        g -= np.sum(g * wz) / (np.sum(wz) * nx * ny)
    f = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    f[1:-1, 1:-1, 1:-1] = g
    f = _r32(f).astype(np.float64)  # the FP32 right-hand side, exactly
    print(f"solve_f32 {ng} {cbc} alpha={alpha}", flush=True)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc)
    pd = api.DeviceArray.from_host(_r32(f))
    rc = api.lib.flutas_psolve_solve_f32(C.c_void_p(plan._h), _fp(api, pd))
    assert rc == 0, api.lib.flutas_last_error().decode()
    got = pd.to_host().astype(np.float64)
    # the FP64 residual of the FP32 solution
    fd, gd = api.DeviceArray.from_host(f), api.DeviceArray.from_host(got)
    rel = plan.residual(fd, gd)
    plan.destroy()
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc)
    ref = f.copy(order="F")
    s.solve(ref)
    err = relmax(gauge_free(interior(got)), gauge_free(interior(ref)))
    print(f"  relres {rel:.2e}, oracle difference {err:.2e}", flush=True)
    assert rel <= 1e-5 and err <= 1e-4


@pytest.mark.parametrize("cc", [True, False])
def test_stencils_f32_against_oracle(api, oracle, cc):
    """fillps, correc, pres_tw_src, pres_sp_src, chkdiv, boundp (halo 1 and
    2), updt_rhs_b and bounduvw on FP32 fields."""
    ng = (21, 14, 9)
    nx, ny, nz = ng
    nh_d, nh_u = 2, 2
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, 1.3, 1.0, nh_d)
    dli = (nx / 1.1, ny / 0.7, float(nz))
    dl = tuple(1.0 / d for d in dli)
    rng = np.random.default_rng(12)
    sh_u = tuple(n + 2 * nh_u for n in ng)
    sh_p = tuple(n + 2 for n in ng)
    u, v, w = (_r32(rng.uniform(-1, 1, sh_u)).astype(np.float64) for _ in range(3))
    p, pold = (_r32(rng.standard_normal(sh_p)).astype(np.float64) for _ in range(2))
    rho = _r32(1.0 + 9.0 * rng.uniform(0, 1, sh_p)).astype(np.float64)
    D = api.DeviceArray.from_host
    dzfi_d, dzci_d = D(_r32(dzfi)), D(_r32(dzci))
    dzfi, dzci = (_r32(a).astype(np.float64) for a in (dzfi, dzci))
    lib = api.lib
    # fillps
    pr = np.zeros(sh_p, order="F")
    oracle.fillps(ng, nh_d, nh_u, dli, dzfi, 1 / 0.3, 2.5, u, v, w, pr)
    ud, vd, wd, pd = D(_r32(u)), D(_r32(v)), D(_r32(w)), D(np.zeros(sh_p, np.float32, order="F"))
    assert lib.flutas_fillps_f32(nx, ny, nz, nh_d, nh_u, F(dli[0]), F(dli[1]), F(dli[2]),
                                 _fp(api, dzfi_d), F(1 / 0.3), F(2.5), _fp(api, ud),
                                 _fp(api, vd), _fp(api, wd), _fp(api, pd)) == 0
    assert relmax(interior(pd.to_host().astype(np.float64)), interior(pr)) <= 1e-5
    # correc
    ur, vr, wr = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
    oracle.correc(ng, nh_d, nh_u, dli, dzci, 0.3, 2.5, p, ur, vr, wr, rho=rho, const_coeff=cc)
    ud, vd, wd, pd, rd = D(_r32(u)), D(_r32(v)), D(_r32(w)), D(_r32(p)), D(_r32(rho))
    assert lib.flutas_correc_f32(nx, ny, nz, nh_d, nh_u, F(dli[0]), F(dli[1]), F(dli[2]),
                                 _fp(api, dzci_d), F(0.3), F(2.5), _fp(api, pd), _fp(api, ud),
                                 _fp(api, vd), _fp(api, wd), _fp(api, rd), 1 if cc else 0) == 0
    for a, b in ((ud, ur), (vd, vr), (wd, wr)):
        assert relmax(a.to_host().astype(np.float64), b) <= 1e-5
    # pres_tw_src and pres_sp_src
    ur, vr, wr = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
    oracle.pres_tw_src(ng, dli, nh_d, nh_u, dzci, 1 / 2.5, 0.3, 0.2, p, pold, rho, ur, vr, wr,
                       const_coeff=cc)
    ud, vd, wd, pod = D(_r32(u)), D(_r32(v)), D(_r32(w)), D(_r32(pold))
    assert lib.flutas_pres_tw_src_f32(nx, ny, nz, F(dli[0]), F(dli[1]), F(dli[2]), nh_d, nh_u,
                                      _fp(api, dzci_d), F(1 / 2.5), F(0.3), F(0.2), _fp(api, pd),
                                      _fp(api, pod), _fp(api, rd), _fp(api, ud), _fp(api, vd),
                                      _fp(api, wd), 1 if cc else 0) == 0
    for a, b in ((ud, ur), (vd, vr), (wd, wr)):
        assert relmax(a.to_host().astype(np.float64), b) <= 1e-5
    ur, vr, wr = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
    oracle.pres_sp_src(ng, 0.3, dli, nh_d, nh_u, dzci, 1 / 2.5, pold, ur, vr, wr)
    ud, vd, wd = D(_r32(u)), D(_r32(v)), D(_r32(w))
    assert lib.flutas_pres_sp_src_f32(nx, ny, nz, F(0.3), F(dli[0]), F(dli[1]), F(dli[2]), nh_d,
                                      nh_u, _fp(api, dzci_d), F(1 / 2.5), _fp(api, pod),
                                      _fp(api, ud), _fp(api, vd), _fp(api, wd)) == 0
    for a, b in ((ud, ur), (vd, vr), (wd, wr)):
        assert relmax(a.to_host().astype(np.float64), b) <= 1e-5
    # chkdiv
    tr, mr = oracle.chkdiv(ng, dli, nh_d, nh_u, dzfi, u, v, w)
    tot, mx = F(), F()
    ud, vd, wd = D(_r32(u)), D(_r32(v)), D(_r32(w))
    assert lib.flutas_chkdiv_f32(nx, ny, nz, F(dli[0]), F(dli[1]), F(dli[2]), nh_d, nh_u,
                                 _fp(api, dzfi_d), _fp(api, ud), _fp(api, vd), _fp(api, wd),
                                 C.byref(tot), C.byref(mx)) == 0
    assert mx.value == pytest.approx(mr, rel=1e-5)
    assert abs(tot.value - tr) <= 1e-5 * mr * nx * ny * nz
    # boundp with halo 1 and nonzero D/N values
    cbc = "DNNDNN"
    bc = (0.2, -0.1, 0.3, 0.4, 0.7, -0.3)
    pr = p.copy(order="F")
    oracle.boundp(cbc, ng, bc, nh_d, dl, dzc, dzf, pr)
    pd = D(_r32(p))
    bcv, dlv = (F * 6)(*bc), (F * 3)(*dl)
    ngv = (C.c_int * 3)(*ng)
    assert lib.flutas_boundp_f32(cbc.encode(), ngv, bcv, dlv, F(dzc[nh_d - 1]),
                                 F(dzc[nz + nh_d - 1]), _fp(api, pd)) == 0
    assert relmax(pd.to_host().astype(np.float64), pr) <= 1e-5
    # bounduvw: periodic x, walls in y and z
    cbcvel = "PPDDDD" * 3
    bcvel = np.zeros(18)
    ur, vr, wr = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
    oracle.bounduvw(cbcvel, ng, bcvel, nh_d, nh_u, [0] * 6, dl, dzc, dzf, ur, vr, wr)
    ud, vd, wd = D(_r32(u)), D(_r32(v)), D(_r32(w))
    dzc32, dzf32 = np.ascontiguousarray(dzc, np.float32), np.ascontiguousarray(dzf, np.float32)
    assert lib.flutas_bounduvw_f32(cbcvel.encode(), ngv, (F * 18)(*bcvel), nh_d, nh_u,
                                   (C.c_int * 6)(*[0] * 6), dlv,
                                   dzc32.ctypes.data_as(C.POINTER(F)),
                                   dzf32.ctypes.data_as(C.POINTER(F)),
                                   _fp(api, ud), _fp(api, vd), _fp(api, wd)) == 0
    for a, b in ((ud, ur), (vd, vr), (wd, wr)):
        assert relmax(a.to_host().astype(np.float64), b) <= 1e-5
