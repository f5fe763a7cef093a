"""GPU parity of the HIP path against the C FP64 oracle on the same seeded
inputs, through the C-ABI (flutas_amd.api).

Tolerances (written per test):
  * transforms: <= 1e-13 relative max-norm against golden scipy vectors;
  * FP64 pressure solve: <= 1e-12 relative max-norm against the oracle with
    the same gauge pin (the z pivots are bit-identical by construction);
  * stencil kernels (fillps, correc, pres_*_src, boundp, updt_rhs_b): the
    same operation order with contraction off, so bit-identical;
  * chkdiv after a projection: divmax <= 1e-12 of the pre-projection value.
"""
import os

import numpy as np
import pytest

from cases import gauge_free, interior, relmax, z_metrics

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "r2r_golden.npz")


@pytest.mark.parametrize("kind", ["R2HC", "HC2R", "REDFT10", "REDFT01", "RODFT10", "RODFT01",
                                  "REDFT11", "RODFT11", "REDFT00", "RODFT00"])
def test_fft_lines_match_golden(api, kind):
    gold = np.load(GOLD)
    for n in gold["sizes"]:
        # DCT-I / DST-I lines may have any length (REDFT00 needs two points)
        if (n % 2 and kind not in ("REDFT00", "RODFT00")) or (kind == "REDFT00" and n < 2):
            continue
        x = gold[f"x_{n}"]
        ref = gold[f"{kind}_{n}"]
        # three copies of the line: contiguous batch
        data = np.concatenate([x, 2 * x, -x])
        d = api.DeviceArray.from_host(data)
        api.fft_lines(kind, d, n, 3)
        got = d.to_host().reshape(3, n)
        scale = np.max(np.abs(ref))
        for c, f in enumerate((1.0, 2.0, -1.0)):
            err = np.max(np.abs(got[c] - f * ref)) / (abs(f) * scale)
            assert err < 1e-13, (kind, n, c, err)


def test_fft_lines_strided(api, oracle):
    """Strided lines (the y-direction geometry) against the oracle."""
    rng = np.random.default_rng(1)
    nx, ny, nz = 12, 48, 3
    a = rng.standard_normal((nx, ny, nz))
    flat = np.asfortranarray(a).ravel(order="F")
    for kind in ("R2HC", "REDFT10", "RODFT01"):
        d = api.DeviceArray.from_host(flat.copy())
        for k in range(nz):  # lines i at plane k: stride nx, dist 1
            sub = api.DeviceArray.from_host(flat[k * nx * ny:(k + 1) * nx * ny].copy())
            api.fft_lines(kind, sub, ny, nx, stride=nx, dist=1)
            got = sub.to_host().reshape((nx, ny), order="F")
            for i in range(nx):
                ref = oracle.r2r(kind, a[i, :, k])
                assert relmax(got[i], ref) < 1e-13


SOLVE_CASES = [
    # (ng, lengths, cbc, alpha)
    ((32, 24, 40), (1.0, 1.3, 0.9), "PPPPNN", 0.0),
    ((32, 24, 40), (2 * np.pi, np.pi, 1.0), "PPPPNN", 1.5),
    ((16, 20, 24), (1.0, 1.0, 1.0), "NNNNNN", 0.0),
    ((16, 20, 24), (1.0, 1.0, 1.0), "NNNNNN", 2.0),
    ((24, 16, 32), (1.0, 1.0, 2.0), "PPPPPP", 0.0),
    ((24, 16, 32), (1.0, 1.0, 1.0), "PPPPDD", 1.2),
    ((20, 12, 16), (1.0, 1.0, 1.0), "DDNNPP", 0.0),
    ((12, 20, 18), (1.0, 0.5, 1.0), "NNDDNN", 1.0),
    ((48, 96, 30), (1.0, 1.0, 1.0), "PPPPNN", 1.5),   # radix-3 lengths
    ((40, 10, 14), (1.0, 1.0, 1.0), "DDPPDD", 0.0),   # radix-5 lengths
    ((14, 22, 12), (1.0, 1.0, 1.0), "PPNNNN", 0.0),   # generic radix 7, 11
    # four-pass fused pipeline (compiled lengths in x and y)
    ((64, 32, 40), (2 * np.pi, np.pi, 1.0), "PPPPNN", 1.5),
    ((32, 64, 26), (1.0, 1.0, 1.0), "NNNNNN", 0.0),
    ((96, 48, 20), (1.0, 1.0, 1.0), "PPPPDD", 1.2),
    ((128, 192, 18), (1.0, 2.0, 1.0), "DDNNND", 0.0),
    ((256, 64, 34), (1.0, 1.0, 1.0), "NNPPDN", 1.1),
    ((64, 64, 100), (2 * np.pi, np.pi, 1.0), "PPPPNN", 1.5),  # many plane groups
    # fused periodic z (cyclic closure of gaussel_periodic in P2)
    ((64, 32, 40), (1.0, 1.0, 2.0), "PPPPPP", 0.0),
    ((32, 64, 26), (1.0, 1.0, 1.0), "PPPPPP", 1.3),
    ((64, 96, 24), (1.0, 1.0, 1.0), "NNDDPP", 0.0),
    ((128, 64, 100), (1.0, 1.0, 2.0), "PPPPPP", 0.0),  # partial last group
    ((32, 32, 4), (1.0, 1.0, 1.0), "PPPPPP", 0.0),     # smallest even cyclic system
    # 512-point lines (the radix-8 plan; 512-thread backward slab pass)
    ((512, 64, 16), (2 * np.pi, np.pi, 1.0), "PPPPNN", 1.5),
    ((48, 512, 12), (1.0, 1.0, 1.0), "NNNNNN", 0.0),
    ((64, 512, 10), (1.0, 1.0, 1.0), "DDDDPP", 0.0),
    # edge sizes: the smallest grid, nz = 2, 768- and 1024-point lines
    ((2, 2, 2), (1.0, 1.0, 1.0), "NNNNNN", 0.0),
    ((4, 2, 6), (1.0, 1.0, 1.0), "PPPPDD", 0.0),
    ((6, 4, 2), (1.0, 1.0, 1.0), "PPNNNN", 0.0),
    ((768, 32, 6), (2 * np.pi, np.pi, 1.0), "PPPPNN", 1.5),
    ((32, 1024, 4), (1.0, 1.0, 1.0), "NNPPNN", 0.0),
    # 1024-point x-lines (512-thread FP64 tiles) with R2HC and RODFT kinds,
    # several periodic-z plane groups with 1024-point y-lines, and
    # 2048-point x-lines (the largest compiled plan)
    ((1024, 32, 6), (2 * np.pi, np.pi, 1.0), "PPPPNN", 1.5),
    ((1024, 32, 6), (1.0, 1.0, 1.0), "DDPPNN", 0.0),
    ((32, 1024, 12), (1.0, 1.0, 1.0), "PPPPPP", 0.0),
    ((2048, 32, 4), (1.0, 1.0, 1.0), "NNPPNN", 0.0),
    # radix-5 lengths on the fused pipeline (80, 160, 320, 640 points)
    ((80, 160, 12), (2 * np.pi, np.pi, 1.0), "PPPPNN", 1.5),
    ((320, 80, 8), (1.0, 1.0, 1.0), "NNNNNN", 0.0),
    ((640, 32, 6), (1.0, 1.0, 1.0), "DDPPNN", 0.0),
    ((32, 640, 8), (1.0, 1.0, 2.0), "PPPPPP", 0.0),
    ((160, 320, 10), (1.0, 1.0, 1.0), "NNDDDD", 1.2),
    # two radix-5 passes (100, 200, 400, 800 points)
    ((200, 100, 12), (2 * np.pi, np.pi, 1.0), "PPPPNN", 1.5),
    ((400, 32, 8), (1.0, 1.0, 1.0), "NNNNNN", 0.0),
    ((32, 400, 8), (1.0, 1.0, 2.0), "PPPPPP", 0.0),
    ((800, 32, 6), (1.0, 1.0, 1.0), "DDPPNN", 0.0),
    ((100, 200, 10), (1.0, 1.0, 1.0), "NDDNNN", 1.2),
    # Neumann-Dirichlet pairs in x or y (REDFT11 / RODFT11, line pipeline)
    ((32, 24, 20), (1.0, 1.0, 1.0), "NDPPNN", 1.3),
    ((24, 32, 16), (1.0, 2.0, 1.0), "PPDNDD", 0.0),
    ((20, 36, 12), (1.0, 1.0, 1.0), "DNNDPP", 0.0),
    # the same kinds on the fused pipeline (DCT-IV / DST-IV tiles): x and y,
    # periodic z, radix-5 lengths, 512-point y-lines in the blocked field,
    # and z chunks (64 x-slabs, 100 planes)
    ((64, 32, 20), (1.0, 1.0, 1.0), "NDPPNN", 1.3),
    ((32, 64, 16), (1.0, 2.0, 1.0), "PPDNDD", 0.0),
    ((64, 48, 24), (1.0, 1.0, 1.0), "DNNDPP", 0.0),
    ((80, 160, 12), (1.0, 1.0, 1.0), "NDDNNN", 0.0),
    ((32, 512, 12), (1.0, 1.0, 1.0), "PPNDNN", 1.5),
    ((512, 32, 10), (1.0, 1.0, 1.0), "DNPPDD", 0.0),
    ((64, 64, 100), (1.0, 1.0, 1.0), "NDPPNN", 1.5),
]


def _oracle_solve(oracle, ng, lengths, cbc, alpha, f_halo, c_or_f="ccc"):
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, ng[2], alpha, lengths[2])
    dli = tuple(n / l for n, l in zip(ng, lengths))
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc, c_or_f=c_or_f)
    p = f_halo.copy(order="F")
    s.solve(p, pin_gauge=True)
    return s, p, dzci, dzfi, dzf, dli


def _rhs(ng, dzf, singular, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(ng)
    if singular:
        w = dzf[1:ng[2] + 1][None, None, :]
        f = f - np.sum(f * w) / (np.sum(w) * ng[0] * ng[1])
    p = np.zeros(tuple(n + 2 for n in ng), order="F")
    p[1:-1, 1:-1, 1:-1] = f
    return p


@pytest.mark.parametrize("ng,lengths,cbc,alpha", SOLVE_CASES)
def test_fp64_solve_matches_oracle(api, oracle, ng, lengths, cbc, alpha):
    _, dzf, _, _, _, _ = z_metrics(oracle, ng[2], alpha, lengths[2])
    singular = all(c in "PN" for c in cbc)
    f = _rhs(ng, dzf, singular, seed=sum(ng))
    s, pref, dzci, dzfi, dzf, dli = _oracle_solve(oracle, ng, lengths, cbc, alpha, f)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc)
    # initsolver outputs are the CPU path's bit for bit
    np.testing.assert_array_equal(plan.lambdaxy, s.lambdaxy)
    np.testing.assert_array_equal(plan.a, s.a)
    np.testing.assert_array_equal(plan.b, s.b)
    np.testing.assert_array_equal(plan.c, s.c)
    assert plan.normfft == s.normfft
    d = api.DeviceArray.from_host(f)
    plan.solve(d)
    got = d.to_host()
    # halo untouched
    np.testing.assert_array_equal(got[0], f[0])
    np.testing.assert_array_equal(got[:, :, -1], f[:, :, -1])
    err = relmax(interior(got), interior(pref))
    assert err <= 1e-12, err
    if singular:
        assert relmax(gauge_free(interior(got)), gauge_free(interior(pref))) <= 1e-12
    fd = api.DeviceArray.from_host(f)
    assert plan.residual(fd, d) < 1e-12


FACE_CASES = [
    # (ng, lengths, cbc, c_or_f, alpha); face-centred z: tridmatrix 'f', and
    # with a D top face nz - 1 planes in the system (src/solver_cpu.f90:71-77)
    ((32, 24, 20), (1.0, 1.0, 1.0), "PPPPNN", "ccf", 1.3),   # four-pass
    ((32, 24, 20), (1.0, 1.0, 1.0), "PPPPDD", "ccf", 1.0),   # four-pass, q = 1
    ((32, 24, 20), (1.0, 1.0, 1.0), "NNPPND", "ccf", 0.0),
    ((32, 24, 20), (1.0, 1.0, 1.0), "NNPPDN", "ccf", 0.0),
    ((32, 24, 18), (1.0, 1.0, 2.0), "PPPPPP", "fff", 0.0),   # periodic 'f'
    ((512, 64, 16), (2 * np.pi, np.pi, 1.0), "PPPPDD", "ccf", 1.5),  # blocked W, q = 1
    # face-centred ND in x (REDFT10 / REDFT01, four-pass) and DN (RODFT01
    # forward, line pipeline)
    ((32, 48, 16), (1.0, 1.0, 1.0), "NDPPDD", "fcf", 1.2),
    ((24, 32, 12), (1.0, 1.0, 1.0), "DNNDNN", "ffc", 0.0),
    ((20, 12, 10), (1.0, 1.0, 1.0), "DNPPDD", "fcf", 0.0),   # line pipeline, q = 1
]


@pytest.mark.parametrize("ng,lengths,cbc,c_or_f,alpha", FACE_CASES)
def test_face_centred_solve_matches_oracle(api, oracle, ng, lengths, cbc, c_or_f, alpha):
    """c_or_f = 'f' (src/initsolver.f90:164-178, :210-236, :277-287;
    src/fft.f90:265-289; src/solver_cpu.f90:71-77): initsolver outputs bit
    for bit, the solve to 1e-12 of the oracle."""
    _, dzf, _, _, _, _ = z_metrics(oracle, ng[2], alpha, lengths[2])
    f = _rhs(ng, dzf, False, seed=sum(ng) + 7)
    print(f"face-centred solve {ng} {cbc} {c_or_f}", flush=True)
    s, pref, dzci, dzfi, dzf, dli = _oracle_solve(oracle, ng, lengths, cbc, alpha, f, c_or_f)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc, c_or_f=c_or_f)
    np.testing.assert_array_equal(plan.lambdaxy, s.lambdaxy)
    np.testing.assert_array_equal(plan.a, s.a)
    np.testing.assert_array_equal(plan.b, s.b)
    np.testing.assert_array_equal(plan.c, s.c)
    assert plan.normfft == s.normfft
    d = api.DeviceArray.from_host(f)
    plan.solve(d)
    got = interior(d.to_host())
    assert relmax(got, interior(pref)) <= 1e-12
    plan.destroy()


@pytest.mark.parametrize("cbc", ["PPPPNN", "PPPPPP"])
@pytest.mark.parametrize("precision", [64, 32])
def test_fused_and_line_pipelines_agree(api, oracle, monkeypatch, precision, cbc):
    """The four-pass pipeline and the five-pass line pipeline (forced with
    FLUTAS_UNFUSED=1 at plan creation) solve the same problem alike."""
    ng, lengths, alpha = (64, 48, 40), (2 * np.pi, np.pi, 1.0), 1.5
    _, dzf, dzci, dzfi, _, _ = z_metrics(oracle, ng[2], alpha, lengths[2])
    dli = tuple(n / l for n, l in zip(ng, lengths))
    f = _rhs(ng, dzf, True, seed=4)
    out = []
    for unfused in ("0", "1"):
        monkeypatch.setenv("FLUTAS_UNFUSED", unfused)
        plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc, precision=precision, tol=1e-10)
        d = api.DeviceArray.from_host(f)
        plan.solve(d)
        out.append(interior(d.to_host()))
        plan.destroy()
    tol = 1e-13 if precision == 64 else 1e-8
    assert relmax(gauge_free(out[0]), gauge_free(out[1])) <= tol


def _fields(ng, seed, nh_u=1):
    rng = np.random.default_rng(seed)
    sh_u = tuple(n + 2 * nh_u for n in ng)
    sh_p = tuple(n + 2 for n in ng)
    u, v, w = (np.asfortranarray(rng.uniform(-1, 1, sh_u)) for _ in range(3))
    p, pold = (np.asfortranarray(rng.standard_normal(sh_p)) for _ in range(2))
    rho = np.asfortranarray(1.0 + 999.0 * rng.uniform(0, 1, sh_p))
    return u, v, w, p, pold, rho


@pytest.mark.parametrize("nh_u", [1, 2])
def test_stencil_kernels_bit_identical(api, oracle, nh_u):
    ng = (37, 20, 11)
    nx, ny, nz = ng
    nh_d = 1
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, 1.3, 1.0)
    dli = (nx / 1.1, ny / 0.7, nz)
    u, v, w, p, pold, rho = _fields(ng, 3, nh_u)
    D = api.DeviceArray.from_host
    dzci_d, dzfi_d = D(dzci), D(dzfi)
    # fillps
    pr = np.zeros_like(p, order="F")
    oracle.fillps(ng, nh_d, nh_u, dli, dzfi, 1 / 0.3, 2.5, u, v, w, pr)
    pd = D(np.zeros_like(p, order="F"))
    api.fillps(nx, ny, nz, nh_d, nh_u, dli[0], dli[1], dli[2], dzfi_d, 1 / 0.3, 2.5,
               D(u), D(v), D(w), pd)
    np.testing.assert_array_equal(pd.to_host(), pr)
    # correc, both branches
    for cc in (True, False):
        ur, vr, wr = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
        oracle.correc(ng, nh_d, nh_u, dli, dzci, 0.3, 2.5, p, ur, vr, wr, rho=rho, const_coeff=cc)
        ud, vd, wd = D(u), D(v), D(w)
        api.correc(nx, ny, nz, nh_d, nh_u, dli[0], dli[1], dli[2], dzci_d, 0.3, 2.5, D(p),
                   ud, vd, wd, rho=D(rho), const_coeff=cc)
        for a, b in ((ud, ur), (vd, vr), (wd, wr)):
            np.testing.assert_array_equal(a.to_host(), b)
    # pres_tw_src, both branches
    for cc in (True, False):
        ur, vr, wr = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
        oracle.pres_tw_src(ng, dli, nh_d, nh_u, dzci, 1 / 1.0, 0.3, 0.25, p, pold, rho,
                           ur, vr, wr, const_coeff=cc)
        ud, vd, wd = D(u), D(v), D(w)
        api.pres_tw_src(nx, ny, nz, dli[0], dli[1], dli[2], nh_d, nh_u, dzci_d, 1.0, 0.3,
                        0.25, D(p), D(pold), D(rho), ud, vd, wd, const_coeff=cc)
        for a, b in ((ud, ur), (vd, vr), (wd, wr)):
            np.testing.assert_array_equal(a.to_host(), b)
    # pres_sp_src
    ur, vr, wr = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
    oracle.pres_sp_src(ng, 0.3, dli, nh_d, nh_u, dzci, 0.5, pold, ur, vr, wr)
    ud, vd, wd = D(u), D(v), D(w)
    api.pres_sp_src(nx, ny, nz, 0.3, dli[0], dli[1], dli[2], nh_d, nh_u, dzci_d, 0.5,
                    D(pold), ud, vd, wd)
    for a, b in ((ud, ur), (vd, vr), (wd, wr)):
        np.testing.assert_array_equal(a.to_host(), b)
    # chkdiv: same divergence values, reduction order differs
    tr, mr = oracle.chkdiv(ng, dli, nh_d, nh_u, dzfi, u, v, w)
    td, md = api.chkdiv(nx, ny, nz, dli[0], dli[1], dli[2], nh_d, nh_u, dzfi_d, D(u), D(v), D(w))
    assert md == mr
    assert abs(td - tr) <= 1e-12 * np.sum(np.abs(u)) * max(dli)


@pytest.mark.parametrize("cbc", ["PPPPNN", "NNDDPP", "DDNNDN", "PPNNND"])
def test_boundp_updt_rhs_b_bit_identical(api, oracle, cbc):
    ng = (10, 14, 12)
    nx, ny, nz = ng
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, 1.4, 1.0)
    bc = tuple(b if c != "P" else 0.0 for b, c in zip((0.2, -0.1, 0.3, 0.4, 0.7, -0.3), cbc))
    rng = np.random.default_rng(2)
    p = np.asfortranarray(rng.standard_normal((nx + 2, ny + 2, nz + 2)))
    dl = (1 / nx, 1 / ny, 1 / nz)
    pr = p.copy(order="F")
    oracle.boundp(cbc, ng, bc, 1, dl, dzc, dzf, pr)
    pd = api.DeviceArray.from_host(p)
    api.boundp(cbc, ng, bc, dl, dzc[0], dzc[nz], pd)
    np.testing.assert_array_equal(pd.to_host(), pr)
    # updt_rhs_b with the plan's rhsb planes
    dli = (float(nx), float(ny), float(nz))
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc, bc=bc)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc, bc=bc)
    for name in ("rhsbx", "rhsby", "rhsbz"):
        np.testing.assert_array_equal(getattr(plan, name), getattr(s, name))
    pr = p.copy(order="F")
    s.updt_rhs_b(pr)
    pd = api.DeviceArray.from_host(p)
    rx, ry, rz = plan.rhsb_device()
    api.updt_rhs_b(nx, ny, nz, "ccc", cbc, rx, ry, rz, pd)
    np.testing.assert_array_equal(pd.to_host(), pr)


@pytest.mark.parametrize("alpha", [0.0, 1.5])
def test_projection_step_channel_64(api, oracle, alpha):
    """The step of main__two_phase_inc_isot.f90:656-716 on a 64^3 PP/PP/NN
    channel: fillps -> updt_rhs_b -> solve -> boundp -> correc -> chkdiv."""
    ng = (64, 64, 64)
    nx, ny, nz = ng
    lengths = (1.0, 1.0, 1.0)
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, alpha, 1.0)
    dli = (nx, ny, nz)
    rng = np.random.default_rng(8)
    u, v, w = (np.asfortranarray(rng.uniform(-1, 1, (nx + 2, ny + 2, nz + 2))) for _ in range(3))
    for a in (u, v, w):
        a[0], a[-1] = a[-2], a[1]
        a[:, 0], a[:, -1] = a[:, -2], a[:, 1]
    w[:, :, 0] = 0.0
    w[:, :, -2] = 0.0
    D = api.DeviceArray.from_host
    ud, vd, wd = D(u), D(v), D(w)
    dzci_d, dzfi_d = D(dzci), D(dzfi)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, "PPPPNN")
    pd = D(np.zeros((nx + 2, ny + 2, nz + 2), order="F"))
    api.fillps(nx, ny, nz, 1, 1, dli[0], dli[1], dli[2], dzfi_d, 1.0, 1.0, ud, vd, wd, pd)
    _, div0 = api.chkdiv(nx, ny, nz, dli[0], dli[1], dli[2], 1, 1, dzfi_d, ud, vd, wd)
    rx, ry, rz = plan.rhsb_device()
    api.updt_rhs_b(nx, ny, nz, "ccc", "PPPPNN", rx, ry, rz, pd)
    f = pd.to_host()
    plan.solve(pd)
    api.boundp("PPPPNN", ng, (0,) * 6, (1 / nx, 1 / ny, 1 / nz), dzc[0], dzc[nz], pd)
    api.correc(nx, ny, nz, 1, 1, dli[0], dli[1], dli[2], dzci_d, 1.0, 1.0, pd, ud, vd, wd)
    uh, vh, wh = ud.to_host(), vd.to_host(), wd.to_host()
    for a in (uh, vh, wh):
        a[0], a[-1] = a[-2], a[1]
        a[:, 0], a[:, -1] = a[:, -2], a[:, 1]
    ud, vd, wd = D(uh), D(vh), D(wh)
    _, div1 = api.chkdiv(nx, ny, nz, dli[0], dli[1], dli[2], 1, 1, dzfi_d, ud, vd, wd)
    assert div1 <= 1e-12 * div0, (div0, div1)
    # pressure parity with the oracle on the same RHS
    s = oracle.Solver(ng, dli, dzci, dzfi, "PPPPNN")
    pr = f.copy(order="F")
    s.solve(pr)
    assert relmax(interior(pd.to_host()), interior(pr)) <= 1e-12


@pytest.mark.parametrize("cbc,alpha", [("PPPPNN", 1.5), ("NNNNNN", 0.0), ("PPPPPP", 0.0),
                                       ("PPPPDD", 1.3), ("NDDNNN", 1.2)])
def test_fp32_refined_reaches_tolerance(api, oracle, cbc, alpha):
    ng = (64, 512, 40) if cbc == "PPPPDD" else (64, 48, 40)
    lengths = (2 * np.pi, np.pi, 1.0)
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, ng[2], alpha, lengths[2])
    dli = tuple(n / l for n, l in zip(ng, lengths))
    singular = all(c in "PN" for c in cbc)
    f = _rhs(ng, dzf, singular, seed=21)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc, precision=api.FP32_REFINE, tol=1e-10)
    d = api.DeviceArray.from_host(f)
    plan.solve(d)
    iters, rel = plan.stats()
    assert rel <= 1e-10 and 1 <= iters <= 4, (iters, rel)
    fd = api.DeviceArray.from_host(f)
    assert plan.residual(fd, d) <= 1e-10
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc)
    pr = f.copy(order="F")
    s.solve(pr)
    got, ref = interior(d.to_host()), interior(pr)
    if singular:
        got, ref = gauge_free(got), gauge_free(ref)
    assert relmax(got, ref) <= 1e-8


def test_refined_rejects_nonfinite_rhs(api, oracle):
    ng = (16, 16, 16)
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, 16, 0.0, 1.0)
    plan = api.PoissonPlan(ng, (16, 16, 16), dzci, dzfi, "PPPPNN",
                           precision=api.FP32_REFINE, tol=1e-10)
    f = np.zeros((18, 18, 18), order="F")
    f[5, 5, 5] = np.nan
    with pytest.raises(api.FlutasError) as e:
        plan.solve(api.DeviceArray.from_host(f))
    assert e.value.code == api.FLUTAS_ENONFINITE


@pytest.mark.parametrize("unfused", ["0", "1"])
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_fp64_solve_reports_nonfinite(api, oracle, monkeypatch, unfused, bad):
    """A NaN or an infinity in the right-hand side makes the synchronous FP64
    solve return FLUTAS_ENONFINITE (four-pass pipeline: from P1's loads;
    five-pass line pipeline: a flag pass); the next finite solve succeeds."""
    ng, lengths, alpha = (64, 32, 24), (2 * np.pi, np.pi, 1.0), 1.5
    _, dzf, dzci, dzfi, _, _ = z_metrics(oracle, ng[2], alpha, lengths[2])
    dli = tuple(n / l for n, l in zip(ng, lengths))
    monkeypatch.setenv("FLUTAS_UNFUSED", unfused)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, "PPPPNN")
    f = _rhs(ng, dzf, True, seed=8)
    fbad = f.copy(order="F")
    fbad[17, 5, 9] = bad
    d = api.DeviceArray.from_host(fbad)
    with pytest.raises(api.FlutasError) as e:
        plan.solve(d)
    assert e.value.code == api.FLUTAS_ENONFINITE
    d = api.DeviceArray.from_host(f)
    plan.solve(d)
    assert np.all(np.isfinite(d.to_host()))
    plan.destroy()
