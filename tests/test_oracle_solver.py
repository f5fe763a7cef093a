"""Pins the oracle's initsolver pieces, Poisson solve and field kernels with
checks that do not retype the formulas: dense 1-D/3-D operators assembled
independently (Kronecker sums), exact discrete solutions (linear profiles,
discrete eigenfunctions), and cross-checks between independent routines
(fillps -> solve -> correc -> chkdiv; pres_tw_src special cases vs correc)."""
import numpy as np
import pytest

from cases import (dense_operator, gauge_free, halo_field, interior, relmax,
                   second_diff_1d, tri_dense, z_metrics)


@pytest.mark.parametrize("bc", ["PP", "NN", "DD", "ND", "DN"])
@pytest.mark.parametrize("n", [6, 8, 13, 16])
def test_eigenvalues_are_spectrum_of_1d_operator(oracle, bc, n):
    """src/initsolver.f90:122-186 ('c'): the eigenvalues are the spectrum of
    the cell-centred second difference with the BC ghost rules."""
    lam = oracle.eigenvalues(n, bc, "c")
    ev = np.linalg.eigvals(second_diff_1d(n, bc, 1.0)).real
    np.testing.assert_allclose(np.sort(lam), np.sort(ev), rtol=0, atol=1e-12)


def test_tridmatrix_uniform_closed_form(oracle):
    nz, lz = 8, 1.0
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, 0.0, lz)
    h2 = (nz / lz) ** 2
    for bc, b0 in (("NN", -h2), ("DD", -3 * h2), ("PP", -2 * h2)):
        a, b, c = oracle.tridmatrix(bc, nz, 1, dzci, dzfi)
        np.testing.assert_allclose(a, h2, rtol=1e-14)
        np.testing.assert_allclose(c, h2, rtol=1e-14)
        np.testing.assert_allclose(b[1:-1], -2 * h2, rtol=1e-14)
        np.testing.assert_allclose([b[0], b[-1]], [b0, b0], rtol=1e-14)


@pytest.mark.parametrize("alpha", [0.0, 1.5, 2.0])
def test_tridmatrix_stretched_invariants(oracle, alpha):
    """NN rows annihilate constants; dzf-weighted symmetry a/c."""
    nz = 24
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, alpha, 1.0)
    a, b, c = oracle.tridmatrix("NN", nz, 1, dzci, dzfi)
    m = tri_dense(a, b, c, False)
    assert np.max(np.abs(m @ np.ones(nz))) < 1e-10 * np.max(np.abs(b))
    wf = dzf[1:nz + 1]
    np.testing.assert_allclose(wf[:-1] * c[:-1], wf[1:] * a[1:], rtol=1e-13)


@pytest.mark.parametrize("alpha", [0.0, 1.5])
def test_initgrid_invariants(oracle, alpha):
    n, lz = 32, 2.0
    dzc, dzf, zc, zf = oracle.initgrid(n, alpha, lz, 1)
    # arrays cover (0 : n+1)
    assert abs(np.sum(dzf[1:n + 1]) - lz) < 1e-13
    assert abs(zf[n] - lz) < 1e-13 and zf[0] == 0.0
    np.testing.assert_allclose(dzf[1:n + 1], dzf[1:n + 1][::-1], rtol=1e-12)
    np.testing.assert_allclose(zc[1:n + 1], 0.5 * (zf[0:n] + zf[1:n + 1]), rtol=0, atol=1e-14)
    # interior centres only: the ghost extension of zc follows the reference
    # literally (src/initgrid.f90:87-94 sets zc(0) = zf(1) - dzc(1))
    np.testing.assert_allclose(dzc[1:n], zc[2:n + 1] - zc[1:n], rtol=1e-12)
    assert dzf[0] == dzf[1]
    # the top ghost follows src/initgrid.f90:80-83 literally: dzf(n+1) =
    # dzf(2n-k-1) = dzf(n-2) (unused on the projection path, which reads
    # dzfi(k), dzci(k) for k <= n only)
    assert dzf[n + 1] == dzf[n - 2] and dzc[n + 1] == dzc[n - 1]
    if alpha == 0.0:
        np.testing.assert_allclose(dzf, lz / n, rtol=1e-13)
    else:
        assert dzf[1] < dzf[n // 2]  # clustered at both walls


SINGULAR = {"PPPPNN", "NNNNNN", "PPPPPP", "NNPPNN"}
PATTERNS = ["PPPPNN", "NNNNNN", "PPPPDD", "PPPPPP", "NNPPNN", "DDNNPP",
            "PPDDNN", "NDPPDD", "DNNNDD", "NNDDDN"]


def _build(oracle, ng, cbc, alpha, lengths=(1.0, 1.0, 1.0)):
    nx, ny, nz = ng
    dzc, dzf, dzci, dzfi, zc, zf = z_metrics(oracle, nz, alpha, lengths[2])
    dli = (nx / lengths[0], ny / lengths[1], nz / lengths[2])
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc)
    return s, dzf, zc


@pytest.mark.parametrize("cbc", PATTERNS)
@pytest.mark.parametrize("alpha", [0.0, 1.5])
def test_solve_matches_dense_solve(oracle, cbc, alpha):
    ng = (6, 8, 10)
    s, dzf, _ = _build(oracle, ng, cbc, alpha if cbc[4] != "P" else 0.0,
                       (1.0, 1.3, 0.9))
    rng = np.random.default_rng(11)
    f = rng.standard_normal(ng)
    if cbc in SINGULAR:
        w = np.broadcast_to(dzf[1:ng[2] + 1][None, None, :], ng)
        f = f - np.sum(w * f) / np.sum(w)  # compatible (left null vector)
    L = dense_operator(ng, cbc, s.dli, s.a, s.b, s.c)
    fv = f.reshape(-1, order="F")
    ref = np.linalg.lstsq(L, fv, rcond=None)[0].reshape(ng, order="F")
    p = halo_field(ng, 1)
    p[1:-1, 1:-1, 1:-1] = f
    s.solve(p, pin_gauge=True)
    pin = interior(p).copy()
    res = L @ pin.reshape(-1, order="F") - fv
    assert np.max(np.abs(res)) < 1e-11 * np.max(np.abs(fv))
    if cbc in SINGULAR:
        assert relmax(gauge_free(pin), gauge_free(ref)) < 1e-10
    else:
        assert relmax(pin, ref) < 1e-11


def test_literal_and_pinned_gauge_agree_mod_constant(oracle):
    ng = (8, 8, 16)
    s, dzf, _ = _build(oracle, ng, "PPPPNN", 1.5)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(ng)
    w = np.broadcast_to(dzf[1:17][None, None, :], ng)
    f -= np.sum(w * f) / np.sum(w)
    p1, p2 = halo_field(ng, 1), halo_field(ng, 1)
    p1[1:-1, 1:-1, 1:-1] = f
    p2[1:-1, 1:-1, 1:-1] = f
    s.solve(p1, pin_gauge=True)
    s.solve(p2, pin_gauge=False)
    assert relmax(gauge_free(interior(p1)), gauge_free(interior(p2))) < 1e-9


def test_discrete_eigenfunction_is_reproduced(oracle):
    """PP/PP/NN uniform: a product of discrete eigenvectors is an exact
    solution (checks transform pairing, ordering and normfft)."""
    ng = (16, 12, 10)
    nx, ny, nz = ng
    s, _, _ = _build(oracle, ng, "PPPPNN", 0.0, (2.0, 1.0, 1.0))
    i = np.arange(nx)[:, None, None]
    j = np.arange(ny)[None, :, None]
    k = np.arange(nz)[None, None, :]
    mx, my, mz = 3, 2, 4
    pe = (np.cos(2 * np.pi * mx * i / nx + 0.3) * np.sin(2 * np.pi * my * j / ny + 0.1) *
          np.cos(np.pi * mz * (k + 0.5) / nz))
    # eigenvalues of the dense 1-D operators (not the closed form)
    lx = np.linalg.eigvalsh(second_diff_1d(nx, "PP", s.dli[0]))
    ly = np.linalg.eigvalsh(second_diff_1d(ny, "PP", s.dli[1]))
    lzm = tri_dense(s.a, s.b, s.c, False)
    vz = np.cos(np.pi * mz * (np.arange(nz) + 0.5) / nz)
    lz = (lzm @ vz)[0] / vz[0]
    lamx = -4 * np.sin(np.pi * mx / nx) ** 2 * s.dli[0] ** 2
    lamy = -4 * np.sin(np.pi * my / ny) ** 2 * s.dli[1] ** 2
    assert np.min(np.abs(lx - lamx)) < 1e-10 and np.min(np.abs(ly - lamy)) < 1e-10
    f = (lamx + lamy + lz) * pe
    p = halo_field(ng, 1)
    p[1:-1, 1:-1, 1:-1] = f
    s.solve(p)
    assert relmax(interior(p), pe) < 1e-12


@pytest.mark.parametrize("zbc,vals", [("DD", (0.7, -1.3)), ("NN", (0.4, 0.4))])
@pytest.mark.parametrize("alpha", [0.0, 1.5])
def test_nonhomogeneous_bc_gives_linear_profile(oracle, zbc, vals, alpha):
    """bc_rhs + updt_rhs_b + solve (src/initsolver.f90:248-298,
    src/bound.f90:829-944): zero source with Dirichlet face values v0, v1
    (or Neumann slope v) has the exact discrete solution linear in z."""
    ng = (4, 6, 16)
    lz = 1.0
    nx, ny, nz = ng
    dzc, dzf, dzci, dzfi, zc, zf = z_metrics(oracle, nz, alpha, lz)
    dli = (nx, ny, nz)
    bc = (0.0, 0.0, 0.0, 0.0, vals[0], vals[1])
    s = oracle.Solver(ng, dli, dzci, dzfi, "PPNN" + zbc, bc=bc)
    p = halo_field(ng, 1)
    s.updt_rhs_b(p)
    s.solve(p)
    z = zc[1:nz + 1]
    if zbc == "DD":
        exact = vals[0] + (vals[1] - vals[0]) * z / lz
        got = interior(p)
        assert relmax(got, np.broadcast_to(exact, ng)) < 1e-11
    else:
        exact = vals[0] * z
        got = interior(p)
        assert relmax(gauge_free(got), gauge_free(np.broadcast_to(exact, ng))) < 1e-10
    # boundp reproduces the face conditions on the solved field
    oracle.boundp("PPNN" + zbc, ng, bc, 1, (1 / nx, 1 / ny, 1 / nz), dzc, dzf, p)
    if zbc == "DD":
        np.testing.assert_allclose(0.5 * (p[1:-1, 1:-1, 0] + p[1:-1, 1:-1, 1]), vals[0], atol=1e-11)
        np.testing.assert_allclose(0.5 * (p[1:-1, 1:-1, -1] + p[1:-1, 1:-1, -2]), vals[1], atol=1e-11)
    else:
        np.testing.assert_allclose((p[1:-1, 1:-1, 1] - p[1:-1, 1:-1, 0]) / dzc[0], vals[0], atol=1e-9)
        np.testing.assert_allclose((p[1:-1, 1:-1, -1] - p[1:-1, 1:-1, -2]) / dzc[nz], vals[1], atol=1e-9)


def test_boundp_face_rules(oracle):
    ng = (5, 6, 7)
    rng = np.random.default_rng(2)
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, 7, 1.2, 1.0)
    p = halo_field(ng, 1, rng)
    bc = (0.0, 0.0, 0.25, -0.5, 1.5, 2.0)
    dl = (0.2, 0.1, 1 / 7)
    oracle.boundp("PPNDDN", ng, bc, 1, dl, dzc, dzf, p)
    I = slice(1, -1)
    np.testing.assert_array_equal(p[0, I, I], p[5, I, I])
    np.testing.assert_array_equal(p[6, I, I], p[1, I, I])
    np.testing.assert_allclose((p[I, 1, I] - p[I, 0, I]) / dl[1], 0.25, atol=1e-12)
    np.testing.assert_allclose(0.5 * (p[I, -1, I] + p[I, -2, I]), -0.5, atol=1e-12)
    np.testing.assert_allclose(0.5 * (p[I, I, 0] + p[I, I, 1]), 1.5, atol=1e-12)
    np.testing.assert_allclose((p[I, I, -1] - p[I, I, -2]) / dzc[7], 2.0, atol=1e-11)


def _periodic_uvw_halo(u, v, w):
    """bounduvw for PP x, PP y, no-penetration walls in z (w = 0 on the
    wall faces k = 0 and k = nz), nh_u = 1."""
    for a in (u, v, w):
        a[0, :, :] = a[-2, :, :]
        a[-1, :, :] = a[1, :, :]
        a[:, 0, :] = a[:, -2, :]
        a[:, -1, :] = a[:, 1, :]
    w[:, :, 0] = 0.0
    w[:, :, -2] = 0.0


@pytest.mark.parametrize("alpha", [0.0, 1.5])
def test_projection_step_is_divergence_free(oracle, alpha):
    """fillps -> solve -> boundp -> correc -> chkdiv (the reference step,
    main__two_phase_inc_isot.f90:674-695, :773)."""
    ng = (16, 12, 20)
    nx, ny, nz = ng
    lx, ly, lz = 2 * np.pi, np.pi, 1.0
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, alpha, lz)
    dli = (nx / lx, ny / ly, nz / lz)
    s = oracle.Solver(ng, dli, dzci, dzfi, "PPPPNN")
    rng = np.random.default_rng(7)
    u, v, w = (halo_field(ng, 1, rng) for _ in range(3))
    _periodic_uvw_halo(u, v, w)
    dt, rho0 = 0.37, 2.5
    p = halo_field(ng, 1)
    oracle.fillps(ng, 1, 1, dli, dzfi, 1 / dt, rho0, u, v, w, p)
    _, divmax0 = oracle.chkdiv(ng, dli, 1, 1, dzfi, u, v, w)
    assert divmax0 > 1.0
    s.updt_rhs_b(p)
    s.solve(p)
    oracle.boundp("PPPPNN", ng, (0,) * 6, 1, (1 / dli[0], 1 / dli[1], 1 / dli[2]), dzc, dzf, p)
    oracle.correc(ng, 1, 1, dli, dzci, dt, rho0, p, u, v, w)
    _periodic_uvw_halo(u, v, w)
    divtot, divmax = oracle.chkdiv(ng, dli, 1, 1, dzfi, u, v, w)
    assert divmax < 1e-11 * divmax0
    assert abs(divtot) < 1e-9


def test_chkdiv_discrete_streamfunction(oracle):
    ng = (8, 10, 6)
    nx, ny, nz = ng
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, 0.0, 1.0)
    dxi, dyi = 3.0, 5.0
    rng = np.random.default_rng(9)
    psi = rng.standard_normal((nx + 3, ny + 3, nz + 2))
    u = halo_field(ng, 1)
    v = halo_field(ng, 1)
    w = halo_field(ng, 1)
    # edge streamfunction Psi(i,j) = psi[i+1, j+1]:
    # u(i,j) = (Psi(i,j) - Psi(i,j-1)) dyi, v(i,j) = -(Psi(i,j) - Psi(i-1,j)) dxi
    u[:, :, :] = (psi[1:, 1:, :] - psi[1:, :-1, :]) * dyi
    v[:, :, :] = -(psi[1:, 1:, :] - psi[:-1, 1:, :]) * dxi
    tot, mx = oracle.chkdiv(ng, (dxi, dyi, nz), 1, 1, dzfi, u, v, w)
    assert mx < 1e-12 and abs(tot) < 1e-11
    # a uniform expansion u = alpha * x_face has divergence alpha
    xf = np.arange(-1, nx + 1) / dxi
    u[:, :, :] = 0.75 * xf[:, None, None]
    v[:] = 0
    tot, mx = oracle.chkdiv(ng, (dxi, dyi, nz), 1, 1, dzfi, u, v, w)
    assert abs(mx - 0.75) < 1e-12 and abs(tot - 0.75 * nx * ny * nz) < 1e-9


def test_pres_tw_src_reduces_to_correc(oracle):
    """src/source.f90:247-309: with rho == rho0 the splitting term is the
    constant-coefficient gradient; with pold == p and dt == dto it is the
    variable-coefficient gradient -dt grad(p)/rho_face (src/correc.f90)."""
    ng = (6, 7, 8)
    rng = np.random.default_rng(4)
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, 8, 1.5, 1.0)
    dli = (6.0, 7.0, 8.0)
    p, pold = halo_field(ng, 1, rng), halo_field(ng, 1, rng)
    rho = np.asfortranarray(1.0 + 999.0 * rng.uniform(0, 1, size=p.shape))
    u0, v0, w0 = (halo_field(ng, 1, rng) for _ in range(3))
    rho0 = 1.0
    # (a) uniform density
    rho_u = np.full_like(rho, rho0, order="F")
    ua, va, wa = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F")
    ub, vb, wb = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F")
    oracle.pres_tw_src(ng, dli, 1, 1, dzci, 1 / rho0, 0.3, 0.2, p, pold, rho_u, ua, va, wa)
    oracle.correc(ng, 1, 1, dli, dzci, 0.3, rho0, p, ub, vb, wb)
    for x, y in ((ua, ub), (va, vb), (wa, wb)):
        assert relmax(x, y) < 1e-14
    # (b) pold == p, dt == dto
    ua, va, wa = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F")
    ub, vb, wb = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F")
    oracle.pres_tw_src(ng, dli, 1, 1, dzci, 1 / rho0, 0.3, 0.3, p, p, rho, ua, va, wa)
    oracle.correc(ng, 1, 1, dli, dzci, 0.3, rho0, p, ub, vb, wb, rho=rho, const_coeff=False)
    for x, y in ((ua, ub), (va, vb), (wa, wb)):
        assert relmax(interior(x), interior(y)) < 1e-12
    # (c) pres_sp_src == constant-coefficient correc with pold
    ua, va, wa = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F")
    ub, vb, wb = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F")
    oracle.pres_sp_src(ng, 0.3, dli, 1, 1, dzci, 1 / 4.0, pold, ua, va, wa)
    oracle.correc(ng, 1, 1, dli, dzci, 0.3, 4.0, pold, ub, vb, wb)
    for x, y in ((ua, ub), (va, vb), (wa, wb)):
        assert relmax(x, y) < 1e-14


def _checkerboard_rho(shape):
    """rho = 1 or 3 in a checkerboard: every face between two neighbouring
    cells has one of each, so the arithmetic face density is exactly 2 (a
    harmonic mean would give 1.5)."""
    i, j, k = np.indices(shape)
    return np.asfortranarray(1.0 + 2.0 * ((i + j + k) % 2))


@pytest.mark.parametrize("const_coeff", [True, False])
def test_pres_tw_src_linear_in_time(oracle, const_coeff):
    """src/source.f90:268-269, :291-296 on a pressure linear in time,
    p(t) = A + B t, with pold = p(t - dto), p = p(t) and dt != dto: the
    extrapolated pressure f1 p - f2 pold is p(t + dt), so the increment is
    dt [-grad p(t)/rho0 - (1/rho_f - 1/rho0) grad p(t + dt)] with the face
    gradients of A + B t taken directly (constant-coefficient branch), or
    -dt grad p(t)/rho_f (variable-coefficient branch).  rho_f = 2 exactly
    (checkerboard density); a swapped dt/dto ratio fails."""
    ng = (6, 7, 8)
    rng = np.random.default_rng(21)
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, ng[2], 1.5, 1.0)
    dli = (6.0, 3.5, 8.0)
    A, B = halo_field(ng, 1, rng), halo_field(ng, 1, rng)
    t, dt, dto, rho0 = 0.7, 0.3, 0.2, 1.0
    p = np.asfortranarray(A + B * t)
    pold = np.asfortranarray(A + B * (t - dto))
    pnext = A + B * (t + dt)
    rho = _checkerboard_rho(p.shape)
    u0, v0, w0 = (halo_field(ng, 1, rng) for _ in range(3))
    u, v, w = (a.copy(order="F") for a in (u0, v0, w0))
    oracle.pres_tw_src(ng, dli, 1, 1, dzci, 1 / rho0, dt, dto, p, pold, rho, u, v, w,
                       const_coeff=const_coeff)
    I = slice(1, -1)
    nz = ng[2]
    metric = (dli[0], dli[1], dzci[1:nz + 1][None, None, :])
    for d, (new, old) in enumerate(((u, u0), (v, v0), (w, w0))):
        nb = [I, I, I]
        nb[d] = slice(2, None)

        def grad(f):
            return (f[tuple(nb)] - f[I, I, I]) * metric[d]
        if const_coeff:
            want = dt * (-grad(p) / rho0 - (1 / 2.0 - 1 / rho0) * grad(pnext))
        else:
            want = -dt * grad(p) / 2.0
        got = new[I, I, I] - old[I, I, I]
        assert relmax(got, want) < 1e-12, d
    if const_coeff:
        # the same call with the ratio inverted misses p(t + dt)
        us = u0.copy(order="F")
        oracle.pres_tw_src(ng, dli, 1, 1, dzci, 1 / rho0, dto, dt, p, pold, rho, us,
                           v0.copy(order="F"), w0.copy(order="F"))
        want = dto * (-(p[2:, I, I] - p[I, I, I]) * dli[0] / rho0
                      - (1 / 2.0 - 1 / rho0) * (pnext[2:, I, I] - pnext[I, I, I]) * dli[0])
        assert relmax(us[I, I, I] - u0[I, I, I], want) > 1e-3


@pytest.mark.parametrize("alpha", [0.0, 1.5])
def test_fillps_linear_velocity(oracle, alpha):
    """src/fillps.f90:50-57: u = a x_face, v = b y_face, w = c z_face (the
    stretched faces of initgrid) have the divergence a + b + c in every
    cell, so fillps gives rho0 (a + b + c) dti exactly (to rounding); a
    dropped rho0 or dti fails."""
    ng = (5, 6, 9)
    nx, ny, nz = ng
    dzc, dzf, dzci, dzfi, zc, zf = z_metrics(oracle, nz, alpha, 1.3)
    dli = (4.0, 2.5, nz / 1.3)
    ca, cb, cc = 0.7, -1.9, 2.3
    # u(i) sits on the face x = i dx, i = 0..nx+1 (halo 1)
    xf = np.arange(0, nx + 2) / dli[0]
    yf = np.arange(0, ny + 2) / dli[1]
    shape = (nx + 2, ny + 2, nz + 2)
    u = np.asfortranarray(np.broadcast_to(ca * xf[:, None, None], shape))
    v = np.asfortranarray(np.broadcast_to(cb * yf[None, :, None], shape))
    w = np.asfortranarray(np.broadcast_to(cc * zf[None, None, :], shape))
    p = halo_field(ng, 1)
    dti, rho0 = 1 / 0.04, 2.5
    oracle.fillps(ng, 1, 1, dli, dzfi, dti, rho0, u, v, w, p)
    want = rho0 * (ca + cb + cc) * dti
    np.testing.assert_allclose(interior(p), want, rtol=1e-11)


@pytest.mark.parametrize("alpha", [0.0, 1.5])
@pytest.mark.parametrize("const_coeff", [True, False])
def test_correc_linear_pressure(oracle, alpha, const_coeff):
    """src/correc.f90:57-68: p = a x_c + b y_c + c z_c has the face
    gradient (a, b, c) everywhere, so correc lowers u, v, w by
    dt (a, b, c)/rho0 (constant coefficients) or dt (a, b, c)/2
    (checkerboard density, arithmetic face mean 2).  z_c is accumulated
    from initgrid's dzc (stretched), so that its differences are the
    spacings correc divides by, ghost layers included."""
    ng = (6, 5, 8)
    nx, ny, nz = ng
    dzc, dzf, dzci, dzfi, zc, zf = z_metrics(oracle, nz, alpha, 0.8)
    dli = (3.0, 5.0, nz / 0.8)
    ca, cb, cc = 1.1, -0.6, 0.35
    xc = (np.arange(0, nx + 2) - 0.5) / dli[0]
    yc = (np.arange(0, ny + 2) - 0.5) / dli[1]
    zcl = np.concatenate([[0.0], np.cumsum(dzc[:nz + 1])])
    p = np.asfortranarray(ca * xc[:, None, None] + cb * yc[None, :, None] +
                          cc * zcl[None, None, :])
    rng = np.random.default_rng(3)
    u0, v0, w0 = (halo_field(ng, 1, rng) for _ in range(3))
    u, v, w = (a.copy(order="F") for a in (u0, v0, w0))
    dt, rho0 = 0.05, 4.0
    rho = _checkerboard_rho(p.shape)
    oracle.correc(ng, 1, 1, dli, dzci, dt, rho0, p, u, v, w, rho=rho, const_coeff=const_coeff)
    scale = rho0 if const_coeff else 2.0
    for new, old, g in ((u, u0, ca), (v, v0, cb), (w, w0, cc)):
        np.testing.assert_allclose(interior(old) - interior(new), dt * g / scale, rtol=1e-10)


# ---- face-centred ('f') pieces ---------------------------------------------

def test_tridmatrix_face_uniform_closed_form(oracle):
    """src/initsolver.f90:210-236, 'f': a = c = 1/dz^2 on a uniform grid; an
    N face adds a(1) or c(n) to the diagonal, a D face (a boundary value,
    no unknown) leaves it at -2/dz^2."""
    nz, lz = 8, 1.0
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, 0.0, lz)
    h2 = (nz / lz) ** 2
    for bc, (b0, b1) in (("NN", (-h2, -h2)), ("DD", (-2 * h2, -2 * h2)),
                         ("ND", (-h2, -2 * h2)), ("DN", (-2 * h2, -h2)),
                         ("PP", (-2 * h2, -2 * h2))):
        a, b, c = oracle.tridmatrix(bc, nz, 1, dzci, dzfi, "f")
        np.testing.assert_allclose(a, h2, rtol=1e-14)
        np.testing.assert_allclose(c, h2, rtol=1e-14)
        np.testing.assert_allclose(b[1:-1], -2 * h2, rtol=1e-14)
        np.testing.assert_allclose([b[0], b[-1]], [b0, b1], rtol=1e-14)


@pytest.mark.parametrize("alpha", [1.0, 2.0])
def test_tridmatrix_face_stretched_invariants(oracle, alpha):
    """Face unknowns at zf(k): the row of face k is the difference of the
    fluxes through the cells k and k+1 over the centre distance dzc(k), so
    dzc(k) c(k) = dzc(k+1) a(k+1) = 1/dzf(k+1), and the NN rows annihilate
    constants."""
    nz = 20
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, alpha, 1.0)
    a, b, c = oracle.tridmatrix("NN", nz, 1, dzci, dzfi, "f")
    wc = dzc[1:nz + 1]
    np.testing.assert_allclose(wc[:-1] * c[:-1], 1.0 / dzf[2:nz + 1], rtol=1e-13)
    np.testing.assert_allclose(wc[1:] * a[1:], 1.0 / dzf[2:nz + 1], rtol=1e-13)
    m = tri_dense(a, b, c, False)
    assert np.max(np.abs(m @ np.ones(nz))) < 1e-10 * np.max(np.abs(b))


def _face_dn_1d(n, dli):
    """Face-centred DN second difference: unknowns on nodes 1..n, the D
    value 0 on node 0, a mirror p(n+1) = p(n-1) about the N node n."""
    m = np.zeros((n, n))
    for i in range(n):
        m[i, i] = -2.0
        if i > 0:
            m[i, i - 1] = 1.0
        if i < n - 1:
            m[i, i + 1] = 1.0
    m[n - 1, n - 2] = 2.0
    return m * dli * dli


@pytest.mark.parametrize("cbc,c_or_f,alpha", [
    ("PPPPNN", "ccf", 1.3),
    ("PPPPDD", "ccf", 1.0),
    ("NNPPND", "ccf", 0.0),
    ("DDNNDN", "ccf", 1.5),
    ("DNPPNN", "fcc", 0.0),
    ("PPDNDD", "cff", 0.0),
])
def test_face_solve_matches_dense_solve(oracle, cbc, c_or_f, alpha):
    """The whole 'f' solve against a dense solve: z rows from the oracle's
    'f' tridmatrix (pinned above), a D top face taking nz - 1 rows with the
    last plane passed through (src/solver_cpu.f90:71-77), and the face DN
    pair of x or y (RODFT01 forward, RODFT10 backward, src/fft.f90:283-286)
    inverting the node-based DN second difference."""
    ng = (6, 8, 10)
    nx, ny, nz = ng
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, alpha, 0.9)
    dli = (nx / 1.0, ny / 1.3, nz / 0.9)
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc, c_or_f=c_or_f)
    q = 1 if (c_or_f[2] == "f" and cbc[5] == "D") else 0
    m = nz - q

    def op1(n, bc, cf, d):
        if cf == "f" and bc == "DN":
            return _face_dn_1d(n, d)
        assert cf == "c" or bc == "PP"
        return second_diff_1d(n, bc, d)

    lx = op1(nx, cbc[0:2], c_or_f[0], s.dli[0])
    ly = op1(ny, cbc[2:4], c_or_f[1], s.dli[1])
    lz = tri_dense(s.a[:m], s.b[:m], s.c[:m], cbc[4:6] == "PP")
    ix, iy, iz = np.eye(nx), np.eye(ny), np.eye(m)
    L = (np.kron(iz, np.kron(iy, lx)) + np.kron(iz, np.kron(ly, ix)) +
         np.kron(lz, np.kron(iy, ix)))
    rng = np.random.default_rng(3)
    f = rng.standard_normal(ng)
    if all(ch in "PN" for ch in cbc) and q == 0:
        # singular: make f compatible with the left null vector of L
        # (constant in x and y, that of the z rows in z)
        _, _, vt = np.linalg.svd(lz.T)
        wz = vt[-1]
        f -= np.sum(f * wz[None, None, :]) / (nx * ny * np.sum(wz))
    p = halo_field(ng, 1)
    p[1:-1, 1:-1, 1:-1] = f
    s.solve(p, pin_gauge=True)
    got = interior(p)
    fv = f[:, :, :m].reshape(-1, order="F")
    res = L @ got[:, :, :m].reshape(-1, order="F") - fv
    assert np.max(np.abs(res)) < 1e-11 * np.max(np.abs(fv))
    if q:
        # the boundary plane goes through the transform pair unchanged
        np.testing.assert_allclose(got[:, :, nz - 1], f[:, :, nz - 1], rtol=0, atol=1e-13)


def test_face_nd_pair_is_the_references(oracle):
    """The face-centred ND pair as the reference selects it: REDFT10 forward,
    REDFT01 backward, norm (2, 0) (src/fft.f90:279-282), with the ND
    eigenvalues of src/initsolver.f90:180-183.  These are the DCT-II/III of
    the cell-centred NN pair, so the solve inverts no second difference
    (the pairing's own behaviour; its components are pinned by the r2r
    golden vectors and the ND spectrum test)."""
    kf, kb, norm = oracle.find_fft("ND", "f")
    assert (kf, kb, norm) == (oracle.KIND["REDFT10"], oracle.KIND["REDFT01"], (2.0, 0.0))
    np.testing.assert_array_equal(oracle.eigenvalues(12, "ND", "f"),
                                  oracle.eigenvalues(12, "ND", "c"))
    kf, kb, norm = oracle.find_fft("DN", "f")
    assert (kf, kb, norm) == (oracle.KIND["RODFT01"], oracle.KIND["RODFT10"], (2.0, 0.0))
