"""The fused prologue and epilogue of the projection step.

  flutas_fillps_rhs     fillps + updt_rhs_b in one pass
  flutas_correc_update  correc + (pnew = pold + phi) in one pass

Both must equal the oracle's fillps / updt_rhs_b / correc / increment update
(src/fillps.f90, src/bound.f90:829-944, src/correc.f90,
main__single_phase.f90:734-740) bit for bit, on odd and even sizes, halo
widths 1 and 2, nonzero boundary values (nonzero rhsb planes) and both
correc branches; a whole step with buffer swapping (pnew = pold) must equal
the reference's step order.
"""
import numpy as np
import pytest

from cases import interior, relmax, z_metrics

pytestmark = pytest.mark.gpu


def _fields(ng, seed, nh_u):
    rng = np.random.default_rng(seed)
    sh_u = tuple(n + 2 * nh_u for n in ng)
    sh_p = tuple(n + 2 for n in ng)
    u, v, w = (np.asfortranarray(rng.uniform(-1, 1, sh_u)) for _ in range(3))
    phi, pold = (np.asfortranarray(rng.standard_normal(sh_p)) for _ in range(2))
    rho = np.asfortranarray(1.0 + 999.0 * rng.uniform(0, 1, sh_p))
    return u, v, w, phi, pold, rho


@pytest.mark.parametrize("cbc,nh_u", [("PPPPNN", 1), ("NNDDPP", 2), ("DDNNDN", 1),
                                      ("PPNNND", 2)])
def test_fillps_rhs_equals_fillps_then_updt_rhs_b(api, oracle, cbc, nh_u):
    ng = (38, 22, 12)
    nx, ny, nz = ng
    nh_d = 1
    _, _, dzci, dzfi, _, _ = z_metrics(oracle, nz, 1.3, 1.0)
    dli = (nx / 1.1, ny / 0.7, float(nz))
    bc = tuple(b if c != "P" else 0.0 for b, c in zip((0.2, -0.1, 0.3, 0.4, 0.7, -0.3), cbc))
    u, v, w, _, _, _ = _fields(ng, 5, nh_u)
    print(f"fillps_rhs {ng} {cbc} nh_u={nh_u}", flush=True)
    # oracle: fillps, then updt_rhs_b with the oracle's own rhsb planes
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc, bc=bc)
    pr = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    oracle.fillps(ng, nh_d, nh_u, dli, dzfi, 1 / 0.3, 2.5, u, v, w, pr)
    s.updt_rhs_b(pr)
    assert np.any(s.rhsbx != 0) or np.any(s.rhsby != 0) or np.any(s.rhsbz != 0)
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc, bc=bc)
    rx, ry, rz = plan.rhsb_device()
    D = api.DeviceArray.from_host
    pd = D(np.zeros_like(pr))
    api.fillps_rhs(nx, ny, nz, nh_d, nh_u, dli[0], dli[1], dli[2], D(dzfi), 1 / 0.3, 2.5,
                   D(u), D(v), D(w), "ccc", cbc, rx, ry, rz, pd)
    np.testing.assert_array_equal(pd.to_host(), pr)
    plan.destroy()


@pytest.mark.parametrize("nh_u", [1, 2])
@pytest.mark.parametrize("cc", [True, False])
def test_correc_update_equals_correc_then_update(api, oracle, nh_u, cc):
    ng = (37, 20, 11)
    nx, ny, nz = ng
    nh_d = 1
    _, _, dzci, _, _, _ = z_metrics(oracle, nz, 1.3, 1.0)
    dli = (nx / 1.1, ny / 0.7, float(nz))
    u, v, w, phi, pold, rho = _fields(ng, 6, nh_u)
    print(f"correc_update {ng} nh_u={nh_u} const_coeff={cc}", flush=True)
    ur, vr, wr = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
    oracle.correc(ng, nh_d, nh_u, dli, dzci, 0.3, 2.5, phi, ur, vr, wr, rho=rho, const_coeff=cc)
    pnew_r = pold.copy(order="F")
    pnew_r[1:-1, 1:-1, 1:-1] = pold[1:-1, 1:-1, 1:-1] + phi[1:-1, 1:-1, 1:-1]
    D = api.DeviceArray.from_host
    dzci_d = D(dzci)
    for in_place in (True, False):
        ud, vd, wd, phid, poldd = D(u), D(v), D(w), D(phi), D(pold)
        pnewd = poldd if in_place else D(np.zeros_like(pold))
        api.correc_update(nx, ny, nz, nh_d, nh_u, dli[0], dli[1], dli[2], dzci_d, 0.3, 2.5,
                          phid, poldd, pnewd, ud, vd, wd, rho=D(rho), const_coeff=cc)
        for a, b in ((ud, ur), (vd, vr), (wd, wr)):
            np.testing.assert_array_equal(a.to_host(), b)
        np.testing.assert_array_equal(interior(pnewd.to_host()), interior(pnew_r))
        np.testing.assert_array_equal(phid.to_host(), phi)  # the increment is only read
        if not in_place:
            np.testing.assert_array_equal(poldd.to_host(), pold)


@pytest.mark.parametrize("cbc,nh_u,cc", [("PPPPNN", 1, True), ("NNDDPP", 2, False),
                                          ("DDNNDN", 1, True), ("PPNNND", 2, True),
                                          ("DNPPDD", 1, False), ("PPPPPP", 1, True)])
def test_correc_update_bc_equals_boundp_correc_update_boundp(api, oracle, cbc, nh_u, cc):
    """flutas_correc_update_bc against boundp(phi), correc_update,
    boundp(pnew) (the step's order around the solve), bit for bit on every
    point of u, v, w and pnew (halos, edges and corners included), with
    nonzero boundary values, in place and out of place; phi stays as it
    was, garbage halo included."""
    ng = (37, 20, 11)
    nx, ny, nz = ng
    nh_d = 1
    dzc, _, dzci, _, _, _ = z_metrics(oracle, nz, 1.3, 1.0)
    dli = (nx / 1.1, ny / 0.7, float(nz))
    dl = tuple(1.0 / d for d in dli)
    bc = tuple(b if c != "P" else 0.0 for b, c in zip((0.2, -0.1, 0.3, 0.4, 0.7, -0.3), cbc))
    u, v, w, phi, pold, rho = _fields(ng, 7, nh_u)
    print(f"correc_update_bc {ng} {cbc} nh_u={nh_u} const_coeff={cc}", flush=True)
    D = api.DeviceArray.from_host
    dzci_d = D(dzci)
    for in_place in (True, False):
        # the three calls
        ua, va, wa, phia, polda = D(u), D(v), D(w), D(phi), D(pold)
        pnewa = polda if in_place else D(np.zeros_like(pold))
        api.boundp(cbc, ng, bc, dl, dzc[0], dzc[nz], phia)
        api.correc_update(nx, ny, nz, nh_d, nh_u, dli[0], dli[1], dli[2], dzci_d, 0.3, 2.5,
                          phia, polda, pnewa, ua, va, wa, rho=D(rho), const_coeff=cc)
        api.boundp(cbc, ng, bc, dl, dzc[0], dzc[nz], pnewa)
        # one pass
        ub, vb, wb, phib, poldb = D(u), D(v), D(w), D(phi), D(pold)
        pnewb = poldb if in_place else D(np.zeros_like(pold))
        api.correc_update_bc(nx, ny, nz, nh_d, nh_u, dli[0], dli[1], dli[2], dzci_d, 0.3, 2.5,
                             phib, poldb, pnewb, ub, vb, wb, cbc, bc, dl, dzc[0], dzc[nz],
                             rho=D(rho), const_coeff=cc)
        for a, b in ((ub, ua), (vb, va), (wb, wa), (pnewb, pnewa)):
            np.testing.assert_array_equal(a.to_host(), b.to_host())
        np.testing.assert_array_equal(phib.to_host(), phi)


def test_correc_update_rejects_aliased_increment(api):
    a = api.DeviceArray((6, 6, 6))
    u = api.DeviceArray((6, 6, 6))
    dz = api.DeviceArray.from_host(np.ones(6))
    with pytest.raises(api.FlutasError) as e:
        api.correc_update(4, 4, 4, 1, 1, 1.0, 1.0, 1.0, dz, 1.0, 1.0, a, u, a, u, u, u)
    assert e.value.code == api.FLUTAS_EINVAL


def test_fused_step_with_swapped_buffers(api, oracle):
    """Two projection steps of main__single_phase.f90:692-747 on a 64^3
    PP/PP/NN channel, the reference order (pold = p, fillps, updt_rhs_b,
    solve, boundp, correc, p = pold + p, boundp) through the oracle against
    the fused order with two swapped buffers (fillps_rhs into the scratch
    buffer, solve, boundp, correc_update into the pressure buffer, boundp)."""
    ng = (64, 64, 64)
    nx, ny, nz = ng
    cbc = "PPPPNN"
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, 1.5, 1.0)
    dli = (nx / (2 * np.pi), ny / np.pi, float(nz))
    dl = tuple(1 / d for d in dli)
    rng = np.random.default_rng(9)
    u, v, w = (np.asfortranarray(rng.uniform(-1, 1, (nx + 2, ny + 2, nz + 2))) for _ in range(3))
    p0 = np.asfortranarray(rng.standard_normal((nx + 2, ny + 2, nz + 2)))

    def velocity_halos(a_u, a_v, a_w):
        for a in (a_u, a_v, a_w):
            a[0], a[-1] = a[-2], a[1]
            a[:, 0], a[:, -1] = a[:, -2], a[:, 1]
        a_w[:, :, 0] = 0.0
        a_w[:, :, -2] = 0.0

    velocity_halos(u, v, w)
    oracle.boundp(cbc, ng, (0.0,) * 6, 1, dl, dzc, dzf, p0)
    dt = 1e-2
    # reference order through the oracle
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc)
    uo, vo, wo, po = (a.copy(order="F") for a in (u, v, w, p0))
    pold_o = np.zeros_like(po)
    for _ in range(2):
        pold_o[...] = po                       # pold = p (+ boundp(pold))
        oracle.fillps(ng, 1, 1, dli, dzfi, 1 / dt, 1.0, uo, vo, wo, po)
        s.updt_rhs_b(po)
        s.solve(po)
        oracle.boundp(cbc, ng, (0.0,) * 6, 1, dl, dzc, dzf, po)
        oracle.correc(ng, 1, 1, dli, dzci, dt, 1.0, po, uo, vo, wo)
        velocity_halos(uo, vo, wo)
        po[1:-1, 1:-1, 1:-1] = pold_o[1:-1, 1:-1, 1:-1] + po[1:-1, 1:-1, 1:-1]
        oracle.boundp(cbc, ng, (0.0,) * 6, 1, dl, dzc, dzf, po)
    # fused order on the GPU: P holds the pressure, S is the scratch buffer
    D = api.DeviceArray.from_host
    ud, vd, wd = D(u), D(v), D(w)
    dzci_d, dzfi_d = D(dzci), D(dzfi)
    P, S = D(p0), D(np.zeros_like(p0))
    plan = api.PoissonPlan(ng, dli, dzci, dzfi, cbc)
    rx, ry, rz = plan.rhsb_device()
    print(f"fused step {ng} {cbc}", flush=True)
    for _ in range(2):
        api.fillps_rhs(nx, ny, nz, 1, 1, dli[0], dli[1], dli[2], dzfi_d, 1 / dt, 1.0,
                       ud, vd, wd, "ccc", cbc, rx, ry, rz, S)
        plan.solve(S)
        api.boundp(cbc, ng, (0.0,) * 6, dl, dzc[0], dzc[nz], S)
        api.correc_update(nx, ny, nz, 1, 1, dli[0], dli[1], dli[2], dzci_d, dt, 1.0,
                          S, P, P, ud, vd, wd)
        uh, vh, wh = ud.to_host(), vd.to_host(), wd.to_host()
        velocity_halos(uh, vh, wh)
        ud, vd, wd = D(uh), D(vh), D(wh)
        api.boundp(cbc, ng, (0.0,) * 6, dl, dzc[0], dzc[nz], P)
    for g, o in ((ud, uo), (vd, vo), (wd, wo)):
        assert relmax(interior(g.to_host()), interior(o)) <= 1e-12
    assert relmax(P.to_host(), po) <= 1e-12
    plan.destroy()
