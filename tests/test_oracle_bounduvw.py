"""Pins the oracle's bounduvw (src/bound.f90:17-144, set_bc :227-646,
outflow :649-773) before it is used as the GPU reference.

With one halo layer, every face rule is fixed by the boundary condition it
discretises: a centred Dirichlet ghost averages with its neighbour to the
value, a centred Neumann ghost differs from it by dr * value, a face-normal
Dirichlet puts the value on the boundary face and mirrors the next ghost, a
periodic ghost copies the opposite interior, and an outflow face makes the
divergence of its cell vanish.  With several layers, the overwrite order of
the reference's layer loop is checked against an independent pure-Python
transcription of set_bc on a tiny grid."""
import numpy as np
import pytest

from cases import z_metrics

CB = lambda s, d, c: s + 2 * d + 6 * c  # (0:1,3,3) memory order


def _fields(ng, nh, seed):
    rng = np.random.default_rng(seed)
    shape = tuple(n + 2 * nh for n in ng)
    return [np.asfortranarray(rng.standard_normal(shape)) for _ in range(3)]


def _run(oracle, cbc, bcv, ng, nh, out, dl, dzc, dzf, nh_d, seed=3):
    u, v, w = _fields(ng, nh, seed)
    u0, v0, w0 = u.copy(), v.copy(), w.copy()
    oracle.bounduvw(cbc, ng, bcv, nh_d, nh, out, dl, dzc, dzf, u, v, w)
    return (u0, v0, w0), (u, v, w)


def test_one_layer_face_rules(oracle):
    """x: u face-normal D, v centred N, w centred D; y: u centred D, v
    face-normal N, w centred N; z walls: u, v centred N / D, w face-normal D;
    checked on the tangential interior, where no later face writes."""
    ng, nh, nh_d = (8, 6, 10), 1, 1
    lz = 1.0
    dzc, dzf, _, _, _, _ = z_metrics(oracle, ng[2], 1.3, lz, nh_d)
    dl = (0.3, 0.25, lz / ng[2])
    cbc = list("P" * 18)
    bcv = np.zeros(18)
    spec = {  # (side, dir, comp): (type, value)
        (0, 0, 0): ("D", 0.7), (1, 0, 0): ("D", -0.4),
        (0, 0, 1): ("N", 1.5), (1, 0, 1): ("N", -2.0),
        (0, 0, 2): ("D", 0.3), (1, 0, 2): ("D", 0.9),
        (0, 1, 0): ("D", 0.2), (1, 1, 0): ("D", -0.6),
        (0, 1, 1): ("N", 0.8), (1, 1, 1): ("N", 1.1),
        (0, 1, 2): ("N", -0.5), (1, 1, 2): ("N", 0.45),
        (0, 2, 0): ("N", 0.35), (1, 2, 0): ("N", -0.75),
        (0, 2, 1): ("D", 1.25), (1, 2, 1): ("D", -1.5),
        (0, 2, 2): ("D", 0.05), (1, 2, 2): ("D", -0.15),
    }
    for (s, d, c), (t, val) in spec.items():
        cbc[CB(s, d, c)] = t
        bcv[CB(s, d, c)] = val
    (u0, v0, w0), (u, v, w) = _run(oracle, "".join(cbc), bcv, ng, nh, [0] * 6, dl, dzc, dzf, nh_d)
    nx, ny, nz = ng
    I, J, K = slice(1, nx + 1), slice(1, ny + 1), slice(1, nz + 1)
    dzc0, dzcn = dzc[0 + nh_d - 1], dzc[nz + nh_d - 1]
    # x faces (tangential interior J, K)
    assert np.all(u[0, J, K] == 0.7)                                  # lower face value
    assert np.all(u[nx, J, K] == -0.4)                                # upper face value
    np.testing.assert_array_equal(u[nx + 1, J, K], u0[nx - 1, J, K])  # mirrored ghost
    np.testing.assert_allclose((v[1, J, K] - v[0, J, K]) / dl[0], 1.5, rtol=1e-12)
    np.testing.assert_allclose((v[nx + 1, J, K] - v[nx, J, K]) / dl[0], -2.0, rtol=1e-12)
    # (w(:,:,nz) is the upper z face, which the later z-face rule rewrites
    # over the whole tangential extent, x halos included)
    Kw = slice(1, nz)
    np.testing.assert_allclose(0.5 * (w[0, J, Kw] + w[1, J, Kw]), 0.3, rtol=1e-12)
    np.testing.assert_allclose(0.5 * (w[nx + 1, J, Kw] + w[nx, J, Kw]), 0.9, rtol=1e-12)
    # y faces
    np.testing.assert_allclose(0.5 * (u[I, 0, K] + u[I, 1, K]), 0.2, rtol=1e-12)
    np.testing.assert_allclose(0.5 * (u[I, ny + 1, K] + u[I, ny, K]), -0.6, rtol=1e-12)
    np.testing.assert_allclose(v[I, 1, K] - v[I, 0, K], dl[1] * 0.8, rtol=1e-12)
    np.testing.assert_allclose(v[I, ny, K] - v[I, ny - 1, K], dl[1] * 1.1, rtol=1e-12)
    np.testing.assert_allclose(v[I, ny + 1, K] - v[I, ny - 1, K], 2 * dl[1] * 1.1, rtol=1e-12)
    np.testing.assert_allclose((w[I, 1, Kw] - w[I, 0, Kw]) / dl[1], -0.5, rtol=1e-12)
    np.testing.assert_allclose((w[I, ny + 1, Kw] - w[I, ny, Kw]) / dl[1], 0.45, rtol=1e-12)
    # z walls: centred rules with dzc at the wall, face-normal with dzf
    np.testing.assert_allclose((u[I, J, 1] - u[I, J, 0]) / dzc0, 0.35, rtol=1e-11)
    np.testing.assert_allclose((u[I, J, nz + 1] - u[I, J, nz]) / dzcn, -0.75, rtol=1e-11)
    np.testing.assert_allclose(0.5 * (v[I, J, 0] + v[I, J, 1]), 1.25, rtol=1e-12)
    np.testing.assert_allclose(0.5 * (v[I, J, nz + 1] + v[I, J, nz]), -1.5, rtol=1e-12)
    assert np.all(w[I, J, 0] == 0.05) and np.all(w[I, J, nz] == -0.15)
    np.testing.assert_array_equal(w[I, J, nz + 1], w0[I, J, nz - 1])
    # interior untouched
    np.testing.assert_array_equal(u[1:nx, J, K], u0[1:nx, J, K])
    np.testing.assert_array_equal(v[I, 1:ny, K], v0[I, 1:ny, K])
    np.testing.assert_array_equal(w[I, J, 1:nz], w0[I, J, 1:nz])


def test_periodic_copies(oracle):
    ng, nh, nh_d = (6, 8, 4), 2, 2
    dzc, dzf, _, _, _, _ = z_metrics(oracle, ng[2], 0.0, 1.0, nh_d)
    (_, _, _), (u, v, w) = _run(oracle, "P" * 18, np.zeros(18), ng, nh, [0] * 6,
                                (0.1, 0.1, 0.25), dzc, dzf, nh_d)
    nx, ny, nz = ng
    for a in (u, v, w):
        for q in range(nh):  # array index = position - 1 + nh
            np.testing.assert_array_equal(a[nh - 1 - q], a[nx + nh - 1 - q])
            np.testing.assert_array_equal(a[nx + nh + q], a[nh + q])
            np.testing.assert_array_equal(a[:, nh - 1 - q], a[:, ny + nh - 1 - q])
            np.testing.assert_array_equal(a[:, :, nz + nh + q], a[:, :, nh + q])


@pytest.mark.parametrize("side,d", [(1, 0), (0, 0), (1, 1), (0, 1), (1, 2), (0, 2)])
def test_outflow_zero_divergence(oracle, side, d):
    """An outflow face makes the divergence of the cell beside it vanish."""
    ng, nh, nh_d = (8, 6, 10), 1, 1
    dzc, dzf, _, dzfi, _, _ = z_metrics(oracle, ng[2], 1.2, 1.0, nh_d)
    dl = (0.2, 0.3, 0.1)
    cbc = "N" * 18
    out = [0] * 6
    out[side + 2 * d] = 1
    (_, _, _), (u, v, w) = _run(oracle, cbc, np.zeros(18), ng, nh, out, dl, dzc, dzf, nh_d)
    nx, ny, nz = ng
    # divergence of cell (i,j,k), array index = position (halo 1)
    def div(i, j, k):
        return ((u[i, j, k] - u[i - 1, j, k]) / dl[0] + (v[i, j, k] - v[i, j - 1, k]) / dl[1] +
                (w[i, j, k] - w[i, j, k - 1]) * dzfi[k + nh_d - 1])
    cell = [None, None, None]
    cell[d] = [1, ng[d]][side]
    others = [r for r in range(3) if r != d]
    worst = 0.0
    for a in range(1, ng[others[0]] + 1):
        for b in range(1, ng[others[1]] + 1):
            idx = list(cell)
            idx[others[0]], idx[others[1]] = a, b
            worst = max(worst, abs(div(*idx)))
    assert worst < 1e-11, worst


def _set_bc_literal(p, n, nh, ctype, ib, idir, centered, rvalue, dr):
    """src/bound.f90:227-646 transcribed loop by loop (1-based positions)."""
    factor = [rvalue] * nh
    sgn = 0.0
    if ctype == "D" and centered:
        factor = [2.0 * f for f in factor]
        sgn = -1.0
    if ctype == "N":
        factor = [(-dr[q] * factor[q]) if ib == 0 else (dr[q] * factor[q]) for q in range(nh)]
        sgn = 1.0
    nd = n[idir]

    def at(a, t):
        idx = [None, None, None]
        idx[idir] = a - 1 + nh
        o = [r for r in range(3) if r != idir]
        idx[o[0]], idx[o[1]] = t
        return tuple(idx)
    o = [r for r in range(3) if r != idir]
    for q in range(nh):
        f = factor[q]
        for t0 in range(n[o[0]] + 2 * nh):
            for t1 in range(n[o[1]] + 2 * nh):
                t = (t0, t1)
                if ctype == "P":
                    p[at(0 - q, t)] = p[at(nd - q, t)]
                    p[at(nd + 1 + q, t)] = p[at(1 + q, t)]
                elif centered:
                    if ib == 0:
                        p[at(0 - q, t)] = f + sgn * p[at(1 + q, t)]
                    else:
                        p[at(nd + 1 + q, t)] = f + sgn * p[at(nd - q, t)]
                elif ctype == "D":
                    if ib == 0:
                        p[at(0 - q, t)] = f
                    else:
                        p[at(nd + q, t)] = f
                        p[at(nd + 1 + q, t)] = p[at(nd - 1 - q, t)]
                else:
                    if ib == 0:
                        p[at(0 - q, t)] = 1.0 * f + p[at(1 + q, t)]
                    else:
                        p[at(nd + q, t)] = 1.0 * f + p[at(nd - 1 - q, t)]
                        p[at(nd + 1 + q, t)] = 2.0 * f + p[at(nd - 1 - q, t)]


def test_multi_layer_matches_literal_transcription(oracle):
    """Three halo layers, non-periodic in x and z, periodic y: the oracle
    equals a loop-by-loop Python transcription of the reference's sequence
    (updthalo in y, x faces, z faces) bit for bit."""
    ng, nh, nh_d = (8, 6, 10), 3, 3
    dzc, dzf, _, _, _, _ = z_metrics(oracle, ng[2], 1.5, 1.0, nh_d)
    dl = (0.2, 0.3, 0.1)
    cbc = list("P" * 18)
    bcv = np.zeros(18)
    types = {0: ("D", "N"), 2: ("N", "D")}
    rng = np.random.default_rng(9)
    for d, (t0, t1) in types.items():
        for c in range(3):
            for s, t in ((0, t0), (1, t1)):
                cbc[CB(s, d, c)] = t if c % 2 == 0 else ("N" if t == "D" else "D")
                bcv[CB(s, d, c)] = rng.uniform(-1, 1)
    cbc = "".join(cbc)
    (u0, v0, w0), got = _run(oracle, cbc, bcv, ng, nh, [0] * 6, dl, dzc, dzf, nh_d)
    ref = [a.copy() for a in (u0, v0, w0)]
    for c in range(3):
        _set_bc_literal(ref[c], ng, nh, "P", 0, 1, True, 0.0, None)
    for s in (0, 1):
        for c in range(3):
            _set_bc_literal(ref[c], ng, nh, cbc[CB(s, 0, c)], s, 0, c != 0,
                            bcv[CB(s, 0, c)], [dl[0]] * nh)
    nz = ng[2]
    for s in (0, 1):
        drc = [dzc[(-q if s == 0 else nz + q) + nh_d - 1] for q in range(nh)]
        drf = [dzf[(-q if s == 0 else nz + q) + nh_d - 1] for q in range(nh)]
        for c in range(2):
            _set_bc_literal(ref[c], ng, nh, cbc[CB(s, 2, c)], s, 2, True, bcv[CB(s, 2, c)], drc)
        _set_bc_literal(ref[2], ng, nh, cbc[CB(s, 2, 2)], s, 2, False, bcv[CB(s, 2, 2)], drf)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)


@pytest.mark.parametrize("cbc", ["PPPPNN", "NNDDPP", "DNNDDN", "PPNNDD"])
def test_boundp_nh_one_layer_equals_boundp(oracle, cbc):
    """boundp with halo nh_p: at nh_p = 1 it equals the halo-1 boundp."""
    ng = (8, 6, 10)
    dzc, dzf, _, _, _, _ = z_metrics(oracle, ng[2], 1.4, 1.0, 2)
    bc = (0.3, -0.2, 0.5, 0.1, -0.4, 0.25)
    bc = tuple(b if c != "P" else 0.0 for b, c in zip(bc, cbc))
    dl = (0.2, 0.3, 0.1)
    rng = np.random.default_rng(4)
    p = np.asfortranarray(rng.standard_normal(tuple(n + 2 for n in ng)))
    a, b = p.copy(order="F"), p.copy(order="F")
    oracle.boundp(cbc, ng, bc, 2, dl, dzc, dzf, a)
    oracle.boundp_nh(cbc, ng, bc, 2, 1, dl, dzc, dzf, b)
    np.testing.assert_array_equal(a, b)


def test_boundp_nh_three_layers_matches_literal_transcription(oracle):
    ng, nh, nh_d = (8, 6, 10), 3, 3
    cbc = "DNPPND"
    bc = (0.3, -0.2, 0.0, 0.0, -0.4, 0.25)
    dzc, dzf, _, _, _, _ = z_metrics(oracle, ng[2], 1.4, 1.0, nh_d)
    dl = (0.2, 0.3, 0.1)
    rng = np.random.default_rng(6)
    p = np.asfortranarray(rng.standard_normal(tuple(n + 2 * nh for n in ng)))
    ref = p.copy(order="F")
    oracle.boundp_nh(cbc, ng, bc, nh_d, nh, dl, dzc, dzf, p)
    _set_bc_literal(ref, ng, nh, "P", 0, 1, True, 0.0, None)
    _set_bc_literal(ref, ng, nh, cbc[0], 0, 0, True, bc[0], [dl[0]] * nh)
    _set_bc_literal(ref, ng, nh, cbc[1], 1, 0, True, bc[1], [dl[0]] * nh)
    nz = ng[2]
    _set_bc_literal(ref, ng, nh, cbc[4], 0, 2, True, bc[4],
                    [dzc[-q + nh_d - 1] for q in range(nh)])
    _set_bc_literal(ref, ng, nh, cbc[5], 1, 2, True, bc[5],
                    [dzc[nz + q + nh_d - 1] for q in range(nh)])
    np.testing.assert_array_equal(p, ref)
