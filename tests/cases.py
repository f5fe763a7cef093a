"""Shared builders for test inputs: grids, BC patterns, seeded fields.

Fields are Fortran-ordered numpy arrays with the halo included, as the
reference passes them (p(0:n+1), u(1-nh_u:n+nh_u)).
"""
import numpy as np


def z_metrics(oracle, nz, alpha, lz, nh_d=1):
    """dzc, dzf, dzci, dzfi over (1-nh_d : nz+nh_d) from the oracle's
    initgrid (src/initgrid.f90:17-97)."""
    dzc, dzf, zc, zf = oracle.initgrid(nz, alpha, lz, nh_d)
    return dzc, dzf, 1.0 / dzc, 1.0 / dzf, zc, zf


def halo_field(ng, h, rng=None, scale=1.0):
    nx, ny, nz = ng
    shape = (nx + 2 * h, ny + 2 * h, nz + 2 * h)
    if rng is None:
        return np.zeros(shape, order="F")
    return np.asfortranarray(rng.uniform(-scale, scale, size=shape))


def interior(a, h=1):
    return a[h:-h, h:-h, h:-h]


def second_diff_1d(n, bc, dli):
    """Dense 1-D cell-centred second difference with ghost rules
    P: wrap, N: p0 = p1, D: p0 = -p1 (homogeneous)."""
    m = np.zeros((n, n))
    for i in range(n):
        m[i, i] = -2.0
        if i > 0:
            m[i, i - 1] = 1.0
        if i < n - 1:
            m[i, i + 1] = 1.0
    for side, row, nb in ((0, 0, n - 1), (1, n - 1, 0)):
        if bc[side] == "P":
            m[row, nb] += 1.0
        elif bc[side] == "N":
            m[row, row] += 1.0
        elif bc[side] == "D":
            m[row, row] -= 1.0
    return m * dli * dli


def tri_dense(a, b, c, periodic):
    n = len(b)
    m = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    if periodic:
        m[0, n - 1] += a[0]
        m[n - 1, 0] += c[n - 1]
    return m


def dense_operator(ng, cbc, dli, a, b, c):
    """7-point operator as a Kronecker sum, index i fastest (Fortran order)."""
    nx, ny, nz = ng
    lx = second_diff_1d(nx, cbc[0:2], dli[0])
    ly = second_diff_1d(ny, cbc[2:4], dli[1])
    lz = tri_dense(a, b, c, cbc[4:6] == "PP")
    ix, iy, iz = np.eye(nx), np.eye(ny), np.eye(nz)
    return (np.kron(iz, np.kron(iy, lx)) + np.kron(iz, np.kron(ly, ix)) +
            np.kron(lz, np.kron(iy, ix)))


def gauge_free(p):
    """Remove the additive constant (the null space of singular problems)."""
    return p - np.mean(p)


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
