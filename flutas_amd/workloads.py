"""Synthetic workloads of the pressure-projection benchmarks: grids, BC
patterns and seeded fields (numpy's counter-based Philox generator, so a
(seed, shape) pair always gives the same field).

Fields are Fortran-ordered float64 arrays with the halo included, in the
layout the reference passes to solver/fillps/correc.
"""
from dataclasses import dataclass, field

import numpy as np


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


@dataclass
class Case:
    name: str
    ng: tuple
    lengths: tuple
    cbc: str
    alpha: float = 0.0          # tanh stretching of z (0 = uniform)
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def dli(self):
        return tuple(n / l for n, l in zip(self.ng, self.lengths))

    @property
    def singular(self):
        return all(c in "PN" for c in self.cbc)

    @property
    def npoints(self):
        return int(np.prod(self.ng))


# The benchmark workloads by name (the uneven batch is batch_problems()).
WORKLOADS = {
    "channel_128_uniform": Case("channel_128_uniform", (128, 128, 128), (1.0, 1.0, 1.0), "PPPPNN", 0.0, 1),
    "channel_512_stretched": Case("channel_512_stretched", (512, 512, 512),
                                  (2 * np.pi, np.pi, 1.0), "PPPPNN", 1.5, 2),
    "twophase_256x256x512_periodic": Case("twophase_256x256x512_periodic", (256, 256, 512),
                                          (1.0, 1.0, 2.0), "PPPPPP", 0.0, 3),
    "neumann_1024x1024x512": Case("neumann_1024x1024x512", (1024, 1024, 512),
                                  (1.0, 1.0, 1.0), "NNNNNN", 0.0, 4),
}


def z_grid(api_or_oracle, case, nh_d=1):
    """dzc, dzf, dzci, dzfi over (1-nh_d : nz+nh_d) from initgrid."""
    dzc, dzf, zc, zf = api_or_oracle.initgrid(case.ng[2], case.alpha, case.lengths[2], nh_d)
    return dzc, dzf, 1.0 / dzc, 1.0 / dzf


def compatible(f, dzf, nh_d=1):
    """Remove the dz-weighted mean (left null vector of the singular
    operator: the z rows are dzf-weighted symmetric)."""
    nz = f.shape[2]
    w = dzf[nh_d:nh_d + nz][None, None, :]
    return f - np.sum(f * w) / (np.sum(w) * f.shape[0] * f.shape[1])


def rhs_normal(case, dzf, nh_d=1):
    """Halo-1 field whose interior is N(0,1) (seeded), made compatible when
    the BC pattern is singular; halo entries are zero."""
    nx, ny, nz = case.ng
    g = rng(case.seed)
    f = g.standard_normal((nx, ny, nz)).reshape((nx, ny, nz), order="F")
    if case.singular:
        f = compatible(f, dzf, nh_d)
    p = np.zeros((nx + 2, ny + 2, nz + 2), order="F")
    p[1:-1, 1:-1, 1:-1] = f
    return p


def velocity_uniform(case, scale=1.0, nh_u=1):
    """u, v, w with i.i.d. U[-scale, scale] entries including the halo."""
    nx, ny, nz = case.ng
    g = rng(case.seed + 1000)
    shape = (nx + 2 * nh_u, ny + 2 * nh_u, nz + 2 * nh_u)
    return [np.asfortranarray(g.uniform(-scale, scale, size=shape)) for _ in range(3)]


def sphere_density(case, rho1=1000.0, rho2=1.0, radius_frac=0.25, width=1.5):
    """rho = psi*rho1 + (1-psi)*rho2 for a tanh-smoothed sphere of radius
    0.25*Lx at the box centre (two-phase config), halo included
    (src/vof.f90:240-271 rule)."""
    nx, ny, nz = case.ng
    lx, ly, lz = case.lengths
    dx = lx / nx
    x = (np.arange(nx + 2) - 0.5) * dx
    y = (np.arange(ny + 2) - 0.5) * (ly / ny)
    z = (np.arange(nz + 2) - 0.5) * (lz / nz)
    r = np.sqrt((x[:, None, None] - lx / 2) ** 2 + (y[None, :, None] - ly / 2) ** 2 +
                (z[None, None, :] - lz / 2) ** 2)
    psi = 0.5 * (1.0 - np.tanh((r - radius_frac * lx) / (width * dx)))
    return np.asfortranarray(psi * rho1 + (1.0 - psi) * rho2)


def batch_problems(seed=5, count=64):
    """The batched uneven set: nx, ny in {64,96,128,192,256}, nz in
    {64,128,256}, BCs in {PP/PP/NN, NN/NN/NN, PP/PP/DD}, half stretched."""
    g = rng(seed)
    sizes_xy = [64, 96, 128, 192, 256]
    sizes_z = [64, 128, 256]
    bcs = ["PPPPNN", "NNNNNN", "PPPPDD"]
    out = []
    for q in range(count):
        nx = int(g.choice(sizes_xy))
        ny = int(g.choice(sizes_xy))
        nz = int(g.choice(sizes_z))
        cbc = bcs[int(g.integers(0, 3))]
        alpha = float(g.uniform(1.0, 2.0)) if (q % 2 == 1) else 0.0
        out.append(Case(f"batch{q}", (nx, ny, nz), (1.0, 1.0, 1.0), cbc, alpha,
                        seed * 1000 + q))
    return out
