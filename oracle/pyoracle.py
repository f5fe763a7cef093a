"""ctypes view of the C FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Fields are numpy
arrays in Fortran order with the halo included, exactly as the C functions
expect them (see flutas_oracle.h).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

KIND = {"R2HC": 0, "HC2R": 1, "REDFT00": 3, "REDFT01": 4, "REDFT10": 5,
        "REDFT11": 6, "RODFT00": 7, "RODFT01": 8, "RODFT10": 9, "RODFT11": 10}

_lib = None
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    assert a.dtype == np.float64
    return a.ctypes.data_as(_D)


def _f64(a):
    return np.require(a, np.float64, ["F", "A", "W"])


def r2r(kind, x, direct=False):
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    fn = lib().or_r2r_direct if direct else lib().or_r2r
    rc = fn(C.c_int(KIND[kind] if isinstance(kind, str) else kind),
            C.c_int(x.size), _p(x), _p(out))
    if rc:
        raise ValueError(f"or_r2r failed for kind={kind} n={x.size}")
    return out


def find_fft(bc, c_or_f="c"):
    """find_fft (src/fft.f90:233-291) of the pair bc ("PP", "NN", ...):
    (kind_fwd, kind_bwd, norm) with FFTW's kind numbers."""
    kf, kb = C.c_int(), C.c_int()
    norm = np.zeros(2)
    rc = lib().or_find_fft(C.c_char(bc[0].encode()), C.c_char(bc[1].encode()),
                           C.c_char(c_or_f.encode()), C.byref(kf), C.byref(kb), _p(norm))
    if rc:
        raise ValueError(f"find_fft: unknown pair {bc!r} {c_or_f!r}")
    return kf.value, kb.value, (float(norm[0]), float(norm[1]))


def eigenvalues(n, bc, c_or_f="c"):
    lam = np.zeros(n)
    lib().or_eigenvalues(C.c_int(n), C.c_char(bc[0].encode()),
                         C.c_char(bc[1].encode()), C.c_char(c_or_f.encode()),
                         _p(lam))
    return lam


def initgrid(n, gr, lz, nh_d=1, inivel="zer"):
    arrs = [np.zeros(n + 2 * nh_d) for _ in range(4)]
    lib().or_initgrid_inivel(C.c_char_p(inivel.encode()), C.c_int(n), C.c_double(gr),
                             C.c_double(lz), C.c_int(nh_d), *[_p(a) for a in arrs])
    return arrs  # dzc, dzf, zc, zf over (1-nh_d : n+nh_d)


def tridmatrix(bc, n, nh_d, dzci, dzfi, c_or_f="c"):
    a, b, c = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().or_tridmatrix(C.c_char(bc[0].encode()), C.c_char(bc[1].encode()),
                        C.c_int(n), C.c_double(0.0), C.c_int(nh_d),
                        _p(np.ascontiguousarray(dzci)), _p(np.ascontiguousarray(dzfi)),
                        C.c_char(c_or_f.encode()), _p(a), _p(b), _p(c))
    return a, b, c


class Solver:
    """initsolver + solver_cpu restated at one rank (src/initsolver.f90:21,
    src/solver_cpu.f90:20).  cbc is a 6-char string in Fortran (0:1,3)
    order, e.g. 'PPPPNN'."""

    def __init__(self, ng, dli, dzci, dzfi, cbc, bc=(0.0,) * 6, c_or_f="ccc",
                 nh_d=1):
        self.ng = tuple(int(v) for v in ng)
        nx, ny, nz = self.ng
        self.cbc = cbc
        self.c_or_f = c_or_f
        self.nh_d = nh_d
        self.dli = np.asarray(dli, np.float64)
        self.dzci = np.ascontiguousarray(dzci, np.float64)
        self.dzfi = np.ascontiguousarray(dzfi, np.float64)
        self.bc = np.asarray(bc, np.float64)
        self.lambdaxy = np.zeros((nx, ny), order="F")
        self.a, self.b, self.c = np.zeros(nz), np.zeros(nz), np.zeros(nz)
        self.kinds = np.zeros(4, np.int32)
        self.sizes = np.zeros(2, np.int32)
        nf = C.c_double()
        self.rhsbx = np.zeros((ny, nz, 2), order="F")
        self.rhsby = np.zeros((nx, nz, 2), order="F")
        self.rhsbz = np.zeros((nx, ny, 2), order="F")
        ngc = (C.c_int * 3)(*self.ng)
        rc = lib().or_initsolver(
            ngc, _p(self.dli), C.c_int(nh_d), _p(self.dzci), _p(self.dzfi),
            C.c_char_p(cbc.encode()), _p(self.bc), C.c_char_p(c_or_f.encode()),
            _p(self.lambdaxy), _p(self.a), _p(self.b), _p(self.c),
            self.kinds.ctypes.data_as(_I), self.sizes.ctypes.data_as(_I),
            C.byref(nf), _p(self.rhsbx), _p(self.rhsby), _p(self.rhsbz))
        if rc:
            raise ValueError("or_initsolver: unsupported BC combination")
        self.normfft = nf.value

    def solve(self, p, pin_gauge=True):
        """In-place on p(0:nx+1,0:ny+1,0:nz+1) (Fortran order)."""
        assert p.flags.f_contiguous and p.shape == tuple(v + 2 for v in self.ng)
        ngc = (C.c_int * 3)(*self.ng)
        rc = lib().or_solver_cpu(
            ngc, self.kinds.ctypes.data_as(_I), self.sizes.ctypes.data_as(_I),
            C.c_double(self.normfft), _p(self.lambdaxy), _p(self.a),
            _p(self.b), _p(self.c), C.c_char(self.cbc[4].encode()),
            C.c_char(self.cbc[5].encode()), C.c_char(self.c_or_f[2].encode()),
            _p(p), C.c_int(1 if pin_gauge else 0))
        if rc:
            raise RuntimeError("or_solver_cpu failed")
        return p

    def laplacian(self, pdense):
        """L p for a dense (nx,ny,nz) F-order block; see or_apply_laplacian."""
        pdense = np.asfortranarray(pdense, np.float64)
        out = np.zeros_like(pdense, order="F")
        ngc = (C.c_int * 3)(*self.ng)
        lib().or_apply_laplacian(ngc, C.c_char_p(self.cbc.encode()), _p(self.dli),
                                 _p(self.a), _p(self.b), _p(self.c),
                                 _p(pdense), _p(out))
        return out

    def updt_rhs_b(self, p):
        nx, ny, nz = self.ng
        lib().or_updt_rhs_b(C.c_int(nx), C.c_int(ny), C.c_int(nz),
                            C.c_char_p(self.c_or_f.encode()),
                            C.c_char_p(self.cbc.encode()), _p(self.rhsbx),
                            _p(self.rhsby), _p(self.rhsbz), _p(p))
        return p


def fillps(ng, nh_d, nh_u, dli, dzfi, dti, rho0, u, v, w, p):
    nx, ny, nz = ng
    lib().or_fillps(C.c_int(nx), C.c_int(ny), C.c_int(nz), C.c_int(nh_d),
                    C.c_int(nh_u), C.c_double(dli[0]), C.c_double(dli[1]),
                    C.c_double(dli[2]), _p(dzfi), C.c_double(dti),
                    C.c_double(rho0), _p(u), _p(v), _p(w), _p(p))
    return p


def correc(ng, nh_d, nh_u, dli, dzci, dt, rho0, p, u, v, w, rho=None,
           const_coeff=True):
    nx, ny, nz = ng
    if rho is None:
        rho = p
    lib().or_correc(C.c_int(nx), C.c_int(ny), C.c_int(nz), C.c_int(nh_d),
                    C.c_int(nh_u), C.c_double(dli[0]), C.c_double(dli[1]),
                    C.c_double(dli[2]), _p(dzci), C.c_double(dt),
                    C.c_double(rho0), _p(p), _p(u), _p(v), _p(w), _p(rho),
                    C.c_int(1 if const_coeff else 0))


def pres_tw_src(ng, dli, nh_d, nh_u, dzci, rho0i, f_t12, f_t12_o, p, pold, rho,
                u, v, w, const_coeff=True):
    nx, ny, nz = ng
    lib().or_pres_tw_src(C.c_int(nx), C.c_int(ny), C.c_int(nz),
                         C.c_double(dli[0]), C.c_double(dli[1]),
                         C.c_double(dli[2]), C.c_int(nh_d), C.c_int(nh_u),
                         _p(dzci), C.c_double(rho0i), C.c_double(f_t12),
                         C.c_double(f_t12_o), _p(p), _p(pold), _p(rho), _p(u),
                         _p(v), _p(w), C.c_int(1 if const_coeff else 0))


def pres_sp_src(ng, f_t12, dli, nh_d, nh_u, dzci, rho0i, pold, u, v, w):
    nx, ny, nz = ng
    lib().or_pres_sp_src(C.c_int(nx), C.c_int(ny), C.c_int(nz),
                         C.c_double(f_t12), C.c_double(dli[0]),
                         C.c_double(dli[1]), C.c_double(dli[2]), C.c_int(nh_d),
                         C.c_int(nh_u), _p(dzci), C.c_double(rho0i), _p(pold),
                         _p(u), _p(v), _p(w))


def chkdiv(ng, dli, nh_d, nh_u, dzfi, u, v, w):
    nx, ny, nz = ng
    tot, mx = C.c_double(), C.c_double()
    lib().or_chkdiv(C.c_int(nx), C.c_int(ny), C.c_int(nz), C.c_double(dli[0]),
                    C.c_double(dli[1]), C.c_double(dli[2]), C.c_int(nh_d),
                    C.c_int(nh_u), _p(dzfi), _p(u), _p(v), _p(w), C.byref(tot),
                    C.byref(mx))
    return tot.value, mx.value


def boundp(cbc, ng, bc, nh_d, dl, dzc, dzf, p):
    ngc = (C.c_int * 3)(*ng)
    bcv = np.asarray(bc, np.float64)
    dlv = np.asarray(dl, np.float64)
    lib().or_boundp(C.c_char_p(cbc.encode()), ngc, _p(bcv), C.c_int(nh_d),
                    _p(dlv), _p(np.ascontiguousarray(dzc)),
                    _p(np.ascontiguousarray(dzf)), _p(p))
    return p


def bounduvw(cbc, ng, bc, nh_d, nh_u, isoutflow, dl, dzc, dzf, u, v, w):
    """bounduvw at one rank (src/bound.f90:17-144) in place on the halo-nh_u
    Fortran arrays u, v, w.  cbc: 18 characters (or a (2,3,3) array) and bc:
    18 values in the (0:1,3,3) memory order [side + 2*dir + 6*comp];
    isoutflow: 6 flags [side + 2*dir]."""
    cbc = cbc if isinstance(cbc, str) else "".join(np.asarray(cbc).ravel(order="F"))
    assert len(cbc) == 18
    ngc = (C.c_int * 3)(*ng)
    bcv = np.ascontiguousarray(np.asarray(bc, np.float64).ravel(order="F"))
    outv = (C.c_int * 6)(*[int(bool(x)) for x in np.asarray(isoutflow).ravel(order="F")])
    dlv = np.asarray(dl, np.float64)
    for a in (u, v, w):
        assert a.flags.f_contiguous and a.dtype == np.float64
    lib().or_bounduvw(C.c_char_p(cbc.encode()), ngc, _p(bcv), C.c_int(nh_d), C.c_int(nh_u),
                      outv, _p(dlv), _p(np.ascontiguousarray(dzc)),
                      _p(np.ascontiguousarray(dzf)), _p(u), _p(v), _p(w))
    return u, v, w


def boundp_nh(cbc, ng, bc, nh_d, nh_p, dl, dzc, dzf, p):
    """boundp (src/bound.f90:146-225) on a field with halo nh_p, in place."""
    ngc = (C.c_int * 3)(*ng)
    bcv = np.asarray(bc, np.float64)
    dlv = np.asarray(dl, np.float64)
    assert p.flags.f_contiguous and p.shape == tuple(n + 2 * nh_p for n in ng)
    lib().or_boundp_nh(C.c_char_p(cbc.encode()), ngc, _p(bcv), C.c_int(nh_d), C.c_int(nh_p),
                       _p(dlv), _p(np.ascontiguousarray(dzc)), _p(np.ascontiguousarray(dzf)),
                       _p(p))
    return p
