"""Python host mirror of the reference's projection-step interface, bound to
libflutas_amd.so through its C-ABI (ctypes; no torch types).

The functions keep the reference's names, argument order and meaning:

    initsolver(n, dims, dims_z, dli, nh_d, dzci, dzfi, cbc, bc, c_or_f)
        -> (lambdaxy, a, b, c, arrplan, normfft, rhsbx, rhsby, rhsbz)
                                                (src/initsolver.f90:21)
    solver_gpu(n, dims, arrplan, normfft, lambdaxy, a, b, c, bc, c_or_f, p)
                                                (src/solver_gpu.f90:31)
    fillps(nx,ny,nz,nh_d,nh_u,dxi,dyi,dzi,dzfi,dti,rho0,u,v,w,p)
                                                (src/fillps.f90:16)
    correc(nx,ny,nz,nh_d,nh_u,dxi,dyi,dzi,dzci,dt,rho0,p,u,v,w,rho)
                                                (src/correc.f90:16)
    pres_tw_src / pres_sp_src / chkdiv / boundp / bounduvw / updt_rhs_b / initgrid

Field arguments are DeviceArray objects (Fortran-ordered device
allocations with the halo included).  Failures raise FlutasError carrying
the library's status code; a missing or unbuilt library raises at import
time -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflutas_amd.so")

FLUTAS_OK = 0
FLUTAS_ENOMEM = 12
FLUTAS_EINVAL = 22
FLUTAS_EDEVICE = 1001
FLUTAS_ENONFINITE = 1002
FLUTAS_ENOTCONV = 1003
FLUTAS_EUNSUPPORTED = 1004
FP64 = 64
FP32_REFINE = 32


class FlutasError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_D = C.POINTER(C.c_double)
_VP = C.c_void_p


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m flutas_amd.build` "
            "(hipcc, gfx950)")
    lib = C.CDLL(LIB_PATH)
    lib.flutas_last_error.restype = C.c_char_p
    lib.flutas_version.restype = C.c_char_p
    lib.flutas_psolve_stage_name.restype = C.c_char_p
    lib.flutas_psolve_stage_name.argtypes = [_VP, C.c_int]
    for name in ("flutas_device_free", "flutas_psolve_destroy",
                 "flutas_event_record", "flutas_event_destroy"):
        getattr(lib, name).argtypes = [_VP]
    lib.flutas_device_alloc.argtypes = [C.c_size_t, C.POINTER(_VP)]
    lib.flutas_memcpy_h2d.argtypes = [_VP, _VP, C.c_size_t]
    lib.flutas_memcpy_d2h.argtypes = [_VP, _VP, C.c_size_t]
    lib.flutas_memset_d.argtypes = [_VP, C.c_int, C.c_size_t]
    lib.flutas_set_stream.argtypes = [_VP]
    return lib


lib = _load()


def _check(rc):
    if rc != FLUTAS_OK:
        raise FlutasError(rc, lib.flutas_last_error().decode())


def synchronize():
    _check(lib.flutas_synchronize())


class DeviceArray:
    """A Fortran-ordered device allocation of a given shape and dtype."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        ptr = _VP()
        _check(lib.flutas_device_alloc(C.c_size_t(self.nbytes), C.byref(ptr)))
        self.ptr = ptr.value

    @classmethod
    def from_host(cls, a):
        a = np.asfortranarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.asfortranarray(a, dtype=self.dtype)
        assert a.shape == self.shape
        _check(lib.flutas_memcpy_h2d(_VP(self.ptr), a.ctypes.data_as(_VP),
                                     C.c_size_t(self.nbytes)))

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        _check(lib.flutas_memcpy_d2h(out.ctypes.data_as(_VP), _VP(self.ptr),
                                     C.c_size_t(self.nbytes)))
        return out

    def zero(self):
        _check(lib.flutas_memset_d(_VP(self.ptr), 0, C.c_size_t(self.nbytes)))

    def free(self):
        if self.ptr:
            lib.flutas_device_free(_VP(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def c(self):
        return C.cast(C.c_void_p(self.ptr), _D)


def _dptr(a):
    if isinstance(a, DeviceArray):
        return a.c
    raise TypeError("field arguments must be DeviceArray objects")


def _chars(s, n):
    s = "".join(s) if not isinstance(s, str) else s
    if len(s) != n:
        raise ValueError(f"expected {n} characters, got {s!r}")
    return C.c_char_p(s.encode())


def _bc6(cbc):
    """Accepts 'PPPPNN' or a (2,3)-shaped Fortran (0:1,3) array of chars."""
    if isinstance(cbc, str):
        return cbc
    arr = np.asarray(cbc)
    if arr.shape == (2, 3):
        return "".join(arr[:, d][s] for d in range(3) for s in range(2))
    return "".join(arr.ravel())


class PoissonPlan:
    """initsolver + fftini (src/initsolver.f90:21, src/fft.f90:24): owns
    the device plan; `arrplan` of the reference is this object."""

    def __init__(self, ng, dli, dzci, dzfi, cbc, bc=(0.0,) * 6, c_or_f="ccc",
                 nh_d=1, precision=FP64, tol=1e-10):
        self.ng = tuple(int(v) for v in ng)
        self.cbc = _bc6(cbc)
        self.nh_d = int(nh_d)
        self.precision = int(precision)
        dzci = np.ascontiguousarray(dzci, np.float64)
        dzfi = np.ascontiguousarray(dzfi, np.float64)
        if dzci.size < self.ng[2] + 2 * nh_d or dzfi.size < self.ng[2] + 2 * nh_d:
            raise ValueError("dzci/dzfi must cover (1-nh_d : nz+nh_d)")
        h = _VP()
        ngc = (C.c_int * 3)(*self.ng)
        bcv = (C.c_double * 6)(*[float(v) for v in np.asarray(bc, np.float64).ravel()])
        dliv = (C.c_double * 3)(*[float(v) for v in dli])
        _check(lib.flutas_psolve_create(
            ngc, _chars(self.cbc, 6), bcv, _chars(c_or_f, 3), dliv,
            C.c_int(nh_d), dzci.ctypes.data_as(_D), dzfi.ctypes.data_as(_D),
            C.c_int(self.precision), C.c_double(tol), C.byref(h)))
        self._h = h.value
        nx, ny, nz = self.ng
        self.lambdaxy = np.zeros((nx, ny), order="F")
        self.a, self.b, self.c = np.zeros(nz), np.zeros(nz), np.zeros(nz)
        self.rhsbx = np.zeros((ny, nz, 2), order="F")
        self.rhsby = np.zeros((nx, nz, 2), order="F")
        self.rhsbz = np.zeros((nx, ny, 2), order="F")
        nf = C.c_double()
        _check(lib.flutas_psolve_coeffs(
            _VP(self._h), self.lambdaxy.ctypes.data_as(_D), self.a.ctypes.data_as(_D),
            self.b.ctypes.data_as(_D), self.c.ctypes.data_as(_D), C.byref(nf),
            self.rhsbx.ctypes.data_as(_D), self.rhsby.ctypes.data_as(_D),
            self.rhsbz.ctypes.data_as(_D)))
        self.normfft = nf.value

    def solve(self, p):
        """solver_gpu on the interior of a halo-1 DeviceArray (in place)."""
        if p.shape != tuple(v + 2 for v in self.ng) or p.dtype != np.float64:
            raise ValueError("p must be a float64 DeviceArray of shape ng+2")
        _check(lib.flutas_psolve_solve(_VP(self._h), p.c))

    def solve_async(self, p):
        """solve, enqueued on the library stream (flutas_psolve_solve_async)."""
        if p.shape != tuple(v + 2 for v in self.ng) or p.dtype != np.float64:
            raise ValueError("p must be a float64 DeviceArray of shape ng+2")
        _check(lib.flutas_psolve_solve_async(_VP(self._h), p.c))

    def solve_into(self, f, p):
        """Out-of-place solve: RHS from f's interior, solution into p's."""
        for a in (f, p):
            if a.shape != tuple(v + 2 for v in self.ng) or a.dtype != np.float64:
                raise ValueError("f and p must be float64 DeviceArrays of shape ng+2")
        _check(lib.flutas_psolve_solve_into(_VP(self._h), f.c, p.c))

    def stats(self):
        it, rr = C.c_int(), C.c_double()
        _check(lib.flutas_psolve_stats(_VP(self._h), C.byref(it), C.byref(rr)))
        return it.value, rr.value

    def residual(self, f, p):
        rr = C.c_double()
        _check(lib.flutas_psolve_residual(_VP(self._h), f.c, p.c, C.byref(rr)))
        return rr.value

    def profile(self, f, p, reps):
        """Mean per-stage durations (ms) of `reps` solves f -> p."""
        ms = (C.c_double * 32)()
        ns = C.c_int()
        _check(lib.flutas_psolve_profile(_VP(self._h), f.c, p.c, C.c_int(reps),
                                         C.c_int(32), ms, C.byref(ns)))
        return [(lib.flutas_psolve_stage_name(_VP(self._h), s).decode(), ms[s])
                for s in range(ns.value)]

    def time_stage(self, f, p, stage, reps=8):
        """Average duration (ms) of one launch of stage 0..3 of the FP64
        four-pass pipeline, launched `reps` times back to back after a
        solve f -> p (flutas_psolve_time_stage)."""
        ms = C.c_double()
        _check(lib.flutas_psolve_time_stage(_VP(self._h), f.c, p.c, C.c_int(stage),
                                            C.c_int(reps), C.byref(ms)))
        return ms.value

    def set_zchunks(self, n):
        """Chunks per x-slab of the z-chunked slab passes (1 = whole slabs,
        0 = the plan's default); returns the count in use
        (flutas_psolve_set_zchunks)."""
        _check(lib.flutas_psolve_set_zchunks(_VP(self._h), C.c_int(n)))
        return self.zchunks

    @property
    def zchunks(self):
        n = C.c_int()
        _check(lib.flutas_psolve_zchunks(_VP(self._h), C.byref(n)))
        return n.value

    def tune(self, f, p, reps=5):
        """Times `reps` solves f -> p at 1..8 z chunks, keeps the fastest
        (flutas_psolve_tune); returns (chosen, {count: ms})."""
        chosen = C.c_int()
        ms = (C.c_double * 8)()
        _check(lib.flutas_psolve_tune(_VP(self._h), f.c, p.c, C.c_int(reps),
                                      C.byref(chosen), ms))
        return chosen.value, {v + 1: ms[v] for v in range(8) if ms[v] >= 0}

    def rhsb_device(self):
        rx, ry, rz = _D(), _D(), _D()
        _check(lib.flutas_psolve_rhsb_device(_VP(self._h), C.byref(rx),
                                             C.byref(ry), C.byref(rz)))
        return rx, ry, rz

    def destroy(self):
        if getattr(self, "_h", None):
            lib.flutas_psolve_destroy(_VP(self._h))
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Graph:
    """A recorded sequence of asynchronous library calls (flutas_graph_*):

        with api.Graph() as g:       # calls inside are recorded, not run
            plan.solve_async(p)
            api.boundp(...)
        g.launch()                   # replays them as one launch
    """

    def __init__(self):
        self._h = None

    def __enter__(self):
        _check(lib.flutas_graph_begin())
        return self

    def __exit__(self, exc_type, exc, tb):
        h = _VP()
        rc = lib.flutas_graph_end(C.byref(h))
        if exc_type is None:
            _check(rc)
            self._h = h.value
        return False

    def launch(self):
        _check(lib.flutas_graph_launch(_VP(self._h)))

    def destroy(self):
        if self._h:
            lib.flutas_graph_destroy(_VP(self._h))
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class BatchPlan:
    """Many independent problems solved together (flutas_batch_*): one
    persistent launch over a work queue of every pass of every problem,
    captured in a hipGraph.

    problems: sequence of (ng, dli, dzci, dzfi, cbc) with nh_d = 1.
    precision: FP64, or FP32_REFINE with the relative residual target tol."""

    def __init__(self, problems, nh_d=1, precision=FP64, tol=1e-10):
        self.problems = [(tuple(int(v) for v in ng), tuple(float(v) for v in dli),
                          np.ascontiguousarray(dzci, np.float64),
                          np.ascontiguousarray(dzfi, np.float64), _bc6(cbc))
                         for ng, dli, dzci, dzfi, cbc in problems]
        n = len(self.problems)
        ng = (C.c_int * (3 * n))(*[v for pr in self.problems for v in pr[0]])
        cbc = "".join(pr[4] for pr in self.problems).encode()
        dli = (C.c_double * (3 * n))(*[v for pr in self.problems for v in pr[1]])
        self._keep = [pr[2] for pr in self.problems] + [pr[3] for pr in self.problems]
        dzci = (_D * n)(*[pr[2].ctypes.data_as(_D) for pr in self.problems])
        dzfi = (_D * n)(*[pr[3].ctypes.data_as(_D) for pr in self.problems])
        h = _VP()
        _check(lib.flutas_batch_create_ex(C.c_int(n), ng, C.c_char_p(cbc), dli, C.c_int(nh_d),
                                          dzci, dzfi, C.c_int(precision), C.c_double(tol),
                                          C.byref(h)))
        self._h = h.value
        self.precision = precision
        self._fields = None

    def set_fields(self, fs, ps):
        """Halo-1 DeviceArrays of every problem: right-hand sides and outputs."""
        n = len(self.problems)
        if len(fs) != n or len(ps) != n:
            raise ValueError("one f and one p per problem")
        for (ng, *_), f, p in zip(self.problems, fs, ps):
            for a in (f, p):
                if a.shape != tuple(v + 2 for v in ng) or a.dtype != np.float64:
                    raise ValueError("fields must be float64 DeviceArrays of shape ng+2")
        fv = (_D * n)(*[f.c for f in fs])
        pv = (_D * n)(*[p.c for p in ps])
        _check(lib.flutas_batch_set_fields(_VP(self._h), fv, pv))
        self._fields = (list(fs), list(ps))

    def solve(self):
        _check(lib.flutas_batch_solve(_VP(self._h)))

    def stats(self):
        """Per problem: (correction rounds, relative residual) of the last solve."""
        n = len(self.problems)
        it, rr = (C.c_int * n)(), (C.c_double * n)()
        _check(lib.flutas_batch_stats(_VP(self._h), it, rr))
        return list(it), list(rr)

    def tune(self, reps=3):
        _check(lib.flutas_batch_tune(_VP(self._h), C.c_int(reps)))

    def tuning(self):
        """Per problem: (x-tile variant, plane-group variant), 0 full, 1 half,
        2 quarter."""
        n = len(self.problems)
        th, gh = (C.c_int * n)(), (C.c_int * n)()
        _check(lib.flutas_batch_tuning(_VP(self._h), th, gh))
        return list(th), list(gh)

    def launch_shape(self):
        """(shape, planes per x-tile item); shape 0: one launch per pass and
        line length, 1: one persistent launch per pass."""
        v, k = C.c_int(), C.c_int()
        _check(lib.flutas_batch_launch_shape(_VP(self._h), C.byref(v), C.byref(k)))
        return v.value, k.value

    def launch_info(self):
        """(workgroups, queue items, workgroups per CU, dynamic LDS bytes) of
        the persistent launch."""
        wg, it, pc, lds = C.c_int(), C.c_int(), C.c_int(), C.c_size_t()
        _check(lib.flutas_batch_launch_info(_VP(self._h), C.byref(wg), C.byref(it),
                                            C.byref(pc), C.byref(lds)))
        return wg.value, it.value, pc.value, lds.value

    def destroy(self):
        if getattr(self, "_h", None):
            lib.flutas_batch_destroy(_VP(self._h))
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---- reference-named entry points --------------------------------------

def initsolver(n, dims, dims_z, dli, nh_d, dzci, dzfi, cbc, bc, c_or_f,
               precision=FP64, tol=1e-10):
    """src/initsolver.f90:21-44 at one GPU (dims = dims_z = (1,1,1))."""
    if tuple(dims) != (1, 1, 1) and tuple(dims)[:2] != (1, 1):
        raise FlutasError(FLUTAS_EUNSUPPORTED, "one GPU: dims must be 1")
    plan = PoissonPlan(n, dli, dzci, dzfi, cbc, bc, c_or_f, nh_d, precision, tol)
    return (plan.lambdaxy, plan.a, plan.b, plan.c, plan, plan.normfft,
            plan.rhsbx, plan.rhsby, plan.rhsbz)


def fftini(plan):  # the plan is created by initsolver/PoissonPlan
    return plan


def fftend(arrplan):
    """src/fft.f90:159-179."""
    arrplan.destroy()


def solver_gpu(n, dims, arrplan, normfft, lambdaxy, a, b, c, bc, c_or_f, p):
    """src/solver_gpu.f90:31-47: the coefficients live in the plan; they are
    accepted for signature parity."""
    arrplan.solve(p)


def fillps(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzfi, dti, rho0, u, v, w, p):
    _check(lib.flutas_fillps(nx, ny, nz, nh_d, nh_u, C.c_double(dxi), C.c_double(dyi),
                             C.c_double(dzi), _dptr(dzfi), C.c_double(dti),
                             C.c_double(rho0), _dptr(u), _dptr(v), _dptr(w), _dptr(p)))


def pres_tw_src(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzci, rho0i, f_t12,
                f_t12_o, p, pold, rho, u, v, w, const_coeff=True):
    _check(lib.flutas_pres_tw_src(nx, ny, nz, C.c_double(dxi), C.c_double(dyi),
                                  C.c_double(dzi), nh_d, nh_u, _dptr(dzci),
                                  C.c_double(rho0i), C.c_double(f_t12),
                                  C.c_double(f_t12_o), _dptr(p), _dptr(pold),
                                  _dptr(rho), _dptr(u), _dptr(v), _dptr(w),
                                  1 if const_coeff else 0))


def pres_sp_src(nx, ny, nz, f_t12, dxi, dyi, dzi, nh_d, nh_u, dzci, rho0i,
                pold, u, v, w):
    _check(lib.flutas_pres_sp_src(nx, ny, nz, C.c_double(f_t12), C.c_double(dxi),
                                  C.c_double(dyi), C.c_double(dzi), nh_d, nh_u,
                                  _dptr(dzci), C.c_double(rho0i), _dptr(pold),
                                  _dptr(u), _dptr(v), _dptr(w)))


def correc(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzci, dt, rho0, p, u, v, w,
           rho=None, const_coeff=True):
    rp = _dptr(rho) if rho is not None else _D()
    _check(lib.flutas_correc(nx, ny, nz, nh_d, nh_u, C.c_double(dxi), C.c_double(dyi),
                             C.c_double(dzi), _dptr(dzci), C.c_double(dt),
                             C.c_double(rho0), _dptr(p), _dptr(u), _dptr(v),
                             _dptr(w), rp, 1 if const_coeff else 0))


def chkdiv(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzfi, u, v, w):
    tot, mx = C.c_double(), C.c_double()
    _check(lib.flutas_chkdiv(nx, ny, nz, C.c_double(dxi), C.c_double(dyi),
                             C.c_double(dzi), nh_d, nh_u, _dptr(dzfi), _dptr(u),
                             _dptr(v), _dptr(w), C.byref(tot), C.byref(mx)))
    return tot.value, mx.value


def chkdiv_noalloc(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzfi, u, v, w):
    """chkdiv through the allocation-free entry point (fixed device scratch)."""
    tot, mx = C.c_double(), C.c_double()
    _check(lib.flutas_chkdiv_noalloc(nx, ny, nz, C.c_double(dxi), C.c_double(dyi),
                                     C.c_double(dzi), nh_d, nh_u, _dptr(dzfi), _dptr(u),
                                     _dptr(v), _dptr(w), C.byref(tot), C.byref(mx)))
    return tot.value, mx.value


def boundp(cbc, n, bc, dl, dzc_0, dzc_n, p):
    ngc = (C.c_int * 3)(*[int(v) for v in n])
    bcv = (C.c_double * 6)(*[float(v) for v in np.asarray(bc, np.float64).ravel()])
    dlv = (C.c_double * 3)(*[float(v) for v in dl])
    _check(lib.flutas_boundp(_chars(_bc6(cbc), 6), ngc, bcv, dlv,
                             C.c_double(dzc_0), C.c_double(dzc_n), _dptr(p)))


def boundp_halo(cbc, n, bc, nh_d, nh_p, dl, dzc, dzf, p):
    """boundp (src/bound.f90:146-225) on a DeviceArray with halo nh_p;
    dzc, dzf: host arrays over (1-nh_d : nz+nh_d)."""
    ngc = (C.c_int * 3)(*[int(v_) for v_ in n])
    bcv = (C.c_double * 6)(*[float(x) for x in np.asarray(bc, np.float64).ravel()])
    dlv = (C.c_double * 3)(*[float(x) for x in dl])
    dzc = np.ascontiguousarray(dzc, np.float64)
    dzf = np.ascontiguousarray(dzf, np.float64)
    _check(lib.flutas_boundp_halo(_chars(_bc6(cbc), 6), ngc, bcv, C.c_int(nh_d), C.c_int(nh_p),
                                  dlv, dzc.ctypes.data_as(_D), dzf.ctypes.data_as(_D), _dptr(p)))


def bounduvw(cbc, n, bc, nh_d, nh_u, isoutflow, dl, dzc, dzf, u, v, w):
    """bounduvw (src/bound.f90:17-144) on halo-nh_u DeviceArrays u, v, w.
    cbc: 18 characters (or a (2,3,3) array) and bc: 18 values in the
    reference's (0:1,3,3) memory order; isoutflow: 6 flags (0:1,3);
    dzc, dzf: host arrays over (1-nh_d : nz+nh_d)."""
    cbc = cbc if isinstance(cbc, str) else "".join(np.asarray(cbc).ravel(order="F"))
    ngc = (C.c_int * 3)(*[int(v_) for v_ in n])
    bcv = (C.c_double * 18)(*[float(x) for x in np.asarray(bc, np.float64).ravel(order="F")])
    outv = (C.c_int * 6)(*[int(bool(x)) for x in np.asarray(isoutflow).ravel(order="F")])
    dlv = (C.c_double * 3)(*[float(x) for x in dl])
    dzc = np.ascontiguousarray(dzc, np.float64)
    dzf = np.ascontiguousarray(dzf, np.float64)
    _check(lib.flutas_bounduvw(_chars(cbc, 18), ngc, bcv, C.c_int(nh_d), C.c_int(nh_u), outv,
                               dlv, dzc.ctypes.data_as(_D), dzf.ctypes.data_as(_D),
                               _dptr(u), _dptr(v), _dptr(w)))


def updt_rhs_b(nx, ny, nz, c_or_f, cbc, rhsbx, rhsby, rhsbz, p):
    """rhsb* are device pointers (PoissonPlan.rhsb_device()) or DeviceArrays."""
    ptrs = [(_dptr(r) if isinstance(r, DeviceArray) else r) for r in (rhsbx, rhsby, rhsbz)]
    _check(lib.flutas_updt_rhs_b(nx, ny, nz, _chars(c_or_f, 3), _chars(_bc6(cbc), 6),
                                 *ptrs, _dptr(p)))


def fillps_rhs(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzfi, dti, rho0, u, v, w,
               c_or_f, cbc, rhsbx, rhsby, rhsbz, p):
    """fillps followed by updt_rhs_b in one pass (flutas_fillps_rhs)."""
    ptrs = [(_dptr(r) if isinstance(r, DeviceArray) else r) for r in (rhsbx, rhsby, rhsbz)]
    _check(lib.flutas_fillps_rhs(nx, ny, nz, nh_d, nh_u, C.c_double(dxi), C.c_double(dyi),
                                 C.c_double(dzi), _dptr(dzfi), C.c_double(dti),
                                 C.c_double(rho0), _dptr(u), _dptr(v), _dptr(w),
                                 _chars(c_or_f, 3), _chars(_bc6(cbc), 6), *ptrs, _dptr(p)))


def correc_update(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzci, dt, rho0, phi, pold,
                  pnew, u, v, w, rho=None, const_coeff=True):
    """correc with the increment phi, then pnew = pold + phi, in one pass
    (flutas_correc_update); pnew may be pold, not phi."""
    rp = _dptr(rho) if rho is not None else _D()
    _check(lib.flutas_correc_update(nx, ny, nz, nh_d, nh_u, C.c_double(dxi),
                                    C.c_double(dyi), C.c_double(dzi), _dptr(dzci),
                                    C.c_double(dt), C.c_double(rho0), _dptr(phi),
                                    _dptr(pold), _dptr(pnew), _dptr(u), _dptr(v),
                                    _dptr(w), rp, 1 if const_coeff else 0))


def correc_update_bc(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzci, dt, rho0, phi, pold,
                     pnew, u, v, w, cbc, bc, dl, dzc_0, dzc_n, rho=None, const_coeff=True):
    """boundp(phi), correc_update and boundp(pnew) in one pass
    (flutas_correc_update_bc); phi's halo is neither read nor written."""
    rp = _dptr(rho) if rho is not None else _D()
    bcv = (C.c_double * 6)(*[float(x) for x in np.asarray(bc, np.float64).ravel()])
    dlv = (C.c_double * 3)(*[float(x) for x in dl])
    _check(lib.flutas_correc_update_bc(nx, ny, nz, nh_d, nh_u, C.c_double(dxi),
                                       C.c_double(dyi), C.c_double(dzi), _dptr(dzci),
                                       C.c_double(dt), C.c_double(rho0), _dptr(phi),
                                       _dptr(pold), _dptr(pnew), _dptr(u), _dptr(v),
                                       _dptr(w), rp, 1 if const_coeff else 0,
                                       _chars(_bc6(cbc), 6), bcv, dlv, C.c_double(dzc_0),
                                       C.c_double(dzc_n)))


def copy_interior(nx, ny, nz, src, dst):
    _check(lib.flutas_copy_interior(nx, ny, nz, _dptr(src), _dptr(dst)))


def add_interior(nx, ny, nz, pold, p):
    _check(lib.flutas_add_interior(nx, ny, nz, _dptr(pold), _dptr(p)))


def load(io, filename, n, fld, halo=1):
    """load (src/load.f90:21): 'r' reads / 'w' writes the raw global field
    file from / into the interior of a DeviceArray with halo `halo`."""
    ngc = (C.c_int * 3)(*[int(v) for v in n])
    _check(lib.flutas_load(C.c_char(io.encode()), C.c_char_p(str(filename).encode()), ngc,
                           C.c_int(halo), _dptr(fld)))


def load_ns(io, filename, n, u, v, w, p, time=0.0, istep=0.0, halo_u=1):
    """load_ns (src/load.f90:132): u, v, w, p and (time, istep); returns
    (time, istep) (read back on 'r')."""
    ngc = (C.c_int * 3)(*[int(x) for x in n])
    t, it = C.c_double(time), C.c_double(istep)
    _check(lib.flutas_load_ns(C.c_char(io.encode()), C.c_char_p(str(filename).encode()), ngc,
                              C.c_int(halo_u), _dptr(u), _dptr(v), _dptr(w), _dptr(p),
                              C.byref(t), C.byref(it)))
    return t.value, it.value


def load_scalar(io, filename, time=0.0, istep=0, dto=0.0):
    """load_scalar (src/load.f90:92): returns (time, istep, dto)."""
    t, it, d = C.c_double(time), C.c_int(istep), C.c_double(dto)
    _check(lib.flutas_load_scalar(C.c_char(io.encode()), C.c_char_p(str(filename).encode()),
                                  C.byref(t), C.byref(it), C.byref(d)))
    return t.value, it.value, d.value


R2R_KIND = {"R2HC": 0, "HC2R": 1, "REDFT00": 3, "REDFT01": 4, "REDFT10": 5, "REDFT11": 6,
            "RODFT00": 7, "RODFT01": 8, "RODFT10": 9, "RODFT11": 10}


def fft_lines(kind, data, n, howmany, stride=1, dist=None):
    """fft (src/fft.f90:181-193): in-place r2r on device lines."""
    k = R2R_KIND[kind] if isinstance(kind, str) else int(kind)
    dist = n if dist is None else dist
    _check(lib.flutas_fft_lines(C.c_int(k), C.c_int(n), C.c_int(howmany),
                                C.c_longlong(stride), C.c_longlong(dist), _dptr(data)))


def fft_guru(kind, data, n, stride, howmany1, dist1, howmany2=1, dist2=0):
    """fft over the two-level batch of fftini's guru plans
    (src/fft.f90:75-86, :113-124)."""
    k = R2R_KIND[kind] if isinstance(kind, str) else int(kind)
    _check(lib.flutas_fft_guru(C.c_int(k), C.c_int(n), C.c_longlong(stride),
                               C.c_int(howmany1), C.c_longlong(dist1),
                               C.c_int(howmany2), C.c_longlong(dist2), _dptr(data)))


def initgrid(n, gr, lz, nh_d=1, inivel="zer"):
    """src/initgrid.f90:17-97 (inivel "hcl"/"hcp" cluster at the lower end);
    returns dzc, dzf, zc, zf over (1-nh_d:n+nh_d)."""
    arrs = [np.zeros(n + 2 * nh_d) for _ in range(4)]
    _check(lib.flutas_initgrid_inivel(C.c_char_p(inivel.encode()), n, C.c_double(gr),
                                      C.c_double(lz), nh_d,
                                      *[a.ctypes.data_as(_D) for a in arrs]))
    return arrs


class Event:
    def __init__(self):
        e = _VP()
        _check(lib.flutas_event_create(C.byref(e)))
        self._e = e.value

    def record(self):
        _check(lib.flutas_event_record(_VP(self._e)))

    def elapsed_ms(self, end):
        ms = C.c_float()
        _check(lib.flutas_event_elapsed_ms(_VP(self._e), _VP(end._e), C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib.flutas_event_destroy(_VP(self._e))
        except Exception:
            pass
