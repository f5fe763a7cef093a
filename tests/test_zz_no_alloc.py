"""The allocation-free chkdiv (flutas_chkdiv_noalloc, fixed per-GPU device
scratch) against flutas_chkdiv (partials in a per-call allocation): the
same kernels and reduction order, so both results agree bit for bit.  Two
calls in a row, the second on a smaller grid, so a stale partial left in
the shared scratch by the first call would show."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(api, oracle, ng, alpha, seed):
    nx, ny, nz = ng
    nh_d, nh_u = 1, 1
    dzc, dzf, _, _ = oracle.initgrid(nz, alpha, 1.0, nh_d)
    dzfi = 1.0 / dzf
    rng = np.random.default_rng(seed)
    u, v, w = (np.asfortranarray(rng.uniform(-1, 1, (nx + 2, ny + 2, nz + 2))) for _ in range(3))
    D = api.DeviceArray.from_host
    return dzfi, D(dzfi), D(u), D(v), D(w)


@pytest.mark.parametrize("sizes", [((96, 80, 72), (40, 24, 16)), ((512, 64, 128), (8, 8, 8))])
def test_chkdiv_noalloc_matches_chkdiv(api, oracle, sizes):
    lib = api.lib
    for call, ng in enumerate(sizes):
        nx, ny, nz = ng
        dli = (float(nx), float(ny) / 2.0, float(nz))
        dzfi, dzfi_d, ud, vd, wd = _fields(api, oracle, ng, 1.2, 30 + call)
        print(f"chkdiv_noalloc call {call}: grid {ng}", flush=True)
        ref = api.chkdiv(nx, ny, nz, dli[0], dli[1], dli[2], 1, 1, dzfi_d, ud, vd, wd)
        tot, mx = C.c_double(), C.c_double()
        rc = lib.flutas_chkdiv_noalloc(nx, ny, nz, C.c_double(dli[0]), C.c_double(dli[1]),
                                       C.c_double(dli[2]), 1, 1, api._dptr(dzfi_d),
                                       api._dptr(ud), api._dptr(vd), api._dptr(wd),
                                       C.byref(tot), C.byref(mx))
        assert rc == api.FLUTAS_OK, lib.flutas_last_error().decode()
        assert tot.value == ref[0]
        assert mx.value == ref[1]
        # and through the Python mirror
        assert api.chkdiv_noalloc(nx, ny, nz, dli[0], dli[1], dli[2], 1, 1, dzfi_d, ud, vd,
                                  wd) == ref
