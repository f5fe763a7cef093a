"""Restart files in the reference's formats (src/load.f90) from and into
device fields: the file is the raw global FP64 array in Fortran order, so
it must equal numpy's Fortran-order bytes of the interior; reading it back
fills the interior and leaves the halo alone; a file of the wrong size is
rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("halo", [1, 3])
def test_load_roundtrip(api, tmp_path, halo):
    n = (12, 10, 6)
    rng = np.random.default_rng(3)
    shape = tuple(v + 2 * halo for v in n)
    a = np.asfortranarray(rng.standard_normal(shape))
    inner = (slice(halo, -halo),) * 3
    d = api.DeviceArray.from_host(a)
    f = tmp_path / "fldp.bin"
    api.load("w", f, n, d, halo)
    raw = np.fromfile(f, np.float64)
    np.testing.assert_array_equal(raw, a[inner].ravel(order="F"))
    b = np.asfortranarray(np.full(shape, 7.0))
    db = api.DeviceArray.from_host(b)
    api.load("r", f, n, db, halo)
    got = db.to_host()
    np.testing.assert_array_equal(got[inner], a[inner])
    got[inner] = 7.0
    assert np.all(got == 7.0)  # halo untouched
    with pytest.raises(api.FlutasError) as e:
        api.load("r", f, (12, 10, 7), db, halo)
    assert e.value.code == api.FLUTAS_EINVAL


def test_load_ns_roundtrip(api, tmp_path):
    n = (8, 6, 4)
    rng = np.random.default_rng(4)
    fields = [np.asfortranarray(rng.standard_normal(tuple(v + 2 for v in n))) for _ in range(4)]
    devs = [api.DeviceArray.from_host(x) for x in fields]
    f = tmp_path / "fld.bin"
    api.load_ns("w", f, n, *devs, time=1.25, istep=40.0)
    raw = np.fromfile(f, np.float64)
    npts = int(np.prod(n))
    for q, x in enumerate(fields):
        np.testing.assert_array_equal(raw[q * npts:(q + 1) * npts], x[1:-1, 1:-1, 1:-1].ravel(order="F"))
    assert list(raw[4 * npts:]) == [1.25, 40.0]
    outs = [api.DeviceArray(x.shape) for x in fields]
    for o in outs:
        o.zero()
    t, it = api.load_ns("r", f, n, *outs)
    assert (t, it) == (1.25, 40.0)
    for x, o in zip(fields, outs):
        np.testing.assert_array_equal(o.to_host()[1:-1, 1:-1, 1:-1], x[1:-1, 1:-1, 1:-1])
