"""CPU checks of the product library: it loads without a GPU, exports every
symbol include/flutas_amd.h declares, rejects bad arguments before touching
the device, and its host-side initgrid matches the oracle's."""
import os
import re

import numpy as np
import pytest

from cases import z_metrics

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flutas_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(flutas_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for name in ("flutas_psolve_create", "flutas_psolve_solve", "flutas_fillps",
                 "flutas_correc", "flutas_chkdiv", "flutas_pres_tw_src"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    from flutas_amd import api
    missing = [s for s in declared_symbols() if not hasattr(api.lib, s)]
    assert not missing, missing


def test_version_string():
    from flutas_amd import api
    assert b"gfx950" in api.lib.flutas_version()


@pytest.mark.parametrize("alpha", [0.0, 1.5, 2.0])
def test_initgrid_matches_oracle(oracle, alpha):
    from flutas_amd import api
    for n, nh in ((16, 1), (64, 2)):
        got = api.initgrid(n, alpha, 1.7, nh)
        ref = oracle.initgrid(n, alpha, 1.7, nh)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)


@pytest.mark.parametrize("inivel", ["hcl", "hcp", "zer", "poi"])
def test_initgrid_inivel_matches_oracle(oracle, inivel):
    """initgrid's inivel selection (src/initgrid.f90:32-39): 'hcl'/'hcp'
    cluster at the lower wall, other values at both; the library's host
    code equals the oracle bit for bit."""
    from flutas_amd import api
    for alpha in (0.0, 1.0, 2.0):
        for n, nh in ((16, 1), (64, 3)):
            got = api.initgrid(n, alpha, 1.3, nh, inivel=inivel)
            ref = oracle.initgrid(n, alpha, 1.3, nh, inivel=inivel)
            for g, r in zip(got, ref):
                np.testing.assert_array_equal(g, r)


@pytest.mark.parametrize("alpha", [0.5, 1.0, 2.0])
def test_oracle_one_end_grid_properties(oracle, alpha):
    """What the mathematics fixes for gridpoint_cluster_one_end: faces 0 and
    n at 0 and lz exactly, spacing growing away from the lower wall, and the
    first-to-last spacing ratio tending to sech^2(alpha) (the derivative
    ratio of 1 + tanh((z0-1) alpha)/tanh(alpha) at z0 = 0 and 1); the
    two-end grid instead has equal first and last spacings."""
    n, nh, lz = 512, 1, 2.0
    dzc, dzf, zc, zf = oracle.initgrid(n, alpha, lz, nh, inivel="hcl")
    F = lambda a, k: a[k + nh - 1]  # (1-nh : n+nh) indexing
    assert F(zf, 0) == 0.0 and abs(F(zf, n) - lz) <= 1e-15 * lz
    inner = np.array([F(dzf, k) for k in range(1, n + 1)])
    assert np.all(np.diff(inner) > 0)
    assert abs(inner[0] / inner[-1] - 1.0 / np.cosh(alpha) ** 2) < 1e-2
    _, dzf2, _, _ = oracle.initgrid(n, alpha, lz, nh, inivel="zer")
    assert abs(F(dzf2, 1) - F(dzf2, n)) < 1e-14


def test_fortran_mod_initgrid():
    """The drop-in mod_initgrid, called as the apps call it, returns the
    library's grid."""
    import subprocess
    from flutas_amd import api
    demo = os.path.join(ROOT, "flutas_amd", "fortran", "_build", "grid_demo")
    if not os.path.exists(demo):
        pytest.skip("Fortran modules not built (make -C flutas_amd/fortran)")
    for inivel, n, gr, nh in (("hcl", 24, 1.5, 2), ("zer", 16, 1.2, 1)):
        r = subprocess.run([demo, inivel, str(n), str(gr), "1.25", str(nh)],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        rows = [np.array([float(x) for x in line.split()]) for line in r.stdout.splitlines()]
        ref = api.initgrid(n, gr, 1.25, nh, inivel=inivel)
        for g, e in zip(rows, ref):
            np.testing.assert_allclose(g, e, rtol=1e-16, atol=0)


def test_invalid_arguments_rejected_without_device():
    """Argument validation happens before any device call."""
    from flutas_amd import api
    dz = np.ones(20)
    with pytest.raises(api.FlutasError) as e:
        api.PoissonPlan((8, 8, 7), (1, 1, 1), dz, dz, "PPPPNN")  # odd nz
    assert e.value.code == api.FLUTAS_EINVAL
    with pytest.raises(api.FlutasError) as e:
        api.PoissonPlan((8, 8, 8), (1, 1, 1), dz, dz, "PXPPNN")
    assert e.value.code == api.FLUTAS_EINVAL
    with pytest.raises(api.FlutasError) as e:
        api.PoissonPlan((8, 8, 8), (1, 1, 1), dz, dz, "PNPPNN")  # unpaired P
    assert e.value.code == api.FLUTAS_EINVAL
    # face-centred NN / DD in x or y (REDFT00/RODFT00 solves) are not built,
    # and the refined mode is cell-centred only
    for cbc, cf in (("NNPPNN", "fcc"), ("PPDDNN", "cfc")):
        with pytest.raises(api.FlutasError) as e:
            api.PoissonPlan((8, 8, 8), (1, 1, 1), dz, dz, cbc, c_or_f=cf)
        assert e.value.code == api.FLUTAS_EUNSUPPORTED
    with pytest.raises(api.FlutasError) as e:
        api.PoissonPlan((8, 8, 8), (1, 1, 1), dz, dz, "PPPPNN", c_or_f="ccf",
                        precision=api.FP32_REFINE, tol=1e-10)
    assert e.value.code == api.FLUTAS_EUNSUPPORTED
    with pytest.raises(api.FlutasError) as e:
        api.PoissonPlan((8, 8, 8), (1, 1, 1), dz, dz, "PPPPNN", c_or_f="ccx")
    assert e.value.code == api.FLUTAS_EINVAL
    with pytest.raises(api.FlutasError) as e:
        api.PoissonPlan((8, 8, 8), (1, 1, 1), dz, dz, "PPPPNN", precision=16)
    assert e.value.code == api.FLUTAS_EINVAL
    # odd sizes (src/sanity.f90:155) and a periodic z shorter than the
    # cyclic closure needs
    with pytest.raises(api.FlutasError) as e:
        api.PoissonPlan((8, 7, 8), (1, 1, 1), dz, dz, "PPPPNN")
    assert e.value.code == api.FLUTAS_EINVAL
    dz2 = np.ones(4)
    with pytest.raises(api.FlutasError) as e:
        api.PoissonPlan((8, 8, 2), (1, 1, 1), dz2, dz2, "PPPPPP")
    assert e.value.code == api.FLUTAS_EINVAL


def test_fortran_dropin_fails_loudly_without_gpu(tmp_path):
    """The amdflang-built drop-in stops with the library's message when no
    gfx950 device is present (no silent CPU fallback)."""
    import subprocess
    demo = os.path.join(ROOT, "flutas_amd", "fortran", "_build", "projection_demo")
    if not os.path.exists(demo):
        pytest.skip("Fortran modules not built (make -C flutas_amd/fortran)")
    ng, h = (8, 8, 8), 1
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as fh:
        fh.write(np.array(list(ng) + [h], np.int32).tobytes())
        fh.write(np.array([1.0, 1.0, 1.0, 0.0, 0.01, 1.0]).tobytes())
        fh.write(b"PPPPNN")
        fh.write(np.zeros(3 * 10 ** 3).tobytes())
    r = subprocess.run([demo, str(fin), str(tmp_path / "out.bin"), "host"],
                       capture_output=True, text=True, timeout=60)
    if r.returncode == 0:
        pytest.skip("a gfx950 device is present")
    assert "flutas_amd: initsolver failed" in r.stdout + r.stderr


def test_load_scalar_reference_format(tmp_path):
    """load_scalar (src/load.f90:92-130): written with (2E15.7, 1I9.8),
    read back list-directed."""
    from flutas_amd import api
    f = tmp_path / "scalar.out"
    api.load_scalar("w", f, 1.0, 10, 1e-3)
    assert f.read_text() == "  0.1000000E+01  0.1000000E-02 00000010\n"
    assert api.load_scalar("r", f) == (1.0, 10, 1e-3)
    api.load_scalar("w", f, -123.456789, 12345, 0.0)
    assert f.read_text() == " -0.1234568E+03  0.0000000E+00 00012345\n"
    t, it, dto = api.load_scalar("r", f)
    assert (t, it, dto) == (-123.4568, 12345, 0.0)
    # rounding that carries into the exponent, and three-digit exponents,
    # which Ew.d writes without the E
    api.load_scalar("w", f, 9.99999996, 1, 1.5e-120)
    assert f.read_text() == "  0.1000000E+02  0.1500000-119 00000001\n"
    api.load_scalar("w", f, -2.5e100, 3, 0.123456749999)
    assert f.read_text() == " -0.2500000+101  0.1234567E+00 00000003\n"
    # a list-directed line with commas, as a hand-edited restart might hold
    f.write_text("2.5, 1.0E-4, 7\n")
    assert api.load_scalar("r", f) == (2.5, 7, 1e-4)
    with pytest.raises(api.FlutasError):
        api.load_scalar("r", tmp_path / "missing.out")


def test_fortran_mod_load_round_trip(tmp_path):
    """The drop-in mod_load (src/load.f90: load, load_ns, load_scalar,
    cmpt_it_chkpt) on host arrays, called as the reference's apps call it
    (interior sections of halo arrays): round trips inside the Fortran demo,
    and the files it writes hold the reference's layouts (raw FP64 global
    arrays in Fortran order; u, v, w, p then time, istep)."""
    import subprocess
    demo = os.path.join(ROOT, "flutas_amd", "fortran", "_build", "restart_demo")
    if not os.path.exists(demo):
        pytest.skip("Fortran modules not built (make -C flutas_amd/fortran)")
    r = subprocess.run([demo, str(tmp_path) + "/"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "restart_demo ok 3" in r.stdout
    n = (6, 5, 4)
    i, j, k = np.meshgrid(*(np.arange(1, m + 1) for m in n), indexing="ij")
    u = i + 10.0 * j + 100.0 * k + 0.25
    got = np.fromfile(tmp_path / "fldu.bin", dtype=np.float64).reshape(n, order="F")
    np.testing.assert_array_equal(got, u)
    raw = np.fromfile(tmp_path / "fld.bin", dtype=np.float64)
    npts = u.size
    assert raw.size == 4 * npts + 2
    for q, ref in enumerate((u, -u, 0.5 * u, 1.0 / (1.0 + u))):
        np.testing.assert_array_equal(raw[q * npts:(q + 1) * npts].reshape(n, order="F"), ref)
    assert tuple(raw[-2:]) == (1.5, 42.0)
    assert (tmp_path / "scalar.out").read_text() == "  0.1250000E+00  0.1000000E-02 00001234\n"


FFT_DEMO = os.path.join(ROOT, "flutas_amd", "fortran", "_build", "fft_demo")


@pytest.mark.parametrize("bcs", ["PPNNcc", "NNDDcc", "NDDNcc", "DDPPcc", "PPNDff", "DNPPfc",
                                 "NDNNfc", "NNPPfc", "PPDDcf", "DDNNff"])
def test_fortran_fftini_normfft(oracle, bcs):
    """The drop-in mod_fft's fftini (src/fft.f90:24-157) gives the oracle's
    fftini normalisation for nx = 12, ny = 10, face-centred pairs included;
    it does no device work."""
    import subprocess
    if not os.path.exists(FFT_DEMO):
        pytest.skip("Fortran modules not built (make -C flutas_amd/fortran)")
    r = subprocess.run([FFT_DEMO, "-n", bcs], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    got = float(r.stdout.split()[0])
    nz = 4
    dzc, dzf, dzci, dzfi, _, _ = z_metrics(oracle, nz, 0.0, 1.0)
    s = oracle.Solver((12, 10, nz), (12.0, 10.0, 4.0), dzci, dzfi, bcs[:4] + "NN",
                      c_or_f=bcs[4:] + "c")
    assert got == s.normfft, (got, s.normfft)


