"""The Fortran drop-in modules (flutas_amd/fortran) on the GPU: one
projection step (initsolver, fillps, chkdiv, updt_rhs_b, solver_gpu, boundp,
correc, bounduvw, fftend) run by the amdflang-built projection_demo, checked against
the same step through the FP64 oracle.  "host" passes plain Fortran arrays
(staged through device buffers by the modules), "device" passes HIP device
allocations."""
import os
import subprocess

import numpy as np
import pytest

from cases import relmax

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "flutas_amd", "fortran", "_build", "projection_demo")


def _velocity(ng, h, cbc, rng):
    """Random (1-h:n+h)^3 velocities with periodic halos in the P directions
    and w = 0 on the z walls when z is not periodic."""
    nx, ny, nz = ng
    out = []
    for _ in range(3):
        a = np.asfortranarray(rng.uniform(-1, 1, (nx + 2 * h, ny + 2 * h, nz + 2 * h)))
        for ax, n in enumerate(ng):
            if cbc[2 * ax] != "P":
                continue
            sl = [slice(None)] * 3
            src = [slice(None)] * 3
            sl[ax], src[ax] = slice(0, h), slice(n, n + h)
            a[tuple(sl)] = a[tuple(src)]
            sl[ax], src[ax] = slice(n + h, n + 2 * h), slice(h, 2 * h)
            a[tuple(sl)] = a[tuple(src)]
        out.append(a)
    if cbc[4] != "P":
        out[2][:, :, h - 1] = 0.0       # w(0)
        out[2][:, :, nz + h - 1] = 0.0  # w(nz)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("cbc,alpha,nh_u,mode,solver,const_coeff", [
    ("PPPPNN", 1.5, 1, "host", "gpu", True),
    ("PPPPNN", 1.5, 1, "device", "gpu", True),
    ("PPPPDD", 0.0, 2, "host", "gpu", True),
    ("NNPPNN", 1.2, 3, "device", "gpu", True),
    # mod_solver_cpu (src/solver_cpu.f90:20-31) backed by the same plan
    ("PPPPNN", 1.5, 1, "host", "cpu", True),
    ("DDPPNN", 0.0, 1, "device", "cpu", True),
    # modules built without _CONSTANT_COEFFS_POISSON: fillps without the
    # rho0 factor (src/fillps.f90:55-57), correc with the face density
    ("PPPPNN", 1.5, 1, "host", "gpu", False),
    ("NNPPNN", 0.0, 2, "device", "gpu", False),
    # modules built with _SINGLE_PRECISION (rp = FP32, src/types.f90:6-12)
    ("PPPPNN", 1.5, 1, "host", "gpu", "sp"),
    ("PPPPDD", 0.0, 2, "device", "gpu", "sp"),
    ("NNPPNN", 1.2, 1, "device", "cpu", "sp"),
])
def test_fortran_projection_step(oracle, tmp_path, cbc, alpha, nh_u, mode, solver,
                                 const_coeff):
    """The drop-in step against the oracle's FP64 step: the FP64 builds to
    1e-14 (right-hand side) and 1e-12 (pressure, velocities); the FP32
    build (its input rounded to FP32, FP32 stencils, FP32 transforms with
    an FP64 z elimination) to FP32 accuracy: 1e-5 for the right-hand side
    and the corrected velocities, 1e-4 for the pressure."""
    sp = const_coeff == "sp"
    demo = DEMO + "_sp" if sp else (DEMO if const_coeff else DEMO + "_vc")
    const_coeff = True if sp else const_coeff
    if not os.path.exists(demo):
        pytest.fail(f"{os.path.basename(demo)} is not built (make -C flutas_amd/fortran)")
    ng = (32, 24, 16)
    nx, ny, nz = ng
    lengths = (2 * np.pi, np.pi, 1.0)
    # rho0 != 1 so that the rho0 factor of fillps is visible
    dt, rho0 = 0.01, 2.5
    rng = np.random.default_rng(5)
    u, v, w = _velocity(ng, nh_u, cbc, rng)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        fh.write(np.array([nx, ny, nz, nh_u], np.int32).tobytes())
        fh.write(np.array(list(lengths) + [alpha, dt, rho0], np.float64).tobytes())
        fh.write(cbc.encode())
        for a in (u, v, w):
            fh.write(a.tobytes(order="F"))
    print(f"fortran step {cbc} alpha={alpha} nh_u={nh_u} mode={mode} solver={solver} "
          f"const_coeff={const_coeff}", flush=True)
    r = subprocess.run([demo, str(fin), str(fout), mode, solver], capture_output=True,
                       text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    # chkdiv reports as the reference does (src/chkdiv.f90:66)
    assert "Total divergence = " in r.stdout and "| Maximum divergence = " in r.stdout

    hp = (nx + 2, ny + 2, nz + 2)
    hu = u.shape
    raw = np.fromfile(fout, np.float32 if sp else np.float64).astype(np.float64)
    if sp:
        # the FP32 build sees the velocities rounded to FP32
        u, v, w = (a.astype(np.float32).astype(np.float64) for a in (u, v, w))
    sizes = [np.prod(hp)] * 2 + [np.prod(hu)] * 3
    parts = np.split(raw, np.cumsum(sizes)[:-1])
    f_g, p_g = (x.reshape(hp, order="F") for x in parts[:2])
    u_g, v_g, w_g = (x.reshape(hu, order="F") for x in parts[2:])
    divtot_g, divmax_g = np.loadtxt(str(fout) + ".div")

    # the same step through the oracle, with nh_d = max(nh_u, 1) and the
    # velocity BCs the demo derives from the pressure pattern
    nh_d = max(nh_u, 1)
    vel = {"P": "P", "N": "D", "D": "N"}
    cbcvel = "".join(vel[c] for _ in range(3) for c in cbc)
    dzc, dzf, zc, zf = oracle.initgrid(nz, alpha, lengths[2], nh_d)
    dzci, dzfi = 1.0 / dzc, 1.0 / dzf
    dli = tuple(n / l for n, l in zip(ng, lengths))
    dl = tuple(1.0 / d for d in dli)
    s = oracle.Solver(ng, dli, dzci, dzfi, cbc, nh_d=nh_d)
    p = np.zeros(hp, order="F")
    uo, vo, wo = (a.copy(order="F") for a in (u, v, w))
    oracle.fillps(ng, nh_d, nh_u, dli, dzfi, 1.0 / dt, rho0 if const_coeff else 1.0,
                  uo, vo, wo, p)
    divtot_o, divmax_o = oracle.chkdiv(ng, dli, nh_d, nh_u, dzfi, uo, vo, wo)
    s.updt_rhs_b(p)
    f_o = p.copy(order="F")
    s.solve(p)
    oracle.boundp(cbc, ng, (0.0,) * 6, nh_d, dl, dzc, dzf, p)
    rho = np.full(hp, rho0, order="F")
    oracle.correc(ng, nh_d, nh_u, dli, dzci, dt, rho0, p, uo, vo, wo, rho=rho,
                  const_coeff=const_coeff)
    oracle.bounduvw(cbcvel, ng, np.zeros(18), nh_d, nh_u, [0] * 6, dl, dzc, dzf, uo, vo, wo)

    inner = (slice(1, -1),) * 3
    tol_f, tol_p, tol_u = (1e-5, 1e-4, 1e-5) if sp else (1e-14, 1e-12, 1e-12)
    assert relmax(f_g[inner], f_o[inner]) <= tol_f
    assert relmax(p_g, p) <= tol_p
    for g, o in ((u_g, uo), (v_g, vo), (w_g, wo)):
        assert relmax(g, o) <= tol_u
    tol_d = 1e-5 if sp else 1e-12
    assert divmax_g == pytest.approx(divmax_o, rel=tol_d)
    assert abs(divtot_g - divtot_o) <= tol_d * divmax_o * nx * ny * nz


FFT_DEMO = os.path.join(ROOT, "flutas_amd", "fortran", "_build", "fft_demo")


@pytest.mark.gpu
@pytest.mark.parametrize("ng,bcs", [
    ((12, 10, 3), "PPNNcc"),
    ((32, 24, 4), "NNDDcc"),
    ((16, 12, 2), "NDDNcc"),
    ((12, 16, 3), "PPNDff"),
    ((20, 12, 2), "DNPPfc"),
    ((12, 10, 3), "NNDDff"),
    ((14, 16, 2), "DDNNff"),
])
def test_fortran_fft_drop_in(oracle, tmp_path, ng, bcs):
    """The drop-in mod_fft as solver_cpu drives it (src/solver_cpu.f90:59-89)
    on host arrays: fft(arrplan(1,1)) is the x transform of every line and
    fft(arrplan(1,2)) the y transform, each equal to the oracle's r2r of the
    kind find_fft selects; the backward pair times normfft returns the
    input."""
    if not os.path.exists(FFT_DEMO):
        pytest.fail("fft_demo is not built (make -C flutas_amd/fortran)")
    nx, ny, nz = ng
    a = np.asfortranarray(np.random.default_rng(sum(ng)).standard_normal(ng))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        fh.write(np.array(ng, np.int32).tobytes())
        fh.write(bcs.encode())
        fh.write(a.tobytes(order="F"))
    print(f"fft_demo {ng} {bcs}", flush=True)
    r = subprocess.run([FFT_DEMO, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fft_demo ok" in r.stdout
    raw = np.fromfile(fout, dtype=np.float64)
    npts = a.size
    normfft = raw[0]
    bx, bxy, back = (raw[1 + q * npts:1 + (q + 1) * npts].reshape(ng, order="F") for q in range(3))
    kx, _, _ = oracle.find_fft(bcs[0:2], bcs[4])
    ky, _, _ = oracle.find_fft(bcs[2:4], bcs[5])
    # a face-centred DD direction transforms n - 1 points; the last face
    # (the boundary) is left as it is (src/fft.f90:71-73, :110-112)
    mx = nx - (1 if bcs[0:2] == "DD" and bcs[4] == "f" else 0)
    my = ny - (1 if bcs[2:4] == "DD" and bcs[5] == "f" else 0)
    ref_x = a.copy()
    for j in range(ny):
        for k in range(nz):
            ref_x[:mx, j, k] = oracle.r2r(kx, a[:mx, j, k])
    assert relmax(bx, ref_x) <= 1e-13
    ref_xy = bx.copy()
    for i in range(nx):
        for k in range(nz):
            ref_xy[i, :my, k] = oracle.r2r(ky, bx[i, :my, k])
    assert relmax(bxy, ref_xy) <= 1e-13
    assert relmax(back[:mx, :my], a[:mx, :my]) <= 1e-13
    assert normfft > 0.0
