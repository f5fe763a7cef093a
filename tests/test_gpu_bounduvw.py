"""bounduvw on the GPU (src/bound.f90:17-144, set_bc :227-646, outflow
:649-773) against the oracle, bit for bit: every halo value of u, v, w,
edges and corners included, for periodic, wall and outflow patterns and
halo widths 1 to 3, in one launch (every n > 2 nh) and as the per-face
launches (small grids)."""
import numpy as np
import pytest

from cases import z_metrics

CB = lambda s, d, c: s + 2 * d + 6 * c  # (0:1,3,3) memory order


def _pattern(name, rng):
    """cbc (18 chars), bc (18 values), isoutflow (6 flags)."""
    cbc, bc, out = ["P"] * 18, np.zeros(18), [0] * 6
    if name == "channel":        # periodic x, y; no-slip walls in z
        for c in range(3):
            for s in (0, 1):
                cbc[CB(s, 2, c)] = "D"
        bc[CB(1, 2, 0)] = 1.0    # moving upper wall (Couette)
    elif name == "cavity":       # walls everywhere, non-zero values
        for c in range(3):
            for d in range(3):
                for s in (0, 1):
                    cbc[CB(s, d, c)] = "D" if (c + d + s) % 3 else "N"
                    bc[CB(s, d, c)] = rng.uniform(-1, 1)
    elif name == "inflow_outflow":  # x inflow / outflow, periodic y, walls z
        for c in range(3):
            cbc[CB(0, 0, c)] = "D"
            cbc[CB(1, 0, c)] = "N"
            cbc[CB(0, 2, c)] = cbc[CB(1, 2, c)] = "D"
        bc[CB(0, 0, 0)] = 1.0
        out[1] = 1               # outflow at the upper x face
    elif name == "all_outflow":  # every non-periodic face an outflow
        for c in range(3):
            for d in range(3):
                for s in (0, 1):
                    cbc[CB(s, d, c)] = "N"
        out = [1] * 6
    return "".join(cbc), bc, out


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["channel", "cavity", "inflow_outflow", "all_outflow"])
@pytest.mark.parametrize("nh_u,ng", [(1, (12, 10, 16)), (2, (12, 10, 16)), (3, (12, 10, 16)),
                                     # some n <= 2 nh: the per-face launches
                                     (3, (4, 6, 5)), (2, (6, 4, 3))])
def test_bounduvw_matches_oracle(api, oracle, name, nh_u, ng):
    nh_d = nh_u
    rng = np.random.default_rng(nh_u * 7 + len(name))
    cbc, bc, out = _pattern(name, rng)
    dzc, dzf, _, _, _, _ = z_metrics(oracle, ng[2], 1.3, 1.0, nh_d)
    dl = (0.21, 0.17, 1.0 / ng[2])
    shape = tuple(n + 2 * nh_u for n in ng)
    host = [np.asfortranarray(rng.standard_normal(shape)) for _ in range(3)]
    ref = [a.copy(order="F") for a in host]
    dev = [api.DeviceArray.from_host(a) for a in host]
    print(f"bounduvw {name} nh_u={nh_u} {ng}", flush=True)
    for _ in range(2):  # applied twice: the second pass starts from set halos
        oracle.bounduvw(cbc, ng, bc, nh_d, nh_u, out, dl, dzc, dzf, *ref)
        api.bounduvw(cbc, ng, bc, nh_d, nh_u, out, dl, dzc, dzf, *dev)
        api.synchronize()
        for c in range(3):
            np.testing.assert_array_equal(dev[c].to_host(), ref[c], err_msg=f"component {c}")


def test_bounduvw_rejects_bad_arguments(api):
    """Checked on the host before any device work."""
    ng = (8, 8, 8)
    dz = np.ones(8 + 4)
    fake = api.DeviceArray.__new__(api.DeviceArray)
    fake.ptr = 1  # never dereferenced: every case fails validation first
    bad = [("P" * 17 + "X", 1, 1), ("D" + "P" * 17, 1, 1), ("P" * 18, 2, 3), ("P" * 18, 1, 9)]
    for cbc, nh_d, nh_u in bad:
        with pytest.raises(api.FlutasError) as e:
            api.bounduvw(cbc, ng, np.zeros(18), nh_d, nh_u, [0] * 6, (1, 1, 1), dz, dz,
                         fake, fake, fake)
        assert e.value.code == api.FLUTAS_EINVAL
    fake.ptr = None


@pytest.mark.gpu
@pytest.mark.parametrize("cbc", ["PPPPNN", "NNDDPP", "DNNDDN", "PPPPPP"])
@pytest.mark.parametrize("nh_p", [1, 2, 3])
def test_boundp_halo_matches_oracle(api, oracle, cbc, nh_p):
    """boundp for fields with any halo (temperature, VOF) against the
    oracle, bit for bit, applied twice."""
    ng = (12, 10, 16)
    nh_d = nh_p
    bc = tuple(b if c != "P" else 0.0 for b, c in zip((0.3, -0.2, 0.5, 0.1, -0.4, 0.25), cbc))
    dzc, dzf, _, _, _, _ = z_metrics(oracle, ng[2], 1.2, 1.0, nh_d)
    dl = (0.21, 0.17, 1.0 / ng[2])
    rng = np.random.default_rng(nh_p + 11)
    p = np.asfortranarray(rng.standard_normal(tuple(n + 2 * nh_p for n in ng)))
    ref = p.copy(order="F")
    dev = api.DeviceArray.from_host(p)
    print(f"boundp_halo {cbc} nh_p={nh_p}", flush=True)
    for _ in range(2):
        oracle.boundp_nh(cbc, ng, bc, nh_d, nh_p, dl, dzc, dzf, ref)
        api.boundp_halo(cbc, ng, bc, nh_d, nh_p, dl, dzc, dzf, dev)
        api.synchronize()
        np.testing.assert_array_equal(dev.to_host(), ref)
