"""The oracle's r2r transforms against golden vectors (scipy.fft / numpy.fft
outputs committed under tests/golden) and against its own long-double direct
sums.  FFTW is un-vendored; these fixtures pin FFTW's unnormalised r2r
definitions (src/fft.f90:233-291 selects the kinds)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "r2r_golden.npz")
KINDS = ["R2HC", "HC2R", "REDFT00", "REDFT10", "REDFT01", "REDFT11",
         "RODFT00", "RODFT10", "RODFT01", "RODFT11"]


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("kind", KINDS)
def test_fast_matches_golden(oracle, gold, kind):
    for n in gold["sizes"]:
        x = gold[f"x_{n}"]
        ref = gold[f"{kind}_{n}"]
        got = oracle.r2r(kind, x)
        err = np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300)
        assert err < 1e-13, (kind, n, err)


@pytest.mark.parametrize("kind", KINDS)
def test_direct_matches_golden(oracle, gold, kind):
    for n in gold["sizes"]:
        if n > 200:
            continue
        x = gold[f"x_{n}"]
        ref = gold[f"{kind}_{n}"]
        got = oracle.r2r(kind, x, direct=True)
        err = np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300)
        assert err < 1e-13, (kind, n, err)


@pytest.mark.parametrize("pair,norm", [(("R2HC", "HC2R"), lambda n: n),
                                       (("REDFT10", "REDFT01"), lambda n: 2 * n),
                                       (("RODFT10", "RODFT01"), lambda n: 2 * n),
                                       (("REDFT11", "REDFT11"), lambda n: 2 * n),
                                       (("RODFT11", "RODFT11"), lambda n: 2 * n),
                                       (("REDFT00", "REDFT00"), lambda n: 2 * (n - 1)),
                                       (("RODFT00", "RODFT00"), lambda n: 2 * (n + 1))])
def test_round_trip_normalisation(oracle, pair, norm):
    """fwd then bwd multiplies by the factor normfft undoes
    (src/fft.f90:87,125,150: norm(1)*(n+norm(2)))."""
    rng = np.random.default_rng(3)
    for n in [4, 6, 10, 16, 30, 64, 96]:
        x = rng.standard_normal(n)
        y = oracle.r2r(pair[1], oracle.r2r(pair[0], x))
        np.testing.assert_allclose(y / norm(n), x, rtol=0, atol=1e-13 * np.max(np.abs(x)))


def test_unknown_kind_rejected(oracle):
    with pytest.raises(ValueError):
        oracle.r2r(2, np.ones(8))  # DHT is not used by FluTAS
