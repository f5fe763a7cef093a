"""Writes tests/golden/r2r_golden.npz: seeded inputs and the outputs of the
unnormalised FFTW r2r kinds as scipy.fft / numpy.fft (pocketfft) compute them.

FFTW itself is absent here; scipy's unnormalised DCT/DST types I-IV are the
same definitions as FFTW's REDFT00/10/01/11 and RODFT00/10/01/11, and
numpy.fft.fft reordered to halfcomplex order is FFTW's R2HC.  Run from the
repository root:  python tests/golden/gen_golden.py
"""
import os

import numpy as np
import scipy.fft as sf

SIZES = [2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 15, 16, 17, 24, 30, 31, 32, 48, 64,
         96, 97, 100, 128, 192, 256, 512]


def r2hc(x):
    n = len(x)
    X = np.fft.fft(x)
    out = np.zeros(n)
    out[: n // 2 + 1] = X[: n // 2 + 1].real
    for k in range(1, (n + 1) // 2):
        out[n - k] = X[k].imag
    return out


def hc2r(h):
    n = len(h)
    X = np.zeros(n, complex)
    X[0] = h[0]
    for k in range(1, (n + 1) // 2):
        X[k] = h[k] + 1j * h[n - k]
        X[n - k] = h[k] - 1j * h[n - k]
    if n % 2 == 0:
        X[n // 2] = h[n // 2]
    return (np.fft.ifft(X) * n).real


KINDS = {
    "R2HC": r2hc, "HC2R": hc2r,
    "REDFT00": lambda x: sf.dct(x, 1), "REDFT10": lambda x: sf.dct(x, 2),
    "REDFT01": lambda x: sf.dct(x, 3), "REDFT11": lambda x: sf.dct(x, 4),
    "RODFT00": lambda x: sf.dst(x, 1), "RODFT10": lambda x: sf.dst(x, 2),
    "RODFT01": lambda x: sf.dst(x, 3), "RODFT11": lambda x: sf.dst(x, 4),
}


def main():
    rng = np.random.default_rng(20240607)
    data = {}
    for n in SIZES:
        x = rng.standard_normal(n)
        data[f"x_{n}"] = x
        for k, f in KINDS.items():
            if k == "REDFT00" and n < 2:
                continue
            data[f"{k}_{n}"] = f(x)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "r2r_golden.npz")
    np.savez_compressed(out, sizes=np.array(SIZES), **data)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
