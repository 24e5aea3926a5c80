"""Write the FFT golden vectors used to pin the oracle's FP64 FFT.

The reference's FFT arithmetic lives in FFTW3 (not vendored, unpinned;
called at src/fft.c:28-29,174,186), which is absent here.  numpy.fft
(pocketfft, FP64) computes the same unnormalised transforms:
  forward r2c  X = numpy.fft.rfftn(x)               (FFTW_FORWARD, e^{-ikx})
  inverse c2r  x = numpy.fft.irfftn(X, s) * Ng**3   (FFTW c2r is unnormalised)
Inputs are seeded standard normals, so the file is reproducible.

    python tests/golden/make_golden.py   ->  tests/golden/fft_golden.npz
"""
import os

import numpy as np

SIZES = (1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 16)


def main():
    out = {}
    for Ng in SIZES:
        rng = np.random.default_rng(1000 + Ng)
        x = rng.standard_normal((Ng, Ng, Ng))
        out[f"x_{Ng}"] = x
        out[f"X_{Ng}"] = np.fft.rfftn(x)
        # an independent Hermitian spectrum (from another real field) for c2r
        y = rng.standard_normal((Ng, Ng, Ng))
        Y = np.fft.rfftn(y)
        out[f"Y_{Ng}"] = Y
        out[f"y_{Ng}"] = np.fft.irfftn(Y, s=(Ng, Ng, Ng), axes=(0, 1, 2)) * Ng ** 3
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
