"""Writes tests/golden/fft_inverse_unnormalised.npz: seeded complex inputs and
their transforms with the sign and scaling of the reference's FFT stage
(cufftExec* with CUFFT_INVERSE, unnormalised: X[r] = sum_n x[n] e^{+2 pi i n r / N},
src/cuna_kernel.cu:43-49), computed with numpy as ifft * N in float64.

    python tests/golden/make_fft_golden.py
"""
import os

import numpy as np

SIZES = [2, 4, 8, 16, 32, 64, 128, 256, 1024]


def main():
    rng = np.random.default_rng(20240917)
    arrays = {}
    for n in SIZES:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        arrays[f"x{n}"] = x
        arrays[f"X{n}"] = np.fft.ifft(x) * n
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_inverse_unnormalised.npz")
    np.savez_compressed(out, sizes=np.array(SIZES), **arrays)
    print(out)


if __name__ == "__main__":
    main()
