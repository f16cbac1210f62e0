"""Writes fft_vectors.npz: small real grids and their FP64 numpy R2C transforms.

numpy (pocketfft) is not the reference; the vectors pin the FFT sign and
normalisation convention the reference takes from cuFFT C2C (forward e^{-ikr},
unnormalised inverse; pair_style.cu:97, :118).
"""
import pathlib

import numpy as np

SHAPES = [(7,), (21,), (35,), (40,), (6, 10), (9, 14), (5, 6, 7), (8, 6, 10), (21, 3, 4)]


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for k, s in enumerate(SHAPES):
        a = rng.standard_normal(s)
        out[f"in_{k}"] = a
        out[f"fwd_{k}"] = np.fft.fftn(a)
    np.savez(pathlib.Path(__file__).with_name("fft_vectors.npz"), n=len(SHAPES), **out)


if __name__ == "__main__":
    main()
