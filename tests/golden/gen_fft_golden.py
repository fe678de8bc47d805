"""Writes fft_golden.npz: FFT input/output vectors computed with numpy.fft
(pocketfft, FP64) for pinning the oracle and the HIP transforms.

Run from the repository root: python tests/golden/gen_fft_golden.py
"""
import os

import numpy as np

rng = np.random.default_rng(20240611)
out = {}
for n in (2, 3, 4, 8, 16, 32, 48, 64, 96, 128, 192, 256, 512):
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out[f"c1d_{n}_in"] = x
    out[f"c1d_{n}_fwd"] = np.fft.fft(x)
    out[f"c1d_{n}_bwd"] = np.fft.ifft(x) * n
for n in (8, 16):
    r = rng.standard_normal((2, n, n, n))
    out[f"r3d_{n}_in"] = r
    out[f"r3d_{n}_rfftn"] = np.fft.rfftn(r, axes=(1, 2, 3))
x = rng.standard_normal((6, 6, 6)) + 1j * rng.standard_normal((6, 6, 6))
out["c3d_6_in"] = x
out["c3d_6_fftn"] = np.fft.fftn(x)
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz"), **out)
print("wrote", len(out), "arrays")
