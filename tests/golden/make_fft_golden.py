"""Writes the FFT golden fixtures: seeded random real fields and their
numpy.fft.rfftn images (a textbook library routine, independent of the
oracle).  Arrays are indexed [z, y, x] so x is the fastest axis, matching the
lattice layout.  Run from the repository root: python tests/golden/make_fft_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

for N, seed in ((8, 11), (16, 12)):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((N, N, N))
    F = np.fft.rfftn(f)
    np.savez_compressed(os.path.join(HERE, f"fft_N{N}.npz"), f=f, F=F)
    print("wrote", N)
