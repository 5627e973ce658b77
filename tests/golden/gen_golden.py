"""Writes tests/golden/fft_golden.npz: seeded complex128 grids and their 3-D DFTs computed by
numpy.fft (pocketfft), which uses FFTW's sign conventions (forward exp(-2 pi i ...), backward
exp(+2 pi i ...)).  The backward arrays are unnormalised (numpy's ifftn times N), matching
FFTW_BACKWARD as the reference calls it (src/cgrid_fft.c:48-52).

Run:  python tests/golden/gen_golden.py
"""
import os

import numpy as np

# appended shapes keep the seeds (1000 + position) of the earlier ones
SHAPES = [(16, 16, 16), (8, 4, 2), (6, 10, 3), (1, 12, 9), (1, 96, 64), (1, 1, 64), (4, 32, 8), (2, 2, 256),
          (12, 48, 6), (1, 192, 24), (3, 1, 384)]


def main():
    out = {}
    for idx, shp in enumerate(SHAPES):
        rng = np.random.default_rng(1000 + idx)
        a = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        tag = "x".join(map(str, shp))
        out[f"in_{tag}"] = a
        out[f"fwd_{tag}"] = np.fft.fftn(a)
        out[f"bwd_{tag}"] = np.fft.ifftn(a) * a.size
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
