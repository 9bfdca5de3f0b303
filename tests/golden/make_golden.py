"""Writes the committed fixtures under tests/golden/.

* analytic_128_reference.json -- the reference's only recorded output, the
  screenshot image.png shown at Readme.md:58-61 (harness
  GPU_FFT_TEST_SCRIPT/test.cu run with mpiexec -np 2 at 128^3, FP64):
  the eight printed spectral bins after 1/N normalisation and the two
  per-rank round-trip max errors, transcribed as data.
* brute_small.npz -- seeded inputs (splitmix64 uniform[-1,1)) and their 3D
  R2C half spectra computed by explicit O(N^2) DFT sums (oracle.brute_rfftn),
  independent of pocketfft, for small grids including odd Nz and
  2^a 3^b 5^c edges.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import gfft_oracle as O  # noqa: E402

# image.png rows: "rank = r , data at (x,y,z) ------> re,im"
REFERENCE_PRINT = [
    (1, 4, 123, 6, -0.0, -1.0),
    (1, 1, 126, 3, -0.0, -1.0),
    (1, 127, 126, 3, 0.0, 1.0),
    (1, 124, 123, 6, -0.0, 1.0),
    (0, 127, 2, 3, -0.0, -1.0),
    (0, 1, 2, 3, 0.0, 1.0),
    (0, 4, 5, 6, 0.0, 1.0),
    (0, 124, 5, 6, -0.0, -1.0),
]
REFERENCE_MAX_ERRORS = [1.06581e-14, 1.42109e-14]

SHAPES = [(4, 6, 10), (3, 5, 9), (8, 12, 20), (6, 4, 5), (2, 3, 8), (5, 10, 15), (1, 4, 6), (4, 1, 7)]


def main():
    peaks = [{"rank": r, "x": x, "y": y, "z": z, "re": re, "im": im} for r, x, y, z, re, im in REFERENCE_PRINT]
    doc = {
        "source": "reference image.png (Readme.md:58-61); test.cu at 128^3 FP64, mpiexec -np 2",
        "nx": 128, "ny": 128, "nz": 128, "procs": 2, "threshold": 0.1,
        "printed_bins": peaks, "roundtrip_max_error_per_rank": REFERENCE_MAX_ERRORS,
    }
    with open(os.path.join(HERE, "analytic_128_reference.json"), "w") as fh:
        json.dump(doc, fh, indent=1)

    arrays = {}
    for i, (nx, ny, nz) in enumerate(SHAPES):
        f = O.random_field(nx, ny, nz, seed=1000 + i)
        arrays[f"in_{i}"] = f
        arrays[f"out_{i}"] = O.brute_rfftn(f)
        arrays[f"shape_{i}"] = np.array([nx, ny, nz])
    np.savez_compressed(os.path.join(HERE, "brute_small.npz"), count=np.array(len(SHAPES)), **arrays)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
