"""Writes the committed golden fixtures (run from the repo root: python tests/golden/make_fixtures.py).

sobol_scipy.npz: Sobol points from scipy.stats.qmc.Sobol(d, scramble=False, bits=32), the
library generator on the Joe-Kuo D6 table (new-joe-kuo-6.21201), in Gray-code order starting at
point 0, as uint32 (x * 2^32). Kept: the first 300 points of dimensions
{0, 1, 2, 3, 479, 480, 1023, 2047} of a 2048-dimensional sequence, and the first 64 points of
dimensions {21199, 21200} of a 21201-dimensional one, plus points 1000..1063 (fast-forwarded)
of dimensions {0, 1, 2047}.
"""
import os

import numpy as np
from scipy.stats import qmc

HERE = os.path.dirname(os.path.abspath(__file__))


def pts(d, n, skip=0):
    s = qmc.Sobol(d, scramble=False, bits=32)
    if skip:
        s.fast_forward(skip)
    return np.round(s.random(n) * 2.0 ** 32).astype(np.uint64).astype(np.uint32)


def main():
    cols = np.array([0, 1, 2, 3, 479, 480, 1023, 2047])
    a = pts(2048, 300)[:, cols]
    tail_cols = np.array([21199, 21200])
    b = pts(21201, 64)[:, tail_cols]
    ff_cols = np.array([0, 1, 2047])
    c = pts(2048, 64, skip=1000)[:, ff_cols]
    np.savez(os.path.join(HERE, "sobol_scipy.npz"), dims2048_cols=cols, dims2048_first300=a,
             dims21201_cols=tail_cols, dims21201_first64=b, ff1000_cols=ff_cols, dims2048_from1000=c)


if __name__ == "__main__":
    main()
