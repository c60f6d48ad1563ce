"""The piecewise-linear form of the VRCP14PS / VRSQRT14PS tables used on the
device reproduces every entry of the tables measured on this host's
AVX-512 instructions (glimpsw_amd/data/approx14.bin)."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "glimpsw_amd", "data")


def test_pairs_reproduce_tables():
    table = np.fromfile(os.path.join(DATA, "approx14.bin"), dtype=np.uint16).astype(np.int64)
    pairs = np.loadtxt(os.path.join(DATA, "approx14_pwl.txt"), dtype=np.int64).reshape(128, 2)
    idx = np.arange(65536)
    assert (((pairs[idx >> 10, 0] - pairs[idx >> 10, 1] * (idx & 1023)) >> 12) == table[:65536]).all()
    idx = np.arange(32768)
    for odd in (0, 1):
        run = 64 + odd * 32 + (idx >> 10)
        got = (pairs[run, 0] - pairs[run, 1] * (idx & 1023)) >> 16
        assert (got == table[65536 + odd * 32768:65536 + (odd + 1) * 32768]).all()
    # device arithmetic ranges: rcp in int32, rsqrt in uint32
    assert pairs[:64, 0].max() < 2 ** 31 and (pairs[:64, 1] * 1023).max() < 2 ** 31
    assert pairs[64:, 0].max() < 2 ** 32 and (pairs[64:, 0] - pairs[64:, 1] * 1023).min() >= 0
