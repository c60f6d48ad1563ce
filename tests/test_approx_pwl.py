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
    # the device's arithmetic (device_math.h: pwl_entry): -B as a signed
    # 24-bit operand times j, added to A in 32 bits, shifted logically
    assert (pairs[:, 1] < 2 ** 23).all() and (pairs[:, 0] < 2 ** 32).all()
    for lo, hi, n, shift, base in ((0, 64, 65536, 12, 0), (64, 96, 32768, 16, 65536), (96, 128, 32768, 16, 98304)):
        idx = np.arange(n)
        run = lo + (idx >> 10)
        dev = ((-pairs[run, 1]) * (idx & 1023) + pairs[run, 0]) & 0xFFFFFFFF
        assert ((dev >> shift) == table[base:base + n]).all()
