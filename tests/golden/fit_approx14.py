"""Derives a piecewise-linear form of the VRCP14PS / VRSQRT14PS tables.

approx14.bin (written by gen_approx14.c from this host's instructions) holds
the 16-bit result mantissas: rcp[65536] indexed by the top 16 input mantissa
bits, rsqrt[2][32768] by exponent parity and the top 15 bits. Both are
exactly piecewise linear over runs of 1024 entries: entry j of run i equals
floor((A_i - B_i * j) / 2^K) with integers A_i, B_i (K = 12 for rcp, 16 for
rsqrt). This script finds those integers with a linear program (largest
margin inside the floor constraints), checks every entry, and writes
approx14_pwl.txt: 64 rcp lines then 64 rsqrt lines (parity 0 runs, then
parity 1), "A B" each. The device evaluates the approximations from these
128 pairs instead of the 256 KiB tables.

    python tests/golden/fit_approx14.py
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "..", "..", "glimpsw_amd", "data")
RUN = 1024
K_RCP, K_RSQRT = 12, 16


def load_tables():
    t = np.fromfile(os.path.join(DATA, "approx14.bin"), dtype=np.uint16).astype(np.int64)
    assert t.size == 131072
    return t[:65536], t[65536:].reshape(2, 32768)


def _fit_run(vals: np.ndarray, K: int) -> tuple[int, int]:
    from scipy.optimize import linprog

    n = vals.size
    j = np.arange(n, dtype=np.float64)
    lo, hi = vals.astype(np.float64), vals + 1 - 1e-7
    A = np.vstack([np.stack([np.ones(n), -j], 1), np.stack([-np.ones(n), j], 1)])
    A = np.hstack([A, np.ones((2 * n, 1))])
    r = linprog(c=[0, 0, -1], A_ub=A, b_ub=np.concatenate([hi, -lo]),
                bounds=[(None, None), (None, None), (0, None)], method="highs")
    assert r.status == 0
    a, b, _ = r.x
    Ai, Bi = int(round(a * 2 ** K)), int(round(b * 2 ** K))
    assert ((Ai - Bi * np.arange(n, dtype=np.int64)) >> K == vals).all()
    return Ai, Bi


def evaluate(pairs: np.ndarray, K: int, n_entries: int) -> np.ndarray:
    idx = np.arange(n_entries, dtype=np.int64)
    A, B = pairs[idx // RUN, 0], pairs[idx // RUN, 1]
    return (A - B * (idx % RUN)) >> K


def main():
    rcp, rsq = load_tables()
    lines = []
    pr = [_fit_run(rcp[s:s + RUN], K_RCP) for s in range(0, 65536, RUN)]
    ps = [_fit_run(rsq[p][s:s + RUN], K_RSQRT) for p in (0, 1) for s in range(0, 32768, RUN)]
    assert (evaluate(np.array(pr), K_RCP, 65536) == rcp).all()
    assert (evaluate(np.array(ps[:32]), K_RSQRT, 32768) == rsq[0]).all()
    assert (evaluate(np.array(ps[32:]), K_RSQRT, 32768) == rsq[1]).all()
    for A, B in pr + ps:
        assert 0 <= A < 2 ** 32 and 0 <= B < 2 ** 31
        lines.append(f"{A} {B}")
    with open(os.path.join(DATA, "approx14_pwl.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {len(lines)} pairs")


if __name__ == "__main__":
    main()
