"""Builds libglimp.so (HIP, gfx950) in-tree with hipcc.

    python -m glimpsw_amd.build

Device code is compiled without FP contraction (every FMA is written out) and
with packed-FP32 instruction selection turned off.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(HERE, "_gen")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libglimp.so")
SOURCES = ["glimp_kernels.hip", "resolve.hip", "diff.hip", "capi.hip", "meshlets.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GLIMP_ARCH", "gfx950")


def gen_tables() -> str:
    """Writes _gen/approx14_pwl.inc (per run A and -B, as u32) from
    data/approx14_pwl.txt after checking that the pairs reproduce every entry
    of data/approx14.bin."""
    import numpy as np

    os.makedirs(GEN, exist_ok=True)
    out = os.path.join(GEN, "approx14_pwl.inc")
    table = np.fromfile(os.path.join(HERE, "data", "approx14.bin"), dtype=np.uint16).astype(np.int64)
    assert table.size == 131072, "approx14.bin must hold 131072 u16"
    pairs = np.loadtxt(os.path.join(HERE, "data", "approx14_pwl.txt"), dtype=np.int64).reshape(128, 2)
    idx = np.arange(65536)
    rcp = (pairs[idx >> 10, 0] - pairs[idx >> 10, 1] * (idx & 1023)) >> 12
    assert (rcp == table[:65536]).all(), "approx14_pwl.txt does not reproduce the rcp table"
    idx = np.arange(32768)
    for odd in (0, 1):
        run = 64 + odd * 32 + (idx >> 10)
        rsq = (pairs[run, 0] - pairs[run, 1] * (idx & 1023)) >> 16
        assert (rsq == table[65536 + odd * 32768:65536 + (odd + 1) * 32768]).all(), \
            "approx14_pwl.txt does not reproduce the rsqrt table"
    # (A, -B): the device evaluates A + (-B) * j with one 24-bit multiply-add
    text = "\n".join(f"{int(a)}u, {int((-b) & 0xFFFFFFFF)}u," for a, b in pairs) + "\n"
    if not os.path.exists(out) or open(out).read() != text:
        with open(out, "w") as f:
            f.write(text)
    return GEN


def compile_cmd() -> list[str]:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    return [
        HIPCC, "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
        "-ffp-contract=off", "-fno-fast-math",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", GEN,
        *srcs, "-o", LIB,
        "-L", "/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib",
    ]


def build(verbose: bool = False) -> str:
    gen_tables()
    os.makedirs(LIBDIR, exist_ok=True)
    cmd = compile_cmd()
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libglimp.so")
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
