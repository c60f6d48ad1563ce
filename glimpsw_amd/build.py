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
SOURCES = ["glimp_kernels.hip", "resolve.hip", "diff.hip", "capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GLIMP_ARCH", "gfx950")


def gen_tables() -> str:
    os.makedirs(GEN, exist_ok=True)
    out = os.path.join(GEN, "approx14_tables.inc")
    src = os.path.join(HERE, "data", "approx14.bin")
    data = open(src, "rb").read()
    assert len(data) == 262144, "approx14.bin must hold 131072 u16"
    vals = [int.from_bytes(data[i:i + 2], "little") for i in range(0, len(data), 2)]
    lines = []
    for i in range(0, len(vals), 16):
        lines.append(",".join(str(v) for v in vals[i:i + 16]) + ",")
    text = "\n".join(lines) + "\n"
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
