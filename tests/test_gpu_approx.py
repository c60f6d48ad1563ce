"""The device's reproduction of VRCP14PS, VRSQRT14PS and the reference's
fast-math division (glimpsw_amd/csrc/device_math.h) against the CPU oracle,
input by input: every 16-bit rcp / 15-bit rsqrt mantissa index at several
exponents, the exponent edges of the fast path (1, 252, 253, 254 and the
denormal and overflow ranges), zeros, infinities, NaNs and random bit
patterns; the division on random and edge operand pairs. The oracle's
tables are the host CPU's own instructions over all 2^32 inputs
(tests/golden/gen_approx14.c), so this pins the device arithmetic
directly, not only through whole frames."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "libapprox_eval.so")


def _inputs():
    mant = np.arange(1 << 16, dtype=np.uint32) << 7
    parts = []
    for e in (1, 2, 100, 126, 127, 128, 200, 251, 252, 253, 254, 0):
        parts.append((np.uint32(e) << 23) | mant)
    rng = np.random.default_rng(21)
    parts.append(rng.integers(0, 2 ** 32, 1 << 20, dtype=np.uint64).astype(np.uint32))
    special = np.array([0, 0x80000000, 0x7F800000, 0xFF800000, 0x7FC00000, 0xFFC00001, 0x7F800001, 1, 0x807FFFFF,
                        0x00800000, 0x7F7FFFFF, 0x7E800000, 0x7E7FFFFF, 0x3F800000, 0xBF800000], np.uint32)
    parts.append(special)
    bits = np.concatenate(parts)
    signs = np.concatenate([bits, bits[: 1 << 18] | np.uint32(0x80000000)])
    return signs.view(np.float32).copy()


def _device(op, a, b=None):
    import glimpsw_amd

    glimpsw_amd.load()          # one HIP runtime per process: bind to the one torch loads
    lib = C.CDLL(LIB)
    out = np.zeros_like(a)
    f32p = C.POINTER(C.c_float)
    b = a if b is None else b
    assert lib.approx_eval(op, a.ctypes.data_as(f32p), b.ctypes.data_as(f32p), out.ctypes.data_as(f32p),
                           C.c_uint32(a.size)) == 0
    return out


def _oracle(op, a, b=None):
    out = np.zeros_like(a)
    f32p = C.POINTER(C.c_float)
    b = a if b is None else b
    oracle.lib().orc_approx_n(C.c_int32(op), a.ctypes.data_as(f32p), b.ctypes.data_as(f32p),
                              out.ctypes.data_as(f32p), C.c_uint64(a.size))
    return out


def _same(got, want):
    g, w = got.view(np.uint32), want.view(np.uint32)
    nan = np.isnan(got) & np.isnan(want)      # NaN payloads: both quiet NaNs suffice
    return (g == w) | nan


@pytest.mark.parametrize("op", [0, 1])
def test_rcp14_rsqrt14(op):
    if not os.path.exists(LIB):
        pytest.fail("tests/cpp/_build/libapprox_eval.so is not built (run __graft_entry__.build())")
    a = _inputs()
    print(f"[approx] op {op}: {a.size} inputs", flush=True)
    got, want = _device(op, a), _oracle(op, a)
    ok = _same(got, want)
    bad = np.nonzero(~ok)[0]
    assert ok.all(), [(hex(int(a.view(np.uint32)[i])), float(got[i]), float(want[i])) for i in bad[:5]]


def test_fast_math_division():
    if not os.path.exists(LIB):
        pytest.fail("tests/cpp/_build/libapprox_eval.so is not built (run __graft_entry__.build())")
    rng = np.random.default_rng(22)
    b = _inputs()
    a = rng.normal(size=b.size).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, b.size).astype(np.float32)
    print(f"[approx] fdiv: {a.size} pairs", flush=True)
    got, want = _device(2, a, b), _oracle(2, a, b)
    ok = _same(got, want)
    bad = np.nonzero(~ok)[0]
    assert ok.all(), [(float(a[i]), float(b[i]), float(got[i]), float(want[i])) for i in bad[:5]]
