"""Pins the oracle's floating-point operation order to what clang emits for
the reference's expression shapes under the reference's flags
(-O2 -march=native -mprefer-vector-width=512 -ffast-math,
src/SwRast/CMakeLists.txt:6).

oracle/probes/fastmath_probe.cpp holds those expression shapes compiled with
those flags; each test feeds both sides the same random inputs and requires
bit-identical outputs. A dropped Newton step, a different FMA contraction or
a swapped operand changes low bits and fails here.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S

f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u16p = C.POINTER(C.c_uint16)


def _has_avx512():
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return "avx512f" in flags


pytestmark = pytest.mark.skipif(not _has_avx512() or not os.path.exists("/opt/rocm/llvm/bin/clang++"),
                                reason="probe needs clang and an AVX-512 host")


@pytest.fixture(scope="module")
def probe():
    return oracle.probe_lib()


def P(a):
    return np.ascontiguousarray(a).ctypes.data_as(f32p if a.dtype == np.float32 else i32p)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def test_rcp14_rsqrt14_random(probe):
    rng = np.random.default_rng(1)
    x = rng.integers(0, 2 ** 32, 1 << 16, dtype=np.uint64).astype(np.uint32).view(np.float32).copy()
    out = np.zeros_like(x)
    probe.probe_rcp14(P(x), P(out), len(x))
    L = oracle.lib()
    mine = np.array([L.orc_rcp14(float(v)) for v in x[:4096]], dtype=np.float32)
    assert (_bits(mine) == _bits(out[:4096])).all()
    probe.probe_rsqrt14(P(x), P(out), len(x))
    mine = np.array([L.orc_rsqrt14(float(v)) for v in x[:4096]], dtype=np.float32)
    nan = np.isnan(out[:4096])
    assert (_bits(mine)[~nan] == _bits(out[:4096])[~nan]).all() and np.isnan(mine[nan]).all()


def test_vector_division(probe):
    rng = np.random.default_rng(2)
    n = 4096
    a = (rng.standard_normal(n) * 10 ** rng.uniform(-3, 3, n)).astype(np.float32)
    b = (rng.standard_normal(n) * 10 ** rng.uniform(-3, 3, n)).astype(np.float32)
    out = np.zeros(n, np.float32)
    probe.probe_div(P(a), P(b), P(out), n)
    L = oracle.lib()
    mine = np.array([L.orc_div(float(x), float(y)) for x, y in zip(a, b)], dtype=np.float32)
    assert (_bits(mine) == _bits(out)).all()
    probe.probe_recip(P(b), P(out), n)
    mine = np.array([L.orc_div(1.0, float(y)) for y in b], dtype=np.float32)
    assert (_bits(mine) == _bits(out)).all()
    # the approximation is not IEEE division: it must differ somewhere
    assert (_bits(out) != _bits((np.float32(1) / b).astype(np.float32))).any()


def _random_clip(rng, n):
    """Triangles in front of the camera with random clip coordinates."""
    v = np.empty((3, 4, n), np.float32)
    w = rng.uniform(0.5, 50.0, (3, n))
    v[:, 3] = w
    v[:, 0] = rng.uniform(-1.2, 1.2, (3, n)) * w
    v[:, 1] = rng.uniform(-1.2, 1.2, (3, n)) * w
    v[:, 2] = 0.01
    return v


@pytest.mark.parametrize("cull_mode", [0, 1, 2])
def test_triangle_setup(probe, cull_mode):
    rng = np.random.default_rng(3 + cull_mode)
    n = 16 * 64
    v = _random_clip(rng, n)
    W, H = 1920, 1080
    X = np.zeros((3, n), np.int32); Y = np.zeros_like(X)
    Z = np.zeros((3, n), np.float32); Wr = np.zeros_like(Z); det = np.zeros(n, np.float32)
    probe.probe_setup(P(v), n, C.c_float(W // 2 * 16), C.c_float(H // 2 * 16), cull_mode, P(X), P(Y), P(Z), P(Wr),
                      P(det))
    L = oracle.lib()
    for p0 in range(0, n, 16):
        clip = np.ascontiguousarray(v[:, :, p0:p0 + 16])
        x = np.zeros((3, 16), np.int32); y = np.zeros_like(x)
        z = np.zeros((3, 16), np.float32); w = np.zeros_like(z); d = np.zeros(16, np.float32)
        mask = np.zeros(1, np.uint16)
        L.orc_setup16(P(clip), W, H, cull_mode, P(x), P(y), P(z), P(w), P(d), mask.ctypes.data_as(u16p))
        assert (x == X[:, p0:p0 + 16]).all() and (y == Y[:, p0:p0 + 16]).all()
        assert (_bits(z) == _bits(Z[:, p0:p0 + 16])).all()
        assert (_bits(w) == _bits(Wr[:, p0:p0 + 16])).all()
        assert (_bits(d) == _bits(det[p0:p0 + 16])).all()


def test_edge_setup_float_part(probe):
    rng = np.random.default_rng(5)
    n = 16 * 32
    det = rng.integers(1, 1 << 30, n).astype(np.int32)
    Z = rng.uniform(0.0, 1.0, (3, n)).astype(np.float32)
    W = rng.uniform(0.0, 1.0, (3, n)).astype(np.float32)
    z10 = np.zeros(n, np.float32); z20 = np.zeros_like(z10)
    w0 = np.zeros_like(z10); w1 = np.zeros_like(z10); w2 = np.zeros_like(z10)
    probe.probe_edge(P(det), P(Z), P(W), n, P(z10), P(z20), P(w0), P(w1), P(w2))
    L = oracle.lib()
    for p0 in range(0, n, 16):
        a = np.zeros(16, np.float32); b = np.zeros(16, np.float32)
        L.orc_edge16(P(det[p0:p0 + 16].copy()), P(np.ascontiguousarray(Z[:, p0:p0 + 16])), P(a), P(b))
        assert (_bits(a) == _bits(z10[p0:p0 + 16])).all() and (_bits(b) == _bits(z20[p0:p0 + 16])).all()


def test_intersect_triangle(probe):
    rng = np.random.default_rng(6)
    n = 16 * 32
    v = _random_clip(rng, n)
    suv = rng.uniform(-1, 1, (2, n)).astype(np.float32)
    bary = np.zeros((3, n), np.float32); ddx = np.zeros_like(bary); ddy = np.zeros_like(bary)
    probe.probe_intersect(P(v), P(suv), n, C.c_float(1920), C.c_float(1080), P(bary), P(ddx), P(ddy))
    L = oracle.lib()
    for p0 in range(0, n, 16):
        b = np.zeros((3, 16), np.float32); dx = np.zeros_like(b); dy = np.zeros_like(b)
        L.orc_intersect_triangle16(P(np.ascontiguousarray(v[:, :, p0:p0 + 16])), P(np.ascontiguousarray(suv[:, p0:p0 + 16])),
                                   C.c_float(1920), C.c_float(1080), P(b), P(dx), P(dy))
        for mine, ref in ((b, bary), (dx, ddx), (dy, ddy)):
            r = ref[:, p0:p0 + 16]
            ok = (_bits(mine) == _bits(r)) | (np.isnan(mine) & np.isnan(r))
            assert ok.all()


def test_uv_gradients_and_mip(probe):
    rng = np.random.default_rng(7)
    n = 16 * 32
    t = rng.uniform(0, 8, (3, 2, n)).astype(np.float32)
    bary = rng.uniform(0, 1, (3, n)).astype(np.float32)
    ddx = (rng.standard_normal((3, n)) * 1e-3).astype(np.float32)
    ddy = (rng.standard_normal((3, n)) * 1e-3).astype(np.float32)
    uv = np.zeros((2, n), np.float32); mip = np.zeros(n, np.int32)
    probe.probe_uvgrad(P(t), P(bary), P(ddx), P(ddy), n, C.c_float(256), C.c_float(256), P(uv), P(mip))
    L = oracle.lib()
    for p0 in range(0, n, 16):
        tt = np.ascontiguousarray(t[:, :, p0:p0 + 16])
        u = np.zeros((2, 16), np.float32); m = np.zeros(16, np.int32)
        L.orc_uvgrad16(P(tt), P(np.ascontiguousarray(bary[:, p0:p0 + 16])), P(np.ascontiguousarray(ddx[:, p0:p0 + 16])),
                       P(np.ascontiguousarray(ddy[:, p0:p0 + 16])), C.c_float(256), C.c_float(256), P(u), P(m))
        assert (_bits(u) == _bits(uv[:, p0:p0 + 16])).all()
        assert (m == mip[p0:p0 + 16]).all()


def test_exhaustive_approx_tables():
    """The committed VRCP14/VRSQRT14 tables reproduce this host's
    instructions over all 2^32 inputs (the tables describe Intel's
    implementation; another vendor's AVX-512 may round differently)."""
    cpu = open("/proc/cpuinfo").read()
    if "GenuineIntel" not in cpu:
        pytest.skip("tables describe Intel's VRCP14PS/VRSQRT14PS")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = os.path.join(root, "tests", "golden", "gen_approx14.c")
    exe = os.path.join(root, "oracle", "_build", "gen_approx14")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    subprocess.run(["gcc", "-O2", "-mavx512f", src, "-o", exe], check=True)
    r = subprocess.run([exe, "--verify", os.path.join(root, "glimpsw_amd", "data", "approx14.bin")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "rcp14 mismatches 0, rsqrt14 mismatches 0" in r.stdout


def _fragment_inputs(rng):
    # a plausible triangle: edge values over a tile, 1/w terms scaled by 16/area
    e1 = rng.integers(-2000, 60000, 16).astype(np.int32)
    e2 = rng.integers(-2000, 60000, 16).astype(np.int32)
    w = rng.uniform(0.05, 2.0, 3).astype(np.float32)
    rcp_area = np.float32(16.0) / np.float32(rng.integers(1 << 10, 1 << 24))
    coef = np.array([rng.uniform(0, 1), rng.uniform(-1, 1) * 1e-5, rng.uniform(-1, 1) * 1e-5,
                     w[0], w[0] * rcp_area, w[1] * rcp_area, w[2] * rcp_area], np.float32)
    return e1, e2, coef


def test_fragment_barycentrics(probe):
    rng = np.random.default_rng(8)
    L = oracle.lib()
    for _ in range(64):
        e1, e2, coef = _fragment_inputs(rng)
        depth = np.zeros(16, np.float32); bu = np.zeros_like(depth); bv = np.zeros_like(depth)
        probe.probe_fragment(P(e1), P(e2), P(coef), 16, P(depth), P(bu), P(bv))
        d2 = np.zeros(16, np.float32); u2 = np.zeros_like(d2); v2 = np.zeros_like(d2)
        L.orc_fragment16(P(e1), P(e2), P(coef), P(d2), P(u2), P(v2))
        for a, b in ((d2, depth), (u2, bu), (v2, bv)):
            assert (_bits(a) == _bits(b)).all()


@pytest.mark.parametrize("clipped", [False, True])
def test_alpha_test_uvs_and_mip(probe, clipped):
    rng = np.random.default_rng(9 + clipped)
    L = oracle.lib()
    for _ in range(64):
        e1, e2, coef = _fragment_inputs(rng)
        w4 = np.ascontiguousarray(coef[3:7])
        remap = rng.uniform(-0.5, 1.5, 6).astype(np.float32) if clipped else None
        uv = rng.uniform(0, 8, 6).astype(np.float32)
        U = np.zeros(16, np.float32); V = np.zeros_like(U); mip = np.zeros(16, np.int32)
        rp = P(remap) if clipped else None
        probe.probe_alpha_uv(P(e1), P(e2), P(w4), rp, P(uv), C.c_float(256 * 256.0), C.c_float(256 * 256.0),
                             P(U), P(V), P(mip))
        U2 = np.zeros(16, np.float32); V2 = np.zeros_like(U2); mip2 = np.zeros(16, np.int32)
        L.orc_alpha_uv16(P(e1), P(e2), P(coef), rp, P(uv), C.c_float(256 * 256.0), C.c_float(256 * 256.0),
                         P(U2), P(V2), P(mip2))
        assert (_bits(U2) == _bits(U)).all() and (_bits(V2) == _bits(V)).all()
        assert (mip2 == mip).all()


def test_normal_mapping(probe):
    rng = np.random.default_rng(11)
    L = oracle.lib()
    for trial in range(64):
        n = rng.normal(size=(3, 16))
        n /= np.linalg.norm(n, axis=0)
        t = rng.normal(size=(3, 16))
        t -= n * (t * n).sum(axis=0)
        t /= np.linalg.norm(t, axis=0)
        n = np.ascontiguousarray(n, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        hand = np.where(rng.random(16) < 0.5, 0, 1 << 31).astype(np.uint32)
        # normal-map bytes inside the unit disc, plus some outside it (NaN z)
        if trial % 4 == 3:
            xy = rng.integers(0, 256, (2, 16))
        else:
            ang = rng.uniform(0, 2 * np.pi, 16)
            rad = rng.uniform(0, 0.99, 16)
            xy = np.rint((np.stack([np.cos(ang), np.sin(ang)]) * rad + 1) * 127.5).astype(np.int64)
        nmr = (xy[0] | (xy[1] << 8) | (rng.integers(0, 256, 16) << 16) | (rng.integers(0, 256, 16) << 24))
        nmr = nmr.astype(np.uint32)
        out = np.zeros((3, 16), np.float32)
        probe.probe_normal_map(P(n), P(t), P(hand.view(np.int32)), P(nmr.view(np.int32)), P(out))
        out2 = np.zeros((3, 16), np.float32)
        L.orc_normal_map16(P(n), P(t), P(hand), P(nmr), P(out2))
        same = (_bits(out) == _bits(out2)) | (np.isnan(out) & np.isnan(out2))
        assert same.all(), (trial, np.argwhere(~same)[:3])


def test_light_sprite_placement_and_tile(probe):
    """The sprite placement (glm mat4 * vec4, divisions by w, radius, tile
    range) and one tile's distances and colours, against the probe."""
    rng = np.random.default_rng(12)
    L = oracle.lib()
    L.orc_sprite_place.restype = C.c_int32
    placed = 0
    for trial in range(200):
        view = S.make_view(int(rng.choice([320, 640, 1920])), int(rng.choice([240, 480, 1080])),
                           view=S.view_matrix((0.0, 0.0, 0.0), rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2)))
        m = np.ascontiguousarray(view.object_to_clip, np.float32)
        z = -rng.uniform(0.005, 12.0)      # mostly in front of the camera, some past the edges
        pos = np.array([rng.uniform(-0.8, 0.8) * -z, rng.uniform(-0.7, 0.7) * -z, z], np.float32)
        W, H = view.width, view.height
        o1 = np.zeros(4, np.float32); t1 = np.zeros(4, np.int32)
        o2 = np.zeros(4, np.float32); t2 = np.zeros(4, np.int32)
        r1 = probe.probe_sprite_host(P(m), P(pos), W, H, P(o1), P(t1))
        r2 = L.orc_sprite_place(P(m), P(pos), C.c_uint32(W), C.c_uint32(H), P(o2), P(t2))
        assert r1 == r2
        if not r1:
            continue
        placed += 1
        assert (_bits(o1) == _bits(o2)).all() and (t1 == t2).all(), (trial, o1, o2, t1, t2)
        color = rng.uniform(0, 1.5, 3).astype(np.float32)
        for _ in range(4):
            x = int(rng.integers(t1[0], max(t1[2], t1[0] + 1))) & ~3
            y = int(rng.integers(t1[1], max(t1[3], t1[1] + 1))) & ~3
            d1 = np.zeros(16, np.float32); f1 = np.zeros(16, np.int32)
            d2 = np.zeros(16, np.float32); f2 = np.zeros(16, np.uint32)
            probe.probe_sprite_tile(x, y, C.c_float(o1[1]), C.c_float(o1[2]), C.c_float(o1[3]), P(color), P(d1), P(f1))
            L.orc_sprite_tile(x, y, C.c_float(o1[1]), C.c_float(o1[2]), C.c_float(o1[3]), P(color), P(d2), P(f2))
            assert (_bits(d1) == _bits(d2)).all() and (f1.view(np.uint32) == f2).all()
    assert placed > 50


def test_alpha_blend_u8_against_float():
    """AlphaBlendU8 is source-over in 8.8 fixed point: within one unit of
    bg + (fg - bg) * alpha/255 per channel, exact at alpha 0 and 255... up to
    the 1/256 step of its weights."""
    rng = np.random.default_rng(13)
    L = oracle.lib()
    L.orc_alpha_blend_u8.restype = C.c_uint32
    for _ in range(2000):
        bg = int(rng.integers(0, 1 << 32)); fg = int(rng.integers(0, 1 << 32))
        out = L.orc_alpha_blend_u8(C.c_uint32(bg), C.c_uint32(fg))
        a = (fg >> 24) / 255.0
        for k in range(4):
            b, f, o = (bg >> 8 * k) & 255, (fg >> 8 * k) & 255, (out >> 8 * k) & 255
            assert abs(o - (b + (f - b) * a)) <= 1.5, (hex(bg), hex(fg), k, o)


def test_hiz_sphere_projection(probe):
    """ProjectSphere and the occlusion test's box, mip level, texel range and
    sphere depth (Shading.cpp:264-279, 811-832) against the probe."""
    rng = np.random.default_rng(14)
    L = oracle.lib()
    projected = 0
    for trial in range(100):
        W, H = int(rng.choice([320, 640, 1920])), int(rng.choice([240, 480, 1080]))
        Pm = S.proj_matrix(60.0, W / H)
        V = S.view_matrix(tuple(rng.uniform(-1, 1, 3)), rng.uniform(-1, 1), rng.uniform(-0.5, 0.5))
        eye_fwd = np.linalg.inv(V.astype(np.float64).reshape(4, 4).T)
        ctr = eye_fwd[:3, 3][:, None] + eye_fwd[:3, :3] @ np.stack(
            [rng.uniform(-4, 4, 16), rng.uniform(-3, 3, 16), -rng.uniform(-1, 12, 16)])
        sphere = np.ascontiguousarray(np.concatenate([ctr, rng.uniform(0.05, 2.0, (1, 16))]), np.float32)
        scale = np.float32(rng.uniform(0.5, 2.0))
        proj = np.array([Pm[14], Pm[0], Pm[5]], np.float32)
        fs = np.array([W, H], np.float32)
        levels = oracle.hiz_levels(W, H)
        keep = []
        occ = oracle.occlusion(np.zeros(1, np.float32), W, H, V, scale, proj, keep)
        outs1 = [np.zeros(16, np.int32), np.zeros(64, np.float32), np.zeros(16, np.int32), np.zeros(64, np.int32),
                 np.zeros(16, np.float32)]
        outs2 = [np.zeros_like(a) for a in outs1]
        Vf = np.ascontiguousarray(V, np.float32)
        probe.probe_hiz_sphere(P(sphere), P(Vf), C.c_float(scale), P(proj), P(fs), levels, *[P(a) for a in outs1])
        L.orc_hiz_sphere16(P(sphere), C.byref(occ), *[P(a) for a in outs2])
        mask = outs1[0] != 0
        projected += int(mask.sum())
        assert (outs1[0] == outs2[0]).all()
        for a, b in zip(outs1[1:], outs2[1:]):
            if a.dtype == np.float32:
                assert (_bits(a) == _bits(b)).all(), trial
            else:
                assert (a == b).all(), trial
    assert projected > 500


@pytest.mark.parametrize("quad", [0, 1])
def test_overdraw_colormap(probe, quad):
    """ColormapTurbo of the overdraw counts (Shading.cpp:249-260, 759-766):
    every pixel count and helper count up to 300 each, plus random words."""
    pix, frag = np.meshgrid(np.arange(301, dtype=np.uint32), np.arange(301, dtype=np.uint32))
    rng = np.random.default_rng(7)
    words = np.concatenate([(pix << 16 | frag).ravel(), rng.integers(0, 2 ** 32, 4096, dtype=np.uint64).astype(np.uint32)])
    words = np.ascontiguousarray(words[: len(words) // 16 * 16])
    n = len(words)
    want = np.zeros(3 * n, np.float32)
    probe.probe_overdraw_colour(words.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_int(quad), P(want), C.c_int(n))
    want = want.reshape(3, n)
    got = np.zeros((3, n), np.float32)
    L = oracle.lib()
    for i in range(0, n, 16):
        o = np.zeros(48, np.float32)
        L.orc_overdraw_colour16(words[i:i + 16].ctypes.data_as(C.POINTER(C.c_uint32)), C.c_int32(quad), P(o))
        got[:, i:i + 16] = o.reshape(3, 16)
    assert (_bits(got) == _bits(want)).all()


def test_gbuffer_pack(probe):
    """MapOctahedron and the packing of FS_EncodeGBuffer's layer 2
    (Shading.cpp:397-401, Texture.h:282-288) on random unit-ish normals,
    axis-aligned and zero-component normals and random NMR texels."""
    rng = np.random.default_rng(11)
    n = rng.normal(size=(4096, 3)).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True).astype(np.float32)
    n[:64] = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [0.6, 0, -0.8],
                       [0, -0.6, -0.8]] * 8, np.float32)
    n[64:128] *= 1.0003                      # slightly long normals push oct past 1
    nmr = rng.integers(0, 2 ** 32, 4096, dtype=np.uint64).astype(np.uint32)
    L = oracle.lib()
    u32p = C.POINTER(C.c_uint32)
    for i in range(0, 4096, 16):
        blk = np.ascontiguousarray(n[i:i + 16].T)
        want = np.zeros(16, np.int32)
        probe.probe_gbuffer_pack(P(blk), P(nmr[i:i + 16].view(np.int32).copy()), P(want))
        got = np.zeros(16, np.uint32)
        L.orc_gbuffer_pack16(P(blk), nmr[i:i + 16].ctypes.data_as(u32p), got.ctypes.data_as(u32p))
        assert (got == want.view(np.uint32)).all(), i


def test_sky_coordinates(probe):
    """MapOctahedron of unnormalised view directions and SampleLevel's
    scaled, rounded coordinates (Texture.h:282-288, 412-416)."""
    rng = np.random.default_rng(12)
    d = rng.normal(size=(4096, 3)).astype(np.float32) * rng.uniform(0.01, 50, (4096, 1)).astype(np.float32)
    d[:16] = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [1, 1, -1], [-1, -1, 1]] * 2,
                      np.float32)
    L = oracle.lib()
    for slu, slv in ((256.0 * 128, 256.0 * 64), (256.0 * 1024, 256.0 * 1024)):
        for i in range(0, 4096, 16):
            blk = np.ascontiguousarray(d[i:i + 16].T)
            wx, wy = np.zeros(16, np.int32), np.zeros(16, np.int32)
            probe.probe_sky_coords(P(blk), C.c_float(slu), C.c_float(slv), P(wx), P(wy))
            gx, gy = np.zeros(16, np.int32), np.zeros(16, np.int32)
            L.orc_sky_coords16(P(blk), C.c_float(slu), C.c_float(slv), P(gx), P(gy))
            assert (gx == wx).all() and (gy == wy).all(), i


def test_env_bilinear(probe):
    """SampleLinear's float filter on R11G11B10f texels (Texture.h:557-571)."""
    rng = np.random.default_rng(13)
    L = oracle.lib()
    u32p = C.POINTER(C.c_uint32)
    for i in range(256):
        d = rng.integers(0, 2 ** 32, 64, dtype=np.uint64).astype(np.uint32)
        ixf = rng.integers(0, 1 << 20, 16).astype(np.int32)
        iyf = rng.integers(0, 1 << 20, 16).astype(np.int32)
        inx = rng.integers(0, 2, 16).astype(np.int32) * -1
        want = np.zeros(48, np.float32)
        probe.probe_env_bilinear(d.ctypes.data_as(u32p), P(ixf), P(iyf), P(inx), P(want))
        got = np.zeros(48, np.float32)
        L.orc_env_bilinear16(d.ctypes.data_as(u32p), P(ixf), P(iyf), P(inx), P(got))
        assert (_bits(got) == _bits(want)).all(), i


def _mip_inputs(rng, n):
    """n quads of texels: every (a, b) byte pair against random (c, d) pairs
    in the four channels, then random texels."""
    a, b = np.meshgrid(np.arange(256, dtype=np.uint32), np.arange(256, dtype=np.uint32), indexing="ij")
    a, b = np.tile(a.reshape(-1), 16), np.tile(b.reshape(-1), 16)
    c = rng.integers(0, 256, size=a.size, dtype=np.uint32)
    d = rng.integers(0, 256, size=a.size, dtype=np.uint32)
    quad = np.stack([a, b, c, d])                                       # [4, 2^20] one channel each
    shuffled = [np.roll(quad, s, axis=0) for s in range(4)]             # the same values in every position
    texels = sum(q << np.uint32(8 * k) for k, q in enumerate(shuffled))
    rand = rng.integers(0, 1 << 32, size=(4, n), dtype=np.uint64).astype(np.uint32)
    return np.concatenate([texels.astype(np.uint32), rand], axis=1)


def _as_level(quads):
    """[4, n] (c00, c10, c01, c11) -> a (2, 2n) level whose 2x2 blocks are the quads."""
    n = quads.shape[1]
    lv = np.zeros((2, 2 * n), np.uint32)
    lv[0, 0::2], lv[0, 1::2], lv[1, 0::2], lv[1, 1::2] = quads
    return lv


def test_generate_mip_rgba8(probe):
    """GenerateMip's box filter for RGBA8u (Texture.h:577-596) as the
    reference's flags compile it against textures.next_mip, which builds the
    mip chains both the GPU and the oracle sample: bit for bit over every
    byte pair in the first two positions (random partners) in every channel,
    and 2^20 random quads."""
    from glimpsw_amd import textures

    quads = _mip_inputs(np.random.default_rng(3), 1 << 20)
    n = quads.shape[1]
    out = np.zeros(n, np.uint32)
    q = [np.ascontiguousarray(quads[k]) for k in range(4)]
    u32 = C.POINTER(C.c_uint32)
    probe.probe_mip_rgba8(*[x.ctypes.data_as(u32) for x in q], out.ctypes.data_as(u32), n)
    got = textures.next_mip(_as_level(quads))[0]
    bad = got != out
    assert not bad.any(), f"{bad.sum()} quads differ, first {quads[:, np.nonzero(bad)[0][0]]}"


def test_generate_mip_r11g11b10(probe):
    """The same filter over R11G11B10f texels (the skybox's mips) against
    textures.next_mip_hdr, bit for bit on 2^21 random quads."""
    from glimpsw_amd import textures

    rng = np.random.default_rng(4)
    quads = rng.integers(0, 1 << 32, size=(4, 1 << 21), dtype=np.uint64).astype(np.uint32)
    n = quads.shape[1]
    out = np.zeros(n, np.uint32)
    q = [np.ascontiguousarray(quads[k]) for k in range(4)]
    u32 = C.POINTER(C.c_uint32)
    probe.probe_mip_r11g11b10(*[x.ctypes.data_as(u32) for x in q], out.ctypes.data_as(u32), n)
    got = textures.next_mip_hdr(_as_level(quads))[0]
    bad = got != out
    assert not bad.any(), f"{bad.sum()} quads differ"
