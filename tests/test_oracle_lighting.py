"""Pins the oracle's lit colour (EvalLighting, Shading.cpp:602-645; BRDF
Shading.cpp:17-33; GetLightAttenuation Shading.cpp:581-600; Tonemap_Unreal
Shading.cpp:221-226; RGBA8u::Pack Texture.h:55-67) against a float64
evaluation of the published formulas the reference cites:

 - Filament's D_GGX, V_SmithGGXCorrelatedFast (0.5 / mix(2 NoL NoV,
   NoL + NoV, a)), F_Schlick with f0 = mix(0.16 r^2, base, metallic) and
   Lambert diffuse base (1 - metallic) / pi;
 - the KHR_lights_punctual range attenuation max(1 - (d^2 / range^2)^2, 0)^2
   / max(d^2, 1e-4), times the squared spot factor clamp(cos * scale +
   offset, 0, 1)^2;
 - ambient 0.05 base, Unreal's x / (x + 0.155) * 1.019 tone curve, and the
   round-to-nearest 8-bit pack.

The base colour is RGBA8u::UnpackSrgb, an integer approximation of the
squared byte value (Texture.h:37-54); it is pinned separately against
(c/255)^2, and the float64 model evaluates that integer formula exactly.

Bound. The oracle computes in float32 with the reference's VRCP14/VRSQRT14
approximations (relative error < 2^-14) in normalize, D_GGX and the
attenuation. Tone mapping does not increase relative errors (d ln y / d ln x
= 0.155 / (x + 0.155) <= 1), so a relative error e of the linear colour moves
the 8-bit value by at most 255 * 1.019 * e. The largest amplification is
D_GGX's sensitivity to NoH, about 4 / a^2 per unit of NoH; with roughness
bytes >= 128 (a >= 0.25) and NoH perturbed by about 2^-13, e stays below
about 1e-3, i.e. a quarter of an 8-bit step, so the packed values may differ
by one step (values near a rounding boundary) and never by two.
"""
import ctypes as C

import numpy as np
import pytest

import oracle

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)


def _srgb_int(c):
    """UnpackSrgb's colour channels in exact arithmetic: the byte moved to
    the high half, 255 added, squared with a 16-bit high multiply, / 65535."""
    c = np.asarray(c, np.int64)
    return ((256 * c + 255) ** 2 // 65536) / 65535.0


def test_unpack_srgb_is_the_squared_curve():
    """Every byte in every channel: UnpackSrgb's colour channels are the
    squared value (c/255)^2 within 0.0042 (the approximation's own error,
    largest at mid-range), the alpha channel is linear within 1/256; a linear
    colour channel would be off by up to 0.25."""
    L = oracle.lib()
    vals = np.arange(256, dtype=np.uint32)
    packed = np.concatenate([vals, vals << 8, vals << 16, vals << 24]).astype(np.uint32)
    out = np.zeros((packed.size, 4), np.float32)
    L.orc_unpack_srgb_n(packed.ctypes.data_as(u32p), out.ctypes.data_as(f32p), C.c_uint64(packed.size))
    want = (vals / 255.0) ** 2
    for ch in range(3):
        got = out[ch * 256:(ch + 1) * 256, ch].astype(np.float64)
        assert np.abs(got - want).max() <= 0.0042, ch
        # the integer formula, up to the float32 rounding of k * (1/65535)
        assert np.abs(got - _srgb_int(vals)).max() <= 2.0 ** -23
        assert np.abs(vals / 255.0 - want).max() > 0.2     # what a linear channel would miss by
    alpha = out[768:, 3].astype(np.float64)
    assert np.abs(alpha - vals / 255.0).max() <= 1.0 / 256


class _Case:
    """Random tiles of 16 surfaces and 1-3 lights each (seeded)."""

    def __init__(self, seed, tiles=2048, min_rough=128):
        rng = np.random.default_rng(seed)
        n = tiles * 16
        N = rng.normal(size=(n, 3))
        self.N = (N / np.linalg.norm(N, axis=1, keepdims=True)).astype(np.float32)
        self.P = rng.uniform(-4, 4, size=(n, 3)).astype(np.float32)
        self.albedo = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32) | np.uint32(0xFF000000)
        metal = rng.integers(0, 256, size=n).astype(np.uint32)
        rough = rng.integers(min_rough, 256, size=n).astype(np.uint32)
        self.nmr = (metal << 16) | (rough << 24) | rng.integers(0, 1 << 16, size=n).astype(np.uint32)
        self.view_pos = rng.uniform(-6, 6, size=(tiles, 3)).astype(np.float32)
        self.exposure = rng.uniform(0.5, 2.0, size=tiles).astype(np.float32)
        self.lights = []
        for _ in range(tiles):
            ls = []
            for _ in range(int(rng.integers(1, 4))):
                t = int(rng.integers(0, 3))
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                pos = rng.uniform(-6, 6, size=3)
                col = rng.uniform(0.2, 1.0, size=3)
                inten = float(rng.uniform(200, 3000)) if t == 0 else float(rng.uniform(500, 20000))
                radius = float(rng.uniform(4, 40))
                inner, outer = sorted(rng.uniform(0.1, 1.2, size=2))
                scale = 1.0 / max(np.cos(inner) - np.cos(outer), 1e-4)
                ls.append((t, pos.astype(np.float32), d.astype(np.float32), col.astype(np.float32),
                           np.float32(inten), np.float32(1.0 / radius ** 2), np.float32(scale),
                           np.float32(-np.cos(outer) * scale)))
            self.lights.append(ls)

    def oracle_tile(self, t, model=None):
        L = oracle.lib()
        arr = (oracle.Light * len(self.lights[t]))()
        for i, (ty, pos, d, col, inten, irs, ss, so) in enumerate(self.lights[t]):
            arr[i].type = ty
            for k in range(3):
                arr[i].position[k] = float(pos[k])
                arr[i].direction[k] = float(d[k])
                arr[i].color[k] = float(col[k])
            arr[i].intensity = float(inten)
            arr[i].inv_radius_sq = float(irs)
            arr[i].spot_scale = float(ss)
            arr[i].spot_offset = float(so)
        rp = oracle.ResolveParams()
        for k in range(3):
            rp.view_pos[k] = float(self.view_pos[t][k])
        rp.exposure = float(self.exposure[t])
        rp.lights = C.cast(arr, C.POINTER(oracle.Light))
        rp.num_lights = len(self.lights[t])
        sl = slice(t * 16, t * 16 + 16)
        albedo = np.ascontiguousarray(self.albedo[sl])
        nmr = np.ascontiguousarray(self.nmr[sl])
        normal = np.ascontiguousarray(self.N[sl].T)
        wpos = np.ascontiguousarray(self.P[sl].T)
        active = np.ones(16, np.uint8)
        out = np.zeros(16, np.uint32)
        L.orc_eval_lighting16(albedo.ctypes.data_as(u32p), nmr.ctypes.data_as(u32p), normal.ctypes.data_as(f32p),
                              wpos.ctypes.data_as(f32p), active.ctypes.data_as(u8p), C.byref(rp),
                              out.ctypes.data_as(u32p))
        return out


def _norm(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def lit_f64(case, t, swap_lerp=False, no_ambient=False, loh_as_noh=False, linear_base=False):
    """Float64 lit colour of tile t, packed to RGBA8, from the published
    formulas (see the module docstring). The keyword switches give the mutant
    models the check must tell apart."""
    sl = slice(t * 16, t * 16 + 16)
    N = case.N[sl].astype(np.float64)
    P = case.P[sl].astype(np.float64)
    alb = case.albedo[sl].astype(np.int64)
    nmr = case.nmr[sl].astype(np.int64)
    chans = np.stack([alb & 255, (alb >> 8) & 255, (alb >> 16) & 255], axis=1)
    base = chans / 255.0 if linear_base else _srgb_int(chans)
    metal = ((nmr >> 16) & 255) / 255.0
    rough = ((nmr >> 24) & 255) / 255.0
    a = np.maximum(rough * rough, 1e-4)
    f0 = 0.16 * 0.25 * (1 - metal[:, None]) + base * metal[:, None]
    diffuse = base * (1 - metal[:, None])
    V = _norm(case.view_pos[t].astype(np.float64) - P)
    NoV = np.abs((N * V).sum(1)) + 1e-5
    color = np.zeros((16, 3))
    exposure = float(case.exposure[t])
    for (ty, pos, d, col, inten, irs, ss, so) in case.lights[t]:
        pos, d, col = pos.astype(np.float64), d.astype(np.float64), col.astype(np.float64)
        Ld = np.tile(-d, (16, 1)) if ty == 0 else _norm(pos - P)
        NoL = (N * Ld).sum(1)
        if (NoL < 1e-4).all():
            continue
        if ty == 0:
            att = np.ones(16)
        else:
            dv = pos - P
            d2 = (dv * dv).sum(1)
            smooth = np.maximum(1 - (d2 * float(irs)) ** 2, 0.0)
            att = smooth ** 2 / np.maximum(d2, 1e-4)
            if ty == 2:
                cd = (-d * _norm(dv)).sum(1)
                att = att * np.clip(cd * float(ss) + float(so), 0, 1) ** 2
        att = att * float(inten) * exposure * 0.001
        if (NoL * att < 1e-4).all():
            continue
        H = _norm(V + Ld)
        NoH = np.clip((N * H).sum(1), 0, 1)
        LoH = np.clip((Ld * H).sum(1), 0, 1)
        D = (a / (1 - NoH ** 2 + (NoH * a) ** 2)) ** 2 / np.pi
        va, vb = 2 * NoL * NoV, NoL + NoV
        mix = vb + a * (va - vb) if swap_lerp else va + a * (vb - va)
        Vis = 0.5 / mix
        f = (1 - (NoH if loh_as_noh else LoH)) ** 5
        F = f[:, None] + f0 * (1 - f[:, None])
        Fr = (D * Vis)[:, None] * F
        color += (diffuse / np.pi + Fr) * col[None, :] * np.maximum(NoL * att, 0.0)[:, None]
    if not no_ambient:
        color += base * 0.05
    x = color * exposure
    tm = x / (x + 0.155) * 1.019
    b = np.clip(np.rint(tm * 255.0), 0, 255).astype(np.uint32)
    return b[:, 0] | (b[:, 1] << 8) | (b[:, 2] << 16) | np.uint32(255 << 24)


def _diff(a, b):
    a8 = a.view(np.uint8).astype(np.int32)
    b8 = b.view(np.uint8).astype(np.int32)
    return np.abs(a8 - b8).max()


@pytest.fixture(scope="module")
def case():
    return _Case(seed=7)


def test_lighting_against_float64(case):
    """2,048 tiles of random surfaces under random directional, point and
    spot lights: every channel of every lane within one 8-bit step."""
    worst = 0
    lit = 0
    for t in range(len(case.lights)):
        got = case.oracle_tile(t)
        worst = max(worst, _diff(got, lit_f64(case, t)))
        lit += int(((got & 0xFFFFFF) != 0).sum())
    print(f"lighting vs float64: max difference {worst} / 255 over {lit} lit lanes")
    assert lit > 20000
    assert worst <= 1


@pytest.mark.parametrize("mutant", ["swap_lerp", "no_ambient", "loh_as_noh", "linear_base"])
def test_lighting_check_discriminates(case, mutant):
    """The float64 models with V's lerp operands swapped, the ambient term
    dropped, NoH in place of LoH in the Fresnel term, or a linear base colour
    all miss the oracle by far more than the bar: the check above would
    catch any of these in the oracle."""
    worst = 0
    for t in range(256):
        worst = max(worst, _diff(case.oracle_tile(t), lit_f64(case, t, **{mutant: True})))
    assert worst >= 4, (mutant, worst)


def test_tile_skips_are_tile_wide():
    """A light that every lane of the tile faces away from (NoL < 1e-4) adds
    nothing, while one lane facing it lights only that lane: the per-light
    skip is taken over the tile, not per lane (Shading.cpp:622, 625)."""
    case = _Case(seed=11, tiles=1)
    case.N[:] = np.array([0.0, 0.0, 1.0], np.float32)
    case.P[:] = 0.0
    case.P[:, 0] = np.arange(16, dtype=np.float32) * 0.01
    case.view_pos[0] = np.array([3.0, 1.0, 5.0], np.float32)
    down = (0, np.zeros(3, np.float32), np.array([0.0, 0.0, 1.0], np.float32), np.ones(3, np.float32),
            np.float32(3000), np.float32(0), np.float32(0), np.float32(0))
    case.lights[0] = [down]           # shines along +z onto normals facing +z: NoL = -1 everywhere
    dark = case.oracle_tile(0)
    assert _diff(dark, lit_f64(case, 0)) <= 1
    amb = lit_f64(case, 0)
    assert (dark == amb).all()        # ambient only
    case.N[3] = np.array([0.0, 0.0, -1.0], np.float32)
    one = case.oracle_tile(0)
    assert _diff(one, lit_f64(case, 0)) <= 1
    changed = one != dark
    assert changed[3] and changed.sum() == 1
