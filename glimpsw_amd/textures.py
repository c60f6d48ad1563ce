"""RGBA8 textures in the reference's storage layout.

``CreateTexture2D`` (Texture.h:599-636): power-of-two sizes, ``RowShift =
ctz(max(width, 8))``, mip level m kept while both sides are >= 4, each level
padded to 64 texels (256 bytes), layers ``LayerStride`` apart. Texels are
stored TiledY8 (Texture.h:482-501): offset ``(y & 7) | (x << 3) | ((y & ~7)
<< stride)`` with ``stride = RowShift - m``.

Mips are built with the reference's box filter (Texture.h:577-596): unpack
to float (byte * (1/255)), average the 2x2 block as ``(c00 + c10 + c01 +
c11) * 0.25``, pack with round-to-nearest-even and saturation
(Texture.h:55-67). Mips are inputs to the path: they are built once here and
uploaded, never recomputed on the device.
"""
from __future__ import annotations

import dataclasses

import numpy as np

LERP_FRAC_BITS = 8


@dataclasses.dataclass
class Texture:
    width: int
    height: int
    mip_levels: int
    num_layers: int
    row_shift: int
    layer_stride: int
    mip_offsets: np.ndarray   # (16,) u32
    data: np.ndarray          # u32, layer_stride * num_layers + 64 slack texels
    format: int = 0           # 0 RGBA8u, 1 R11G11B10f (HDR, the skybox)

    @property
    def scale_u(self) -> float:
        return float(self.width)

    @property
    def scale_v(self) -> float:
        return float(self.height)

    @property
    def mask_lerp_u(self) -> int:
        return (self.width << LERP_FRAC_BITS) - 1

    @property
    def mask_lerp_v(self) -> int:
        return (self.height << LERP_FRAC_BITS) - 1

    def level(self, m: int, layer: int = 0) -> np.ndarray:
        """Row-major (h, w) view of one mip level (for tests)."""
        w, h = self.width >> m, self.height >> m
        ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
        off = tiled_offset(xs, ys, self.row_shift - m)
        base = layer * self.layer_stride + int(self.mip_offsets[m])
        return self.data[base + off]


def tiled_offset(x, y, stride):
    x = np.asarray(x, dtype=np.int64)
    y = np.asarray(y, dtype=np.int64)
    return (y & 7) | (x << 3) | ((y & ~7) << stride)


def layout(width: int, height: int, max_levels: int = 16):
    assert width & (width - 1) == 0 and height & (height - 1) == 0
    row_shift = (max(width, 8) & -max(width, 8)).bit_length() - 1
    offsets = np.zeros(16, dtype=np.uint32)
    stride = 0
    mip = 0
    while mip < max_levels:
        if (width >> mip) < 4 or (height >> mip) < 4:
            break
        offsets[mip] = stride
        stride += ((width >> mip) * (height >> mip) + 63) & ~63
        mip += 1
    return row_shift, offsets, stride, mip


def unpack_f32(texels: np.ndarray) -> np.ndarray:
    t = texels.astype(np.uint32)
    scale = np.float32(1.0) / np.float32(255.0)
    ch = [((t >> np.uint32(8 * k)) & np.uint32(255)).astype(np.float32) * scale for k in range(4)]
    return np.stack(ch, axis=-1)


def pack_f32(c: np.ndarray) -> np.ndarray:
    v = np.rint(c.astype(np.float32) * np.float32(255.0))   # RNE
    v = np.clip(v, 0, 255).astype(np.uint32)
    return v[..., 0] | (v[..., 1] << 8) | (v[..., 2] << 16) | (v[..., 3] << 24)


def next_mip(level: np.ndarray) -> np.ndarray:
    """GenerateMip's box filter for RGBA8u (Texture.h:577-596) in the operation
    order the reference's optimised build gives `Pack((Unpack(c00) + ... +
    Unpack(c11)) * 0.25f)`: the 1/255 of Unpack, the 0.25 and the 255 of Pack
    fold into one multiply by 0.25 and the four bytes of a channel are summed
    as integers, so each channel is round-half-even(sum / 4)."""
    t = level.astype(np.uint32)
    q = (t[0::2, 0::2], t[0::2, 1::2], t[1::2, 0::2], t[1::2, 1::2])
    out = np.zeros(q[0].shape, np.uint32)
    for k in range(4):
        s = sum(((x >> np.uint32(8 * k)) & np.uint32(255)) for x in q).astype(np.float32)
        out |= np.clip(np.rint(s * np.float32(0.25)), 0, 255).astype(np.uint32) << np.uint32(8 * k)
    return out


def pack_r11g11b10(c: np.ndarray) -> np.ndarray:
    """R11G11B10f::Pack (Texture.h:135-157): clamp to [2^-15, the largest
    value], then cut the float's bits (no rounding)."""
    c = c.astype(np.float32)
    lo = np.float32(1.0 / (1 << 15))
    r = np.minimum(np.maximum(c[..., 0], lo), np.float32(130048.0)).view(np.uint32)
    g = np.minimum(np.maximum(c[..., 1], lo), np.float32(130048.0)).view(np.uint32)
    b = np.minimum(np.maximum(c[..., 2], lo), np.float32(129024.0)).view(np.uint32)
    pr = ((r >> 17) & 0x3FFF) - 0x1C00
    pg = ((g >> 17) & 0x3FFF) - 0x1C00
    pb = ((b >> 18) & 0x1FFF) - 0x0E00
    return ((pr << 21) | (pg << 10) | pb).astype(np.uint32)


def unpack_r11g11b10(t: np.ndarray) -> np.ndarray:
    """R11G11B10f::Unpack (Texture.h:157-182)."""
    t = t.astype(np.uint32)
    r = ((((t >> 21) << 17) & 0x0FFE0000) + 0x38000000).astype(np.uint32).view(np.float32)
    g = ((((t >> 10) << 17) & 0x0FFE0000) + 0x38000000).astype(np.uint32).view(np.float32)
    b = (((t << 18) & 0x0FFC0000) + 0x38000000).astype(np.uint32).view(np.float32)
    return np.stack([r, g, b], axis=-1)


def next_mip_hdr(level: np.ndarray) -> np.ndarray:
    c = unpack_r11g11b10(level)
    c00, c10 = c[0::2, 0::2], c[0::2, 1::2]
    c01, c11 = c[1::2, 0::2], c[1::2, 1::2]
    return pack_r11g11b10((((c00 + c10) + c01) + c11) * np.float32(0.25))


def hdr_texture(rgb: np.ndarray, max_levels: int = 16) -> Texture:
    """An R11G11B10f texture (HdrTexture2D, Texture.h:216) from float RGB
    level 0 of shape (h, w, 3), mips by the same box filter."""
    return create([pack_r11g11b10(rgb)], max_levels, fmt=1)


def create(levels0: list, max_levels: int = 16, fmt: int = 0) -> Texture:
    """levels0: one (h, w) u32 array per layer (level 0), RGBA8 texels
    (fmt 0) or R11G11B10f texels (fmt 1)."""
    h, w = levels0[0].shape
    row_shift, offsets, layer_stride, nmips = layout(w, h, max_levels)
    data = np.zeros(layer_stride * len(levels0) + 64, dtype=np.uint32)
    for layer, lvl in enumerate(levels0):
        cur = lvl.astype(np.uint32)
        for m in range(nmips):
            hh, ww = cur.shape
            ys, xs = np.meshgrid(np.arange(hh), np.arange(ww), indexing="ij")
            off = tiled_offset(xs, ys, row_shift - m)
            data[layer * layer_stride + int(offsets[m]) + off] = cur
            if m + 1 < nmips:
                cur = next_mip(cur) if fmt == 0 else next_mip_hdr(cur)
    return Texture(w, h, nmips, len(levels0), row_shift, layer_stride, offsets, data, fmt)


def checker_texture(size: int = 256) -> Texture:
    """Procedural checker: 32-texel cells, with red/green ramps inside the
    cells so bilinear filtering has gradients to show."""
    y, x = np.meshgrid(np.arange(size, dtype=np.uint32), np.arange(size, dtype=np.uint32), indexing="ij")
    cell = ((x >> 5) ^ (y >> 5)) & 1
    r = (x * 255 // (size - 1)) & 255
    g = (y * 255 // (size - 1)) & 255
    b = np.where(cell == 1, 230, 40).astype(np.uint32)
    img = r | (g << 8) | (b << 16) | (np.uint32(255) << 24)
    return create([img.astype(np.uint32)])


def noise_texture(size: int, seed: int) -> Texture:
    from .scene import SplitMix64

    rng = SplitMix64(seed)
    bits = rng.next_u64(size * size // 2)
    img = bits.view(np.uint32).reshape(size, size) | np.uint32(0xFF000000)
    return create([img])


def cutout_texture(size: int, seed: int, period: int = 16) -> Texture:
    """Seeded noise colour with a smooth alpha pattern for the alpha test:
    alpha = 255 * (1/2 + 1/2 sin(2 pi x/period) sin(2 pi y/period)), rounded,
    so every alpha value occurs and a cutoff carves holes with texel-scale
    gradients around them."""
    from .scene import SplitMix64

    rng = SplitMix64(seed)
    rgb = rng.next_u64(size * size // 2).view(np.uint32).reshape(size, size) & np.uint32(0x00FFFFFF)
    y, x = np.meshgrid(np.arange(size, dtype=np.float64), np.arange(size, dtype=np.float64), indexing="ij")
    a = np.rint(255.0 * (0.5 + 0.5 * np.sin(2 * np.pi * x / period) * np.sin(2 * np.pi * y / period)))
    img = rgb | (a.astype(np.uint32) << np.uint32(24))
    return create([img.astype(np.uint32)])


def nmr_texture(size: int, seed: int, period: int = 32) -> Texture:
    """Two layers as the reference's material textures have them: layer 0
    seeded noise albedo, layer 1 normal/metallic/roughness (bytes 0-1 the
    tangent-space normal xy of a bump pattern, inside the unit disc; byte 2
    metallic and byte 3 roughness, constant over 8x8 blocks)."""
    from .scene import SplitMix64

    rng = SplitMix64(seed)
    albedo = rng.next_u64(size * size // 2).view(np.uint32).reshape(size, size) | np.uint32(0xFF000000)
    y, x = np.meshgrid(np.arange(size, dtype=np.float64), np.arange(size, dtype=np.float64), indexing="ij")
    nx = 0.6 * np.sin(2 * np.pi * x / period)
    ny = 0.6 * np.cos(2 * np.pi * y / period)
    bx = np.rint((nx + 1.0) * 127.5).astype(np.uint32)
    by = np.rint((ny + 1.0) * 127.5).astype(np.uint32)
    blocks = rng.next_u64((size // 8) * (size // 8)).astype(np.uint64)
    mr = (blocks & np.uint64(0xFFFF)).astype(np.uint32).reshape(size // 8, size // 8)
    mr = np.repeat(np.repeat(mr, 8, axis=0), 8, axis=1)
    nmr = bx | (by << np.uint32(8)) | (mr << np.uint32(16))
    return create([albedo.astype(np.uint32), nmr.astype(np.uint32)])


def unmap_octahedron64(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Float64 inverse of the octahedral map (Texture.h:289-296): (u, v) in
    [0, 1] -> unit direction."""
    x, y = u * 2 - 1, v * 2 - 1
    z = 1 - np.abs(x) - np.abs(y)
    t = np.maximum(-z, 0)
    x = x - np.where(x < 0, -t, t)
    y = y - np.where(y < 0, -t, t)
    d = np.stack([x, y, z], axis=-1)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def sky_rgb(d: np.ndarray) -> np.ndarray:
    """A smooth HDR sky over directions d (..., 3): a blue gradient by
    elevation, a warm horizon band and a bright sun (values up to ~40)."""
    sun = np.array([0.3, 0.55, -0.78])
    sun = sun / np.linalg.norm(sun)
    up = np.clip(d[..., 1], -1, 1)
    base = np.stack([0.35 + 0.25 * up, 0.55 + 0.3 * up, 0.9 + 0.6 * up], axis=-1).clip(0.02)
    band = np.exp(-8 * up ** 2)[..., None] * np.array([1.6, 0.9, 0.4])
    s = np.clip((d * sun).sum(axis=-1), -1, 1)
    glow = (np.maximum(s, 0) ** 64)[..., None] * np.array([40.0, 34.0, 26.0])
    return base + band + glow


def sky_texture(size: int = 256) -> Texture:
    """An octahedral HDR sky map (the layout LoadOctahedronFromPanoramaHDR
    produces, ImageHelpers.cpp:73-105) with its mips, built from sky_rgb
    with a +-30% checker of 2x2-texel cells so that every mip level looks
    different from its neighbours."""
    j, i = np.meshgrid(np.arange(size), np.arange(size), indexing="ij")
    d = unmap_octahedron64((i + 0.5) / size, (j + 0.5) / size)
    cells = 1.0 + 0.3 * ((((i >> 1) + (j >> 1)) & 1) * 2 - 1)
    return hdr_texture((sky_rgb(d) * cells[..., None]).astype(np.float32))
