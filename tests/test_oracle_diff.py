"""Pins the FP64 differentiable-resolve oracle on the CPU.

- its barycentrics (textbook perspective-correct formula in double) against
  the fp32 IntersectTriangle restatement and against vertex positions;
- its forward against closed forms: a constant texture returns the constant,
  a unit texture returns the perspective-correct interpolation of the
  vertex colours, a linear ramp texture returns the bilinear coordinate
  (texel centres at +0.5, clamped at the left edge);
- its backward against central finite differences (h = 1e-3) of its own
  forward, for texels, vertex UVs and vertex colours, with the discrete
  choices (surface, level, filter, texel cell) held fixed.
"""
import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from glimpsw_amd import diff as D


def _setup(n=512, w=160, h=120, seed=7):
    sc, view = S.checker_scene(n, w, h)
    fb, _, b = oracle.render_frame(sc, view)
    chains = D.chains_from_textures(sc.textures)
    uv = D.vertex_uv(sc)
    rng = np.random.default_rng(seed)
    color = rng.uniform(0.5, 1.5, size=(sc.count * 64, 4)).astype(np.float32)
    return sc, view, fb, b, chains, uv, color


def test_bary_f64_matches_fp32_intersect():
    rng = np.random.default_rng(1)
    worst = 0.0
    for _ in range(200):
        clip = np.zeros((3, 4), np.float32)
        clip[:, 3] = rng.uniform(0.5, 20.0, 3)
        clip[:, :2] = rng.uniform(-1, 1, (3, 2)) * clip[:, 3:4]
        clip[:, 2] = 0.01
        su, sv = rng.uniform(-0.5, 0.5, 2)
        b64 = oracle.bary_f64(clip, su, sv)
        # fp32 restatement: one lane of the 16-wide entry
        c16 = np.repeat(clip.reshape(3, 4, 1), 16, axis=2).astype(np.float32)
        suv = np.array([[su] * 16, [sv] * 16], np.float32)
        bary = np.zeros(48, np.float32)
        ddx = np.zeros(48, np.float32)
        ddy = np.zeros(48, np.float32)
        L = oracle.lib()
        L.orc_intersect_triangle16(oracle._ptr(c16, oracle.f32p), oracle._ptr(suv, oracle.f32p),
                                   oracle.C.c_float(640.0), oracle.C.c_float(480.0), oracle._ptr(bary, oracle.f32p),
                                   oracle._ptr(ddx, oracle.f32p), oracle._ptr(ddy, oracle.f32p))
        b32 = bary.reshape(3, 16)[:, 0].astype(np.float64)
        scale = max(1.0, np.abs(b64).max())
        worst = max(worst, np.abs(b32 - b64).max() / scale)
        assert abs(b64.sum() - 1.0) < 1e-12
    assert worst < 2e-4, worst


def test_bary_f64_at_vertices():
    clip = np.array([[0.2, 0.1, 0.01, 2.0], [-1.5, 0.3, 0.01, 3.0], [0.4, -2.0, 0.01, 4.0]], np.float32)
    for k in range(3):
        c = clip.astype(np.float64)
        b = oracle.bary_f64(clip, c[k, 0] / c[k, 3], c[k, 1] / c[k, 3])
        want = np.zeros(3)
        want[k] = 1.0
        assert np.allclose(b, want, atol=1e-12)


def test_forward_constant_texture():
    sc, view, fb, b, chains, uv, color = _setup()
    texels = np.full_like(chains.data, 0.375)
    ones = np.ones_like(color)
    out = oracle.diff_forward(fb, b, view.object_to_clip, chains, texels, uv, ones)
    covered = fb.depth() > 0
    assert covered.sum() > 1000
    assert np.allclose(out[covered], 0.375, rtol=0, atol=1e-12)
    assert (out[~covered] == 0).all()


def test_forward_vertex_colors_interpolate():
    """Unit texture: out is the perspective-correct interpolation of the
    vertex colours at the pixel centre. The fp32 barycentrics (IntersectTriangle's
    formula) agree with the textbook double formula to 1e-5 on 99% of the
    pixels; the rest are near-degenerate triangles (5e-4 bound)."""
    sc, view, fb, b, chains, uv, color = _setup()
    texels = np.ones_like(chains.data)
    out = oracle.diff_forward(fb, b, view.object_to_clip, chains, texels, uv, color)
    ids = fb.ids()
    depth = fb.depth()
    ys, xs = np.nonzero(depth > 0)
    M = view.object_to_clip.astype(np.float64).reshape(4, 4).T
    errs = []
    for y, x in zip(ys, xs):
        mi, tri = divmod(int(ids[y, x]), 128)
        vis = [int(sc.indices[mi, k, tri]) for k in range(3)]
        clip = np.zeros((3, 4), np.float32)
        for k, vi in enumerate(vis):
            p = np.array([*sc.positions[mi, :, vi].astype(np.float64), 1.0])
            clip[k] = (M @ p).astype(np.float32)
        su = (2 * x + 1) / view.width - 1
        sv = (2 * y + 1) / view.height - 1
        bw = oracle.bary_f64(clip, su, sv)
        want = sum(bw[k] * color[mi * 64 + vis[k]].astype(np.float64) for k in range(3))
        errs.append(np.abs(out[y, x] - want).max())
    errs = np.array(errs)
    assert errs.size > 5000
    assert np.percentile(errs, 99) < 1e-5, np.percentile(errs, [50, 99, 100])
    assert errs.max() < 5e-4


def test_forward_bilinear_ramp():
    """Level-0 texture whose red channel is the texel's x: bilinear pixels
    return max(fu*W - 0.5, 0) (until the right edge clamp)."""
    sc, view, fb, b, chains, uv, color = _setup()
    uv = uv * np.float32(0.004)   # magnify: a few texels per patch
    W, H, L = chains.sizes[0]
    texels = np.zeros_like(chains.data)
    lvl0 = chains.level(texels, 0, 0)
    lvl0[..., 0] = np.arange(W, dtype=np.float32)[None, :]
    lvl0[..., 3] = 1.0
    for m in range(1, L):       # mark minified pixels
        chains.level(texels, 0, m)[..., 3] = -1.0
    ones = np.ones_like(color)
    out = oracle.diff_forward(fb, b, view.object_to_clip, chains, texels, uv, ones)
    bil = (out[..., 3] == 1.0) & (fb.depth() > 0)
    assert bil.sum() > 200
    # U at the pixel from the double barycentrics of the forward's own definition
    ys, xs = np.nonzero(bil)
    ids = fb.ids()
    M = view.object_to_clip.astype(np.float64).reshape(4, 4).T
    checked = 0
    for y, x in zip(ys[:300], xs[:300]):
        mi, tri = divmod(int(ids[y, x]), 128)
        vis = [int(sc.indices[mi, k, tri]) for k in range(3)]
        clip = np.array([(M @ np.array([*sc.positions[mi, :, vi].astype(np.float64), 1.0])) for vi in vis],
                        np.float32)
        bw = oracle.bary_f64(clip, (2 * x + 1) / view.width - 1, (2 * y + 1) / view.height - 1)
        U = sum(bw[k] * float(uv[mi * 64 + vis[k], 0]) for k in range(3))
        xf = (U - np.floor(U)) * W - 0.5
        if not (0.01 < xf < W - 1.01):
            continue                   # near a clamp or a wrap: choice taken in fp32
        assert abs(out[y, x, 0] - xf) < 1e-3, (x, y, out[y, x, 0], xf)
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("magnify", [False, True])
@pytest.mark.parametrize("which", ["texels", "uv", "color"])
def test_backward_matches_central_differences(which, magnify):
    sc, view, fb, b, chains, uv, color = _setup()
    if magnify:                   # bilinear pixels (UV gradients through the weights)
        uv = uv * np.float32(0.004)
    rng = np.random.default_rng(11)
    texels = chains.data
    target = oracle.diff_forward(fb, b, view.object_to_clip, chains,
                                 texels + rng.normal(0, 0.05, texels.shape).astype(np.float32), uv, color)
    out = oracle.diff_forward(fb, b, view.object_to_clip, chains, texels, uv, color)
    gt, gu, gc = oracle.diff_backward(fb, b, view.object_to_clip, chains, texels, uv, color, 2.0 * (out - target))
    g = {"texels": gt, "uv": gu, "color": gc}[which]
    if which == "uv" and not magnify:
        assert (g == 0).all()    # nearest texels: piecewise constant in U, V
        return
    base = {"texels": texels, "uv": uv, "color": color}[which].astype(np.float64)
    flat = np.abs(g).reshape(-1)
    cand = np.nonzero(flat > 1e-3 * flat.max())[0]
    assert cand.size > 20
    h = 1e-3

    def loss(arr):
        kw = {"texels_d": None, "uv_d": None, "color_d": None}
        kw[which + "_d"] = arr
        o = oracle.diff_forward(fb, b, view.object_to_clip, chains, texels, uv, color, **kw)
        return float(((o - target) ** 2).sum())

    for i in rng.choice(cand, 12, replace=False):
        p = base.copy().reshape(-1)
        p[i] += h
        lp = loss(p.reshape(base.shape))
        p[i] -= 2 * h
        lm = loss(p.reshape(base.shape))
        fd = (lp - lm) / (2 * h)
        an = g.reshape(-1)[i]
        tol = 1e-6 * max(1.0, abs(an)) if which != "uv" else 1e-4 * max(1.0, abs(an))
        assert abs(fd - an) <= tol, (which, i, fd, an)
