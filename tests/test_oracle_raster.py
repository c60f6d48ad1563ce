"""Pins the oracle's raster against what the mathematics fixes.

The reference has no tests or golden vectors, so the oracle's integer
raster is checked here against independent statements of the same rules:
 - coverage: Pineda edge functions evaluated exactly (Python integers) at
   pixel centres on the 1/16-pixel snapped vertices, with the top-left tie
   rule, must equal the oracle's coverage; the only allowed differences are
   pixels whose centre lies exactly on the render box's rounded edge;
 - depth: the interpolated z at each covered pixel equals the float64 plane
   through the three vertices within the floor-rounding bound;
 - watertightness: a tessellated surface covers every pixel exactly once;
 - clipping: the fan triangles cover exactly the part of the triangle
   inside the planes (float64 Sutherland-Hodgman in barycentric space).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from glimpsw_amd import textures
from tests import scenes as T

f32p = C.POINTER(C.c_float)


def one_triangle_scene(verts_xyz: np.ndarray):
    """Meshlets of one triangle each, vertices given in clip space with w=1
    (object_to_clip = identity)."""
    n = verts_xyz.shape[0]
    sc = S.build_patches(np.zeros((n, 3)), np.tile(np.eye(3), (n, 1, 1)), np.ones(n), np.zeros(n, np.uint32),
                         [(0, True, 255)], [textures.checker_texture(8)])
    pos = np.zeros((n, 3, 64), np.float32)
    for k in range(3):
        pos[:, :, k] = verts_xyz[:, k, :]
    idx = np.zeros((n, 3, 128), np.uint8)
    idx[:, 0, 0], idx[:, 1, 0], idx[:, 2, 0] = 0, 1, 2
    import dataclasses
    return dataclasses.replace(sc, positions=pos, indices=idx, num_triangles=np.ones(n, np.uint8),
                               num_vertices=np.full(n, 3, np.uint8))


IDENT = np.eye(4, dtype=np.float32).reshape(16)


def textbook_coverage(X, Y, W, H):
    """Exact edge functions at pixel centres with the top-left rule."""
    hw, hh = W // 2, H // 2
    x = [int(v) for v in X]
    y = [int(v) for v in Y]
    px = np.arange(W)
    py = np.arange(H)
    sx = (px - hw) * 16 + 8
    sy = (py - hh) * 16 + 8
    A = {(0, 1): y[1] - y[0], (1, 2): y[2] - y[1], (2, 0): y[0] - y[2]}
    B = {(0, 1): x[0] - x[1], (1, 2): x[1] - x[2], (2, 0): x[2] - x[0]}
    det = B[(2, 0)] * A[(0, 1)] - B[(0, 1)] * A[(2, 0)]
    sgn = -1 if det < 0 else 1
    inside = np.ones((H, W), bool)
    for (i, j) in ((1, 2), (2, 0), (0, 1)):
        a, b = sgn * A[(i, j)], sgn * B[(i, j)]
        w = a * (sx[None, :].astype(np.int64) - x[i]) + b * (sy[:, None].astype(np.int64) - y[i])
        top_left = a > 0 or (a == 0 and b > 0)
        inside &= (w > 0) | ((w == 0) & top_left)
    return inside, det


def snapped(v, W, H):
    fsx, fsy = np.float32(W // 2 * 16), np.float32(H // 2 * 16)
    X = np.rint(v[:, 0].astype(np.float32) * fsx).astype(np.int64)
    Y = np.rint(v[:, 1].astype(np.float32) * fsy).astype(np.int64)
    return X, Y


def draw_single(tri, W, H):
    sc = one_triangle_scene(tri[None])
    b = oracle.SceneBinding(sc)
    fb = oracle.OracleFramebuffer(W, H)
    fb.clear(0, 0.0)
    ctr, cov = oracle.draw(fb, b, IDENT, 1, coverage=True)
    return cov, fb.depth(), ctr


@pytest.mark.parametrize("seed", range(4))
def test_coverage_matches_exact_edge_functions(seed):
    rng = np.random.default_rng(100 + seed)
    W, H = 64, 48
    checked = 0
    for _ in range(150):
        c = rng.uniform(-0.8, 0.8, 2)
        size = 10 ** rng.uniform(-2.5, -0.3)
        tri = np.zeros((3, 3), np.float32)
        tri[:, :2] = (c + rng.uniform(-1, 1, (3, 2)) * size).astype(np.float32)
        tri[:, 2] = 0.5
        cov, depth, ctr = draw_single(tri, W, H)
        X, Y = snapped(tri, W, H)
        tb, det = textbook_coverage(X, Y, W, H)
        if ctr["triangles_rasterized"] == 0:
            # culled by the float area test or an empty tile box: nothing drawn
            assert cov.sum() == 0
            continue
        diff = (cov > 0) != tb
        if diff.any():
            # only pixel centres lying exactly on the max vertex x/y (the
            # render box rounds the max edge to an exclusive bound) may differ
            ys, xs = np.nonzero(diff)
            cx = (xs - W // 2) * 16 + 8
            cy = (ys - H // 2) * 16 + 8
            ok = (cx == X.max()) | (cy == Y.max()) | (cy == Y.min())
            assert ok.all(), f"coverage differs at {list(zip(xs, ys))[:4]}"
            assert (tb[diff]).all(), "oracle covers a pixel the edge functions reject"
        checked += 1
    assert checked > 50


def test_top_left_rule_on_shared_edges():
    """Two triangles sharing an edge through pixel centres: every centre on
    the shared edge goes to exactly one of them."""
    W, H = 32, 32
    # vertices exactly on pixel centres: ndc = (16*(px-16)+8)/(16*16)
    def ndc(p):
        return (16 * (p - 16) + 8) / 256.0
    a, b, c, d = (2, 2), (20, 2), (20, 20), (2, 20)
    for quad in ((a, b, c, d), (b, c, d, a)):   # both diagonals
        t1 = np.array([[ndc(quad[0][0]), ndc(quad[0][1]), 0.5], [ndc(quad[1][0]), ndc(quad[1][1]), 0.5],
                       [ndc(quad[2][0]), ndc(quad[2][1]), 0.5]], np.float32)
        t2 = np.array([[ndc(quad[0][0]), ndc(quad[0][1]), 0.5], [ndc(quad[2][0]), ndc(quad[2][1]), 0.5],
                       [ndc(quad[3][0]), ndc(quad[3][1]), 0.5]], np.float32)
        c1, _, _ = draw_single(t1, W, H)
        c2, _, _ = draw_single(t2, W, H)
        assert ((c1 + c2) <= 1).all(), "a pixel on the shared edge was covered twice"
        X1, Y1 = snapped(t1, W, H)
        X2, Y2 = snapped(t2, W, H)
        tb1, _ = textbook_coverage(X1, Y1, W, H)
        tb2, _ = textbook_coverage(X2, Y2, W, H)
        assert not (tb1 & tb2).any()
        assert ((c1 + c2) == 1).sum() == 18 * 18, "the quad's 18x18 pixel centres are covered exactly once"


def test_depth_plane_closed_form():
    rng = np.random.default_rng(7)
    W, H = 64, 64
    for _ in range(60):
        tri = np.zeros((3, 3), np.float32)
        tri[:, :2] = rng.uniform(-0.9, 0.9, (3, 2))
        tri[:, 2] = rng.uniform(0.05, 0.95, 3)
        cov, depth, ctr = draw_single(tri, W, H)
        if cov.sum() == 0:
            continue
        X, Y = snapped(tri, W, H)
        ys, xs = np.nonzero(cov)
        sx = (xs - W // 2) * 16 + 8.0
        sy = (ys - H // 2) * 16 + 8.0
        x0, x1, x2 = X.astype(float)
        y0, y1, y2 = Y.astype(float)
        den = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0)
        l1 = ((sx - x0) * (y2 - y0) - (x2 - x0) * (sy - y0)) / den
        l2 = ((x1 - x0) * (sy - y0) - (sx - x0) * (y1 - y0)) / den
        z = tri[:, 2].astype(float)
        ref = z[0] + l1 * (z[1] - z[0]) + l2 * (z[2] - z[0])
        area = abs(den) / 512.0
        tol = 2e-6 + 2.0 * (abs(z[1] - z[0]) + abs(z[2] - z[0])) / (16.0 * max(area, 0.5))
        err = np.abs(depth[ys, xs].astype(float) - ref)
        assert (err <= tol).all(), (err.max(), tol)


def test_watertight_surface():
    sc = T.fullscreen_grid(200, 152, cells=6)
    view = S.make_view(200, 152, count=sc.count)
    b = oracle.SceneBinding(sc)
    fb = oracle.OracleFramebuffer(200, 152)
    fb.clear(0, 0.0)
    _, cov = oracle.draw(fb, b, view.object_to_clip, sc.count, coverage=True)
    assert (cov == 1).all()


def _clip_area_reference(v, outcodes, gbx, gby):
    """float64 Sutherland-Hodgman of the triangle in its own barycentric
    coordinates: a point l (l0+l1+l2=1) maps to clip space as sum l_k v_k,
    so each plane is a linear function of l."""
    poly = [np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])]
    V = v.astype(np.float64)
    for plane in range(6):
        if not (outcodes >> plane) & 1:
            continue
        comp = plane // 2
        sign = -1.0 if plane % 2 else 1.0
        scale = (gbx, gby, 1.0)[comp]

        def dist(l):
            p = l @ V
            return sign * p[comp] + p[3] * scale

        out = []
        for i in range(len(poly)):
            a, b = poly[i], poly[(i + 1) % len(poly)]
            da, db = dist(a), dist(b)
            if da >= 0:
                out.append(a)
            if (da >= 0) != (db >= 0):
                t = da / (da - db)
                out.append(a + t * (b - a))
        poly = out
        if len(poly) < 3:
            return 0.0
    pts = np.array([[p[1], p[2]] for p in poly])
    x, y = pts[:, 0], pts[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def _bary_of(v, q):
    """Barycentric coordinates of clip-space point q w.r.t. triangle v."""
    M = v.astype(np.float64).T  # 4x3
    l, *_ = np.linalg.lstsq(M, q.astype(np.float64), rcond=None)
    return l


def test_clipper_geometry():
    rng = np.random.default_rng(11)
    L = oracle.lib()
    checked = 0
    for _ in range(300):
        v = np.zeros((3, 4), np.float32)
        v[:, 3] = rng.uniform(-0.5, 3.0, 3)        # some vertices behind the eye
        v[:, 0] = rng.uniform(-4, 4, 3)
        v[:, 1] = rng.uniform(-4, 4, 3)
        v[:, 2] = 0.01
        codes = 0
        for k in range(3):
            x, y, z, w = v[k]
            codes |= (x < -w) << 0 | (x > w) << 1 | (y < -w) << 2 | (y > w) << 3 | (z < -w) << 4 | (z > w) << 5
        gb = 1.5
        out = np.zeros((128, 3, 4), np.float32)
        n = L.orc_clip_triangle(v.ctypes.data_as(f32p), C.c_uint32(codes), C.c_float(gb), C.c_float(gb),
                                out.ctypes.data_as(f32p))
        ref = _clip_area_reference(v, codes, gb, gb)
        area = 0.0
        for i in range(n):
            ls = [_bary_of(v, out[i, k]) for k in range(3)]
            a = np.array([ls[0][1:], ls[1][1:], ls[2][1:]])
            area += 0.5 * abs((a[1, 0] - a[0, 0]) * (a[2, 1] - a[0, 1]) - (a[2, 0] - a[0, 0]) * (a[1, 1] - a[0, 1]))
            for k in range(3):
                x, y, z, w = out[i, k].astype(np.float64)
                for plane in range(6):
                    if (codes >> plane) & 1:
                        comp = plane // 2
                        s = (gb, gb, 1.0)[comp]
                        d = (-1 if plane % 2 else 1) * (x, y, z)[comp] + w * s
                        assert d >= -1e-4 * (1 + abs(w)), "clipped vertex outside a clip plane"
        assert area == pytest.approx(ref, rel=1e-4, abs=1e-6)
        checked += n > 0
    assert checked > 50
