"""Overdraw stress near the fixed-point limit: 4,014,080 triangles in 8
depth-stacked layers (depth complexity about 8, 30% of patches sub-pixel),
16 views of 2880x1620 rendered by one batched launch (identical to per-view
draws; one view against the oracle), then the differentiable resolve of all
16 views in one forward and one backward launch (reproducible; its forward
equals per-view forwards; two views' gradients against the FP64 oracle)."""
import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from glimpsw_amd import diff as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stress(gpu_ctx):
    import glimpsw_amd as g

    sc, views = S.stress_views()
    ds = g.Scene(gpu_ctx, sc)
    fbs = [g.Framebuffer(gpu_ctx, v.width, v.height) for v in views]
    yield sc, views, ds, fbs
    for fb in fbs:
        fb.close()
    ds.close()


def test_stress_batch_matches_per_view_and_oracle(gpu_ctx, stress):
    import glimpsw_amd as g

    sc, views, ds, fbs = stress
    b = g.Batch(gpu_ctx, [(ds, fb, v) for fb, v in zip(fbs, views)])
    print(f"[stress] {len(views)} views {views[0].width}x{views[0].height}, {sc.triangles()} tris", flush=True)
    b.launch_per_view()
    want = [fb.read() for fb in fbs]
    b.launch(graph=True)
    got = [fb.read() for fb in fbs]
    for i, ((wi, wd), (gi, gd)) in enumerate(zip(want, got)):
        assert (wi == gi).all() and (wd.view(np.uint32) == gd.view(np.uint32)).all(), f"view {i}"
    b.close()
    ofb, ctr, _ = oracle.render_frame(sc, views[0])
    print(f"[stress] oracle view 0: {ctr}", flush=True)
    assert (got[0][0] == ofb.ids()).all()
    assert (got[0][1].view(np.uint32) == ofb.depth().view(np.uint32)).all()


def test_stress_diff_views(gpu_ctx, stress):
    import torch
    import glimpsw_amd as g

    sc, views, ds, fbs = stress
    b = g.Batch(gpu_ctx, [(ds, fb, v) for fb, v in zip(fbs, views)])
    b.launch(graph=True)
    b.close()
    chains = D.chains_from_textures(sc.textures)
    uv = D.vertex_uv(sc) * np.float32(1.0 / 32)
    rng = np.random.default_rng(4)
    color = rng.uniform(0.5, 1.5, size=(sc.count * 64, 4)).astype(np.float32)
    dev = torch.device("cuda", 0)
    texels_t, uv_t, color_t = (torch.from_numpy(a).to(dev) for a in (chains.data, uv, color))
    multi = D.DiffResolveViews(gpu_ctx, ds, [(fb, v.object_to_clip) for fb, v in zip(fbs, views)], chains)
    out = multi.forward(texels_t, uv_t, color_t)
    noisy = texels_t + 0.05 * torch.randn(texels_t.shape, generator=torch.Generator(device=dev).manual_seed(3),
                                          device=dev)
    target = multi.forward(noisy, uv_t, color_t)
    gout = 2.0 * (out - target)
    grads = multi.backward(texels_t, uv_t, color_t, gout)
    again = multi.backward(texels_t, uv_t, color_t, gout)
    for a, c in zip(grads, again):
        assert torch.equal(a, c)
    print("[stress] diff 16 views reproducible", flush=True)
    # two views through the FP64 oracle, as a two-view call
    sub = D.DiffResolveViews(gpu_ctx, ds, [(fbs[i], views[i].object_to_clip) for i in (0, 7)], chains)
    sout = sub.forward(texels_t, uv_t, color_t)
    parts = multi.split(out)
    for k, i in enumerate((0, 7)):
        assert torch.equal(sub.split(sout)[k], parts[i])
    sg = sub.split(gout.new_zeros(sout.shape))
    gsub = torch.cat([multi.split(gout)[i].reshape(-1, 4) for i in (0, 7)])
    grads2 = sub.backward(texels_t, uv_t, color_t, gsub)
    ref = [np.zeros((chains.num_texels, 4)), np.zeros((uv.shape[0], 2)), np.zeros((color.shape[0], 4))]
    off = 0
    for i in (0, 7):
        v = views[i]
        ofb, _, bind = oracle.render_frame(sc, v)
        n = v.width * v.height
        g64 = oracle.diff_backward(ofb, bind, v.object_to_clip, chains, chains.data, uv, color,
                                   gsub[off:off + n].cpu().numpy().astype(np.float64).reshape(v.height, v.width, 4))
        off += n
        for r, x in zip(ref, g64):
            r += x
    del sg
    for name, gg, r in zip(("texels", "uv", "color"), grads2, ref):
        got = gg.cpu().numpy().astype(np.float64)
        rel = np.linalg.norm(got - r) / np.linalg.norm(r)
        print(f"[stress] {name}: norm-rel={rel:.3g}", flush=True)
        assert rel <= 1e-4
