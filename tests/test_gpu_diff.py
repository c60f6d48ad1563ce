"""Differentiable resolve on the GPU against the FP64 oracle.

Forward within 2e-5 absolute of the FP64 forward; backward (texels, vertex
UVs, vertex colours) within 1e-4 relative of the FP64 gradient, per tensor
in the 2-norm and per element against the tensor's largest entry; gradients
bitwise identical from run to run; the torch.autograd path equals the
direct backward. The full-size case is the 1M-triangle scene rendered at
1024x1024 with 16 noise materials of 2048^2 (fp32 mip chains), UVs scaled
by 1/128, loss L2 against a target rendered from perturbed texels.
"""
import numpy as np
import pytest

import oracle
from glimpsw_amd import scene as S
from glimpsw_amd import diff as D

pytestmark = pytest.mark.gpu

FWD_ATOL = 2e-5
GRAD_RTOL = 1e-4


def _case(ctx, sc, view, magnify=1.0, seed=5):
    import torch
    import glimpsw_amd as g

    print(f"[diff] case {view.width}x{view.height} meshlets={sc.count} magnify={magnify}", flush=True)
    ds = g.Scene(ctx, sc)
    fb = g.render_frame(ctx, ds, view)
    ofb, _, bind = oracle.render_frame(sc, view)
    chains = D.chains_from_textures(sc.textures)
    uv = D.vertex_uv(sc) * np.float32(magnify)
    rng = np.random.default_rng(seed)
    color = rng.uniform(0.5, 1.5, size=(sc.count * 64, 4)).astype(np.float32)
    noisy = chains.data + rng.normal(0, 0.05, chains.data.shape).astype(np.float32)
    dev = torch.device("cuda", 0)
    T = {k: torch.from_numpy(v).to(dev) for k, v in (("texels", chains.data), ("uv", uv), ("color", color),
                                                       ("noisy", noisy))}
    op = D.DiffResolve(ctx, fb, ds, view.object_to_clip, chains)
    return dict(ds=ds, fb=fb, ofb=ofb, bind=bind, chains=chains, uv=uv, color=color, noisy=noisy, T=T, op=op,
                view=view)


def _close(c):
    c["fb"].close()
    c["ds"].close()


def _check_case(c):
    import torch

    T, op, view = c["T"], c["op"], c["view"]
    out = op.forward(T["texels"], T["uv"], T["color"])
    target = op.forward(T["noisy"], T["uv"], T["color"])
    torch.cuda.synchronize()
    o64 = oracle.diff_forward(c["ofb"], c["bind"], view.object_to_clip, c["chains"], c["chains"].data, c["uv"],
                              c["color"])
    fwd_err = np.abs(out.cpu().numpy().astype(np.float64) - o64).max()
    assert fwd_err <= FWD_ATOL, f"forward differs by {fwd_err}"
    gout = 2.0 * (out - target)
    grads = op.backward(T["texels"], T["uv"], T["color"], gout)
    torch.cuda.synchronize()
    g64 = oracle.diff_backward(c["ofb"], c["bind"], view.object_to_clip, c["chains"], c["chains"].data, c["uv"],
                               c["color"], gout.cpu().numpy().astype(np.float64))
    for name, gg, ref in zip(("texels", "uv", "color"), grads, g64):
        got = gg.cpu().numpy().astype(np.float64)
        nref = np.linalg.norm(ref)
        if nref == 0:
            assert not got.any(), f"{name}: expected zero gradient"
            continue
        rel = np.linalg.norm(got - ref) / nref
        elem = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"[diff]   {name}: |g|={nref:.4g} norm-rel={rel:.3g} elem-rel={elem:.3g}", flush=True)
        assert rel <= GRAD_RTOL and elem <= GRAD_RTOL, (name, rel, elem)
    # bitwise reproducible
    again = op.backward(T["texels"], T["uv"], T["color"], gout)
    for a, b in zip(grads, again):
        assert torch.equal(a, b)
    return out, target, grads


@pytest.mark.parametrize("magnify", [1.0, 0.004])
def test_small_scene(gpu_ctx, magnify):
    sc, view = S.checker_scene(512, 160, 120)
    c = _case(gpu_ctx, sc, view, magnify)
    try:
        _check_case(c)
    finally:
        _close(c)


def test_autograd_function(gpu_ctx):
    import torch

    sc, view = S.checker_scene(256, 96, 64)
    c = _case(gpu_ctx, sc, view, 0.01)
    try:
        T, op = c["T"], c["op"]
        texels = T["texels"].clone().requires_grad_(True)
        uv = T["uv"].clone().requires_grad_(True)
        color = T["color"].clone().requires_grad_(True)
        target = op.forward(T["noisy"], T["uv"], T["color"])
        out = op(texels, uv, color)
        loss = ((out - target) ** 2).sum()
        loss.backward()
        gout = 2.0 * (out.detach() - target)
        direct = op.backward(T["texels"], T["uv"], T["color"], gout)
        for a, b in zip((texels.grad, uv.grad, color.grad), direct):
            assert torch.equal(a, b)
    finally:
        _close(c)


def test_uncovered_and_empty(gpu_ctx):
    """A framebuffer with nothing drawn: zero output, zero gradients."""
    import torch
    import glimpsw_amd as g

    sc, view = S.checker_scene(16, 32, 16)
    ds = g.Scene(gpu_ctx, sc)
    fb = g.Framebuffer(gpu_ctx, 32, 16)
    fb.clear(0, 0.0)
    chains = D.chains_from_textures(sc.textures)
    dev = torch.device("cuda", 0)
    texels = torch.from_numpy(chains.data).to(dev)
    uv = torch.from_numpy(D.vertex_uv(sc)).to(dev)
    color = torch.ones((sc.count * 64, 4), device=dev)
    op = D.DiffResolve(gpu_ctx, fb, ds, view.object_to_clip, chains)
    out = op.forward(texels, uv, color)
    assert not out.any()
    gt, gu, gc = op.backward(texels, uv, color, torch.ones_like(out))
    assert not gt.any() and not gu.any() and not gc.any()
    fb.close()
    ds.close()


def test_full_size_1024(gpu_ctx):
    """The 1M-triangle scene at 1024x1024, 16 x 2048^2 fp32 mip chains."""
    sc, _ = S.frame_1080p()
    view = S.make_view(1024, 1024, count=sc.count)
    # UVs scaled so that patches wider than ~110 px magnify (bilinear) and
    # the rest minify: both filters and the UV gradients are exercised
    c = _case(gpu_ctx, sc, view, magnify=1.0 / 128)
    try:
        _check_case(c)
    finally:
        _close(c)


def test_views_in_one_launch(gpu_ctx):
    """Three framebuffers of one scene in one call: the forward equals the
    single-view forwards bit for bit; the gradients (summed over views) match
    the FP64 sum within 1e-4 and are reproducible."""
    import torch
    import glimpsw_amd as g

    sc, _ = S.checker_scene(512, 160, 120)
    ds = g.Scene(gpu_ctx, sc)
    views = [S.make_view(w, h, count=sc.count, view=S.view_matrix((0.0, 0.0, 0.0), yaw, 0.0))
             for (w, h, yaw) in ((160, 120, 0.0), (96, 64, 0.3), (128, 128, -0.2))]
    fbs = [g.render_frame(gpu_ctx, ds, v) for v in views]
    chains = D.chains_from_textures(sc.textures)
    uv = D.vertex_uv(sc) * np.float32(0.01)
    rng = np.random.default_rng(9)
    color = rng.uniform(0.5, 1.5, size=(sc.count * 64, 4)).astype(np.float32)
    dev = torch.device("cuda", 0)
    texels_t, uv_t, color_t = (torch.from_numpy(a).to(dev) for a in (chains.data, uv, color))
    multi = D.DiffResolveViews(gpu_ctx, ds, [(fb, v.object_to_clip) for fb, v in zip(fbs, views)], chains)
    out = multi.forward(texels_t, uv_t, color_t)
    singles = [D.DiffResolve(gpu_ctx, fb, ds, v.object_to_clip, chains) for fb, v in zip(fbs, views)]
    for part, op in zip(multi.split(out), singles):
        assert torch.equal(part, op.forward(texels_t, uv_t, color_t))
    gout = torch.from_numpy(rng.normal(0, 1, size=tuple(out.shape)).astype(np.float32)).to(dev)
    grads = multi.backward(texels_t, uv_t, color_t, gout)
    again = multi.backward(texels_t, uv_t, color_t, gout)
    for a, b in zip(grads, again):
        assert torch.equal(a, b)
    ref = [np.zeros((chains.num_texels, 4)), np.zeros((uv.shape[0], 2)), np.zeros((color.shape[0], 4))]
    for v, part in zip(views, multi.split(gout)):
        ofb, _, bind = oracle.render_frame(sc, v)
        g64 = oracle.diff_backward(ofb, bind, v.object_to_clip, chains, chains.data, uv, color,
                                   part.cpu().numpy().astype(np.float64))
        for r, x in zip(ref, g64):
            r += x
    for name, gg, r in zip(("texels", "uv", "color"), grads, ref):
        got = gg.cpu().numpy().astype(np.float64)
        rel = np.linalg.norm(got - r) / np.linalg.norm(r)
        elem = np.abs(got - r).max() / np.abs(r).max()
        print(f"[diff] views {name}: norm-rel={rel:.3g} elem-rel={elem:.3g}", flush=True)
        assert rel <= GRAD_RTOL and elem <= GRAD_RTOL
    for fb in fbs:
        fb.close()
    ds.close()
