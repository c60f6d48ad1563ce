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
