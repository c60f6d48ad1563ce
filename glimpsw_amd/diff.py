"""Differentiable texture-sampling resolve behind torch.autograd.

``DiffResolve`` maps fp32 texel mip chains, per-vertex UVs and per-vertex
colours to an RGBA image through a fixed vis-buffer (glimp_diff_forward in
include/glimp.h); its backward is the HIP kernel glimp_diff_backward, whose
gradients are reduced deterministically. Torch only holds the device
tensors and the stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import DiffDesc, DiffTextureDesc, DiffViewDesc, _check, _stream_ptr
from . import textures as tx


@dataclass
class TexelChains:
    """fp32 RGBA mip chains of several textures in one flat texel array.
    Level m of texture i is ``(h >> m, w >> m)`` texels, row-major, starting
    at texel ``offsets[i][m]``."""
    data: np.ndarray            # (num_texels, 4) float32
    sizes: list                 # [(width, height, levels)]
    offsets: list               # [[offset per level]]

    @property
    def num_texels(self) -> int:
        return int(self.data.shape[0])

    def descs(self):
        arr = (DiffTextureDesc * max(1, len(self.sizes)))()
        for i, ((w, h, L), offs) in enumerate(zip(self.sizes, self.offsets)):
            arr[i].width, arr[i].height, arr[i].mip_levels = w, h, L
            for m, o in enumerate(offs):
                arr[i].level_offset[m] = o
        return arr

    def level(self, data: np.ndarray, i: int, m: int) -> np.ndarray:
        w, h, _ = self.sizes[i]
        o = self.offsets[i][m]
        return data[o:o + (w >> m) * (h >> m)].reshape(h >> m, w >> m, 4)


def chains_from_textures(texs: list) -> TexelChains:
    """Byte texels / 255 of every level of the RGBA8 textures (layer 0)."""
    parts, sizes, offsets = [], [], []
    total = 0
    for t in texs:
        offs = []
        for m in range(t.mip_levels):
            lvl = tx.unpack_f32(t.level(m)).reshape(-1, 4)
            offs.append(total)
            parts.append(lvl)
            total += lvl.shape[0]
        sizes.append((t.width, t.height, t.mip_levels))
        offsets.append(offs)
    data = np.concatenate(parts, axis=0).astype(np.float32) if parts else np.zeros((0, 4), np.float32)
    return TexelChains(data, sizes, offsets)


def vertex_uv(scene) -> np.ndarray:
    """Half texcoords of the scene as fp32 [meshlets * 64][2]."""
    tc = scene.texcoords.reshape(-1).astype(np.uint32)
    u = (tc & np.uint32(0xFFFF)).astype(np.uint16).view(np.float16).astype(np.float32)
    v = (tc >> np.uint32(16)).astype(np.uint16).view(np.float16).astype(np.float32)
    return np.stack([u, v], axis=1)


class _Binding:
    """Scene, texel layout and the views (framebuffer, object-to-clip matrix)
    of a differentiable resolve."""

    def __init__(self, ctx, dscene, views, chains: TexelChains):
        self.ctx, self.dscene = ctx, dscene
        self.views = list(views)
        self.chains = chains
        self.tex_descs = chains.descs()
        self.view_descs = (DiffViewDesc * len(self.views))()
        for i, (fb, clip) in enumerate(self.views):
            self.view_descs[i].framebuffer = fb.handle.value if hasattr(fb.handle, "value") else fb.handle
            for k in range(16):
                self.view_descs[i].object_to_clip[k] = float(clip[k])
        self.sizes = [(fb.height, fb.width) for fb, _ in self.views]
        self.total_pixels = sum(h * w for h, w in self.sizes)

    def desc(self, texels, uv, color) -> DiffDesc:
        d = DiffDesc()
        d.textures = C.cast(self.tex_descs, C.POINTER(DiffTextureDesc))
        d.num_textures = len(self.chains.sizes)
        d.num_texels = self.chains.num_texels
        d.texels, d.uv, d.color = texels.data_ptr(), uv.data_ptr(), color.data_ptr()
        return d


def _forward(b: _Binding, texels, uv, color):
    import torch

    out = torch.empty((b.total_pixels, 4), dtype=torch.float32, device=texels.device)
    d = b.desc(texels, uv, color)
    stream = torch.cuda.current_stream(texels.device)
    _check(b.ctx.lib.glimp_diff_forward_views(b.ctx.handle, b.dscene.handle, C.byref(d), b.view_descs,
                                              len(b.views), C.c_void_p(out.data_ptr()), _stream_ptr(stream)),
           "glimp_diff_forward_views")
    return out


def _backward(b: _Binding, texels, uv, color, grad_out):
    import torch

    grad_out = grad_out.contiguous().float()
    gt = torch.empty_like(texels)
    gu = torch.empty_like(uv)
    gc = torch.empty_like(color)
    d = b.desc(texels, uv, color)
    stream = torch.cuda.current_stream(texels.device)
    _check(b.ctx.lib.glimp_diff_backward_views(b.ctx.handle, b.dscene.handle, C.byref(d), b.view_descs,
                                               len(b.views), C.c_void_p(grad_out.data_ptr()),
                                               C.c_void_p(gt.data_ptr()), C.c_void_p(gu.data_ptr()),
                                               C.c_void_p(gc.data_ptr()), _stream_ptr(stream)),
           "glimp_diff_backward_views")
    return gt, gu, gc


def _torch_function():
    import torch

    class DiffResolveFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, texels, uv, color, binding):
            for t in (texels, uv, color):
                if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                    raise ValueError("texels, uv and color must be contiguous float32 device tensors")
            ctx.binding = binding
            ctx.save_for_backward(texels, uv, color)
            return _forward(binding, texels, uv, color)

        @staticmethod
        def backward(ctx, grad_out):
            texels, uv, color = ctx.saved_tensors
            gt, gu, gc = _backward(ctx.binding, texels, uv, color, grad_out)
            return (gt if ctx.needs_input_grad[0] else None, gu if ctx.needs_input_grad[1] else None,
                    gc if ctx.needs_input_grad[2] else None, None)

    return DiffResolveFn


_FN = None


def _apply(binding, texels, uv, color):
    global _FN
    if _FN is None:
        _FN = _torch_function()
    return _FN.apply(texels, uv, color, binding)


class DiffResolveViews:
    """Differentiable resolve of several framebuffers of one scene in one
    launch per stage: ``out = op(texels, uv, color)`` is (sum of H*W, 4), the
    views' row-major images one after another; gradients sum over views."""

    def __init__(self, ctx, dscene, views, chains: TexelChains):
        self.binding = _Binding(ctx, dscene, views, chains)

    def __call__(self, texels, uv, color):
        return _apply(self.binding, texels, uv, color)

    def forward(self, texels, uv, color):
        return _forward(self.binding, texels, uv, color)

    def backward(self, texels, uv, color, grad_out):
        return _backward(self.binding, texels, uv, color, grad_out)

    def split(self, out):
        """Per-view (H, W, 4) views of a (sum of H*W, 4) tensor."""
        parts, o = [], 0
        for h, w in self.binding.sizes:
            parts.append(out[o:o + h * w].view(h, w, 4))
            o += h * w
        return parts


class DiffResolve(DiffResolveViews):
    """Resolve of one framebuffer as a differentiable function:
    ``out = DiffResolve(ctx, fb, scene, object_to_clip, chains)(texels, uv, color)``
    with ``texels`` (num_texels, 4), ``uv`` (meshlets*64, 2) and ``color``
    (meshlets*64, 4) float32 CUDA tensors; ``out`` is (H, W, 4)."""

    def __init__(self, ctx, fb, dscene, object_to_clip, chains: TexelChains):
        super().__init__(ctx, dscene, [(fb, object_to_clip)], chains)
        self.shape = (fb.height, fb.width, 4)

    def __call__(self, texels, uv, color):
        return _apply(self.binding, texels, uv, color).view(self.shape)

    def forward(self, texels, uv, color):
        return _forward(self.binding, texels, uv, color).view(self.shape)

    def backward(self, texels, uv, color, grad_out):
        return _backward(self.binding, texels, uv, color, grad_out.reshape(-1, 4))
