"""Edge gradients, the opacity auxiliary gradient, view-direction encoding and
triangle-id counting.

Mirrors python/diffsoup/rasterize.py (same names, argument meaning and
assertions); the arithmetic runs in libdiffsoup_hip.so.
"""
from __future__ import annotations

import torch

from . import _lib
from ._workspace import raster_workspace
from .multires import feats_at_level


def _views(B: int, H: int, W: int):
    return _lib.int_array([H] * B), _lib.int_array([W] * B)


# ---------------------------------------------------------------------------
# Opacity auxiliary gradient (python/diffsoup/rasterize.py:139-252)
# ---------------------------------------------------------------------------

class _OpacityAuxLossFn(torch.autograd.Function):
    """Zero-valued loss whose backward is the analytic opacity-masking
    gradient.  Fragments are regenerated by the HIP raster bins; nothing
    is stored between forward and backward but the gradient itself."""

    @staticmethod
    def forward(ctx, color, target, rast, pos, tri, level, alpha_src):
        B, H, W, C = color.shape
        T = tri.shape[0]
        V = pos.shape[1]
        dev = color.device
        P = B * H * W
        s = _lib.stream_ptr(dev)
        pix_loss = torch.empty(P, dtype=torch.float32, device=dev)
        _lib.check(_lib.lib().ds_pixel_l1(P, C, _lib.ptr(color), _lib.ptr(target), _lib.ptr(pix_loss), s),
                   "ds_pixel_l1")
        grad_alpha_src = torch.zeros_like(alpha_src)
        ws = raster_workspace(dev)
        wptr, nbytes, cap, Ha, Wa = ws.prepare(B, [H] * B, [W] * B, T)
        _lib.check(_lib.lib().ds_opacity_aux_bwd(
            B, Ha, Wa, V, _lib.ptr(pos), T, _lib.ptr(tri), level, _lib.ptr(alpha_src), 16,
            _lib.ptr(rast), _lib.ptr(pix_loss), _lib.ptr(grad_alpha_src), 0, wptr, nbytes, cap, s),
            "ds_opacity_aux_bwd")
        ws.note_call()
        ctx.save_for_backward(grad_alpha_src)
        ctx.weight = 1.0 / color.numel()
        return torch.zeros(1, dtype=torch.float32, device=dev)

    @staticmethod
    def backward(ctx, grad_loss):
        (grad_alpha_src,) = ctx.saved_tensors
        return None, None, None, None, None, None, ctx.weight * grad_loss * grad_alpha_src


def opacity_aux_loss(color, target, rast, pos, tri, level: int, alpha_src) -> torch.Tensor:
    """Zero-valued scalar whose backward fills alpha_src.grad with the
    stochastic-opacity-masking gradient (python/diffsoup/rasterize.py:199-252)."""
    B, H, W, C = color.shape
    _, V, _ = pos.shape
    T, _ = tri.shape
    _, S, _ = alpha_src.shape
    assert color.is_contiguous() and color.dtype == torch.float32
    assert target.shape == (B, H, W, C) and target.is_contiguous() and target.dtype == torch.float32
    assert rast.shape == (B, H, W, 4) and rast.is_contiguous() and rast.dtype == torch.float32
    assert pos.shape == (B, V, 4) and pos.is_contiguous() and pos.dtype == torch.float32
    assert tri.shape == (T, 3) and tri.is_contiguous() and tri.dtype == torch.int32
    assert alpha_src.shape == (T, S, 1) and alpha_src.dtype == torch.float32 and alpha_src.is_contiguous()
    assert S == feats_at_level(level)
    _lib.require_device(color, "color")
    assert target.device == color.device and rast.device == color.device
    assert pos.device == color.device and tri.device == color.device
    alpha_src = alpha_src.squeeze(-1)
    return _OpacityAuxLossFn.apply(color, target, rast, pos.detach(), tri, level, alpha_src)


# ---------------------------------------------------------------------------
# Edge gradients (python/diffsoup/rasterize.py:259-312)
# ---------------------------------------------------------------------------

class _EdgeGradFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, color, rast, pos, tri):
        ctx.save_for_backward(color, rast, pos, tri)
        return color.view_as(color)

    @staticmethod
    def backward(ctx, grad_color):
        color, rast, pos, tri = ctx.saved_tensors
        grad_color = grad_color.contiguous()
        grad_pos = torch.zeros_like(pos)
        B, H, W, C = color.shape
        Ha, Wa = _views(B, H, W)
        _lib.check(_lib.lib().ds_edge_grad_bwd(
            B, Ha, Wa, C, _lib.ptr(color), _lib.ptr(grad_color), _lib.ptr(rast), pos.shape[1],
            _lib.ptr(pos), _lib.ptr(grad_pos), tri.shape[0], _lib.ptr(tri), _lib.stream_ptr(color.device)),
            "ds_edge_grad_bwd")
        return grad_color, None, grad_pos, None


def edge_grad(color, rast, pos, tri) -> torch.Tensor:
    """Identity on color whose backward adds silhouette / edge gradients to pos."""
    B, H, W, C = color.shape
    _, V, _ = pos.shape
    T, _ = tri.shape
    assert color.is_contiguous() and color.dtype == torch.float32
    assert rast.shape == (B, H, W, 4) and rast.is_contiguous() and rast.dtype == torch.float32
    assert pos.shape == (B, V, 4) and pos.is_contiguous() and pos.dtype == torch.float32
    assert tri.shape == (T, 3) and tri.is_contiguous() and tri.dtype == torch.int32
    _lib.require_device(color, "color")
    assert rast.device == color.device and pos.device == color.device and tri.device == color.device
    return _EdgeGradFn.apply(color, rast, pos, tri)


# ---------------------------------------------------------------------------
# View-direction encoding (python/diffsoup/rasterize.py:319-341)
# ---------------------------------------------------------------------------

def encode_view_dir_sh2(rast: torch.Tensor, inv_mvp: torch.Tensor) -> torch.Tensor:
    B, H, W, _ = rast.shape
    assert rast.shape == (B, H, W, 4) and rast.is_contiguous() and rast.dtype == torch.float32
    assert inv_mvp.shape == (B, 4, 4) and inv_mvp.is_contiguous() and inv_mvp.dtype == torch.float32
    _lib.require_device(rast, "rast")
    assert inv_mvp.device == rast.device
    enc = torch.empty(B, H, W, 9, dtype=torch.float32, device=rast.device)
    Ha, Wa = _views(B, H, W)
    _lib.check(_lib.lib().ds_sh2(B, Ha, Wa, _lib.ptr(rast), _lib.ptr(inv_mvp), _lib.ptr(enc),
                                 _lib.stream_ptr(rast.device)), "ds_sh2")
    return enc


# ---------------------------------------------------------------------------
# Utilities (python/diffsoup/rasterize.py:348-366)
# ---------------------------------------------------------------------------

def count_triangle_ids(rast: torch.Tensor, num_tris: int, out: torch.Tensor | None = None) -> torch.Tensor:
    """Pixels per triangle in a raster (python/diffsoup/rasterize.py:348-366):
    the bincount of the 1-based ids in rast[..., 3], as a (num_tris,) int64
    tensor.  With ``out`` the counts are added to it (ds_count_triangle_ids)."""
    assert rast.shape[-1] == 4 and rast.dtype == torch.float32 and rast.is_contiguous()
    _lib.require_device(rast, "rast")
    acc = out is not None
    if out is None:
        out = torch.empty(num_tris, dtype=torch.int64, device=rast.device)
    assert out.shape == (num_tris,) and out.dtype == torch.int64 and out.is_contiguous()
    _lib.check(_lib.lib().ds_count_triangle_ids(rast.numel() // 4, _lib.ptr(rast), int(num_tris), _lib.ptr(out),
                                                int(acc), _lib.stream_ptr(rast.device)), "ds_count_triangle_ids")
    return out
