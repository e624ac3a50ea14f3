"""Multi-resolution triangle features: opacity rasterisation, colour gather,
cross-level accumulation and the colour MLP.

Mirrors python/diffsoup/multires.py (same names, argument meaning and
assertion behaviour); the arithmetic runs in libdiffsoup_hip.so.
Keyword-only extensions: ``seed`` and ``tile`` for rasterisation, and the
multi-level gather ``multires_feature_gather(..., mode='sum'|'concat')``.
"""
from __future__ import annotations

from typing import Tuple

import torch
from torch import nn

from . import _lib
from ._workspace import raster_workspace


def feats_at_level(level: int) -> int:
    """Lattice samples per triangle at a level: 3 at level 0, else
    (2^(L-1)+1)(2^L+1) (python/diffsoup/multires.py:16-23)."""
    assert level >= 0
    return 3 if level == 0 else ((1 << (level - 1)) + 1) * ((1 << level) + 1)


def feats_in_levels(min_level: int, max_level: int) -> int:
    return sum(feats_at_level(l) for l in range(min_level, max_level + 1))


def build_multires_triangle_color(T: int, min_level: int, max_level: int, feat_dim: int) -> torch.Tensor:
    """Zero (T, S, feat_dim) buffer (python/diffsoup/multires.py:26-45)."""
    return torch.zeros(T, feats_in_levels(min_level, max_level), feat_dim, dtype=torch.float32)


def _draw_seed() -> int:
    # CPU generator: honours torch.manual_seed and never touches the GPU stream
    return int(torch.randint(0, 2**62, (1,), dtype=torch.int64).item())


def rasterize_views(
    resolutions,
    pos: torch.Tensor,
    tri: torch.Tensor,
    level: int,
    alpha_src: torch.Tensor,
    stochastic: bool = True,
    *,
    seed: int | None = None,
    tile: int = 16,
    stats: bool = False,
):
    """Rasterise views of possibly different resolution in one launch.

    resolutions: list of (H, W), one per view; pos (B, V, 4).  Returns the
    packed rast (sum H*W, 4): view b occupies rows [sum_{b'<b} H W, ...).
    With ``stats=True`` also returns the device int64 stats block
    (DS_STAT_COVERED = covered fragments before the mask, DS_STAT_PAIRS, ...).
    """
    B, V, _ = pos.shape
    T = tri.shape[0]
    assert len(resolutions) == B
    assert pos.dtype == torch.float32 and pos.is_contiguous()
    assert tri.shape == (T, 3) and tri.dtype == torch.int32 and tri.is_contiguous()
    assert alpha_src.numel() == T * feats_at_level(level)
    a2 = alpha_src.reshape(T, feats_at_level(level))
    assert a2.dtype == torch.float32 and a2.is_contiguous()
    _lib.require_device(pos, "pos")
    assert tri.device == pos.device and alpha_src.device == pos.device
    H = [int(r[0]) for r in resolutions]
    W = [int(r[1]) for r in resolutions]
    P = sum(h * w for h, w in zip(H, W))
    rast = torch.empty((P, 4), dtype=torch.float32, device=pos.device)
    if seed is None:
        seed = _draw_seed() if stochastic else 0
    ws = raster_workspace(pos.device)
    wptr, nbytes, cap, Ha, Wa = ws.prepare(B, H, W, T, tile)
    _lib.check(_lib.lib().ds_rasterize_fwd(
        B, Ha, Wa, V, _lib.ptr(pos), T, _lib.ptr(tri), level, _lib.ptr(a2), int(bool(stochastic)),
        seed & (2**64 - 1), tile, _lib.ptr(rast), wptr, nbytes, cap, _lib.stream_ptr(pos.device)),
        "ds_rasterize_fwd")
    ws.note_call()
    if stats:
        return rast, ws.stats_tensor().clone()
    return rast


def rasterize_multires_triangle_alpha(
    resolution: Tuple[int, int],
    pos: torch.Tensor,
    tri: torch.Tensor,
    level: int,
    alpha_src: torch.Tensor,
    stochastic: bool = True,
    *,
    seed: int | None = None,
    tile: int = 16,
) -> torch.Tensor:
    """Rasterise with per-fragment multires opacity and depth testing.

    Same contract as python/diffsoup/multires.py:52-102: pos (B, V, 4),
    tri (T, 3) int32, alpha_src (T, S, 1) with S = feats_at_level(level);
    returns (B, H, W, 4) = (b0, b1, z/w, triangle id + 1), zeros on
    background.  ``stochastic`` keeps a fragment when alpha >= u, with u a
    hash of (seed, view, triangle, row, column); ``seed=None`` draws the seed
    from torch's CPU generator.  ``stochastic=False`` uses u = 0.5.
    """
    H, W = resolution
    B, V, _ = pos.shape
    T, _ = tri.shape
    _, S, _ = alpha_src.shape
    assert pos.shape == (B, V, 4) and pos.dtype == torch.float32 and pos.is_contiguous()
    assert tri.shape == (T, 3) and tri.dtype == torch.int32 and tri.is_contiguous()
    assert alpha_src.shape == (T, S, 1) and alpha_src.dtype == torch.float32 and alpha_src.is_contiguous()
    assert S == feats_at_level(level)
    _lib.require_device(pos, "pos")
    rast = rasterize_views([(H, W)] * B, pos, tri, level, alpha_src.detach(), stochastic, seed=seed, tile=tile)
    return rast.view(B, H, W, 4)


# ---------------------------------------------------------------------------
# Colour MLP (plain PyTorch in the reference, kept so that models port)
# ---------------------------------------------------------------------------

class ColorMLP(nn.Module):
    """Per-pixel MLP with residual blend (python/diffsoup/multires.py:109-207):

    output = (1 - x[..., 3]) * x[..., :3] + x[..., 3] * sigmoid(mlp(x)).

    The fused HIP shading kernel (diffsoup_amd.fused) evaluates the same
    network; this module keeps the reference's torch behaviour for the
    drop-in path.
    """

    def __init__(self, input_dim: int, output_dim: int, hidden_dim: int = 16, n_layers: int = 2,
                 zero_last: bool = False):
        super().__init__()
        layers: list[nn.Module] = [nn.Linear(input_dim, hidden_dim), nn.ReLU()]
        for _ in range(n_layers - 1):
            layers += [nn.Linear(hidden_dim, hidden_dim), nn.ReLU()]
        layers += [nn.Linear(hidden_dim, output_dim), nn.Sigmoid()]
        self.input_dim = input_dim
        self.output_dim = output_dim
        self.hidden_dim = hidden_dim
        self.n_layers = n_layers
        self.mlp = nn.Sequential(*layers)
        self._init_weights(zero_last=zero_last)

    def _init_weights(self, zero_last: bool = False):
        linears = [m for m in self.mlp if isinstance(m, nn.Linear)]
        *hidden, last = linears
        for lin in hidden:
            nn.init.kaiming_uniform_(lin.weight, a=0.0, mode="fan_in", nonlinearity="relu")
            nn.init.zeros_(lin.bias)
        if zero_last:
            nn.init.zeros_(last.weight)
            nn.init.zeros_(last.bias)
        else:
            nn.init.xavier_uniform_(last.weight, gain=1.0)
            nn.init.zeros_(last.bias)

    def linears(self):
        return [m for m in self.mlp if isinstance(m, nn.Linear)]

    def forward(self, x: torch.Tensor, mask: torch.Tensor | None = None):
        B, H, W, _ = x.shape
        assert x.shape == (B, H, W, self.input_dim)
        assert mask is None or mask.shape == (B, H, W)
        rgb = x[..., :3]
        res = x[..., 3:4]
        if mask is not None:
            x_flat = x.view(-1, self.input_dim)
            mask_flat = mask.view(-1)
            out = torch.zeros(B * H * W, self.output_dim, device=x.device, dtype=x.dtype)
            if mask_flat.any():
                y = self.mlp(x_flat[mask_flat])
                vr = rgb[mask]
                vs = res[mask]
                out[mask_flat] = (1.0 - vs) * vr + vs * y
            return out.view(B, H, W, self.output_dim)
        y = self.mlp(x.view(-1, self.input_dim)).view(B, H, W, self.output_dim)
        return (1.0 - res) * rgb + res * y


# ---------------------------------------------------------------------------
# Multires colour gather (differentiable in feat)
# ---------------------------------------------------------------------------

class _MultiresFeatureFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rast, min_level, max_level, feat, concat):
        P = rast.numel() // 4
        T, S, C = feat.shape
        width = (max_level - min_level + 1) * C if concat else C
        out = torch.empty(P, width, dtype=torch.float32, device=rast.device)
        _lib.check(_lib.lib().ds_multires_color_fwd(
            P, _lib.ptr(rast), min_level, max_level, T, C, _lib.ptr(feat), int(concat), _lib.ptr(out),
            _lib.stream_ptr(rast.device)), "ds_multires_color_fwd")
        ctx.save_for_backward(rast)
        ctx.meta = (min_level, max_level, concat, feat.shape)
        return out.view(*rast.shape[:-1], width)

    @staticmethod
    def backward(ctx, grad_out):
        (rast,) = ctx.saved_tensors
        lo, hi, concat, shape = ctx.meta
        T, S, C = shape
        grad_out = grad_out.contiguous()
        grad_feat = torch.zeros(shape, dtype=torch.float32, device=rast.device)
        _lib.check(_lib.lib().ds_multires_color_bwd(
            rast.numel() // 4, _lib.ptr(rast), lo, hi, T, C, _lib.ptr(grad_feat), int(concat),
            _lib.ptr(grad_out), _lib.stream_ptr(rast.device)), "ds_multires_color_bwd")
        return None, None, None, grad_feat, None


def multires_feature_gather(rast: torch.Tensor, min_level: int, max_level: int, feat: torch.Tensor,
                            mode: str = "sum") -> torch.Tensor:
    """Gather levels [min_level, max_level] of feat (T, S_tot, C) per pixel.

    mode='sum' adds the levels (the reference's multires_triangle_color
    semantics, src/cuda/multires.cu:10-68); mode='concat' returns one C-wide
    slice per level, each equal to the single-level gather at that level.
    """
    assert mode in ("sum", "concat")
    assert rast.shape[-1] == 4 and rast.dtype == torch.float32 and rast.is_contiguous()
    assert feat.dim() == 3 and feat.dtype == torch.float32 and feat.is_contiguous()
    assert feat.shape[1] == feats_in_levels(min_level, max_level)
    _lib.require_device(rast, "rast")
    assert feat.device == rast.device
    return _MultiresFeatureFn.apply(rast, min_level, max_level, feat, mode == "concat")


def multires_triangle_color(rast: torch.Tensor, level: int, feat: torch.Tensor) -> torch.Tensor:
    """Interpolate per-triangle features at one level (python/diffsoup/multires.py:245-272)."""
    B, H, W, _ = rast.shape
    _, S, _ = feat.shape
    assert rast.shape == (B, H, W, 4) and rast.dtype == torch.float32 and rast.is_contiguous()
    assert feat.dim() == 3 and feat.dtype == torch.float32 and feat.is_contiguous()
    assert S == feats_at_level(level)
    _lib.require_device(rast, "rast")
    return _MultiresFeatureFn.apply(rast, level, level, feat, False)


# ---------------------------------------------------------------------------
# Cross-level accumulation (differentiable)
# ---------------------------------------------------------------------------

class _AccumulateToLevelFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, min_level, max_level, target_level, feat):
        T, S, C = feat.shape
        out = torch.empty(T, feats_at_level(target_level), C, dtype=torch.float32, device=feat.device)
        _lib.check(_lib.lib().ds_accumulate_fwd(
            T, min_level, max_level, target_level, C, _lib.ptr(feat), _lib.ptr(out),
            _lib.stream_ptr(feat.device)), "ds_accumulate_fwd")
        ctx.meta = (min_level, max_level, target_level, feat.shape)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        lo, hi, target, shape = ctx.meta
        T, S, C = shape
        grad_out = grad_out.contiguous()
        grad_feat = torch.zeros(shape, dtype=torch.float32, device=grad_out.device)
        _lib.check(_lib.lib().ds_accumulate_bwd(
            T, lo, hi, target, C, _lib.ptr(grad_feat), _lib.ptr(grad_out),
            _lib.stream_ptr(grad_out.device)), "ds_accumulate_bwd")
        return None, None, None, grad_feat


def accumulate_to_level(min_level: int, max_level: int, feat: torch.Tensor,
                        target_level: int | None = None) -> torch.Tensor:
    """Evaluate the level-summed features at the target lattice
    (python/diffsoup/multires.py:312-339)."""
    assert min_level >= 0
    assert feat.ndim == 3 and feat.dtype == torch.float32
    assert feat.is_contiguous()
    _lib.require_device(feat, "feat")
    assert feat.shape[1] == feats_in_levels(min_level, max_level)
    if target_level is None:
        target_level = max_level
    return _AccumulateToLevelFn.apply(min_level, max_level, target_level, feat)
