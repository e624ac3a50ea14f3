"""The reference's photometric loss on the GPU (examples/02_synthetic.py:376-380):

    loss = 0.8 * mean |gt - color| + 0.2 * (1 - ssim(gt, color)) / 2

``ssim`` has the signature and semantics of pytorch_msssim.ssim (the
reference's dependency, unpinned in its requirements.txt): an 11-tap Gaussian
window of sigma 1.5, "valid" filtering, C1 = (0.01 R)^2, C2 = (0.03 R)^2, the
SSIM map averaged per (image, channel), then over channels and images.  The
filtering, the map and its adjoint run in the library's HIP kernels
(ds_ssim_fwd / ds_ssim_bwd) on channels-last images, which is how the
reference's images are laid out (it passes ``permute(0, 3, 1, 2)`` views of
[B, H, W, 3] tensors): such views are read in place.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_lp = ctypes.POINTER(ctypes.c_int64)

_lib.register("ds_ssim_tiles", [_i, _ip, _ip, _i, _lp], ctypes.c_int64)
_lib.register("ds_ssim_fwd", [_i, _ip, _ip, _i, _vp, _vp, _i, _fp, _f, _f, _vp, _vp, _vp])
_lib.register("ds_ssim_bwd", [_i, _ip, _ip, _i, _vp, _vp, _i, _fp, _vp, _fp, _vp, _f, _vp, _vp, _i, _vp])


def gaussian_window(size: int = 11, sigma: float = 1.5) -> np.ndarray:
    """The normalised 1-D Gaussian of pytorch_msssim._fspecial_gauss_1d,
    evaluated in float32 as it is there."""
    coords = torch.arange(size, dtype=torch.float32) - size // 2
    g = torch.exp(-(coords ** 2) / (2 * sigma ** 2))
    g = g / g.sum()
    return g.numpy().astype(np.float32)


class _Views:
    """Host-side description of B equal views: resolutions, window tiles and
    the view of every tile (for the per-view sums of the partials)."""

    def __init__(self, B, H, W, win_size, device):
        self.B, self.H, self.W = B, H, W
        self.Ha, self.Wa = _lib.int_array([H] * B), _lib.int_array([W] * B)
        base = (ctypes.c_int64 * (B + 1))()
        n = _lib.lib().ds_ssim_tiles(B, self.Ha, self.Wa, win_size, base)
        if n < 0:
            _lib.check(-1, "ds_ssim_tiles")
        self.tiles = int(n)
        counts = np.diff(np.array(base[:], np.int64))
        self.tile_view = torch.from_numpy(np.repeat(np.arange(B), counts)).to(device)
        kh = win_size if H >= win_size else 1
        kw = win_size if W >= win_size else 1
        self.windows = (H - kh + 1) * (W - kw + 1)


MAX_VIEWS = 256   # images per launch of the SSIM kernels; larger batches run in chunks


def _fwd(views, C, X, Y, win, c1, c2):
    """Per-(view, channel) mean SSIM [B, C] (float64) and the derivative maps."""
    P = views.B * views.H * views.W
    maps = torch.empty((3, C, P), dtype=torch.float32, device=X.device)
    partial = torch.empty((max(views.tiles, 1), C), dtype=torch.float32, device=X.device)
    wa = (ctypes.c_float * len(win))(*win.tolist())
    _lib.check(_lib.lib().ds_ssim_fwd(views.B, views.Ha, views.Wa, C, _lib.ptr(X), _lib.ptr(Y), len(win), wa,
                                      float(c1), float(c2), _lib.ptr(maps), _lib.ptr(partial),
                                      _lib.stream_ptr(X.device)), "ds_ssim_fwd")
    sums = torch.zeros((views.B, C), dtype=torch.float64, device=X.device)
    sums.index_add_(0, views.tile_view, partial[: views.tiles].double())
    return sums / views.windows, maps


def _bwd(views, C, X, Y, win, maps, scale, dscale, l1_scale, dl1, grad):
    wa = (ctypes.c_float * len(win))(*win.tolist())
    sa = (ctypes.c_float * (views.B * C))(*[float(s) for s in scale])
    _lib.check(_lib.lib().ds_ssim_bwd(views.B, views.Ha, views.Wa, C, _lib.ptr(X), _lib.ptr(Y), len(win), wa,
                                      _lib.ptr(maps), sa, _lib.ptr(dscale), float(l1_scale), _lib.ptr(dl1),
                                      _lib.ptr(grad), 0, _lib.stream_ptr(Y.device)), "ds_ssim_bwd")


def _fwd_batch(X, Y, win, c1, c2):
    """_fwd over [B, H, W, C] images in chunks of MAX_VIEWS: the per-(image,
    channel) means [B, C] and one (first image, views, maps) per chunk."""
    B, H, W, C = Y.shape
    means, chunks = [], []
    for b0 in range(0, B, MAX_VIEWS):
        b1 = min(B, b0 + MAX_VIEWS)
        views = _Views(b1 - b0, H, W, len(win), Y.device)
        s, maps = _fwd(views, C, X[b0:b1], Y[b0:b1], win, c1, c2)
        means.append(s)
        chunks.append((b0, views, maps))
    return torch.cat(means), chunks


def _bwd_batch(chunks, X, Y, win, scale, dscale, l1_scale, dl1):
    """The gradient w.r.t. Y of sum_{b,c} scale dscale[b, c] S[b, c] (+ the L1
    term), chunk by chunk into one [B, H, W, C] tensor."""
    C = Y.shape[3]
    grad = torch.empty_like(Y)
    for b0, views, maps in chunks:
        b1 = b0 + views.B
        _bwd(views, C, X[b0:b1], Y[b0:b1], win, maps, [scale] * (views.B * C),
             dscale[b0 * C:b1 * C].contiguous(), l1_scale, dl1, grad[b0:b1])
    return grad


class _SsimPerChannel(torch.autograd.Function):
    """[B, H, W, C] channels-last X, Y -> mean SSIM per (image, channel) [B, C]."""

    @staticmethod
    def forward(ctx, X, Y, win, c1, c2):
        s, chunks = _fwd_batch(X, Y, win, c1, c2)
        chunks_x = None
        if ctx.needs_input_grad[0]:
            _, chunks_x = _fwd_batch(Y, X, win, c1, c2)   # S is symmetric in its two images
        ctx.save_for_backward(X, Y)
        ctx.chunks, ctx.chunks_x, ctx.win = chunks, chunks_x, win
        return s.float()

    @staticmethod
    def backward(ctx, g):
        X, Y = ctx.saved_tensors
        windows = ctx.chunks[0][1].windows
        dscale = g.float().contiguous().reshape(-1)
        gx = gy = None
        if ctx.needs_input_grad[1]:
            gy = _bwd_batch(ctx.chunks, X, Y, ctx.win, 1.0 / windows, dscale, 0.0, None)
        if ctx.needs_input_grad[0]:
            gx = _bwd_batch(ctx.chunks_x, Y, X, ctx.win, 1.0 / windows, dscale, 0.0, None)
        return gx, gy, None, None, None


def _channels_last(t: torch.Tensor, name: str) -> torch.Tensor:
    """NCHW tensor -> contiguous [N, H, W, C] (free for a permuted NHWC tensor)."""
    if t.ndim != 4:
        raise ValueError(f"{name}: expected a 4-D (N, C, H, W) image batch, got shape {tuple(t.shape)}")
    _lib.require_device(t, name)
    return t.permute(0, 2, 3, 1).to(torch.float32).contiguous()


def ssim(X: torch.Tensor, Y: torch.Tensor, data_range: float = 255, size_average: bool = True,
         win_size: int = 11, win_sigma: float = 1.5, win=None, K=(0.01, 0.03), nonnegative_ssim: bool = False):
    """pytorch_msssim.ssim on the GPU: X, Y (N, C, H, W), 1 <= C <= 4,
    differentiable in both.  Returns the mean SSIM (size_average) or one
    value per image."""
    if X.shape != Y.shape:
        raise ValueError(f"Input images should have the same dimensions, but got {X.shape} and {Y.shape}.")
    if win is not None:
        w = torch.as_tensor(win).detach().to("cpu", torch.float32).reshape(-1)
        win_size = w.numel()
        wnp = w.numpy().astype(np.float32)
    else:
        if win_size % 2 != 1:
            raise ValueError("Window size should be odd.")
        wnp = gaussian_window(win_size, win_sigma)
    if not np.array_equal(wnp, wnp[::-1]):
        raise ValueError("the window must be symmetric")
    if wnp.size != 11:
        raise ValueError("the SSIM kernels take an 11-tap window (pytorch_msssim's default win_size)")
    c1 = (K[0] * data_range) ** 2
    c2 = (K[1] * data_range) ** 2
    per_channel = _SsimPerChannel.apply(_channels_last(X, "X"), _channels_last(Y, "Y"), wnp, c1, c2)
    if nonnegative_ssim:
        per_channel = torch.relu(per_channel)
    return per_channel.mean() if size_average else per_channel.mean(1)


class _L1Ssim(torch.autograd.Function):
    """w_l1 * mean |t - c| + w_ssim * (1 - ssim(t, c)) / 2 over [B, H, W, C]
    images, with one backward kernel for both terms."""

    @staticmethod
    def forward(ctx, color, target, w_l1, w_ssim, win, c1, c2):
        s, chunks = _fwd_batch(target, color, win, c1, c2)
        l1 = (color - target).abs().mean(dtype=torch.float64)
        ctx.save_for_backward(color, target)
        ctx.chunks, ctx.win, ctx.w = chunks, win, (w_l1, w_ssim)
        return (w_l1 * l1 + w_ssim * 0.5 * (1.0 - s.mean())).float()

    @staticmethod
    def backward(ctx, g):
        color, target = ctx.saved_tensors
        B, _, _, C = color.shape
        w_l1, w_ssim = ctx.w
        windows = ctx.chunks[0][1].windows
        gs = g.float().reshape(1).contiguous()
        dscale = gs.expand(B * C).contiguous()
        grad = _bwd_batch(ctx.chunks, target, color, ctx.win, -0.5 * w_ssim / (B * C * windows), dscale,
                          w_l1 / color.numel(), gs)
        return grad, None, None, None, None, None, None


def l1_ssim_loss(color: torch.Tensor, target: torch.Tensor, l1_weight: float = 0.8, ssim_weight: float = 0.2,
                 data_range: float = 1.0, win_size: int = 11, win_sigma: float = 1.5, K=(0.01, 0.03)):
    """The reference's image loss (examples/02_synthetic.py:376-380) on
    channels-last images color, target [B, H, W, C]: l1_weight * mean|target
    - color| + ssim_weight * (1 - ssim(target, color)) / 2, differentiable in
    ``color``."""
    if color.shape != target.shape or color.ndim != 4:
        raise ValueError("color and target must both be [B, H, W, C]")
    _lib.require_device(color, "color")
    c1 = (K[0] * data_range) ** 2
    c2 = (K[1] * data_range) ** 2
    return _L1Ssim.apply(color.to(torch.float32).contiguous(), target.detach().to(torch.float32).contiguous(),
                         float(l1_weight), float(ssim_weight), gaussian_window(win_size, win_sigma), c1, c2)
