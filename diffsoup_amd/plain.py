"""The plain per-view FP32 path: the reference's own composition, run on the
HIP fragment-list kernels of plain.hip with a torch MLP and autograd.

This is the comparator the fused step is measured against and a second,
structurally different GPU implementation of the same training step:

* rasterization follows python/diffsoup/rasterize.py:40-126 and
  python/diffsoup/multires.py:52-102 (rects -> fragments -> compaction ->
  opacity -> thresholds -> global-atomic depth test), with host
  synchronisations where the reference has them;
* the opacity auxiliary gradient recomputes the fragments as
  python/diffsoup/rasterize.py:148-196 does;
* the pixel MLP is ColorMLP's masked torch evaluation
  (python/diffsoup/multires.py:185-207), one view per launch as in the
  reference's per-batch loop.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .multires import feats_at_level, multires_feature_gather
from .rasterize import edge_grad

_vp, _i, _i64, _u64, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_size_t
_lib.register("ds_plain_scan_tmp_bytes", [_i64], _sz)
_lib.register("ds_plain_triangle_rects", [_i, _i, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i64)
_lib.register("ds_plain_fragments", [_i, _i, _i, _i, _vp, _i, _vp, _vp, _vp, _i64, _vp, _vp, _vp])
_lib.register("ds_plain_compact", [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i64)
_lib.register("ds_plain_alpha", [_i64, _vp, _i, _i, _vp, _vp, _vp])
_lib.register("ds_plain_alpha_bwd", [_i64, _vp, _i, _i, _vp, _vp, _vp])
_lib.register("ds_plain_thresholds", [_i64, _vp, _vp, _i, _u64, _i, _vp, _vp])
_lib.register("ds_plain_depth_test", [_i, _i, _i, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp])
_lib.register("ds_plain_opacity_aux", [_i, _i, _i, _i, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp])


def _scan_tmp(n, dev):
    nbytes = int(_lib.lib().ds_plain_scan_tmp_bytes(max(int(n), 1)))
    return torch.empty(nbytes, dtype=torch.uint8, device=dev)


def compute_fragments(resolution, pos, tri):
    """Valid fragments (frag_pix [N,3], frag_attrs [N,4]) in (view, triangle,
    rectangle position) order (python/diffsoup/rasterize.py:40-76)."""
    H, W = resolution
    B, V, _ = pos.shape
    T = tri.shape[0]
    dev = pos.device
    L = _lib.lib()
    s = _lib.stream_ptr(dev)
    rects = torch.empty(B * T, 4, dtype=torch.int32, device=dev)
    prefix = torch.empty(B * T, dtype=torch.int64, device=dev)
    counts = torch.empty(B * T, dtype=torch.int32, device=dev)
    n = int(L.ds_plain_triangle_rects(H, W, B, V, _lib.ptr(pos), T, _lib.ptr(tri), _lib.ptr(rects),
                                      _lib.ptr(prefix), _lib.ptr(counts), _lib.ptr(_scan_tmp(B * T, dev)), s))
    _lib.check(0 if n >= 0 else n, "ds_plain_triangle_rects")
    frag_pix = torch.empty(n, 3, dtype=torch.int32, device=dev)
    frag_attrs = torch.empty(n, 4, dtype=torch.float32, device=dev)
    _lib.check(L.ds_plain_fragments(H, W, B, V, _lib.ptr(pos), T, _lib.ptr(tri), _lib.ptr(prefix), _lib.ptr(rects),
                                    n, _lib.ptr(frag_pix), _lib.ptr(frag_attrs), s), "ds_plain_fragments")
    pix_out = torch.empty_like(frag_pix)
    attrs_out = torch.empty_like(frag_attrs)
    flags = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    nv = int(L.ds_plain_compact(n, _lib.ptr(frag_pix), _lib.ptr(frag_attrs), _lib.ptr(pix_out), _lib.ptr(attrs_out),
                                _lib.ptr(flags), _lib.ptr(_scan_tmp(n, dev)), s))
    _lib.check(0 if nv >= 0 else nv, "ds_plain_compact")
    return pix_out[:nv].contiguous(), attrs_out[:nv].contiguous()


def _fragment_alpha(frag_attrs, level, alpha_src):
    n = frag_attrs.shape[0]
    out = torch.empty(n, dtype=torch.float32, device=frag_attrs.device)
    _lib.check(_lib.lib().ds_plain_alpha(n, _lib.ptr(frag_attrs), level, level, _lib.ptr(alpha_src), _lib.ptr(out),
                                         _lib.stream_ptr(frag_attrs.device)), "ds_plain_alpha")
    return out


def rasterize(resolution, pos, tri, level, alpha_src, stochastic=True, seed=0, view_base=0):
    """(B,H,W,4) raster through the fragment-list path.  ``view_base`` offsets
    the view index fed to the mask hash (a one-view call for view b of a
    batch passes b, so it draws the same mask as the batched call)."""
    H, W = resolution
    B = pos.shape[0]
    dev = pos.device
    a2 = alpha_src.reshape(tri.shape[0], feats_at_level(level)).contiguous()
    fp, fa = compute_fragments(resolution, pos, tri)
    n = fp.shape[0]
    alpha = _fragment_alpha(fa, level, a2)
    thr = torch.empty(n, dtype=torch.float32, device=dev)
    s = _lib.stream_ptr(dev)
    L = _lib.lib()
    _lib.check(L.ds_plain_thresholds(n, _lib.ptr(fp), _lib.ptr(fa), int(bool(stochastic)), int(seed) & (2**64 - 1),
                                     view_base, _lib.ptr(thr), s), "ds_plain_thresholds")
    zbuf = torch.empty(B * H * W, dtype=torch.int64, device=dev)
    rast = torch.empty(B, H, W, 4, dtype=torch.float32, device=dev)
    _lib.check(L.ds_plain_depth_test(B, H, W, n, _lib.ptr(fp), _lib.ptr(fa), _lib.ptr(alpha), _lib.ptr(thr),
                                     _lib.ptr(zbuf), _lib.ptr(rast), s), "ds_plain_depth_test")
    return rast


class _PlainOpacityAuxFn(torch.autograd.Function):
    """python/diffsoup/rasterize.py:139-196 on the fragment-list kernels."""

    @staticmethod
    def forward(ctx, color, target, rast, pos, tri, level, alpha_src):
        B, H, W, C = color.shape
        dev = color.device
        s = _lib.stream_ptr(dev)
        L = _lib.lib()
        fp, fa = compute_fragments((H, W), pos, tri)
        n = fp.shape[0]
        a2 = alpha_src.reshape(tri.shape[0], -1).contiguous()
        alpha = _fragment_alpha(fa, level, a2)
        g = torch.zeros(n, dtype=torch.float32, device=dev)
        _lib.check(L.ds_plain_opacity_aux(B, H, W, C, _lib.ptr(color), _lib.ptr(target), _lib.ptr(rast), n,
                                          _lib.ptr(fp), _lib.ptr(fa), _lib.ptr(alpha), _lib.ptr(g), s),
                   "ds_plain_opacity_aux")
        ga = torch.zeros_like(a2)
        _lib.check(L.ds_plain_alpha_bwd(n, _lib.ptr(fa), level, level, _lib.ptr(ga), _lib.ptr(g), s), "ds_plain_alpha_bwd")
        ctx.save_for_backward(ga)
        ctx.weight = 1.0 / color.numel()
        ctx.shape = alpha_src.shape
        return torch.zeros(1, dtype=torch.float32, device=dev)

    @staticmethod
    def backward(ctx, grad_loss):
        (ga,) = ctx.saved_tensors
        return None, None, None, None, None, None, (ctx.weight * grad_loss * ga).view(ctx.shape)


def project_ordered(verts, m):
    """clip = m @ [v, 1] for one row-major 4x4 m, summed as
    ((m0 x + m1 y) + m2 z) + m3 with every product and sum rounded to fp32 -
    the same order as the fused step's projection kernel, so both paths see
    bit-identical clip positions.  Differentiable in verts."""
    x, y, z = verts[:, 0], verts[:, 1], verts[:, 2]
    cols = [m[i, 0] * x + m[i, 1] * y + m[i, 2] * z + m[i, 3] for i in range(4)]
    return torch.stack(cols, dim=-1)


class PlainStep:
    """The training step of the fused path, composed per view from the
    reference's building blocks; gradients land in ``.grad`` of the leaves."""

    def __init__(self, soup_params, mlp_layers, levels, alpha_level, background=1.0):
        self.p = soup_params
        self.layers = mlp_layers          # list of torch.nn.Linear (3)
        self.lo, self.hi = levels
        self.alpha_level = alpha_level
        self.background = background

    def _mlp(self, x):
        h = x
        for i, lin in enumerate(self.layers):
            h = lin(h)
            h = torch.relu(h) if i < len(self.layers) - 1 else torch.sigmoid(h)
        return h

    def run(self, mvp, target, resolutions, seed):
        """One forward + backward over all views, one view per launch.
        target is packed [P,3]; returns the (detached) mean-squared loss."""
        p = self.p
        tri = p.tri
        alpha = torch.sigmoid(p.alpha_logits)
        P = sum(h * w for h, w in resolutions)
        total = torch.zeros((), device=p.verts.device)
        off = 0
        for b, (H, W) in enumerate(resolutions):
            pos = project_ordered(p.verts, mvp[b]).unsqueeze(0).contiguous()
            rast = rasterize((H, W), pos.detach(), tri, self.alpha_level, alpha.detach(), True, seed, view_base=b)
            x = multires_feature_gather(rast, self.lo, self.hi, p.feat, "concat")
            mask = rast[..., 3] > 0
            xf = x.view(-1, x.shape[-1])
            mf = mask.view(-1)
            out = torch.zeros(H * W, 3, device=x.device)
            if mf.any():
                xv = xf[mf]
                y = self._mlp(xv)
                res = xv[:, 3:4]
                out = out.index_put((mf.nonzero().squeeze(1),), (1.0 - res) * xv[:, :3] + res * y)
            m3 = mask.view(-1, 1).float()
            color = (m3 * out + (1.0 - m3) * self.background).view(1, H, W, 3)
            tgt = target[off:off + H * W].view(1, H, W, 3)
            aux = _PlainOpacityAuxFn.apply(color.detach().contiguous(), tgt.contiguous(), rast, pos.detach(), tri,
                                           self.alpha_level, alpha)
            # the aux weight is 1/numel of this view's colour; the batched loss averages over all views
            aux = aux * (H * W / P)
            color = edge_grad(color.contiguous(), rast, pos, tri)
            loss = ((color - tgt) ** 2).sum() / (3 * P)
            total = total + loss + aux
            off += H * W
        total.backward()
        return total.detach()
