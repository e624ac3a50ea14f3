"""Drop-in for the reference's compiled ``_core`` module (src/main.cpp:23-304).

Every rasterization, gradient and feature function that the reference's
Python layer binds is here, under the same name and with the same arguments,
and each writes its outputs in place just as the reference does. Each one
calls libdiffsoup_hip.so on torch's current stream. This module allocates
the scratch that the C-ABI leaves to its caller (counts, flags, scan space,
the 64-bit z-buffer), the same scratch the reference's CUDA functions
allocate internally. As with nanobind, tensors of the wrong dtype, rank or
layout raise TypeError. A feature tensor whose size does not match the
level range raises RuntimeError("Invalid feature size.").

The two soup splitters (split_triangle_soup, split_triangle_soup_clip,
src/main.cpp:308-385) take and return NumPy arrays, as the reference's do,
and run on the host (ds_split_soup / ds_split_soup_clip).

The reference's Python layer needs one line changed to run on MI355X:
``from diffsoup_amd import _core``.
"""
from __future__ import annotations

import torch

from . import _lib
from . import plain as _plain  # registers the ds_plain_* signatures

__version__ = "0.1.0"

__all__ = [
    "compute_triangle_rects", "compute_fragments", "depth_test", "filter_valid_fragments",
    "backward_edge_grad", "backward_opacity_aux_loss", "encode_view_dir_sh2",
    "multires_triangle_alpha", "backward_multires_triangle_alpha",
    "multires_triangle_color", "backward_multires_triangle_color",
    "accumulate_to_level_forward", "accumulate_to_level_backward",
    "split_triangle_soup", "split_triangle_soup_clip",
]

_INT32_MAX = 2**31 - 1


def _arr(t, dtype, shape, name):
    """nanobind's ndarray<dtype, pytorch, shape<...>, c_contig> check; -1 is
    any extent.  Also requires a GPU tensor (the library has no CPU path)."""
    ok = (isinstance(t, torch.Tensor) and t.dtype == dtype and t.dim() == len(shape)
          and all(s == -1 or s == d for s, d in zip(shape, t.shape)) and t.is_contiguous())
    if not ok:
        desc = f"{t.dtype}, shape {tuple(t.shape)}" if isinstance(t, torch.Tensor) else type(t).__name__
        raise TypeError(f"{name}: expected a contiguous {dtype} tensor of shape {shape}, got {desc}")
    _lib.require_device(t, name)
    return t


def _feats(lo, hi):
    if lo < 0 or hi < lo:
        raise RuntimeError("Invalid feature size.")
    return sum(int(_lib.lib().ds_feats_at_level(l)) for l in range(lo, hi + 1))


def _hw(B, H, W):
    return _lib.int_array([H] * B), _lib.int_array([W] * B)


def _stream(t):
    return _lib.stream_ptr(t.device)


# --------------------------------------------------------------------------
# Rasterisation (src/main.cpp:30-108)
# --------------------------------------------------------------------------

def compute_triangle_rects(H, W, pos, tri, triangle_rects, frag_prefix_sum) -> int:
    """Conservative per-(view, triangle) rectangles (h0, h_len, w0, w_len) and
    the inclusive int32 prefix of their pixel counts; returns the total
    (src/main.cpp:30-48, src/cuda/rasterize.cu:138-166)."""
    _arr(pos, torch.float32, (-1, -1, 4), "pos")
    _arr(tri, torch.int32, (-1, 3), "tri")
    B, V, T = pos.shape[0], pos.shape[1], tri.shape[0]
    _arr(triangle_rects, torch.int32, (B * T, 4), "triangle_rects")
    _arr(frag_prefix_sum, torch.int32, (B * T,), "frag_prefix_sum")
    if B * T == 0:
        return 0
    dev = pos.device
    prefix = torch.empty(B * T, dtype=torch.int64, device=dev)
    counts = torch.empty(B * T, dtype=torch.int32, device=dev)
    n = int(_lib.lib().ds_plain_triangle_rects(int(H), int(W), B, V, _lib.ptr(pos), T, _lib.ptr(tri),
                                               _lib.ptr(triangle_rects), _lib.ptr(prefix), _lib.ptr(counts),
                                               _lib.ptr(_plain._scan_tmp(B * T, dev)), _stream(pos)))
    _lib.check(0 if n >= 0 else n, "compute_triangle_rects")
    if n > _INT32_MAX:
        raise RuntimeError(f"compute_triangle_rects: {n} fragments overflow the int32 prefix sum")
    frag_prefix_sum.copy_(prefix)
    return n


def compute_fragments(H, W, pos, tri, frag_prefix_sum, triangle_rects, frag_pix, frag_attrs) -> None:
    """Candidate fragments, one per rectangle pixel; pixels the triangle does
    not cover keep frag_pix = -1 (src/main.cpp:50-76,
    src/cuda/rasterize.cu:231-305)."""
    _arr(pos, torch.float32, (-1, -1, 4), "pos")
    _arr(tri, torch.int32, (-1, 3), "tri")
    B, V, T = pos.shape[0], pos.shape[1], tri.shape[0]
    _arr(frag_prefix_sum, torch.int32, (B * T,), "frag_prefix_sum")
    _arr(triangle_rects, torch.int32, (B * T, 4), "triangle_rects")
    _arr(frag_pix, torch.int32, (-1, 3), "frag_pix")
    _arr(frag_attrs, torch.float32, (frag_pix.shape[0], 4), "frag_attrs")
    n = frag_pix.shape[0]
    if n == 0 or B * T == 0:
        return
    prefix = frag_prefix_sum.to(torch.int64)
    _lib.check(_lib.lib().ds_plain_fragments(int(H), int(W), B, V, _lib.ptr(pos), T, _lib.ptr(tri),
                                             _lib.ptr(prefix), _lib.ptr(triangle_rects), n, _lib.ptr(frag_pix),
                                             _lib.ptr(frag_attrs), _stream(pos)), "compute_fragments")


def filter_valid_fragments(frag_pix, frag_attrs, frag_pix_out, frag_attrs_out) -> int:
    """Compacts the fragments with frag_pix[:,0] >= 0 to the front of the
    outputs and returns their count (src/main.cpp:96-108).  The output order
    is the input order (the reference's order depends on an atomic race)."""
    _arr(frag_pix, torch.int32, (-1, 3), "frag_pix")
    n = frag_pix.shape[0]
    _arr(frag_attrs, torch.float32, (n, 4), "frag_attrs")
    _arr(frag_pix_out, torch.int32, (n, 3), "frag_pix_out")
    _arr(frag_attrs_out, torch.float32, (n, 4), "frag_attrs_out")
    if n == 0:
        return 0
    dev = frag_pix.device
    flags = torch.empty(n, dtype=torch.int32, device=dev)
    nv = int(_lib.lib().ds_plain_compact(n, _lib.ptr(frag_pix), _lib.ptr(frag_attrs), _lib.ptr(frag_pix_out),
                                         _lib.ptr(frag_attrs_out), _lib.ptr(flags),
                                         _lib.ptr(_plain._scan_tmp(n, dev)), _stream(frag_pix)))
    _lib.check(0 if nv >= 0 else nv, "filter_valid_fragments")
    return nv


def depth_test(frag_pix, frag_attrs, frag_alpha, alpha_thresh, rast_out) -> None:
    """Keeps fragments with alpha >= thresh and depth in (-1, 1], resolves the
    nearest per pixel (ties to the first fragment) and writes
    (b0, b1, depth, id + 1) to covered pixels of rast_out
    (src/main.cpp:78-94, src/cuda/rasterize.cu:307-400).  Pixels without a
    fragment are written as zeros, which is what the reference's callers
    get from their zero-filled rast_out."""
    _arr(frag_pix, torch.int32, (-1, 3), "frag_pix")
    n = frag_pix.shape[0]
    _arr(frag_attrs, torch.float32, (n, 4), "frag_attrs")
    _arr(frag_alpha, torch.float32, (n,), "frag_alpha")
    _arr(alpha_thresh, torch.float32, (n,), "alpha_thresh")
    _arr(rast_out, torch.float32, (-1, -1, -1, 4), "rast_out")
    B, H, W, _ = rast_out.shape
    zbuf = torch.empty(B * H * W, dtype=torch.int64, device=rast_out.device)
    _lib.check(_lib.lib().ds_plain_depth_test(B, H, W, n, _lib.ptr(frag_pix), _lib.ptr(frag_attrs),
                                              _lib.ptr(frag_alpha), _lib.ptr(alpha_thresh), _lib.ptr(zbuf),
                                              _lib.ptr(rast_out), _stream(rast_out)), "depth_test")


# --------------------------------------------------------------------------
# Gradients (src/main.cpp:112-154)
# --------------------------------------------------------------------------

def backward_edge_grad(color, grad_color, rast, pos, grad_pos, tri) -> None:
    """Silhouette gradients w.r.t. clip positions, accumulated into grad_pos
    (src/main.cpp:112-130, src/cuda/rasterize.cu:489-921)."""
    _arr(color, torch.float32, (-1, -1, -1, -1), "color")
    B, H, W, C = color.shape
    _arr(grad_color, torch.float32, (B, H, W, C), "grad_color")
    _arr(rast, torch.float32, (B, H, W, 4), "rast")
    _arr(pos, torch.float32, (B, -1, 4), "pos")
    _arr(grad_pos, torch.float32, tuple(pos.shape), "grad_pos")
    _arr(tri, torch.int32, (-1, 3), "tri")
    Ha, Wa = _hw(B, H, W)
    _lib.check(_lib.lib().ds_edge_grad_bwd(B, Ha, Wa, C, _lib.ptr(color), _lib.ptr(grad_color), _lib.ptr(rast),
                                           pos.shape[1], _lib.ptr(pos), _lib.ptr(grad_pos), tri.shape[0],
                                           _lib.ptr(tri), _stream(color)), "backward_edge_grad")


def backward_opacity_aux_loss(color, target, rast, frag_pix, frag_attrs, frag_alpha, grad_frag_alpha) -> None:
    """Per-fragment score-function gradient of the masked loss, accumulated
    into grad_frag_alpha (src/main.cpp:134-154, src/cuda/rasterize.cu:923-992)."""
    _arr(color, torch.float32, (-1, -1, -1, -1), "color")
    B, H, W, C = color.shape
    _arr(target, torch.float32, (B, H, W, C), "target")
    _arr(rast, torch.float32, (B, H, W, 4), "rast")
    _arr(frag_pix, torch.int32, (-1, 3), "frag_pix")
    n = frag_pix.shape[0]
    _arr(frag_attrs, torch.float32, (n, 4), "frag_attrs")
    _arr(frag_alpha, torch.float32, (n,), "frag_alpha")
    _arr(grad_frag_alpha, torch.float32, (n,), "grad_frag_alpha")
    _lib.check(_lib.lib().ds_plain_opacity_aux(B, H, W, C, _lib.ptr(color), _lib.ptr(target), _lib.ptr(rast), n,
                                               _lib.ptr(frag_pix), _lib.ptr(frag_attrs), _lib.ptr(frag_alpha),
                                               _lib.ptr(grad_frag_alpha), _stream(color)),
               "backward_opacity_aux_loss")


# --------------------------------------------------------------------------
# Encoding and multires features (src/main.cpp:158-304)
# --------------------------------------------------------------------------

def encode_view_dir_sh2(rast, inv_mvp, encoding) -> None:
    """Degree-2 real SH of each covered pixel's view direction
    (src/main.cpp:158-170, src/cuda/encoding.cu:9-130)."""
    _arr(rast, torch.float32, (-1, -1, -1, 4), "rast")
    B, H, W, _ = rast.shape
    _arr(inv_mvp, torch.float32, (B, 4, 4), "inv_mvp")
    _arr(encoding, torch.float32, (B, H, W, 9), "encoding")
    Ha, Wa = _hw(B, H, W)
    _lib.check(_lib.lib().ds_sh2(B, Ha, Wa, _lib.ptr(rast), _lib.ptr(inv_mvp), _lib.ptr(encoding), _stream(rast)),
               "encode_view_dir_sh2")


def multires_triangle_alpha(frag_attrs, min_level, max_level, alpha_src, frag_alpha) -> None:
    """Per-fragment opacity summed over levels (src/main.cpp:174-191)."""
    _arr(frag_attrs, torch.float32, (-1, 4), "frag_attrs")
    _arr(alpha_src, torch.float32, (-1, -1), "alpha_src")
    _arr(frag_alpha, torch.float32, (frag_attrs.shape[0],), "frag_alpha")
    if alpha_src.shape[1] != _feats(min_level, max_level):
        raise RuntimeError("Invalid feature size.")
    _lib.check(_lib.lib().ds_plain_alpha(frag_attrs.shape[0], _lib.ptr(frag_attrs), int(min_level), int(max_level),
                                         _lib.ptr(alpha_src), _lib.ptr(frag_alpha), _stream(frag_attrs)),
               "multires_triangle_alpha")


def backward_multires_triangle_alpha(frag_attrs, min_level, max_level, grad_alpha_src, grad_frag_alpha) -> None:
    """Adjoint of multires_triangle_alpha, accumulated into grad_alpha_src
    (src/main.cpp:193-210)."""
    _arr(frag_attrs, torch.float32, (-1, 4), "frag_attrs")
    _arr(grad_alpha_src, torch.float32, (-1, -1), "grad_alpha_src")
    _arr(grad_frag_alpha, torch.float32, (frag_attrs.shape[0],), "grad_frag_alpha")
    if grad_alpha_src.shape[1] != _feats(min_level, max_level):
        raise RuntimeError("Invalid feature size.")
    _lib.check(_lib.lib().ds_plain_alpha_bwd(frag_attrs.shape[0], _lib.ptr(frag_attrs), int(min_level),
                                             int(max_level), _lib.ptr(grad_alpha_src), _lib.ptr(grad_frag_alpha),
                                             _stream(frag_attrs)), "backward_multires_triangle_alpha")


def multires_triangle_color(rast, min_level, max_level, features, out) -> None:
    """Per-pixel features summed over levels; background pixels get zeros
    (src/main.cpp:212-232)."""
    _arr(rast, torch.float32, (-1, -1, -1, 4), "rast")
    B, H, W, _ = rast.shape
    _arr(features, torch.float32, (-1, -1, -1), "features")
    T, S, C = features.shape
    _arr(out, torch.float32, (B, H, W, C), "out")
    if S != _feats(min_level, max_level):
        raise RuntimeError("Invalid feature size.")
    _lib.check(_lib.lib().ds_multires_color_fwd(B * H * W, _lib.ptr(rast), int(min_level), int(max_level), T, C,
                                                _lib.ptr(features), 0, _lib.ptr(out), _stream(rast)),
               "multires_triangle_color")


def backward_multires_triangle_color(rast, min_level, max_level, grad_features, grad_out) -> None:
    """Adjoint of multires_triangle_color, accumulated into grad_features
    (src/main.cpp:234-254)."""
    _arr(rast, torch.float32, (-1, -1, -1, 4), "rast")
    B, H, W, _ = rast.shape
    _arr(grad_features, torch.float32, (-1, -1, -1), "grad_features")
    T, S, C = grad_features.shape
    _arr(grad_out, torch.float32, (B, H, W, C), "grad_out")
    if S != _feats(min_level, max_level):
        raise RuntimeError("Invalid feature size.")
    _lib.check(_lib.lib().ds_multires_color_bwd(B * H * W, _lib.ptr(rast), int(min_level), int(max_level), T, C,
                                                _lib.ptr(grad_features), 0, _lib.ptr(grad_out), _stream(rast)),
               "backward_multires_triangle_color")


def accumulate_to_level_forward(min_level, max_level, target_level, feat_all, feat_max) -> None:
    """Level-summed features evaluated on the target lattice
    (src/main.cpp:258-280)."""
    _arr(feat_all, torch.float32, (-1, -1, -1), "feat_all")
    T, S, C = feat_all.shape
    _arr(feat_max, torch.float32, (T, -1, C), "feat_max")
    if S != _feats(min_level, max_level) or target_level < 0 or feat_max.shape[1] != _feats(target_level, target_level):
        raise RuntimeError("Invalid feature size.")
    _lib.check(_lib.lib().ds_accumulate_fwd(T, int(min_level), int(max_level), int(target_level), C,
                                            _lib.ptr(feat_all), _lib.ptr(feat_max), _stream(feat_all)),
               "accumulate_to_level_forward")


def accumulate_to_level_backward(min_level, max_level, target_level, grad_feat_all, grad_feat_max) -> None:
    """Adjoint of accumulate_to_level_forward, accumulated into grad_feat_all
    (src/main.cpp:282-304)."""
    _arr(grad_feat_all, torch.float32, (-1, -1, -1), "grad_feat_all")
    T, S, C = grad_feat_all.shape
    _arr(grad_feat_max, torch.float32, (T, -1, C), "grad_feat_max")
    if (S != _feats(min_level, max_level) or target_level < 0
            or grad_feat_max.shape[1] != _feats(target_level, target_level)):
        raise RuntimeError("Invalid feature size.")
    _lib.check(_lib.lib().ds_accumulate_bwd(T, int(min_level), int(max_level), int(target_level), C,
                                            _lib.ptr(grad_feat_all), _lib.ptr(grad_feat_max), _stream(grad_feat_all)),
               "accumulate_to_level_backward")


def _np_arr(a, dtype, shape, name):
    """nanobind's ndarray<dtype, numpy, shape<...>, c_contig> check."""
    import numpy as np
    if not isinstance(a, np.ndarray) or a.dtype != dtype or a.ndim != len(shape) or not a.flags.c_contiguous or any(
            s != -1 and s != d for s, d in zip(shape, a.shape)):
        raise TypeError(f"{name}: expected a C-contiguous {np.dtype(dtype).name} NumPy array of shape "
                        f"{tuple('*' if s == -1 else s for s in shape)}")
    return a


def split_triangle_soup(verts, faces, numSplits, tau):
    """Longest-edge bisection in world space (src/main.cpp:308-345): NumPy in,
    (verts (N',3) f32, faces (M',3) i32, face_mapping (M',), face_flags (M',))
    NumPy out."""
    import numpy as np
    from .remesh import split_triangle_soup as _split
    v = _np_arr(verts, np.float32, (-1, 3), "verts")
    f = _np_arr(faces, np.int32, (-1, 3), "faces")
    out = _split(torch.from_numpy(v), torch.from_numpy(f), int(numSplits), float(tau))
    return tuple(t.numpy() for t in out)


def split_triangle_soup_clip(mvp, verts, faces, valid_faces, numSplits, tau_ratio, aspectWH):
    """Longest-edge bisection by NDC length through one camera
    (src/main.cpp:347-385), NumPy in and out."""
    import ctypes

    import numpy as np
    from . import remesh as _remesh
    m = _np_arr(mvp, np.float32, (4, 4), "mvp")
    v = _np_arr(verts, np.float32, (-1, 3), "verts")
    f = _np_arr(faces, np.int32, (-1, 3), "faces")
    ok = _np_arr(valid_faces, np.int32, (-1,), "valid_faces")
    if ok.shape[0] != f.shape[0]:
        raise TypeError("valid_faces: one entry per face expected")
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().ds_split_soup_clip(_remesh._ptr(m), _remesh._ptr(v), v.shape[0], _remesh._ptr(f),
                                             f.shape[0], _remesh._ptr(ok), int(numSplits), float(tau_ratio),
                                             float(aspectWH), ctypes.byref(h)), "split_triangle_soup_clip")
    return tuple(t.numpy() for t in _remesh._collect(h, torch.device("cpu")))
