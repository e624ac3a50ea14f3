"""Soup splitting: longest-edge bisection between training phases
(python/diffsoup/remesh.py, examples/utils.py:493-545).

The bisection itself is a sequential priority-queue walk and runs on the
host in libdiffsoup_hip.so (``ds_split_soup`` / ``ds_split_soup_clip``), with
outputs bit-identical to the reference's splitters and in the same order.
What surrounds it on the GPU is the library's: the training-view driver
finds each view's visible triangles with the batched rasterizer and the
triangle-id count kernel instead of a per-view ``unique``.

Every function takes and returns torch tensors on the input's device, with
the reference's names, arguments and four-tuple results.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_pp = ctypes.POINTER(ctypes.c_void_p)

_lib.register("ds_split_soup", [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, _pp])
_lib.register("ds_split_soup_clip", [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_float, _pp])
_lib.register("ds_split_soup_counts", [_vp, _ip, _ip])
_lib.register("ds_split_soup_export", [_vp, _vp, _vp, _vp, _vp])
_lib.register("ds_split_soup_free", [_vp], None)


def _host(t, dtype, shape_tail, name):
    if not torch.is_tensor(t):
        t = torch.as_tensor(t)
    a = t.detach().to("cpu", dtype).contiguous().numpy()
    if a.ndim != 1 + len(shape_tail) or tuple(a.shape[1:]) != tuple(shape_tail):
        raise AssertionError(f"{name} must be (N, {', '.join(map(str, shape_tail))})" if shape_tail
                             else f"{name} must be 1-D")
    return a


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _collect(handle, dev):
    L = _lib.lib()
    try:
        nv, nt = ctypes.c_int(), ctypes.c_int()
        _lib.check(L.ds_split_soup_counts(handle, ctypes.byref(nv), ctypes.byref(nt)), "ds_split_soup_counts")
        v = np.empty((nv.value, 3), np.float32)
        f = np.empty((nt.value, 3), np.int32)
        fmap = np.empty(nt.value, np.int32)
        flag = np.empty(nt.value, np.int32)
        _lib.check(L.ds_split_soup_export(handle, _ptr(v), _ptr(f), _ptr(fmap), _ptr(flag)), "ds_split_soup_export")
    finally:
        L.ds_split_soup_free(handle)
    return (torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev),
            torch.from_numpy(fmap).to(dev), torch.from_numpy(flag).to(dev))


def split_triangle_soup(verts: torch.Tensor, faces: torch.Tensor, num_splits: int, tau: float = 0.0):
    """Bisect the longest world-space edge up to ``num_splits`` times (-1: no
    cap, with ``tau`` > 0), stopping once the longest edge is below ``tau``
    (0: no threshold).  Returns (verts (N',3) f32, faces (M',3) i32,
    face_mapping (M',) i32 source face, face_flags (M',) i32 1 = unchanged)
    (python/diffsoup/remesh.py:9-51)."""
    assert verts.ndim == 2 and verts.shape[1] == 3, "verts must be (N, 3)"
    assert faces.ndim == 2 and faces.shape[1] == 3, "faces must be (M, 3)"
    v = _host(verts, torch.float32, (3,), "verts")
    f = _host(faces, torch.int32, (3,), "faces")
    L = _lib.lib()
    h = ctypes.c_void_p()
    _lib.check(L.ds_split_soup(_ptr(v), v.shape[0], _ptr(f), f.shape[0], int(num_splits), float(tau),
                               ctypes.byref(h)), "ds_split_soup")
    return _collect(h, verts.device)


def split_triangle_soup_until(verts: torch.Tensor, faces: torch.Tensor, tau: float, hard_cap: int | None = None):
    """Split until every edge is below ``tau`` (python/diffsoup/remesh.py:54-75)."""
    ns = -1 if hard_cap is None else int(hard_cap)
    return split_triangle_soup(verts, faces, ns, tau=float(tau))


def split_triangle_soup_clip(resolution, mvp: torch.Tensor, verts: torch.Tensor, faces: torch.Tensor,
                             valid_faces: torch.Tensor, num_splits: int, tau_ratio: float = 0.0):
    """Bisect the longest edges as seen through one camera: lengths in NDC
    with x scaled by W/H (unit = one image height), edges with an end outside
    the NDC cube skipped, only ``valid_faces`` considered at the start, split
    points perspective-correct in world space (python/diffsoup/remesh.py:78-143)."""
    H, W = resolution
    assert tuple(mvp.shape) == (4, 4)
    assert verts.ndim == 2 and verts.shape[1] == 3
    assert faces.ndim == 2 and faces.shape[1] == 3
    assert tuple(valid_faces.shape) == (faces.shape[0],)
    m = _host(mvp, torch.float32, (4,), "mvp")
    v = _host(verts, torch.float32, (3,), "verts")
    f = _host(faces, torch.int32, (3,), "faces")
    ok = _host(valid_faces, torch.int32, (), "valid_faces")
    L = _lib.lib()
    h = ctypes.c_void_p()
    _lib.check(L.ds_split_soup_clip(_ptr(m), _ptr(v), v.shape[0], _ptr(f), f.shape[0], _ptr(ok), int(num_splits),
                                    float(tau_ratio), float(W) / float(H), ctypes.byref(h)), "ds_split_soup_clip")
    return _collect(h, verts.device)


def split_triangle_soup_clip_until(resolution, mvp, verts, faces, valid_faces, tau_ratio: float,
                                   hard_cap: int | None = None):
    """Split until every image-space edge is below ``tau_ratio`` image heights
    (python/diffsoup/remesh.py:146-175)."""
    ns = -1 if hard_cap is None else int(hard_cap)
    return split_triangle_soup_clip(resolution, mvp, verts, faces, valid_faces, ns, tau_ratio)


def expand_by_index(source: torch.Tensor, index_map: torch.Tensor) -> torch.Tensor:
    """result[j] = source[index_map[j]], for carrying per-face parameters
    through a split or a prune (python/diffsoup/remesh.py:178-197)."""
    if not torch.is_tensor(source):
        raise TypeError("`source` must be a torch.Tensor.")
    if not torch.is_tensor(index_map):
        raise TypeError("`index_map` must be a torch.Tensor.")
    index_map = index_map.to(device=source.device, dtype=torch.long)
    n = source.size(0)
    if bool(((index_map < 0) | (index_map >= n)).any()):
        raise ValueError("`index_map` contains out-of-range indices.")
    return source.index_select(0, index_map)


def remove_unreferenced_vertices_from_soup(verts: torch.Tensor, faces: torch.Tensor):
    """Drop vertices no face uses; faces are renumbered, keeping vertex order
    (python/diffsoup/point3d.py:115-143)."""
    assert verts.ndim == 2 and verts.shape[1] == 3
    assert faces.ndim == 2 and faces.shape[1] == 3
    used = torch.unique(faces.contiguous())
    remap = torch.full((verts.shape[0],), -1, dtype=torch.long, device=faces.device)
    remap[used] = torch.arange(used.numel(), device=faces.device, dtype=torch.long)
    return verts.contiguous()[used], remap[faces.long()].int()


def split_edges_from_training_views(resolution, mvps: torch.Tensor, verts: torch.Tensor, faces: torch.Tensor,
                                    level: int, alpha_acc: torch.Tensor, tau_ratio: float, num_views_cap: int,
                                    generator: torch.Generator | None = None):
    """Split long screen-space edges seen from a random subset of up to
    ``num_views_cap`` training views (examples/utils.py:493-545).

    All chosen views are rasterized in one batched launch with the
    deterministic mask; for each view in turn, the triangles it sees (by the
    GPU triangle-id count, on the faces of the input soup) select which of
    the current faces are split in that view's clip space.  Returns (verts,
    faces, face_map) with face_map giving each output face's source face in
    the input ``faces``.
    """
    from .multires import rasterize_views
    from .rasterize import count_triangle_ids
    from .visibility import project_clip

    H, W = resolution
    dev = verts.device
    n_src = faces.shape[0]
    perm = torch.randperm(mvps.shape[0], generator=generator, device=generator.device if generator is not None else dev)
    chosen = mvps[perm[: min(mvps.shape[0], num_views_cap)].to(mvps.device)]
    pos = project_clip(verts, chosen)
    rast = rasterize_views([(H, W)] * chosen.shape[0], pos, faces.int().contiguous(), level,
                           alpha_acc.reshape(n_src, -1).contiguous(), stochastic=False)
    fmap = torch.arange(n_src, device=dev, dtype=torch.long)
    per_view = rast.reshape(chosen.shape[0], H * W, 4)
    for i in range(chosen.shape[0]):
        seen = count_triangle_ids(per_view[i].contiguous(), n_src) > 0
        valid = seen[fmap].to(torch.int32).contiguous()
        verts, faces, fnext, _ = split_triangle_soup_clip_until((H, W), chosen[i], verts, faces, valid,
                                                                tau_ratio=tau_ratio)
        fmap = fmap[fnext.long()].contiguous()
    return verts, faces, fmap
