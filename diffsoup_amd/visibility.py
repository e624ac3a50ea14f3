"""Visibility counting and pruning (examples/utils.py:445-490).

count_visible_triangles rasterizes the training views with the deterministic
mask (threshold 0.5), batched and of any mix of resolutions, and sums every
view's per-triangle pixel count on the GPU (ds_count_triangle_ids);
build_keep_map keeps all but the ``remove`` least visible triangles, in index
order, with the reference's stable tie-break.
"""
from __future__ import annotations

import torch

from .multires import rasterize_views
from .rasterize import count_triangle_ids


def project_clip(verts: torch.Tensor, mvps: torch.Tensor) -> torch.Tensor:
    """clip[b, v] = mvps[b] @ [verts[v], 1] (examples/utils.py:412-426), summed
    in the fused step's fixed order so both see the same clip positions."""
    from .plain import project_ordered

    return torch.stack([project_ordered(verts, m) for m in mvps]).contiguous()


def count_visible_triangles(resolutions, mvps: torch.Tensor, verts: torch.Tensor, tri: torch.Tensor, level: int,
                            alpha_src: torch.Tensor, batch_size: int = 16, tile: int = 16) -> torch.Tensor:
    """Pixels per triangle summed over all views (int64 [T]).  ``resolutions``
    is one (H, W) for every view or a list with one per view."""
    n = mvps.shape[0]
    if isinstance(resolutions, tuple) and len(resolutions) == 2 and isinstance(resolutions[0], int):
        resolutions = [resolutions] * n
    T = tri.shape[0]
    count = torch.zeros(T, dtype=torch.int64, device=verts.device)
    a2 = alpha_src.reshape(T, -1).contiguous()
    for start in range(0, n, batch_size):
        end = min(n, start + batch_size)
        pos = project_clip(verts, mvps[start:end])
        rast = rasterize_views(resolutions[start:end], pos, tri, level, a2, stochastic=False, tile=tile)
        count_triangle_ids(rast, T, out=count)
    return count


def build_keep_map(counts: torch.Tensor, remove: int) -> torch.Tensor:
    """Indices of the triangles kept after removing the ``remove`` least
    visible ones (stable order on ties), sorted ascending."""
    order = torch.argsort(counts, stable=True)
    keep, _ = torch.sort(order[remove:])
    return keep
