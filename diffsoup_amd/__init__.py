"""diffsoup_amd: MI355X-native differentiable triangle-soup rasterizer.

Drop-in for the operator API of python/diffsoup (kenji-tojo/diffsoup): the same
names, arguments and autograd behaviour, with the arithmetic in hand-written
HIP kernels for gfx950 behind the C-ABI of libdiffsoup_hip.so
(include/diffsoup_hip.h).  There is no CPU execution path for the GPU
operators; the soup splitters are host code in the same library, as they are
in the reference.  ``losses`` holds the reference's L1 + SSIM image loss and
``export`` the on-disk formats of its checkpoints and viewers.
"""
from __future__ import annotations

__version__ = "0.1.0"

from . import export, losses, optimize
from .multires import (
    ColorMLP,
    accumulate_to_level,
    build_multires_triangle_color,
    feats_at_level,
    feats_in_levels,
    multires_feature_gather,
    multires_triangle_color,
    rasterize_multires_triangle_alpha,
    rasterize_views,
)
from .rasterize import count_triangle_ids, edge_grad, encode_view_dir_sh2, opacity_aux_loss
from .remesh import (
    expand_by_index,
    remove_unreferenced_vertices_from_soup,
    split_edges_from_training_views,
    split_triangle_soup,
    split_triangle_soup_clip,
    split_triangle_soup_clip_until,
    split_triangle_soup_until,
)
from .visibility import build_keep_map, count_visible_triangles


def project_vertices(verts, mvp):
    """(V,3) world vertices -> (B,V,4) clip space, clip = mvp @ [v, 1]
    (examples/utils.py:412-426)."""
    import torch

    vh = torch.cat([verts, torch.ones_like(verts[:, :1])], dim=-1)
    return torch.matmul(mvp.unsqueeze(1), vh.unsqueeze(0).unsqueeze(-1)).squeeze(-1).contiguous()
