"""Writes the golden fixtures of tests/golden from the CPU oracle.

    python tests/golden/make_golden.py        (from the repository root)

Two workloads, both small enough to commit:

* ``random_soup_1k.npz``: the 1,000-triangle random soup with one 256x256
  view (level-0 opacity and features, MLP 8->16->16->3, MSE to a seeded
  uniform target), run through the oracle's full training step with seed 3:
  clip positions, raster (stochastic and deterministic masks), colour, loss
  and every parameter gradient (MLP in float64, stored as float32 except the
  loss).
* ``tiny_scenes.npz``: hand-placed triangles in two 24x32 views, built so
  that every branch of the raster path has a case: two overlapping
  triangles at different depths, a triangle crossing the near plane
  (w <= 0 culled), one outside the frustum, one clipped by the image border,
  a sliver, and a shared edge; raster at opacity levels 0 and 2, level-sum
  and concatenated feature gathers over levels 0..2, edge and opacity
  gradients.

The fixtures hold inputs and expected outputs only.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def tiny_scene():
    """Clip-space triangles for two 24x32 views (view 1 is view 0 shifted and
    with the depth order of the overlapping pair swapped)."""
    def tri(a, b, c, w=1.0):
        return [[*a, w], [*b, w], [*c, w]]

    tris0 = [
        tri((-0.8, -0.7, 0.2), (0.6, -0.6, 0.2), (-0.1, 0.8, 0.2)),          # big, depth 0.2
        tri((-0.3, -0.9, -0.3), (0.9, 0.1, -0.3), (-0.6, 0.5, -0.3)),        # overlapping, in front
        [[0.2, 0.2, 0.5, 1.0], [0.4, 0.3, 0.5, -0.5], [0.3, 0.6, 0.5, 1.0]],  # crosses w = 0: culled
        tri((1.2, 1.1, 0.0), (1.5, 1.3, 0.0), (1.3, 1.6, 0.0)),              # outside the frustum
        tri((0.7, -1.3, 0.1), (1.4, -0.2, 0.1), (0.5, 0.1, 0.1)),            # clipped by the border
        tri((-0.9, 0.9, 0.0), (0.9, 0.88, 0.0), (-0.9, 0.93, 0.0)),          # sliver
        tri((-0.8, -0.7, 0.6), (0.6, -0.6, 0.6), (0.0, -0.95, 0.6)),         # shares an edge with #0
        tri((-0.5, -0.2, 1.5), (0.2, -0.1, 1.5), (-0.2, 0.4, 1.5)),          # beyond far (z/w > 1)
    ]
    t0 = np.asarray(tris0, np.float32)
    T = t0.shape[0]
    # homogeneous scaling by w keeps the NDC position but exercises the 1/w paths
    scale = np.linspace(0.7, 1.6, T, dtype=np.float32)[:, None, None]
    t0 = t0 * np.where(t0[..., 3:4] > 0, scale, 1.0)
    t1 = t0.copy()
    t1[..., 0] += 0.15 * t1[..., 3]
    t1[0, :, 2] = -0.4 * t1[0, :, 3]
    t1[1, :, 2] = 0.3 * t1[1, :, 3]
    pos = np.stack([t0.reshape(-1, 4), t1.reshape(-1, 4)]).astype(np.float32)
    tri_idx = np.arange(3 * T, dtype=np.int32).reshape(T, 3)
    return pos, tri_idx


def main():
    import torch

    from diffsoup_amd import scenes
    from oracle import oracle as O

    O.build()
    rng = np.random.Generator(np.random.PCG64(1234))

    # --- config-1 random soup ---------------------------------------------
    s = scenes.small_random_soup(0)
    seed = 3
    r = O.reference_step(s, seed, dtype=torch.float64)
    rast_det = O.rasterize(s.resolutions[0], r["pos"], s.tri, s.alpha_level, r["alpha"], False, 0)
    np.savez_compressed(
        os.path.join(HERE, "random_soup_1k.npz"),
        seed=np.int64(seed), verts=s.verts, tri=s.tri, alpha_logits=s.alpha_logits, feat=s.feat, mvp=s.mvp,
        target=s.target, w1=s.mlp[0][0], b1=s.mlp[0][1], w2=s.mlp[1][0], b2=s.mlp[1][1], w3=s.mlp[2][0],
        b3=s.mlp[2][1], resolution=np.asarray(s.resolutions[0], np.int32),
        pos=r["pos"], alpha=r["alpha"], rast=r["rast"].reshape(1, 256, 256, 4), rast_det=rast_det,
        covered_fragments=np.int64(O.covered_fragments(256, 256, r["pos"], s.tri)),
        color=r["color"].astype(np.float32), loss=np.float64(r["loss"]),
        grad_feat=r["grad_feat"], grad_verts=r["grad_verts"], grad_alpha_logits=r["grad_alpha_logits"],
        grad_mlp=r["grad_mlp"])

    # --- tiny hand-placed scenes ------------------------------------------
    pos, tri = tiny_scene()
    H, W = 24, 32
    T = tri.shape[0]
    out = dict(pos=pos, tri=tri, resolution=np.asarray((H, W), np.int32))
    for level in (0, 2):
        a = rng.uniform(0.05, 0.95, size=(T, O.feats_at_level(level))).astype(np.float32)
        out[f"alpha_l{level}"] = a
        out[f"rast_l{level}_det"] = O.rasterize((H, W), pos, tri, level, a, False, 0)
        out[f"rast_l{level}_seed11"] = O.rasterize((H, W), pos, tri, level, a, True, 11)
    fp, fa = O.fragments(H, W, pos, tri)
    out["frag_pix"], out["frag_attrs"] = fp, fa
    rects, counts, total = O.triangle_rects(H, W, pos, tri)
    out["rects"], out["rect_counts"] = rects, counts
    rast = out["rast_l2_seed11"]
    feat = rng.normal(0, 0.5, size=(T, O.feats_in_levels(0, 2), 3)).astype(np.float32)
    out["feat_l012"] = feat
    out["gather_sum"] = O.multires_color(rast, 0, 2, feat, False)
    out["gather_concat"] = O.multires_color(rast, 0, 2, feat, True)
    color = rng.uniform(0, 1, size=(2, H, W, 3)).astype(np.float32)
    gcol = rng.normal(0, 1, size=(2, H, W, 3)).astype(np.float32)
    target = rng.uniform(0, 1, size=(2, H, W, 3)).astype(np.float32)
    out["color"], out["grad_color"], out["target"] = color, gcol, target
    out["edge_grad_pos"] = O.edge_grad(color, gcol, rast, pos, tri)
    out["aux_grad_alpha_l2"] = O.opacity_aux_grad(color, target, rast, pos, tri, 2, out["alpha_l2"])
    np.savez_compressed(os.path.join(HERE, "tiny_scenes.npz"), **out)
    for f in ("random_soup_1k.npz", "tiny_scenes.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
