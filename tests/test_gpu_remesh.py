"""The training-view splitting driver (examples/utils.py:493-545) on the GPU:
views rasterized in one batched launch, visible faces from the GPU
triangle-id count, then the host splitter per view.  Checked against the
same composition on the CPU (the oracle's deterministic raster of each view,
the visible set from numpy, the splitter run view by view) - the splitter
itself is pinned to the reference's by the CPU tests."""
import numpy as np
import pytest
import torch

from diffsoup_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_split_edges_from_training_views(dev):
    import diffsoup_amd as ds
    H, W = 120, 160
    s = scenes.sphere_torus_soup(4, n_surface=2000, n_float=500, views=5, res=(H, W))
    alpha = (1 / (1 + np.exp(-s.alpha_logits.astype(np.float64)))).astype(np.float32)
    cap, tau_ratio, gseed = 3, 0.008, 3
    V, F, fmap = ds.split_edges_from_training_views(
        (H, W), torch.from_numpy(s.mvp).to(dev), torch.from_numpy(s.verts).to(dev), torch.from_numpy(s.tri).to(dev),
        s.alpha_level, torch.from_numpy(alpha).to(dev), tau_ratio, cap, generator=torch.Generator().manual_seed(gseed))
    assert V.is_cuda and F.is_cuda and fmap.is_cuda

    perm = torch.randperm(s.B, generator=torch.Generator().manual_seed(gseed)).numpy()[:cap]
    chosen = s.mvp[perm]
    pos = O.project(s.verts, chosen)
    rv, rf = torch.from_numpy(s.verts), torch.from_numpy(s.tri)
    rmap = np.arange(s.T)
    total_valid = 0
    for i in range(cap):
        r = O.rasterize((H, W), pos[i:i + 1], s.tri, s.alpha_level, alpha, False, 0)
        ids = r[..., 3].astype(np.int64).ravel()
        seen = np.zeros(s.T, bool)
        seen[ids[ids > 0] - 1] = True
        valid = seen[rmap].astype(np.int32)
        total_valid += int(valid.sum())
        rv, rf, nxt, _ = ds.split_triangle_soup_clip_until((H, W), torch.from_numpy(chosen[i]), rv, rf,
                                                           torch.from_numpy(valid), tau_ratio)
        rmap = rmap[nxt.numpy()]
    assert total_valid > 300 and rf.shape[0] > s.T + 50   # the views saw faces and split some
    np.testing.assert_array_equal(V.cpu().numpy().view(np.uint32), rv.numpy().view(np.uint32))
    np.testing.assert_array_equal(F.cpu().numpy(), rf.numpy())
    np.testing.assert_array_equal(fmap.cpu().numpy(), rmap)
