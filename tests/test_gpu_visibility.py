"""Visibility counting (examples/utils.py:445-490) on the native counter:
count_triangle_ids against numpy's bincount of the oracle's raster (exact),
count_visible_triangles over batched views of mixed resolution against the
oracle's deterministic rasters view by view, and build_keep_map against a
stable numpy argsort."""
import numpy as np
import pytest
import torch

from diffsoup_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_count_triangle_ids_matches_bincount(dev):
    import diffsoup_amd as ds
    s = scenes.sphere_torus_soup(4, n_surface=4000, n_float=1000, views=3, res=(200, 260))
    pos = O.project(s.verts, s.mvp)
    alpha = (1 / (1 + np.exp(-s.alpha_logits.astype(np.float64)))).astype(np.float32)
    rast = O.rasterize(s.resolutions[0], pos, s.tri, s.alpha_level, alpha, True, 9)
    ids = rast[..., 3].astype(np.int64).ravel()
    ref = np.bincount(ids[ids > 0] - 1, minlength=s.T)
    got = ds.count_triangle_ids(torch.from_numpy(rast).to(dev), s.T).cpu().numpy()
    np.testing.assert_array_equal(got, ref)
    acc = torch.full((s.T,), 5, dtype=torch.int64, device=dev)
    ds.count_triangle_ids(torch.from_numpy(rast).to(dev), s.T, out=acc)
    np.testing.assert_array_equal(acc.cpu().numpy(), ref + 5)
    empty = ds.count_triangle_ids(torch.zeros(0, 4, device=dev), s.T).cpu().numpy()
    assert (empty == 0).all()


def test_count_visible_triangles_and_keep_map(dev):
    import diffsoup_amd as ds
    s = scenes.batched_uneven_soup(5, T=4000, views=7, res_scale=0.2, edge_scale=5.0)
    verts = torch.from_numpy(s.verts).to(dev)
    tri = torch.from_numpy(s.tri).to(dev)
    mvps = torch.from_numpy(s.mvp).to(dev)
    alpha = (1 / (1 + np.exp(-s.alpha_logits.astype(np.float64)))).astype(np.float32)
    got = ds.count_visible_triangles(s.resolutions, mvps, verts, tri, s.alpha_level, torch.from_numpy(alpha).to(dev),
                                     batch_size=3).cpu().numpy()
    pos = O.project(s.verts, s.mvp)
    ref = np.zeros(s.T, np.int64)
    for b, (h, w) in enumerate(s.resolutions):
        r = O.rasterize((h, w), pos[b:b + 1], s.tri, s.alpha_level, alpha, False, 0)
        ids = r[..., 3].astype(np.int64).ravel()
        ref += np.bincount(ids[ids > 0] - 1, minlength=s.T)
    np.testing.assert_array_equal(got, ref)
    assert ref.sum() > 1000
    keep = ds.build_keep_map(torch.from_numpy(got).to(dev), 1500).cpu().numpy()
    np.testing.assert_array_equal(keep, np.sort(np.argsort(ref, kind="stable")[1500:]))
