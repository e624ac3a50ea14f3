"""The drop-in ``_core`` module on the GPU: the reference's own composition
(python/diffsoup/rasterize.py:40-126, python/diffsoup/multires.py:52-102,
:139-196) driven through the reference's function names and argument lists,
checked against the oracle."""
import numpy as np
import pytest
import torch

from diffsoup_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _sig(x):
    return (1.0 / (1.0 + np.exp(-x.astype(np.float64)))).astype(np.float32)


def _fragments(core, H, W, pos, tri):
    """python/diffsoup/rasterize.py:40-76 verbatim in structure."""
    B, V, _ = pos.shape
    T = tri.shape[0]
    rects = torch.empty((B * T, 4), dtype=torch.int32, device=pos.device)
    prefix = torch.empty(B * T, dtype=torch.int32, device=pos.device)
    n = core.compute_triangle_rects(H, W, pos, tri, rects, prefix)
    fp = torch.full((n, 3), -1, dtype=torch.int32, device=pos.device)
    fa = torch.empty((n, 4), dtype=torch.float32, device=pos.device)
    core.compute_fragments(H, W, pos, tri, prefix, rects, fp, fa)
    fpo, fao = torch.empty_like(fp), torch.empty_like(fa)
    nv = core.filter_valid_fragments(fp, fa, fpo, fao)
    return fpo[:nv].contiguous(), fao[:nv].contiguous(), rects, prefix, n


@pytest.fixture(scope="module")
def scene(dev):
    s = scenes.sphere_torus_soup(5, n_surface=4000, n_float=1200, views=2, res=(240, 304))
    pos = scenes.project_vertices_np(s.verts, s.mvp)
    return s, pos, _sig(s.alpha_logits)


def test_core_raster_composition_bit_exact(dev, scene):
    from diffsoup_amd import _core as core
    s, pos, alpha = scene
    H, W = s.resolutions[0]
    pos_d = torch.from_numpy(pos).to(dev)
    tri_d = torch.from_numpy(s.tri).to(dev)
    fp, fa, rects, prefix, n = _fragments(core, H, W, pos_d, tri_d)
    r_rects, r_counts, r_total = O.triangle_rects(H, W, pos, s.tri)
    assert n == r_total
    np.testing.assert_array_equal(rects.cpu().numpy(), r_rects)
    np.testing.assert_array_equal(prefix.cpu().numpy(), np.cumsum(r_counts))
    rfp, rfa = O.fragments(H, W, pos, s.tri)
    np.testing.assert_array_equal(fp.cpu().numpy(), rfp)
    np.testing.assert_array_equal(fa.cpu().numpy().view(np.uint32), rfa.view(np.uint32))

    frag_alpha = torch.zeros(fp.shape[0], dtype=torch.float32, device=dev)
    a2 = torch.from_numpy(alpha.reshape(s.T, -1)).to(dev)
    core.multires_triangle_alpha(fa, s.alpha_level, s.alpha_level, a2, frag_alpha)
    ralpha = O.multires_alpha(rfa, s.alpha_level, alpha)
    np.testing.assert_array_equal(frag_alpha.cpu().numpy().view(np.uint32), ralpha.view(np.uint32))

    for stochastic in (False, True):
        thr = O.thresholds(rfp, rfa, stochastic, 99)
        rast = torch.zeros(s.B, H, W, 4, dtype=torch.float32, device=dev)
        core.depth_test(fp, fa, frag_alpha, torch.from_numpy(thr).to(dev), rast)
        ref = O.rasterize((H, W), pos, s.tri, s.alpha_level, alpha, stochastic, 99)
        np.testing.assert_array_equal(rast.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_core_alpha_level_range(dev, scene):
    """min_level < max_level sums the levels; checked against the oracle's
    level-summed gather with one channel, and its adjoint against the oracle's
    scatter."""
    from diffsoup_amd import _core as core
    s, pos, _ = scene
    H, W = s.resolutions[0]
    fp, fa, *_ = _fragments(core, H, W, torch.from_numpy(pos).to(dev), torch.from_numpy(s.tri).to(dev))
    lo, hi = 1, 3
    S = O.feats_in_levels(lo, hi)
    rng = np.random.default_rng(3)
    src = rng.uniform(size=(s.T, S)).astype(np.float32)
    out = torch.zeros(fp.shape[0], dtype=torch.float32, device=dev)
    core.multires_triangle_alpha(fa, lo, hi, torch.from_numpy(src).to(dev), out)
    fa_np = fa.cpu().numpy()
    ref = O.multires_color(fa_np.reshape(1, 1, -1, 4), lo, hi, src.reshape(s.T, S, 1)).reshape(-1)
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    g = rng.normal(size=fp.shape[0]).astype(np.float32)
    gsrc = torch.zeros(s.T, S, dtype=torch.float32, device=dev)
    core.backward_multires_triangle_alpha(fa, lo, hi, gsrc, torch.from_numpy(g).to(dev))
    rg = O.multires_color_bwd(fa_np.reshape(1, 1, -1, 4), lo, hi, (s.T, S, 1), g.reshape(1, 1, -1, 1))
    np.testing.assert_allclose(gsrc.cpu().numpy(), rg.reshape(s.T, S), rtol=1e-4, atol=1e-5)
    with pytest.raises(RuntimeError, match="Invalid feature size"):
        core.multires_triangle_alpha(fa, lo, hi, torch.from_numpy(src[:, :-1].copy()).to(dev), out)


def test_core_gradients_and_features(dev, scene):
    from diffsoup_amd import _core as core
    s, pos, alpha = scene
    H, W = s.resolutions[0]
    B = s.B
    pos_d = torch.from_numpy(pos).to(dev)
    tri_d = torch.from_numpy(s.tri).to(dev)
    rast_np = O.rasterize((H, W), pos, s.tri, s.alpha_level, alpha, True, 5)
    rast = torch.from_numpy(rast_np).to(dev)
    lo, hi = s.levels
    rng = np.random.default_rng(4)

    # multires colour (level sum) and its adjoint
    feat = torch.from_numpy(s.feat).to(dev)
    col = torch.zeros(B, H, W, s.feat.shape[2], dtype=torch.float32, device=dev)
    core.multires_triangle_color(rast, lo, hi, feat, col)
    np.testing.assert_array_equal(col.cpu().numpy().view(np.uint32),
                                  O.multires_color(rast_np, lo, hi, s.feat).view(np.uint32))
    g = rng.normal(size=col.shape).astype(np.float32)
    gf = torch.zeros_like(feat)
    core.backward_multires_triangle_color(rast, lo, hi, gf, torch.from_numpy(g).to(dev))
    np.testing.assert_allclose(gf.cpu().numpy(), O.multires_color_bwd(rast_np, lo, hi, s.feat.shape, g),
                               rtol=1e-4, atol=1e-4)

    # accumulate to a level
    fm = torch.empty(s.T, O.feats_at_level(2), s.feat.shape[2], dtype=torch.float32, device=dev)
    core.accumulate_to_level_forward(lo, hi, 2, feat, fm)
    ref_fm = O.accumulate_fwd(lo, hi, 2, s.feat)
    np.testing.assert_array_equal(fm.cpu().numpy().view(np.uint32), ref_fm.view(np.uint32))
    y = rng.normal(size=ref_fm.shape).astype(np.float32)
    gacc = torch.zeros_like(feat)
    core.accumulate_to_level_backward(lo, hi, 2, gacc, torch.from_numpy(y).to(dev))
    np.testing.assert_allclose(gacc.cpu().numpy(), O.accumulate_bwd(lo, hi, 2, s.feat.shape, y), rtol=1e-4, atol=1e-4)

    # SH2
    inv = np.linalg.inv(s.mvp.astype(np.float64)).astype(np.float32)
    enc = torch.empty(B, H, W, 9, dtype=torch.float32, device=dev)
    core.encode_view_dir_sh2(rast, torch.from_numpy(inv).to(dev), enc)
    np.testing.assert_allclose(enc.cpu().numpy(), O.sh2(rast_np, inv), atol=2e-6, rtol=0)

    # edge gradients
    color = rng.uniform(size=(B, H, W, 3)).astype(np.float32)
    gcol = rng.normal(size=(B, H, W, 3)).astype(np.float32)
    gpos = torch.zeros_like(pos_d)
    core.backward_edge_grad(torch.from_numpy(color).to(dev), torch.from_numpy(gcol).to(dev), rast, pos_d, gpos, tri_d)
    ref = O.edge_grad(color, gcol, rast_np, pos, s.tri)
    np.testing.assert_allclose(gpos.cpu().numpy(), ref, rtol=0, atol=1e-4 * np.abs(ref).max())

    # opacity auxiliary gradient through fragments, then the alpha adjoint
    target = rng.uniform(size=(B, H, W, 3)).astype(np.float32)
    fp, fa, *_ = _fragments(core, H, W, pos_d, tri_d)
    a2 = torch.from_numpy(alpha.reshape(s.T, -1)).to(dev)
    fal = torch.zeros(fp.shape[0], dtype=torch.float32, device=dev)
    core.multires_triangle_alpha(fa, s.alpha_level, s.alpha_level, a2, fal)
    gfa = torch.zeros_like(fal)
    core.backward_opacity_aux_loss(torch.from_numpy(color).to(dev), torch.from_numpy(target).to(dev), rast,
                                   fp, fa, fal, gfa)
    ga = torch.zeros_like(a2)
    core.backward_multires_triangle_alpha(fa, s.alpha_level, s.alpha_level, ga, gfa)
    ref = O.opacity_aux_grad(color, target, rast_np, pos, s.tri, s.alpha_level, alpha)
    np.testing.assert_allclose(ga.cpu().numpy(), ref, rtol=0, atol=1e-4 * np.abs(ref).max())


def test_core_empty_inputs(dev):
    from diffsoup_amd import _core as core
    pos = torch.zeros(1, 3, 4, dtype=torch.float32, device=dev)
    tri = torch.zeros(0, 3, dtype=torch.int32, device=dev)
    rects = torch.empty(0, 4, dtype=torch.int32, device=dev)
    prefix = torch.empty(0, dtype=torch.int32, device=dev)
    assert core.compute_triangle_rects(8, 8, pos, tri, rects, prefix) == 0
    fp = torch.empty(0, 3, dtype=torch.int32, device=dev)
    fa = torch.empty(0, 4, dtype=torch.float32, device=dev)
    assert core.filter_valid_fragments(fp, fa, torch.empty_like(fp), torch.empty_like(fa)) == 0
    rast = torch.ones(1, 8, 8, 4, dtype=torch.float32, device=dev)
    e = torch.empty(0, dtype=torch.float32, device=dev)
    core.depth_test(fp, fa, e, e, rast)
    torch.cuda.synchronize()
    assert float(rast.abs().sum()) == 0.0
