"""GPU parity of the per-operator HIP kernels (the drop-in API) against the
CPU oracle: feature gather (bit-exact), its adjoint, cross-level
accumulation, SH2, edge gradients and the opacity auxiliary gradient."""
import numpy as np
import pytest
import torch

from diffsoup_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _sig(x):
    return (1.0 / (1.0 + np.exp(-x.astype(np.float64)))).astype(np.float32)


@pytest.fixture(scope="module")
def scene(dev):
    import diffsoup_amd as ds
    s = scenes.sphere_torus_soup(2, n_surface=6000, n_float=1500, views=2, res=(320, 400))
    pos = scenes.project_vertices_np(s.verts, s.mvp)
    alpha = _sig(s.alpha_logits)
    rast = O.rasterize(s.resolutions[0], pos, s.tri, s.alpha_level, alpha, True, 17)
    return s, pos, alpha, rast


def test_feature_gather_bit_exact(dev, scene):
    import diffsoup_amd as ds
    s, pos, alpha, rast = scene
    lo, hi = s.levels
    for mode in ("sum", "concat"):
        ref = O.multires_color(rast, lo, hi, s.feat, mode == "concat")
        got = ds.multires_feature_gather(torch.from_numpy(rast).to(dev), lo, hi,
                                         torch.from_numpy(s.feat).to(dev), mode).cpu().numpy()
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_feature_gather_backward(dev, scene):
    import diffsoup_amd as ds
    s, pos, alpha, rast = scene
    lo, hi = s.levels
    rng = np.random.default_rng(0)
    for mode in ("sum", "concat"):
        feat = torch.from_numpy(s.feat).to(dev).requires_grad_(True)
        out = ds.multires_feature_gather(torch.from_numpy(rast).to(dev), lo, hi, feat, mode)
        g = rng.normal(size=out.shape).astype(np.float32)
        out.backward(torch.from_numpy(g).to(dev))
        ref = O.multires_color_bwd(rast, lo, hi, s.feat.shape, g, mode == "concat")
        np.testing.assert_allclose(feat.grad.cpu().numpy(), ref, rtol=1e-4, atol=1e-4)


def test_single_level_api_and_accumulate(dev, scene):
    import diffsoup_amd as ds
    s, pos, alpha, rast = scene
    T = s.T
    rng = np.random.default_rng(1)
    x = rng.normal(size=(T, ds.feats_in_levels(0, 3), 4)).astype(np.float32)
    acc = ds.accumulate_to_level(0, 3, torch.from_numpy(x).to(dev), target_level=2)
    ref = O.accumulate_fwd(0, 3, 2, x)
    np.testing.assert_array_equal(acc.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    xt = torch.from_numpy(x).to(dev).requires_grad_(True)
    y = rng.normal(size=ref.shape).astype(np.float32)
    ds.accumulate_to_level(0, 3, xt, target_level=2).backward(torch.from_numpy(y).to(dev))
    np.testing.assert_allclose(xt.grad.cpu().numpy(), O.accumulate_bwd(0, 3, 2, x.shape, y), rtol=1e-4, atol=1e-4)
    f2 = torch.from_numpy(ref).to(dev)
    got = ds.multires_triangle_color(torch.from_numpy(rast).to(dev), 2, f2).cpu().numpy()
    np.testing.assert_array_equal(got.view(np.uint32), O.multires_color(rast, 2, 2, ref).view(np.uint32))


def test_sh2(dev, scene):
    import diffsoup_amd as ds
    s, pos, alpha, rast = scene
    inv = np.linalg.inv(s.mvp.astype(np.float64)).astype(np.float32)
    got = ds.encode_view_dir_sh2(torch.from_numpy(rast).to(dev), torch.from_numpy(inv).to(dev)).cpu().numpy()
    np.testing.assert_allclose(got, O.sh2(rast, inv), atol=2e-6, rtol=0)


def test_edge_grad(dev, scene):
    import diffsoup_amd as ds
    s, pos, alpha, rast = scene
    rng = np.random.default_rng(2)
    B, H, W, _ = rast.shape
    color = rng.uniform(0, 1, size=(B, H, W, 3)).astype(np.float32)
    gcol = rng.normal(size=(B, H, W, 3)).astype(np.float32)
    pos_t = torch.from_numpy(pos).to(dev).requires_grad_(True)
    c_t = torch.from_numpy(color).to(dev).requires_grad_(True)
    out = ds.edge_grad(c_t, torch.from_numpy(rast).to(dev), pos_t, torch.from_numpy(s.tri).to(dev))
    out.backward(torch.from_numpy(gcol).to(dev))
    ref = O.edge_grad(color, gcol, rast, pos, s.tri)
    got = pos_t.grad.cpu().numpy()
    scale = np.abs(ref).max()
    assert scale > 0
    np.testing.assert_allclose(got, ref, atol=1e-5 * scale, rtol=1e-4)
    np.testing.assert_array_equal(c_t.grad.cpu().numpy(), gcol)



@pytest.mark.parametrize("name", ["silhouette", "occlusion", "intersection", "interior"])
def test_edge_grad_against_finite_differences(dev, name):
    """The GPU rasterizer and edge-gradient kernel against float64 central
    differences of L = sum g . colour from an independent point-sampling
    renderer: to the pixel discretisation across silhouettes, occlusion edges
    and intersection lines, twice dL/dpos inside a triangle (the estimator visits each
    same-triangle pixel pair from both pixels), at the tolerances of the
    oracle's own check."""
    import diffsoup_amd as ds
    import fd_render as F
    sc = F.scene(name)
    H, W = sc["H"], sc["W"]
    T = sc["tri"].shape[0]
    pos_t = torch.from_numpy(sc["pos"][None].copy()).to(dev)
    tri_t = torch.from_numpy(sc["tri"]).to(dev)
    rast = ds.rasterize_multires_triangle_alpha((H, W), pos_t, tri_t, 0, torch.ones(T, 3, 1, device=dev),
                                                stochastic=False)
    ids, _ = F.render(sc["pos"], sc["tri"], H, W)
    np.testing.assert_array_equal(rast[0, ..., 3].cpu().numpy().astype(int) - 1, ids)
    color, g = F.image(sc)
    pos_t.requires_grad_(True)
    out = ds.edge_grad(torch.from_numpy(color).to(dev), rast, pos_t, tri_t)
    out.backward(torch.from_numpy(g).to(dev))
    got = pos_t.grad[0].cpu().numpy().astype(np.float64)
    want = F.expected(sc)
    np.testing.assert_allclose(got, want, rtol=0, atol=F.TOLERANCE[name] * np.abs(want).max())

def test_opacity_aux(dev, scene):
    import diffsoup_amd as ds
    s, pos, alpha, rast = scene
    rng = np.random.default_rng(3)
    B, H, W, _ = rast.shape
    color = rng.uniform(0, 1, size=(B, H, W, 3)).astype(np.float32)
    target = rng.uniform(0, 1, size=(B, H, W, 3)).astype(np.float32)
    T = s.T
    a_t = torch.from_numpy(alpha).reshape(T, -1, 1).contiguous().to(dev).requires_grad_(True)
    aux = ds.opacity_aux_loss(torch.from_numpy(color).to(dev), torch.from_numpy(target).to(dev),
                              torch.from_numpy(rast).to(dev), torch.from_numpy(pos).to(dev),
                              torch.from_numpy(s.tri).to(dev), s.alpha_level, a_t)
    assert aux.item() == 0.0
    aux.backward()
    ref = O.opacity_aux_grad(color, target, rast, pos, s.tri, s.alpha_level, alpha) / color.size
    got = a_t.grad.cpu().numpy().reshape(T, -1)
    scale = np.abs(ref).max()
    assert scale > 0
    np.testing.assert_allclose(got, ref, atol=1e-5 * scale, rtol=1e-4)
