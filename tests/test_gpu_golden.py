"""The HIP path on the committed golden fixtures: per-operator entry points on
the hand-placed tiny scenes (bit-exact raster, fragments and gathers;
gradients to tolerance) and the fused training step on the 1,000-triangle
random soup."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _u32(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def tiny():
    return dict(np.load(os.path.join(GOLDEN, "tiny_scenes.npz")))


def test_tiny_per_op(dev, tiny):
    import diffsoup_amd as ds
    from diffsoup_amd import _core
    H, W = tiny["resolution"].tolist()
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    pos, tri = d(tiny["pos"]), d(tiny["tri"])
    for level in (0, 2):
        a = d(tiny[f"alpha_l{level}"])
        for tile in (8, 16, 32):
            det = ds.rasterize_views([(H, W)] * 2, pos, tri, level, a, stochastic=False, tile=tile)
            np.testing.assert_array_equal(_u32(det.cpu().numpy()).reshape(2, H, W, 4),
                                          _u32(tiny[f"rast_l{level}_det"]))
            sto = ds.rasterize_views([(H, W)] * 2, pos, tri, level, a, stochastic=True, seed=11, tile=tile)
            np.testing.assert_array_equal(_u32(sto.cpu().numpy()).reshape(2, H, W, 4),
                                          _u32(tiny[f"rast_l{level}_seed11"]))
    # fragment list of the drop-in module
    B, T = 2, tri.shape[0]
    rects = torch.empty(B * T, 4, dtype=torch.int32, device=dev)
    prefix = torch.empty(B * T, dtype=torch.int32, device=dev)
    n = _core.compute_triangle_rects(H, W, pos, tri, rects, prefix)
    fp = torch.full((n, 3), -1, dtype=torch.int32, device=dev)
    fa = torch.empty(n, 4, dtype=torch.float32, device=dev)
    _core.compute_fragments(H, W, pos, tri, prefix, rects, fp, fa)
    fpo, fao = torch.empty_like(fp), torch.empty_like(fa)
    nv = _core.filter_valid_fragments(fp, fa, fpo, fao)
    np.testing.assert_array_equal(rects.cpu().numpy(), tiny["rects"])
    np.testing.assert_array_equal(fpo[:nv].cpu().numpy(), tiny["frag_pix"])
    np.testing.assert_array_equal(_u32(fao[:nv].cpu().numpy()), _u32(tiny["frag_attrs"]))
    # gathers, edge and opacity gradients on the stochastic level-2 raster
    rast = d(tiny["rast_l2_seed11"])
    feat = d(tiny["feat_l012"])
    got = ds.multires_feature_gather(rast, 0, 2, feat, "sum").cpu().numpy()
    np.testing.assert_array_equal(_u32(got), _u32(tiny["gather_sum"]))
    got = ds.multires_feature_gather(rast, 0, 2, feat, "concat").cpu().numpy()
    np.testing.assert_array_equal(_u32(got), _u32(tiny["gather_concat"]))
    gpos = torch.zeros_like(pos)
    _core.backward_edge_grad(d(tiny["color"]), d(tiny["grad_color"]), rast, pos, gpos, tri)
    ref = tiny["edge_grad_pos"]
    np.testing.assert_allclose(gpos.cpu().numpy(), ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    pos_g = pos.clone().requires_grad_(True)
    color = d(tiny["color"]).requires_grad_(True)
    out = ds.edge_grad(color, rast, pos_g, tri)
    out.backward(d(tiny["grad_color"]))
    np.testing.assert_allclose(pos_g.grad.cpu().numpy(), ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    alpha_src = d(tiny["alpha_l2"]).unsqueeze(-1).requires_grad_(True)
    aux = ds.opacity_aux_loss(d(tiny["color"]), d(tiny["target"]), rast, pos, tri, 2, alpha_src)
    aux.backward()
    ref = tiny["aux_grad_alpha_l2"] / tiny["color"].size
    np.testing.assert_allclose(alpha_src.grad.squeeze(-1).cpu().numpy(), ref, rtol=0, atol=1e-5 * np.abs(ref).max())


def test_random_soup_fused_step(dev):
    from diffsoup_amd import fused
    g = dict(np.load(os.path.join(GOLDEN, "random_soup_1k.npz")))
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    H, W = g["resolution"].tolist()
    prm = fused.SoupParams(d(g["verts"]), d(g["tri"]), d(g["alpha_logits"]), d(g["feat"]), d(g["w1"]), d(g["b1"]),
                           d(g["w2"]), d(g["b2"]), d(g["w3"]), d(g["b3"]))
    st = fused.FusedStep([(H, W)], g["tri"].shape[0], g["verts"].shape[0], 0, (0, 0), 8, 16, device=dev)
    loss = float(st.run(prm, d(g["mvp"]), d(g["target"]), seed=int(g["seed"])).item())
    np.testing.assert_array_equal(_u32(st.pos.cpu().numpy()), _u32(g["pos"]))
    # the step's sigmoid is evaluated in double and rounded once, as the
    # oracle's: opacities are bit-identical, so every mask decision is too
    np.testing.assert_array_equal(_u32(st.alpha.cpu().numpy()), _u32(g["alpha"]))
    rast = st.rast.cpu().numpy().reshape(1, H, W, 4)
    np.testing.assert_array_equal(_u32(rast), _u32(g["rast"]))
    assert int(st.stats()[0].item()) == int(g["covered_fragments"])
    np.testing.assert_allclose(st.color.cpu().numpy(), g["color"], rtol=0, atol=1e-5)
    assert abs(loss - float(g["loss"])) <= 1e-5 * float(g["loss"])
    ref = g["grad_feat"]
    np.testing.assert_allclose(st.grad_feat.cpu().numpy(), ref, rtol=0, atol=1e-4 * np.abs(ref).max())
    ref = g["grad_verts"]
    np.testing.assert_allclose(st.grad_verts.cpu().numpy(), ref, rtol=0, atol=1e-3 * np.abs(ref).max())
    ref = g["grad_alpha_logits"]
    np.testing.assert_allclose(st.grad_alpha_logits.cpu().numpy(), ref, rtol=0, atol=1e-4 * np.abs(ref).max())
    ref = g["grad_mlp"]
    got = st.grad_mlp.cpu().numpy()[:ref.shape[0]]   # the step appends the covered squared error
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-4 * np.abs(ref).max())
