"""The plain fragment-list path (the reference's structure on HIP) against the
oracle (bit-exact raster) and against the fused step (same training step,
structurally different kernels: gradients must agree)."""
import numpy as np
import pytest
import torch

from diffsoup_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_plain_raster_bit_exact(dev):
    from diffsoup_amd import plain
    s = scenes.sphere_torus_soup(3, n_surface=4000, n_float=1000, views=3, res=(300, 420))
    pos = scenes.project_vertices_np(s.verts, s.mvp)
    alpha = (1 / (1 + np.exp(-s.alpha_logits.astype(np.float64)))).astype(np.float32)
    for stochastic in (False, True):
        got = plain.rasterize((300, 420), torch.from_numpy(pos).to(dev), torch.from_numpy(s.tri).to(dev),
                              s.alpha_level, torch.from_numpy(alpha).to(dev), stochastic, seed=21).cpu().numpy()
        ref = O.rasterize((300, 420), pos, s.tri, s.alpha_level, alpha, stochastic, 21)
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    fp, fa = plain.compute_fragments((300, 420), torch.from_numpy(pos).to(dev), torch.from_numpy(s.tri).to(dev))
    rfp, rfa = O.fragments(300, 420, pos, s.tri)
    np.testing.assert_array_equal(fp.cpu().numpy(), rfp)
    np.testing.assert_array_equal(fa.cpu().numpy().view(np.uint32), rfa.view(np.uint32))


def test_plain_step_matches_fused_step(dev):
    from diffsoup_amd import fused, plain
    s = scenes.sphere_torus_soup(4, n_surface=5000, n_float=1500, views=3, res=(256, 320))
    prm = fused.SoupParams.from_soup(s, dev)
    mvp = torch.from_numpy(s.mvp).to(dev)
    target = torch.from_numpy(s.target).to(dev)
    st = fused.FusedStep(s.resolutions, s.T, s.verts.shape[0], s.alpha_level, s.levels, 8, 32, device=dev)
    loss_f = st.run(prm, mvp, target, seed=77).item()
    leaves = fused.SoupParams(prm.verts.clone().requires_grad_(True), prm.tri,
                              prm.alpha_logits.clone().requires_grad_(True), prm.feat.clone().requires_grad_(True),
                              *[t.clone() for t in (prm.w1, prm.b1, prm.w2, prm.b2, prm.w3, prm.b3)])
    layers = []
    for w, b in ((leaves.w1, leaves.b1), (leaves.w2, leaves.b2), (leaves.w3, leaves.b3)):
        lin = torch.nn.Linear(w.shape[1], w.shape[0]).to(dev)
        with torch.no_grad():
            lin.weight.copy_(w)
            lin.bias.copy_(b)
        layers.append(lin)
    ps = plain.PlainStep(leaves, layers, s.levels, s.alpha_level)
    loss_p = ps.run(mvp, target, s.resolutions, seed=77).item()
    torch.cuda.synchronize()
    assert abs(loss_f - loss_p) <= 1e-5 * loss_p
    np.testing.assert_array_equal(st.rast.cpu().numpy().reshape(3, 256, 320, 4).view(np.uint32),
                                  O.rasterize((256, 320), st.pos.cpu().numpy(), s.tri, s.alpha_level,
                                              st.alpha.cpu().numpy(), True, 77).view(np.uint32))

    def close(a, b, rel, what):
        a, b = a.detach().cpu().numpy(), b.detach().cpu().numpy()
        scale = np.abs(b).max()
        err = np.abs(a - b).max()
        assert err <= rel * scale, f"{what}: {err:.3e} vs scale {scale:.3e}"

    close(st.grad_feat, leaves.feat.grad, 1e-4, "features")
    close(st.grad_alpha_logits, leaves.alpha_logits.grad, 1e-4, "opacity logits")
    close(st.grad_verts, leaves.verts.grad, 2e-4, "vertices")
    dW1, db1, dW2, db2, dW3, db3 = st.mlp_grads()
    for (dw, db), lin in zip(((dW1, db1), (dW2, db2), (dW3, db3)), layers):
        close(dw, lin.weight.grad, 1e-4, "mlp weight")
        close(db, lin.bias.grad, 1e-4, "mlp bias")
