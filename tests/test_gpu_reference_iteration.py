"""The reference's own training iteration (examples/02_synthetic.py:354-388)
through the fused path: feat_dim 7 features and opacities as multires
sources accumulated to the top level and passed through a sigmoid, the 9
SH2 values of the view direction after the features, ColorMLP(16, 16,
n_layers=2), the L1 + SSIM loss plus the opacity term, edge gradients, and
Adam / VectorAdam with the exp_decay_mult learning-rate schedule inside a
captured graph.  Levels (0, 0) as the driver starts, and (2, 5) as it
continues after lifting (561 samples per triangle at level 5: the shading
reads such rows from global memory instead of staging them).

Tolerances as the fused-step tests: colour and loss 1e-5 (relative for the
loss) of float64, the colour gradient 1e-3 and MLP / feature gradients 2e-3
of their largest magnitude (fp32 SSIM moments), vertex and opacity
gradients 1e-4 on the step's own colour and colour gradient; the raster
bit-exact."""
import numpy as np
import pytest
import torch

from diffsoup_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _close(got, ref, rel, what):
    scale = max(np.abs(ref).max(), 1e-12)
    err = np.abs(got - ref).max()
    assert err <= rel * scale, f"{what}: max err {err:.3e} > {rel:.1e} x {scale:.3e}"


def _step(dev, s, loss="l1_ssim", bf16=False):
    from diffsoup_amd import fused
    return fused.FusedStep(s.resolutions, s.T, s.verts.shape[0], s.alpha_level, s.levels, 7, 16, device=dev,
                           loss=loss, accumulate=True, alpha_lo=s.alpha_lo, sh2=True, bf16=bf16)


def _inputs(dev, s):
    from diffsoup_amd import fused
    return (fused.SoupParams.from_soup(s, dev), torch.from_numpy(s.mvp).to(dev), torch.from_numpy(s.target).to(dev),
            torch.from_numpy(s.inv_mvp).to(dev))


# opacities over the features' levels (the driver's two phases: the
# accumulation carries them as one more channel of the feature passes), and
# over other levels (their own corner-form and restriction passes)
@pytest.mark.parametrize("levels,alpha_levels", [((0, 0), (0, 0)), ((2, 5), (2, 5)), ((2, 5), (1, 3))])
def test_reference_iteration_matches_oracle(dev, levels, alpha_levels):
    s = scenes.reference_training_soup(0, levels=levels, alpha_levels=alpha_levels)
    print(f"reference iteration: levels {levels}, opacity levels {alpha_levels}, feat {s.feat.shape}, "
          f"alpha {s.alpha_logits.shape}, views {s.resolutions}", flush=True)
    prm, mvp, target, inv = _inputs(dev, s)
    st = _step(dev, s)
    loss = float(st.run(prm, mvp, target, 21, inv_mvp=inv).item())
    torch.cuda.synchronize()
    g = {k: getattr(st, k).cpu().numpy() for k in ("pos", "alpha", "rast", "color", "grad_color", "grad_feat",
                                                   "grad_mlp", "grad_verts", "grad_alpha_logits")}
    ref = O.reference_step(s, 21, pos=g["pos"], color_in=g["color"], grad_color_in=g["grad_color"],
                           loss="l1_ssim")
    # the accumulated, sigmoided opacities against the oracle's (double sigmoid rounded once)
    np.testing.assert_allclose(g["alpha"], ref["alpha"], rtol=0, atol=1e-6)
    ref = O.reference_step(s, 21, pos=g["pos"], alpha=g["alpha"], color_in=g["color"],
                           grad_color_in=g["grad_color"], loss="l1_ssim")
    np.testing.assert_array_equal(g["rast"].view(np.uint32), ref["rast"].view(np.uint32))
    err = np.abs(g["color"] - ref["color"]).max()
    print(f"  colour err {err:.3e}; loss {loss:.7f} vs {ref['loss']:.7f}", flush=True)
    assert err <= 1e-5
    assert abs(loss - ref["loss"]) <= 1e-5 * ref["loss"]
    _close(g["grad_color"], ref["grad_color"], 1e-3, "colour grads")
    _close(g["grad_mlp"][:-1], ref["grad_mlp"], 2e-3, "MLP grads")
    _close(g["grad_feat"], ref["grad_feat"], 2e-3, "feature-source grads")
    _close(g["grad_verts"], ref["grad_verts"], 1e-4, "vertex grads")
    _close(g["grad_alpha_logits"], ref["grad_alpha_logits"], 1e-4, "opacity-source grads")


def test_render_autograd_composes_any_loss(dev):
    """render() (ds_step_forward / ds_step_backward behind an autograd
    Function) under two losses composed in torch: the library's L1 + SSIM and
    a plain torch MSE give the gradients of the fused steps with those
    losses."""
    from diffsoup_amd import fused, losses
    s = scenes.reference_training_soup(1, levels=(0, 0), alpha_levels=(0, 0))
    H, W = s.resolutions[0]
    B = s.B
    for name, loss_fn, kind in (
            ("l1_ssim", lambda c, t: losses.l1_ssim_loss(c.view(B, H, W, 3), t.view(B, H, W, 3)), "l1_ssim"),
            ("torch mse", lambda c, t: ((c - t) ** 2).mean(), "mse")):
        prm, mvp, target, inv = _inputs(dev, s)
        leaves = fused.SoupParams(*[t.clone().requires_grad_(t.dtype == torch.float32 and k != "tri")
                                    for k, t in zip(("verts", "tri", "alpha_logits", "feat", "w1", "b1", "w2", "b2",
                                                     "w3", "b3"),
                                                    (prm.verts, prm.tri, prm.alpha_logits, prm.feat, prm.w1, prm.b1,
                                                     prm.w2, prm.b2, prm.w3, prm.b3))])
        st = _step(dev, s, loss="mse")
        color = fused.render(st, leaves, mvp, target, 33, inv_mvp=inv)
        loss = loss_fn(color, target)
        loss.backward()
        torch.cuda.synchronize()
        ref = _step(dev, s, loss=kind)
        ref_loss = float(ref.run(prm, mvp, target, 33, inv_mvp=inv).item())
        torch.cuda.synchronize()
        print(f"  {name}: loss {float(loss):.7f} vs fused {ref_loss:.7f}", flush=True)
        assert abs(float(loss) - ref_loss) <= 1e-5 * ref_loss
        np.testing.assert_array_equal(color.detach().cpu().numpy(), ref.color.cpu().numpy())
        pairs = [(leaves.feat.grad, ref.grad_feat), (leaves.alpha_logits.grad, ref.grad_alpha_logits),
                 (leaves.verts.grad, ref.grad_verts)]
        dW1, db1, dW2, db2, dW3, db3 = ref.mlp_grads()
        pairs += [(leaves.w1.grad, dW1), (leaves.b1.grad, db1), (leaves.w2.grad, dW2), (leaves.b2.grad, db2),
                  (leaves.w3.grad, dW3), (leaves.b3.grad, db3)]
        for i, (a, b) in enumerate(pairs):
            _close(a.cpu().numpy(), b.cpu().numpy(), 1e-4, f"{name}: gradient {i}")


def test_captured_reference_iteration_follows_lr_schedule(dev):
    """Three replays of the captured reference iteration with the fused
    optimiser's on-device exp_decay_mult schedule against three eager
    iterations whose Adam / VectorAdam learning rates are set on the host
    each iteration (examples/02_synthetic.py:341-349), with eager work
    between the replays."""
    from diffsoup_amd import fused
    s = scenes.reference_training_soup(2, levels=(0, 0), alpha_levels=(0, 0))
    steps, final = 10, 0.01
    mult = lambda i: final ** (min(max(i, 1), steps) / float(steps))
    base = dict(lr_feat=5e-2, lr_alpha=5e-2, lr_verts=1e-3, lr_mlp=1e-2)

    pe, mvp, target, inv = _inputs(dev, s)
    se = _step(dev, s)
    oe = fused.soup_optimizer(se, pe, **base)
    base_lrs = [g["lr"] for g in oe.groups]
    eager = []
    for i in range(1, 4):
        for k, lr in enumerate(base_lrs):
            oe.set_lr(k, lr * mult(i))
        se.run(pe, mvp, target, 500 + i - 1, inv_mvp=inv)
        oe.step()
        eager.append({k: getattr(pe, k).detach().clone() for k in ("feat", "alpha_logits", "verts", "w1", "b3")})
    torch.cuda.synchronize()

    pg, _, _, _ = _inputs(dev, s)
    sg = _step(dev, s)
    og = fused.soup_optimizer(sg, pg, decay=(steps, final), **base)
    cap = fused.CapturedStep(sg, pg, mvp, target, og, seed=500, inv_mvp=inv)
    for i in range(3):
        cap.replay()
        torch.cuda.synchronize()
        for k, ref in eager[i].items():
            got = getattr(pg, k)
            off = (got - ref).abs() > 1e-5 * max(1.0, float(ref.abs().max()))
            assert float(off.float().mean()) < 1e-3, f"replay {i} {k}: {int(off.sum())} of {off.numel()} differ"
        _ = torch.randn(4096, device=dev).sum().item()   # eager work between replays
    assert int(og.step_count.item()) == 3


def test_lr_mult_read_at_replay(dev):
    """FusedAdam's device learning-rate multiplier: a captured update follows
    the value stored between replays (0 freezes the parameters, 1 moves
    them)."""
    from diffsoup_amd.optimize import FusedAdam
    p = torch.zeros(1000, device=dev)
    gr = torch.ones(1000, device=dev)
    opt = FusedAdam([dict(param=p, grad=gr, lr=0.1)], lr_mult=True)
    graph = torch.cuda.CUDAGraph()
    opt.step()   # warm-up outside the capture
    torch.cuda.synchronize()
    p.zero_()
    with torch.cuda.graph(graph):
        opt.step()
    opt.set_lr_mult(0.0)
    graph.replay()
    torch.cuda.synchronize()
    assert float(p.abs().max()) == 0.0
    opt.set_lr_mult(1.0)
    graph.replay()
    torch.cuda.synchronize()
    assert float(p.abs().min()) > 0.0
