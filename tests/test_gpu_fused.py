"""GPU parity of the fused training step (ds_train_step) against the CPU
oracle for the raster / gather / edge / opacity parts and a float64 torch
evaluation of the pixel MLP, blend and MSE loss, plus float64 central finite
differences on the continuous parameters.

Tolerances: fp32 MLP - colour within 1e-5 of float64, parameter gradients
within 1e-4 of their largest magnitude; bf16 MLP (fp32 accumulation) -
colour within 1e-3 of float64 (the image-error bar), gradients within 2e-2;
bf16 MLP and bf16 feature rows - colour within 1e-3 of float64 on the
unrounded fp32 features (the kernel reads the four residual-blend channels
of the first level from the fp32 master, everything the MLP contracts from
the bf16 rows), gradients within 2e-2 of float64 on the features as the
kernel reads them.  The raster output must be bit-identical in every
precision."""
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


def _run(dev, s, bf16, seed=5, bf16_features=False, loss="mse", aux_records=0):
    from diffsoup_amd import fused
    prm = fused.SoupParams.from_soup(s, dev)
    if bf16_features:
        prm.with_bf16_features()
    nl = s.levels[1] - s.levels[0] + 1
    st = fused.FusedStep(s.resolutions, s.T, s.verts.shape[0], s.alpha_level, s.levels, s.feat.shape[2],
                         s.mlp_dims[1], bf16=bf16, bf16_features=bf16_features, device=dev, loss=loss,
                         aux_records=aux_records)
    loss = st.run(prm, torch.from_numpy(s.mvp).to(dev), torch.from_numpy(s.target).to(dev), seed)
    torch.cuda.synchronize()
    out = {k: getattr(st, k).cpu().numpy() for k in ("pos", "alpha", "rast", "color", "grad_color",
                                                     "pix_loss", "grad_verts", "grad_alpha_logits",
                                                     "grad_feat", "grad_mlp")}
    out["loss"] = float(loss.item())
    out["stats"] = st.stats().cpu().numpy()
    del nl
    return out


@pytest.fixture(scope="module")
def small():
    return scenes.small_random_soup(0)


@pytest.fixture(scope="module")
def medium():
    return scenes.sphere_torus_soup(0)


@pytest.fixture(scope="module")
def small_l4(small):
    """The small soup with its 8 feature channels at lattice level 4 (153
    samples a triangle): rows large enough that the shading keeps one row
    slot (tiles stop at each triangle's end) to hold two workgroups per CU."""
    import dataclasses
    rng = np.random.Generator(np.random.PCG64(11))
    feat = rng.normal(0, 0.1, size=(small.T, O.feats_at_level(4), 8)).astype(np.float32)
    return dataclasses.replace(small, feat=feat, levels=(4, 4))


def _bf16_round(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.bfloat16).to(torch.float32).numpy()


def _bf16_rows_as_read(feat, lo):
    """The features as the bf16-row shading kernel reads them: rounded to
    bf16, except channels 0-3 of the first level's samples (the residual-blend
    inputs), which it reads from the fp32 master."""
    out = _bf16_round(feat)
    s_lo = O.feats_at_level(lo)
    out[:, :s_lo, :4] = feat[:, :s_lo, :4]
    return out


@pytest.mark.parametrize("which,bf16,bf16_feat", [("small", False, False), ("small_l4", False, False),
                                                  ("medium", False, False), ("medium", True, False),
                                                  ("medium", True, True)])
def test_fused_step_matches_reference(dev, small, small_l4, medium, which, bf16, bf16_feat):
    import dataclasses
    s0 = {"small": small, "small_l4": small_l4, "medium": medium}[which]
    print(f"fused step {s0.name} bf16={bf16} bf16 features={bf16_feat}", flush=True)
    g = _run(dev, s0, bf16, bf16_features=bf16_feat)
    # with bf16 feature rows the reference sees the features as the kernel
    # reads them; the error against the unrounded float64 model is checked
    # separately below
    s = dataclasses.replace(s0, feat=_bf16_rows_as_read(s0.feat, s0.levels[0])) if bf16_feat else s0
    # projection against float64
    pos64 = np.einsum("bij,nj->bni", s.mvp.astype(np.float64),
                      np.concatenate([s.verts, np.ones((s.verts.shape[0], 1), np.float32)], 1).astype(np.float64))
    np.testing.assert_allclose(g["pos"], pos64, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(g["alpha"], 1 / (1 + np.exp(-s.alpha_logits.astype(np.float64))), atol=1e-6)
    # raster: bit-exact with the oracle on the step's own clip positions
    H, W = s.resolutions[0]
    rast = O.rasterize((H, W), g["pos"], s.tri, s.alpha_level, g["alpha"], True, 5).reshape(-1, 4)
    np.testing.assert_array_equal(g["rast"].view(np.uint32), rast.view(np.uint32))
    assert g["stats"][0] == O.covered_fragments(H, W, g["pos"], s.tri)
    ref = O.reference_step(s, 5, pos=g["pos"], alpha=g["alpha"], color_in=g["color"],
                           grad_color_in=g["grad_color"])
    tol_c = 1e-3 if bf16 else 1e-5
    err = np.abs(g["color"] - ref["color"]).max()
    print(f"  colour max abs err vs fp64: {err:.3e}; loss {g['loss']:.6f} vs {ref['loss']:.6f}", flush=True)
    assert err <= tol_c
    assert abs(g["loss"] - ref["loss"]) <= (1e-3 if bf16 else 1e-5) * ref["loss"]
    np.testing.assert_allclose(g["pix_loss"], np.abs(g["color"] - s.target).sum(1), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(g["grad_color"], 2 * (g["color"] - s.target) / g["color"].size, rtol=1e-5,
                               atol=1e-12)
    if bf16_feat:
        # the image-error bar against the float64 model on the unrounded fp32
        # features: the bf16 rows reach the colour only through the MLP
        ref0 = O.reference_step(s0, 5, pos=g["pos"], alpha=g["alpha"])
        err0 = np.abs(g["color"] - ref0["color"]).max()
        print(f"  colour max abs err vs fp64 on unrounded features: {err0:.3e}", flush=True)
        assert err0 <= 1e-3
    rel = 2e-2 if bf16 else 1e-4
    _close(g["grad_mlp"][:-1], ref["grad_mlp"], rel, "MLP grads")
    assert abs(g["grad_mlp"][-1] - ((g["color"][ref["covered"]] - s.target[ref["covered"]]) ** 2).sum()) < 1e-3 * max(1.0, g["grad_mlp"][-1])
    _close(g["grad_feat"], ref["grad_feat"], rel, "feature grads")
    _close(g["grad_verts"], ref["grad_verts"], 1e-4, "vertex grads")
    _close(g["grad_alpha_logits"], ref["grad_alpha_logits"], 1e-4, "opacity logit grads")


def test_opacity_records_overflow_takes_the_bin_walk(dev, medium):
    """The opacity gradient from the resolve's per-tile fragment records and,
    for the tiles whose records overflow deliberately few slots, from a
    second walk of their bins: the same gradient (up to the order of float
    atomics), within the oracle's bar, every other output identical."""
    g = _run(dev, medium, False)
    walked = int(g["stats"][4])
    print(f"tiles walked with the default slots: {walked}", flush=True)
    assert walked == 0   # two slots per pixel hold this scene's records
    g2 = _run(dev, medium, False, aux_records=4)
    walked2 = int(g2["stats"][4])
    print(f"tiles walked with 4 slots: {walked2}", flush=True)
    assert walked2 > 0
    for k in ("rast", "color"):
        np.testing.assert_array_equal(g[k], g2[k], err_msg=k)
    for k in ("grad_feat", "grad_mlp", "grad_verts"):
        _close(g2[k], g[k], 1e-6, k)
    _close(g2["grad_alpha_logits"], g["grad_alpha_logits"], 1e-5, "opacity logit grads (walk vs records)")
    ref = O.reference_step(medium, 5, pos=g["pos"], alpha=g["alpha"], color_in=g["color"],
                           grad_color_in=g["grad_color"])
    _close(g2["grad_alpha_logits"], ref["grad_alpha_logits"], 1e-4, "opacity logit grads (walk)")


def test_fused_fd_continuous_params(dev, small):
    """float64 central differences of the loss w.r.t. features and MLP weights."""
    s = small
    g = _run(dev, s, False, seed=9)
    H, W = s.resolutions[0]
    rast = g["rast"]
    lo, hi = s.levels
    P = rast.shape[0]
    covered = rast[:, 3] > 0
    tgt = s.target.astype(np.float64)

    def loss_of(feat, mlp):
        X = O.multires_color(rast, lo, hi, feat, True).reshape(P, -1).astype(np.float64)
        h = X[covered]
        for i, (w, b) in enumerate(mlp):
            h = h @ w.astype(np.float64).T + b
            h = np.maximum(h, 0) if i < 2 else 1 / (1 + np.exp(-h))
        res = X[covered, 3:4]
        col = np.ones((P, 3))
        col[covered] = (1 - res) * X[covered, :3] + res * h
        return ((col - tgt) ** 2).mean()

    rng = np.random.default_rng(0)
    ts = np.unique(rast[covered, 3].astype(int) - 1)
    eps = 1e-3
    checked = 0
    for _ in range(8):
        t = int(rng.choice(ts))
        k = int(rng.integers(0, s.feat.shape[1]))
        c = int(rng.integers(0, s.feat.shape[2]))
        fp, fm = s.feat.copy(), s.feat.copy()
        fp[t, k, c] += eps
        fm[t, k, c] -= eps
        fd = (loss_of(fp, s.mlp) - loss_of(fm, s.mlp)) / (2 * eps)
        got = g["grad_feat"][t, k, c]
        assert abs(got - fd) <= 2e-3 * abs(fd) + 1e-9, (t, k, c, got, fd)
        checked += 1
    dW1 = g["grad_mlp"][:16 * 8].reshape(16, 8)
    for _ in range(6):
        i, j = int(rng.integers(0, 16)), int(rng.integers(0, 8))
        mp = [(w.copy(), b.copy()) for w, b in s.mlp]
        mm = [(w.copy(), b.copy()) for w, b in s.mlp]
        mp[0][0][i, j] += eps
        mm[0][0][i, j] -= eps
        fd = (loss_of(s.feat, mp) - loss_of(s.feat, mm)) / (2 * eps)
        assert abs(dW1[i, j] - fd) <= 2e-3 * abs(fd) + 1e-8, (i, j, dW1[i, j], fd)
        checked += 1
    assert checked == 14


@pytest.mark.parametrize("which", ["small", "uneven", "medium", "many_views"])
def test_fused_step_l1_ssim_loss(dev, small, medium, which):
    """The reference's image loss 0.8 L1 + 0.2 (1 - SSIM) / 2 inside the fused
    step (colour pass, SSIM kernels, adjoint pass) against the oracle step
    with the loss and its colour gradient in float64 (SSIM per view, averaged
    over views; "many_views" packs 200 views of mixed resolution into one
    step, above the 64 of one SSIM launch of the per-op API).  Tolerances:
    colour 1e-5 and loss 1e-5 relative (fp32), the
    colour gradient 1e-3 of its largest magnitude (fp32 SSIM moments), the
    parameter gradients 2e-3 (they inherit it); edge and opacity gradients
    1e-4 on the step's own colour and colour gradient."""
    if which == "many_views":
        s = scenes.batched_uneven_soup(6, T=3000, views=200, res_scale=0.06, edge_scale=12.0)
    else:
        s = {"small": small, "medium": medium,
             "uneven": scenes.batched_uneven_soup(5, T=4000, views=5, res_scale=0.2, edge_scale=5.0)}[which]
    print(f"fused step with L1 + SSIM: {s.name}, {len(s.resolutions)} views, first {s.resolutions[:4]}", flush=True)
    g = _run(dev, s, False, seed=7, loss="l1_ssim")
    rast = np.concatenate([O.rasterize((h, w), g["pos"][b:b + 1], s.tri, s.alpha_level, g["alpha"], True, 7,
                                       view_base=b).reshape(-1, 4) for b, (h, w) in enumerate(s.resolutions)])
    np.testing.assert_array_equal(g["rast"].view(np.uint32), rast.view(np.uint32))
    ref = O.reference_step(s, 7, pos=g["pos"], alpha=g["alpha"], color_in=g["color"],
                           grad_color_in=g["grad_color"], loss="l1_ssim")
    err = np.abs(g["color"] - ref["color"]).max()
    print(f"  colour err {err:.3e}; loss {g['loss']:.7f} vs {ref['loss']:.7f}", flush=True)
    assert err <= 1e-5
    assert abs(g["loss"] - ref["loss"]) <= 1e-5 * ref["loss"]
    np.testing.assert_allclose(g["pix_loss"], np.abs(g["color"] - s.target).sum(1), rtol=1e-6, atol=1e-6)
    _close(g["grad_color"], ref["grad_color"], 1e-3, "colour grads")
    _close(g["grad_mlp"][:-1], ref["grad_mlp"], 2e-3, "MLP grads")
    _close(g["grad_feat"], ref["grad_feat"], 2e-3, "feature grads")
    _close(g["grad_verts"], ref["grad_verts"], 1e-4, "vertex grads")
    _close(g["grad_alpha_logits"], ref["grad_alpha_logits"], 1e-4, "opacity logit grads")
