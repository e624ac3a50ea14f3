"""The larger workloads on the fused step: batched uneven views (50K-triangle
unbounded scene, mixed resolutions), bf16 features with the 80->64->64->3
MLP (100K triangles, 5 levels x 16 channels), and the stress scene (300K
triangles, 40% stacked depth layers), plus the 80->64->64->3 MLP in fp32.

Each runs scaled down against the oracle - raster bit-exact on every view,
colour and gradients to the fused-step tolerances - and the batched and bf16
scenes also at full size, checked view by view on a subset (raster
bit-exact, colour against float64)."""
import dataclasses

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


def _step(dev, s, bf16, seed):
    from diffsoup_amd import fused
    prm = fused.SoupParams.from_soup(s, dev)
    if bf16:
        prm.with_bf16_features()
    st = fused.FusedStep(s.resolutions, s.T, s.verts.shape[0], s.alpha_level, s.levels, s.feat.shape[2],
                         s.mlp_dims[1], bf16=bf16, bf16_features=bf16, device=dev)
    loss = float(st.run(prm, torch.from_numpy(s.mvp).to(dev), torch.from_numpy(s.target).to(dev), seed).item())
    torch.cuda.synchronize()
    return st, loss


SMALL = {
    "uneven": (lambda: scenes.batched_uneven_soup(1, T=6000, views=4, res_scale=0.25, edge_scale=4.0), False),
    "bf16_large": (lambda: scenes.bf16_large_soup(2, T=5000, views=2, resolutions=[(216, 384)], edge_scale=5.0),
                   True),
    "stress": (lambda: scenes.stress_soup(3, T=9000, views=4, res_scale=0.2, edge_scale=5.0), True),
    # the 80->64->64->3 MLP in fp32: weight operands from global memory, LDS atomics in the scatter
    "large_fp32": (lambda: scenes.bf16_large_soup(4, T=3000, views=2, resolutions=[(216, 384)], edge_scale=5.0),
                   False),
}


@pytest.mark.parametrize("name", list(SMALL))
def test_scaled_down_against_oracle(dev, name):
    make, bf16 = SMALL[name]
    s0 = make()
    print(f"{name}: T={s0.T} views={s0.resolutions} levels={s0.levels} C={s0.feat.shape[2]} "
          f"mlp={s0.mlp_dims} bf16={bf16}", flush=True)
    st, loss = _step(dev, s0, bf16, seed=7)
    pos, alpha = st.pos.cpu().numpy(), st.alpha.cpu().numpy()
    s = dataclasses.replace(s0, feat=_bf16_rows_as_read(s0.feat, s0.levels[0])) if bf16 else s0
    ref = O.reference_step(s, 7, pos=pos, alpha=alpha, color_in=st.color.cpu().numpy(),
                           grad_color_in=st.grad_color.cpu().numpy())
    np.testing.assert_array_equal(st.rast.cpu().numpy().view(np.uint32), ref["rast"].view(np.uint32))
    assert int(st.stats()[0].item()) == sum(
        O.covered_fragments(h, w, pos[b:b + 1], s.tri) for b, (h, w) in enumerate(s.resolutions))
    color = st.color.cpu().numpy()
    err = np.abs(color - ref["color"]).max()
    print(f"  colour max abs err vs fp64: {err:.3e}; loss {loss:.6f} vs {ref['loss']:.6f}", flush=True)
    assert err <= (1e-3 if bf16 else 1e-5)
    assert abs(loss - ref["loss"]) <= (1e-3 if bf16 else 1e-5) * ref["loss"]
    rel = 2e-2 if bf16 else 1e-4
    _close(st.grad_mlp.cpu().numpy()[:-1], ref["grad_mlp"], rel, "MLP grads")
    _close(st.grad_feat.cpu().numpy(), ref["grad_feat"], rel, "feature grads")
    _close(st.grad_verts.cpu().numpy(), ref["grad_verts"], 1e-4, "vertex grads")
    _close(st.grad_alpha_logits.cpu().numpy(), ref["grad_alpha_logits"], 1e-4, "opacity logit grads")


def _view(s, b):
    """The one-view soup of view b (its target slice, its camera)."""
    off = np.cumsum([0] + [h * w for h, w in s.resolutions])
    return dataclasses.replace(s, mvp=s.mvp[b:b + 1], resolutions=[s.resolutions[b]],
                               target=s.target[off[b]:off[b + 1]])


@pytest.mark.parametrize("name", ["uneven", "bf16_large"])
def test_full_size_views(dev, name):
    """Every view of the full-size scene bit-exact with the oracle, and view 0's
    colour against float64 (bf16: on the features as the kernel reads them;
    the end-to-end error against the unrounded model is in
    test_gpu_precision.py)."""
    make = {"uneven": lambda: scenes.batched_uneven_soup(0), "bf16_large": lambda: scenes.bf16_large_soup(0)}[name]
    bf16 = name == "bf16_large"
    s0 = make()
    print(f"{name}: T={s0.T} views={len(s0.resolutions)} pixels={s0.pixels} bf16={bf16}", flush=True)
    st, loss = _step(dev, s0, bf16, seed=11)
    assert np.isfinite(loss)
    pos, alpha = st.pos.cpu().numpy(), st.alpha.cpu().numpy()
    rast = st.rast.cpu().numpy()
    color = st.color.cpu().numpy()
    off = np.cumsum([0] + [h * w for h, w in s0.resolutions])
    # every view's raster is bit-exact with the oracle
    for b, (h, w) in enumerate(s0.resolutions):
        ref = O.rasterize((h, w), pos[b:b + 1], s0.tri, s0.alpha_level, alpha, True, 11, view_base=b)
        np.testing.assert_array_equal(rast[off[b]:off[b + 1]].view(np.uint32), ref.reshape(-1, 4).view(np.uint32),
                                      err_msg=f"view {b}")
    # colour of the first view against float64 (the step's view 0 draws the mask of view 0)
    s = dataclasses.replace(s0, feat=_bf16_rows_as_read(s0.feat, s0.levels[0])) if bf16 else s0
    ref = O.reference_step(_view(s, 0), 11, pos=pos[:1], alpha=alpha)
    err = np.abs(color[:off[1]] - ref["color"]).max()
    print(f"  view 0 colour max abs err vs fp64: {err:.3e}", flush=True)
    assert err <= (1e-3 if bf16 else 1e-5)


def test_stress_full_size_captured(dev):
    """The 300K-triangle stress scene at full size (64 views up to 1920x1080,
    bf16 features and MLP): the autotuned tile, the captured step with the
    optimiser, two replays.  Loss and every gradient finite, the optimiser
    moved the parameters, and the raster of four views spread over the
    batch bit-exact with the oracle under the replay's seed."""
    from diffsoup_amd import fused
    s = scenes.stress_soup(0)
    print(f"stress: T={s.T} views={s.B} pixels={s.pixels}", flush=True)
    mvp = torch.from_numpy(s.mvp).to(dev)
    target = torch.from_numpy(s.target).to(dev)
    st = fused.FusedStep(s.resolutions, s.T, s.verts.shape[0], s.alpha_level, s.levels, s.feat.shape[2],
                         s.mlp_dims[1], bf16=True, bf16_features=True, device=dev)
    params = fused.SoupParams.from_soup(s, dev).with_bf16_features()
    updated = fused.SoupParams.from_soup(s, dev).with_bf16_features()
    tune = st.autotune(params, mvp, target, tiles=(8, 16, 32), reps=1, cache=None)
    print(f"  autotune {tune} -> tile {st.tile}", flush=True)
    opt = fused.soup_optimizer(st, params, out=updated)
    cap = fused.CapturedStep(st, params, mvp, target, opt, seed=500)
    for i in range(2):
        loss = float(cap.replay().item())
        torch.cuda.synchronize()
        print(f"  replay {i}: loss {loss:.6f}", flush=True)
        assert np.isfinite(loss) and loss > 0
    for name in ("grad_feat", "grad_mlp", "grad_verts", "grad_alpha_logits", "color"):
        assert bool(torch.isfinite(getattr(st, name)).all()), name
    assert float((updated.feat - params.feat).abs().max()) > 0
    assert float((updated.verts - params.verts).abs().max()) > 0
    pos, alpha, rast = st.pos.cpu().numpy(), st.alpha.cpu().numpy(), st.rast.cpu().numpy()
    off = np.cumsum([0] + [h * w for h, w in s.resolutions])
    for b in (0, 21, 42, 63):
        h, w = s.resolutions[b]
        ref = O.rasterize((h, w), pos[b:b + 1], s.tri, s.alpha_level, alpha, True, 501, view_base=b)
        np.testing.assert_array_equal(rast[off[b]:off[b + 1]].view(np.uint32), ref.reshape(-1, 4).view(np.uint32),
                                      err_msg=f"view {b}")
        print(f"  view {b} ({h}x{w}) raster bit-exact", flush=True)
