"""Precision report of the bf16 step at the 100K-triangle, 16 x 1920x1080
configuration (5 levels x 16 bf16 channels, MLP 80->64->64->3 on bf16 MFMA
with fp32 accumulation; depth, barycentrics and edge terms fp32): image and
gradient error against the float64 oracle on 4 of the 16 views, at full
resolution with the full soup.

The step runs on views 0-3 (each view's mask is keyed by its index, so these
are the same four views as in the 16-view batch).  Two references:

* the float64 model on the unrounded fp32 features - the end-to-end error
  of the bf16 step: colour within 1e-3 (the image-error bar), MLP and
  feature gradients within 3e-2 and vertex and opacity gradients within
  2e-2 of their largest magnitude;
* the float64 model on the features as the kernel reads them (bf16 rows,
  fp32 residual-blend channels) with the step's own colour and colour
  gradient fed to the edge and opacity estimators - the kernels' own
  arithmetic: colour 1e-3, MLP and feature gradients 2e-2, vertex and
  opacity gradients 1e-4.

Every number is printed (run pytest with -s to see the report)."""
import dataclasses

import numpy as np
import pytest
import torch

from diffsoup_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(got, ref):
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-30))


def _bf16_rows_as_read(feat, lo):
    out = torch.from_numpy(np.ascontiguousarray(feat)).to(torch.bfloat16).to(torch.float32).numpy()
    s_lo = O.feats_at_level(lo)
    out[:, :s_lo, :4] = feat[:, :s_lo, :4]
    return out


def test_bf16_100k_four_views_against_fp64(dev):
    from diffsoup_amd import fused
    full = scenes.bf16_large_soup(0)
    nv = 4
    off = np.cumsum([0] + [h * w for h, w in full.resolutions])
    s0 = dataclasses.replace(full, mvp=full.mvp[:nv], resolutions=full.resolutions[:nv], target=full.target[:off[nv]])
    print(f"bf16 100K: T={s0.T} views 0-{nv - 1} of {full.B} at {s0.resolutions[0]}, levels {s0.levels}, "
          f"C={s0.feat.shape[2]}, MLP {s0.mlp_dims}", flush=True)
    del full
    prm = fused.SoupParams.from_soup(s0, dev).with_bf16_features()
    st = fused.FusedStep(s0.resolutions, s0.T, s0.verts.shape[0], s0.alpha_level, s0.levels, s0.feat.shape[2],
                         s0.mlp_dims[1], bf16=True, bf16_features=True, device=dev)
    seed = 11
    loss = float(st.run(prm, torch.from_numpy(s0.mvp).to(dev), torch.from_numpy(s0.target).to(dev), seed).item())
    torch.cuda.synchronize()
    g = {k: getattr(st, k).cpu().numpy() for k in ("pos", "alpha", "rast", "color", "grad_color", "grad_feat",
                                                   "grad_mlp", "grad_verts", "grad_alpha_logits")}
    assert np.isfinite(loss)
    print(f"  step done: loss {loss:.6f}, covered pixels {int((g['rast'][:, 3] > 0).sum())}", flush=True)

    # end-to-end: float64 model on the unrounded features, its own colour
    ref = O.reference_step(s0, seed, pos=g["pos"], alpha=g["alpha"])
    np.testing.assert_array_equal(g["rast"].view(np.uint32), ref["rast"].view(np.uint32))
    cov = ref["covered"]
    e2e = {"colour_abs": float(np.abs(g["color"] - ref["color"]).max()),
           "loss_rel": abs(loss - ref["loss"]) / ref["loss"],
           "grad_mlp": _rel(g["grad_mlp"][:-1], ref["grad_mlp"]),
           "grad_feat": _rel(g["grad_feat"], ref["grad_feat"]),
           "grad_verts": _rel(g["grad_verts"], ref["grad_verts"]),
           "grad_alpha_logits": _rel(g["grad_alpha_logits"], ref["grad_alpha_logits"])}
    print("  vs float64 on unrounded features: " + ", ".join(f"{k} {v:.3e}" for k, v in e2e.items()), flush=True)
    del ref

    # kernel arithmetic: the features as read, the step's own image into the estimators
    s_read = dataclasses.replace(s0, feat=_bf16_rows_as_read(s0.feat, s0.levels[0]))
    ref = O.reference_step(s_read, seed, pos=g["pos"], alpha=g["alpha"], color_in=g["color"],
                           grad_color_in=g["grad_color"])
    ker = {"colour_abs": float(np.abs(g["color"] - ref["color"]).max()),
           "grad_mlp": _rel(g["grad_mlp"][:-1], ref["grad_mlp"]),
           "grad_feat": _rel(g["grad_feat"], ref["grad_feat"]),
           "grad_verts": _rel(g["grad_verts"], ref["grad_verts"]),
           "grad_alpha_logits": _rel(g["grad_alpha_logits"], ref["grad_alpha_logits"])}
    print("  vs float64 on the features as read: " + ", ".join(f"{k} {v:.3e}" for k, v in ker.items()), flush=True)
    assert int(cov.sum()) > 0

    assert e2e["colour_abs"] <= 1e-3
    assert e2e["loss_rel"] <= 1e-3
    assert e2e["grad_mlp"] <= 3e-2 and e2e["grad_feat"] <= 3e-2
    assert e2e["grad_verts"] <= 2e-2 and e2e["grad_alpha_logits"] <= 2e-2
    assert ker["colour_abs"] <= 1e-3
    assert ker["grad_mlp"] <= 2e-2 and ker["grad_feat"] <= 2e-2
    assert ker["grad_verts"] <= 1e-4 and ker["grad_alpha_logits"] <= 1e-4
