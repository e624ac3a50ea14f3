"""Fused training step: one C-ABI call (ds_train_step) per forward + backward.

The step reproduces the reference's training iteration
(examples/02_synthetic.py:354-388) for a triangle soup with per-level
concatenated multires features, in HIP kernels: projection,
stochastic-opacity rasterization, background shading, gather + pixel MLP +
blend (fp32 or bf16 MFMA), the image loss and its gradient (MSE, or with
``loss="l1_ssim"`` the reference's 0.8 L1 + 0.2 (1 - SSIM) / 2), the MLP
adjoint and feature scatter, edge gradients, the opacity auxiliary gradient
and the world-space vertex gradient.  Views may differ in resolution.

All buffers are allocated once by ``FusedStep`` and reused, so a step
allocates nothing and never synchronises the host: it can be replayed from a
captured HIP graph (``CapturedStep``).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .multires import feats_at_level, feats_in_levels


class StepDesc(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_int), ("H", ctypes.POINTER(ctypes.c_int)), ("W", ctypes.POINTER(ctypes.c_int)),
        ("tile", ctypes.c_int),
        ("V", ctypes.c_int), ("verts", ctypes.c_void_p), ("mvp", ctypes.c_void_p),
        ("T", ctypes.c_int), ("tri", ctypes.c_void_p),
        ("alpha_level", ctypes.c_int), ("alpha_logits", ctypes.c_void_p),
        ("stochastic", ctypes.c_int), ("seed", ctypes.c_uint64),
        ("feat_lo", ctypes.c_int), ("feat_hi", ctypes.c_int), ("channels", ctypes.c_int),
        ("feat", ctypes.c_void_p),
        ("hidden", ctypes.c_int), ("bf16", ctypes.c_int),
        ("w1", ctypes.c_void_p), ("b1", ctypes.c_void_p), ("w2", ctypes.c_void_p), ("b2", ctypes.c_void_p),
        ("w3", ctypes.c_void_p), ("b3", ctypes.c_void_p),
        ("target", ctypes.c_void_p), ("background", ctypes.c_float),
        ("pos", ctypes.c_void_p), ("alpha", ctypes.c_void_p), ("rast", ctypes.c_void_p),
        ("color", ctypes.c_void_p), ("grad_color", ctypes.c_void_p), ("pix_loss", ctypes.c_void_p),
        ("grad_verts", ctypes.c_void_p), ("grad_alpha_logits", ctypes.c_void_p),
        ("grad_feat", ctypes.c_void_p), ("grad_mlp", ctypes.c_void_p), ("loss", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t), ("capacity", ctypes.c_int64),
        ("timed_stage", ctypes.c_int), ("ev_begin", ctypes.c_void_p), ("ev_end", ctypes.c_void_p),
        ("seed_device", ctypes.c_void_p), ("feat_bf16", ctypes.c_void_p),
        ("loss_kind", ctypes.c_int), ("advance_seed", ctypes.c_int),
        ("accumulate", ctypes.c_int), ("alpha_lo", ctypes.c_int), ("sh2", ctypes.c_int),
        ("inv_mvp", ctypes.c_void_p), ("aux_records", ctypes.c_int64),
        ("advance_counter", ctypes.c_void_p),
    ]


LOSS_KINDS = {"mse": 0, "l1_ssim": 1}


STAGES = {"prep": 1, "blobs": 2, "project": 3, "bin": 4, "resolve": 5, "shade": 6, "edge": 7, "aux": 8,
          "final": 9}


_lib.register("ds_step_mlp_params", [ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int64)
_lib.register("ds_step_mlp_params_sh2", [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int64)
_lib.register("ds_step_forward", [ctypes.POINTER(StepDesc), ctypes.c_void_p], ctypes.c_int)
_lib.register("ds_step_backward", [ctypes.POINTER(StepDesc), ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int)
_lib.register("ds_step_workspace_size", [ctypes.POINTER(StepDesc)], ctypes.c_size_t)
_lib.register("ds_train_step", [ctypes.POINTER(StepDesc), ctypes.c_void_p], ctypes.c_int)
_lib.register("ds_autotune_step", [ctypes.POINTER(StepDesc), ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_float), ctypes.c_void_p], ctypes.c_int)


@dataclass
class SoupParams:
    """Trainable tensors of a soup (all fp32 on the GPU)."""
    verts: torch.Tensor          # [V,3]
    tri: torch.Tensor            # [T,3] int32 (not trained)
    alpha_logits: torch.Tensor   # [T,S_a]
    feat: torch.Tensor           # [T,S_tot,C]
    w1: torch.Tensor
    b1: torch.Tensor
    w2: torch.Tensor
    b2: torch.Tensor
    w3: torch.Tensor
    b3: torch.Tensor
    feat_bf16: torch.Tensor | None = None   # bf16 copy of feat read by steps with bf16 features

    @staticmethod
    def from_soup(s, dev) -> "SoupParams":
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        (w1, b1), (w2, b2), (w3, b3) = s.mlp
        return SoupParams(t(s.verts), t(s.tri), t(s.alpha_logits), t(s.feat), t(w1), t(b1), t(w2), t(b2),
                          t(w3), t(b3))

    def tensors(self):
        return [self.verts, self.alpha_logits, self.feat, self.w1, self.b1, self.w2, self.b2, self.w3, self.b3]

    def with_bf16_features(self) -> "SoupParams":
        """Attach (or refresh) the bf16 copy of the features, rounded to
        nearest even from the fp32 master."""
        self.feat_bf16 = self.feat.detach().to(torch.bfloat16).contiguous()
        return self


class FusedStep:
    """Pre-allocated buffers and workspace for ds_train_step on one set of views.

    Two feature compositions:

    * default - the features of ``levels`` (lo, hi) are concatenated into the
      MLP input, opacities are sigmoid(alpha_logits) at ``alpha_level``;
    * ``accumulate=True`` - the reference's own training iteration
      (examples/02_synthetic.py:354-364): ``feat`` [T, S(lo..hi), C] and
      ``alpha_logits`` [T, S(alpha_lo..alpha_level)] are multires sources,
      the step evaluates sigmoid(accumulate_to_level(...)) of both at the
      top level, gathers the C features there, and the gradients are those
      of the sources.

    ``sh2=True`` appends the 9 SH2 values of each pixel's view direction to
    the MLP input (``run`` then takes ``inv_mvp``); the reference's network
    is ``accumulate=True, channels=7, hidden=16, sh2=True``.

    ``aux_records`` is the number of fragment records per 8x8 pixel quad the
    workspace keeps for the opacity gradient (0: 128, two per pixel); a quad
    whose fragments overflow them computes that gradient by a second walk of
    its tile's bins (the same result, slower)."""

    def __init__(self, resolutions, T, V, alpha_level, levels, channels, hidden, *, bf16=False,
                 bf16_features=False, stochastic=True, tile=16, background=1.0, device="cuda", capacity=None,
                 loss="mse", accumulate=False, alpha_lo=None, sh2=False, aux_records=0):
        self.dev = torch.device(device)
        self.res = [(int(h), int(w)) for h, w in resolutions]
        self.B = len(self.res)
        self.P = sum(h * w for h, w in self.res)
        self.T, self.V = int(T), int(V)
        self.alpha_level = alpha_level
        self.lo, self.hi = levels
        self.C, self.hidden = channels, hidden
        self.bf16, self.stochastic, self.tile, self.background = bf16, stochastic, tile, background
        self.bf16_features = bool(bf16_features)
        if loss not in LOSS_KINDS:
            raise ValueError(f"loss must be one of {sorted(LOSS_KINDS)}")
        self.loss_kind = loss
        self.accumulate, self.sh2 = bool(accumulate), bool(sh2)
        self.alpha_lo = alpha_level if alpha_lo is None else int(alpha_lo)
        if self.accumulate and self.bf16_features:
            raise ValueError("accumulated features are fp32: bf16_features needs accumulate=False")
        self.Sa = feats_at_level(alpha_level)
        self.Sa_src = feats_in_levels(self.alpha_lo, alpha_level) if self.accumulate else self.Sa
        self.S_tot = feats_in_levels(self.lo, self.hi)
        self.nl = 1 if self.accumulate else self.hi - self.lo + 1
        self.din = self.nl * channels + (9 if self.sh2 else 0)
        L = _lib.lib()
        self.nparam = int(L.ds_step_mlp_params_sh2(self.nl, channels, hidden, int(self.sh2)))
        f = lambda *shape: torch.empty(*shape, dtype=torch.float32, device=self.dev)
        self.pos = f(self.B, self.V, 4)
        self.alpha = f(self.T, self.Sa)
        self.rast = f(self.P, 4)
        self.color = f(self.P, 3)
        self.grad_color = f(self.P, 3)
        self.pix_loss = f(self.P)
        self.grad_verts = f(self.V, 3)
        self.grad_alpha_logits = f(self.T, self.Sa_src)
        self.grad_feat = f(self.T, self.S_tot, self.C)
        self.grad_mlp = f(self.nparam)
        self.loss = f(1)
        self._Ha = _lib.int_array([h for h, _ in self.res])
        self._Wa = _lib.int_array([w for _, w in self.res])
        self.capacity = int(capacity or max(4 * self.B * self.T, 1 << 16))
        self.aux_records = int(aux_records)
        self.desc = StepDesc()
        self._fill_static()
        nbytes = int(L.ds_step_workspace_size(ctypes.byref(self.desc)))
        if nbytes == 0:
            raise _lib.DiffSoupError("ds_step_workspace_size: " + L.ds_last_error().decode())
        self._ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=self.dev)
        self.desc.workspace = (self._ws.data_ptr() + 255) & ~255
        self.desc.workspace_bytes = nbytes
        self._captured = False   # set by CapturedStep: workspace and tile are then fixed
        self._generation = 0     # forwards run through render()

    def _fill_static(self):
        d = self.desc
        d.B, d.H, d.W, d.tile = self.B, self._Ha, self._Wa, self.tile
        d.V, d.T = self.V, self.T
        d.alpha_level = self.alpha_level
        d.stochastic = int(bool(self.stochastic))
        d.feat_lo, d.feat_hi, d.channels = self.lo, self.hi, self.C
        d.hidden, d.bf16 = self.hidden, int(bool(self.bf16))
        d.background = self.background
        d.loss_kind = LOSS_KINDS[self.loss_kind]
        d.accumulate, d.alpha_lo, d.sh2 = int(self.accumulate), self.alpha_lo, int(self.sh2)
        for name in ("pos", "alpha", "rast", "color", "grad_color", "pix_loss", "grad_verts",
                     "grad_alpha_logits", "grad_feat", "grad_mlp", "loss"):
            setattr(d, name, getattr(self, name).data_ptr())
        d.capacity = self.capacity
        d.aux_records = self.aux_records

    def mlp_grads(self):
        """Views of grad_mlp as (dW1, db1, dW2, db2, dW3, db3)."""
        DH, DIN = self.hidden, self.din
        g = self.grad_mlp
        o = 0
        dW1 = g[o:o + DH * DIN].view(DH, DIN); o += DH * DIN
        dW2 = g[o:o + DH * DH].view(DH, DH); o += DH * DH
        dW3 = g[o:o + 3 * DH].view(3, DH); o += 3 * DH
        db1 = g[o:o + DH]; o += DH
        db2 = g[o:o + DH]; o += DH
        db3 = g[o:o + 3]
        return dW1, db1, dW2, db2, dW3, db3

    def _bind(self, params: SoupParams, mvp: torch.Tensor, target: torch.Tensor, seed: int,
              inv_mvp: torch.Tensor | None = None) -> None:
        assert mvp.shape == (self.B, 4, 4) and target.shape == (self.P, 3)
        for t in params.tensors() + [mvp, target]:
            assert t.is_contiguous() and t.dtype == torch.float32 and t.device == self.dev
        assert params.feat.shape == (self.T, self.S_tot, self.C), "feat shape"
        assert params.alpha_logits.shape == (self.T, self.Sa_src), "alpha_logits shape"
        if self.sh2:
            if inv_mvp is None:
                raise ValueError("sh2: pass inv_mvp [B,4,4]")
            assert inv_mvp.shape == (self.B, 4, 4) and inv_mvp.dtype == torch.float32 and inv_mvp.is_contiguous()
            assert inv_mvp.device == self.dev
            self.desc.inv_mvp = inv_mvp.data_ptr()
        else:
            self.desc.inv_mvp = None
        d = self.desc
        d.verts, d.mvp, d.tri = params.verts.data_ptr(), mvp.data_ptr(), params.tri.data_ptr()
        d.alpha_logits, d.feat = params.alpha_logits.data_ptr(), params.feat.data_ptr()
        d.w1, d.b1, d.w2, d.b2 = params.w1.data_ptr(), params.b1.data_ptr(), params.w2.data_ptr(), params.b2.data_ptr()
        d.w3, d.b3 = params.w3.data_ptr(), params.b3.data_ptr()
        d.target = target.data_ptr()
        d.seed = int(seed) & (2**64 - 1)
        if self.bf16_features:
            fb = params.feat_bf16
            if fb is None or fb.dtype != torch.bfloat16 or fb.shape != params.feat.shape or not fb.is_contiguous():
                raise ValueError("bf16_features: params.feat_bf16 must be a contiguous bfloat16 copy of feat "
                                 "(SoupParams.with_bf16_features)")
            d.feat_bf16 = fb.data_ptr()
        else:
            d.feat_bf16 = None

    def run(self, params: SoupParams, mvp: torch.Tensor, target: torch.Tensor, seed: int,
            inv_mvp: torch.Tensor | None = None) -> torch.Tensor:
        """Enqueue one forward + backward; returns the device loss tensor."""
        self._bind(params, mvp, target, seed, inv_mvp)
        _lib.check(_lib.lib().ds_train_step(ctypes.byref(self.desc), _lib.stream_ptr(self.dev)), "ds_train_step")
        return self.loss

    def forward(self, params: SoupParams, mvp: torch.Tensor, target: torch.Tensor, seed: int,
                inv_mvp: torch.Tensor | None = None) -> torch.Tensor:
        """Enqueue the forward up to the colour (ds_step_forward); returns the
        step's colour buffer [P, 3] (the target is read only by the opacity
        term of the backward)."""
        self._bind(params, mvp, target, seed, inv_mvp)
        _lib.check(_lib.lib().ds_step_forward(ctypes.byref(self.desc), _lib.stream_ptr(self.dev)),
                   "ds_step_forward")
        return self.color

    def backward(self, grad_color: torch.Tensor) -> None:
        """Enqueue the backward of the last ``forward`` on dLoss/dcolor [P, 3]
        (ds_step_backward): every parameter gradient lands in the step's
        gradient buffers."""
        assert grad_color.shape == (self.P, 3) and grad_color.dtype == torch.float32
        assert grad_color.is_contiguous() and grad_color.device == self.dev
        _lib.check(_lib.lib().ds_step_backward(ctypes.byref(self.desc), grad_color.data_ptr(),
                                               _lib.stream_ptr(self.dev)), "ds_step_backward")

    def workspace_bytes(self, tile: int) -> int:
        d = self.desc
        old = d.tile
        d.tile = tile
        n = int(_lib.lib().ds_step_workspace_size(ctypes.byref(d)))
        d.tile = old
        return n

    def _ensure_workspace(self, nbytes: int) -> None:
        if nbytes > self.desc.workspace_bytes:
            if self._captured:
                raise _lib.DiffSoupError(
                    "a CapturedStep replays this step's workspace and tile: the workspace cannot grow after "
                    "capture (choose the tile before capturing, or build a new FusedStep)")
            self._ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=self.dev)
            self.desc.workspace = (self._ws.data_ptr() + 255) & ~255
            self.desc.workspace_bytes = nbytes

    def set_tile(self, tile: int) -> None:
        if self._captured and int(tile) != self.tile:
            raise _lib.DiffSoupError("the tile of a captured step is fixed: the graph replays the tile it recorded")
        n = self.workspace_bytes(tile)
        if n == 0:
            raise _lib.DiffSoupError(f"tile {tile}: " + _lib.lib().ds_last_error().decode())
        self._ensure_workspace(n)
        self.tile = self.desc.tile = int(tile)

    def tune_key(self) -> str:
        """Cache key of the tile choice: device, views, soup size and shading
        configuration (the quantities the best tile depends on)."""
        name = torch.cuda.get_device_name(self.dev).replace(" ", "_")
        res = ",".join(f"{h}x{w}" for h, w in sorted(self.res))
        return (f"{name}|B{self.B}|{res}|T{self.T}|V{self.V}|a{self.alpha_level}|l{self.lo}-{self.hi}|C{self.C}"
                f"|h{self.hidden}|{'bf16' if self.bf16 else 'fp32'}|{'f16feat' if self.bf16_features else 'f32feat'}"
                f"|{'acc' if self.accumulate else 'concat'}{self.alpha_lo}|{'sh2' if self.sh2 else 'nosh'}")

    def autotune(self, params: SoupParams, mvp: torch.Tensor, target: torch.Tensor, tiles=(8, 16, 32),
                 reps: int = 5, seed: int = 0, cache: str | None = "default",
                 inv_mvp: torch.Tensor | None = None) -> dict:
        """Pick the screen tile edge by timing whole steps (ds_autotune_step),
        and remember the choice in a JSON cache keyed by ``tune_key()``
        (``cache="default"``: $DIFFSOUP_AMD_TUNE_CACHE or
        ~/.cache/diffsoup_amd/tiles.json; None: no cache).  Returns
        {tile: ms per step} (from the cache: {tile: None} for the cached
        choice).  The timed steps write this step's outputs."""
        import json
        import os

        path = None
        if cache == "default":
            path = os.environ.get("DIFFSOUP_AMD_TUNE_CACHE",
                                  os.path.join(os.path.expanduser("~"), ".cache", "diffsoup_amd", "tiles.json"))
        elif cache:
            path = cache
        key = self.tune_key()
        table = {}
        if path and os.path.exists(path):
            try:
                with open(path) as fh:
                    table = json.load(fh)
            except (OSError, ValueError):
                table = {}
        if self._captured:
            raise _lib.DiffSoupError("autotune on a captured step: tune before capturing")
        if key in table and int(table[key]) in tiles:
            self.set_tile(int(table[key]))
            return {int(table[key]): None}
        self._ensure_workspace(max(self.workspace_bytes(t) for t in tiles))
        self._bind(params, mvp, target, seed, inv_mvp)
        arr = (ctypes.c_int * len(tiles))(*tiles)
        ms = (ctypes.c_float * len(tiles))()
        _lib.check(_lib.lib().ds_autotune_step(ctypes.byref(self.desc), arr, len(tiles), int(reps), ms,
                                               _lib.stream_ptr(self.dev)), "ds_autotune_step")
        self.tile = int(self.desc.tile)
        result = {int(t): (float(m) if m >= 0 else None) for t, m in zip(tiles, ms)}
        if path:
            table[key] = self.tile
            os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
            tmp = path + ".tmp"
            with open(tmp, "w") as fh:
                json.dump(table, fh, indent=1, sort_keys=True)
            os.replace(tmp, path)
        return result

    def time_stage(self, stage, ev_begin=None, ev_end=None):
        """Bracket one stage of the next run() with two torch.cuda.Events
        (recorded on the launch stream around that stage's kernels)."""
        d = self.desc
        if stage is None:
            d.timed_stage, d.ev_begin, d.ev_end = 0, None, None
            return
        # make sure the events exist before handing their handles to the library
        ev_begin.record()
        ev_end.record()
        d.timed_stage = STAGES[stage]
        d.ev_begin, d.ev_end = ev_begin.cuda_event, ev_end.cuda_event

    def covered_pixels(self) -> int:
        return int((self.rast[:, 3] > 0).sum().item())

    def algorithmic_bytes(self, stage: str) -> float:
        """Compulsory HBM bytes of one launch of a stage (the roofline numerator).

        resolve: rast written for every pixel (16 B), background pixels read the
        target and write colour, colour gradient and L1 (40 B), covered pixels
        append their index (4 B), every (view, triangle) record is read once
        (64 B) and every bin entry once (4 B), and every covered fragment is
        listed for the opacity gradient (16 B; the fragments found behind
        their pixel's surface during the walk, a few per cent, are not, so
        this counts slightly high).
        shade: per covered pixel 60 B (rast, index, target, colour, colour
        gradient, L1) plus one read of the feature table (fp32 or bf16) and
        one write of its fp32 gradient.
        edge: per covered pixel its list record (16 B) and its rast, colour
        and colour gradient (40 B), each read once (a neighbour's values are
        another pixel's own reads), plus the clip-position gradient (16 B
        per view-vertex).
        aux: per listed fragment (counted as every covered fragment) its
        record (16 B) and its pixel's rast and L1 (20 B), plus one read of the
        opacities and one write of their gradient (8 B per lattice sample).
        """
        P = self.P
        cov = self.covered_pixels()
        stats = self.stats()
        pairs = int(stats[1].item())
        frags = int(stats[0].item())
        n_feat = float(self.T * self.S_tot * self.C)
        feat_read = (2.0 if self.bf16_features else 4.0) * n_feat
        if stage == "resolve":
            return (16.0 * P + 40.0 * (P - cov) + 4.0 * cov + 64.0 * self.B * self.T + 4.0 * pairs
                    + 16.0 * frags)
        if stage == "shade":
            return 60.0 * cov + feat_read + 4.0 * n_feat
        if stage == "edge":
            return (16.0 + 40.0) * cov + 16.0 * self.B * self.V
        if stage == "aux":
            return 36.0 * frags + 8.0 * self.T * self.Sa
        raise ValueError(stage)

    def iteration_bytes(self, adam: bool = True) -> float:
        """Compulsory HBM bytes of one training iteration: forward (clip
        positions, triangles, rast, feature table, opacities, colour) and
        backward (target, rast, feature, opacity and position gradients),
        plus 28 B per parameter for Adam / VectorAdam (p, g, m, v read;
        p, m, v written) over features, opacity logits, MLP and vertices."""
        P, T, B, V = float(self.P), float(self.T), float(self.B), float(self.V)
        n_feat = T * self.S_tot * self.C
        n_alpha = T * self.Sa
        s_f = 2.0 if self.bf16_features else 4.0
        fwd = B * V * 16 + T * 12 + P * 16 + n_feat * s_f + n_alpha * 4 + P * 12
        bwd = P * 12 + P * 16 + n_feat * 4 + n_alpha * 4 + B * V * 16
        opt = 28.0 * (n_feat + n_alpha + self.nparam + 3 * V) if adam else 0.0
        return fwd + bwd + opt

    def algorithmic_flops(self, stage: str) -> float:
        """Compulsory floating-point work of one launch of the shading stage.

        Per covered pixel: the MLP forward (din*h + h*h + 3*h multiply-adds),
        the input adjoint (the same count again) and the weight gradients
        (the same count a third time), 2 flops per multiply-add; plus the
        feature gather and its scatter (3 corners x channels per level, each
        way).  With din = 32, h = 32, 4 levels x 8 channels: 13,248 flops.
        """
        if stage != "shade":
            raise ValueError(stage)
        din, h = self.din, self.hidden
        per_px = 6.0 * (din * h + h * h + 3 * h) + 4.0 * 3 * din
        return per_px * self.covered_pixels()

    def stats(self) -> torch.Tensor:
        """int64 raster stats of the last step (DS_STAT_* order)."""
        base = self.desc.workspace - self._ws.data_ptr()
        return self._ws[base:base + 64].view(torch.int64)


class _RenderFn(torch.autograd.Function):
    """The fused forward up to the colour and its backward (the two-pass
    shading around any caller-side loss)."""

    @staticmethod
    def forward(ctx, step, tri, mvp, target, seed, inv_mvp, verts, alpha_logits, feat, w1, b1, w2, b2, w3, b3):
        params = SoupParams(verts.detach(), tri, alpha_logits.detach(), feat.detach(), w1.detach(), b1.detach(),
                            w2.detach(), b2.detach(), w3.detach(), b3.detach())
        color = step.forward(params, mvp, target, seed, inv_mvp).clone()
        step._generation += 1
        ctx.step, ctx.generation = step, step._generation
        return color

    @staticmethod
    def backward(ctx, grad_color):
        step = ctx.step
        if step._generation != ctx.generation:
            raise RuntimeError("render: the FusedStep ran another forward before this backward "
                               "(one forward in flight per step)")
        step.backward(grad_color.contiguous())
        dW1, db1, dW2, db2, dW3, db3 = (g.clone() for g in step.mlp_grads())
        return (None, None, None, None, None, None, step.grad_verts.clone(), step.grad_alpha_logits.clone(),
                step.grad_feat.clone(), dW1, db1, dW2, db2, dW3, db3)


def render(step: FusedStep, params: SoupParams, mvp: torch.Tensor, target: torch.Tensor, seed: int,
           inv_mvp: torch.Tensor | None = None) -> torch.Tensor:
    """Differentiable colour [P, 3] (views packed) of the fused step, for a
    loss composed by the caller: ``loss_fn(render(...)).backward()`` fills
    ``.grad`` of params.verts, alpha_logits, feat and the MLP tensors.

    The forward is ds_step_forward (projection, stochastic-opacity raster,
    gather, MLP, blend, white background); the backward is ds_step_backward
    on the incoming colour gradient: MLP adjoint, feature scatter, edge
    gradients (the reference's edge_grad) and the opacity term of its
    opacity_aux_loss (weight 1, per-pixel L1 against ``target``), as in the
    reference's composition (examples/02_synthetic.py:354-385) whose loss
    adds ``aux_loss``.  One forward in flight per step: call backward before
    the step's next forward."""
    return _RenderFn.apply(step, params.tri, mvp, target, seed, inv_mvp, params.verts, params.alpha_logits,
                           params.feat, params.w1, params.b1, params.w2, params.b2, params.w3, params.b3)


def soup_optimizer(step: FusedStep, params: SoupParams, lr_feat=5e-2, lr_alpha=5e-2, lr_verts=1e-3,
                   lr_mlp=1e-2, out: SoupParams | None = None, decay=None):
    """The reference's optimisers for a soup (examples/02_synthetic.py:310-315):
    Adam on features and opacity logits (lr 5e-2) and on the MLP (1e-2),
    VectorAdam on vertex positions (1e-3), reading the gradients that
    ``step`` writes.  One FusedAdam, so one launch per iteration.  With
    ``out`` the updated parameters are written there and ``params`` stays
    unchanged (every iteration then sees the same scene, as a throughput
    measurement of a fixed configuration needs).  When the parameters carry a
    bf16 feature copy, the update refreshes it in the same launch.  ``decay``
    = (steps, final) applies the reference's learning-rate schedule
    exp_decay_mult (examples/02_synthetic.py:341-349) on the device, so a
    captured iteration follows it replay after replay."""
    from .optimize import FusedAdam

    o = out
    fb = (o if o is not None else params).feat_bf16
    groups = [dict(param=params.feat, grad=step.grad_feat, lr=lr_feat, param_out=o and o.feat, param_bf16_out=fb),
              dict(param=params.alpha_logits, grad=step.grad_alpha_logits, lr=lr_alpha,
                   param_out=o and o.alpha_logits),
              dict(param=params.verts, grad=step.grad_verts, lr=lr_verts, vector_dim=3, param_out=o and o.verts)]
    dW1, db1, dW2, db2, dW3, db3 = step.mlp_grads()
    for name, g in (("w1", dW1), ("b1", db1), ("w2", dW2), ("b2", db2), ("w3", dW3), ("b3", db3)):
        groups.append(dict(param=getattr(params, name), grad=g, lr=lr_mlp, param_out=o and getattr(o, name)))
    return FusedAdam(groups, decay=decay)


class CapturedStep:
    """One training iteration captured in a HIP graph: ds_train_step with the
    mask seed read from device memory and advanced by the step itself, and
    the optimiser step.  Replay i after ``set_seed(s)`` uses seed s + i.

    The graph owns the kernel arguments of the capture: the tensors in
    ``params``, ``mvp``, ``target`` and the step's buffers must stay alive
    and in place (update them with copy_, not by rebinding).  The capture
    holds the step's workspace and fixes its tile: set_tile and autotune on
    the step raise afterwards.
    """

    def __init__(self, step: FusedStep, params: SoupParams, mvp: torch.Tensor, target: torch.Tensor,
                 optimizer=None, seed: int = 0, inv_mvp: torch.Tensor | None = None):
        self.step = step
        self.seed_dev = torch.zeros(1, dtype=torch.int64, device=step.dev)
        # one eager step first: it loads the code objects and sets kernel
        # attributes, which must not happen during capture; it writes only
        # the step's own outputs, not the parameters
        step.run(params, mvp, target, seed, inv_mvp)
        torch.cuda.synchronize(step.dev)
        step.desc.seed_device = self.seed_dev.data_ptr()
        step.desc.advance_seed = 1
        # a FusedAdam's step counter is advanced by the step's last kernel,
        # and the update reads it as advanced: no launch for the counter
        fused_counter = getattr(optimizer, "step_count", None) is not None and hasattr(optimizer, "step_advanced")
        if fused_counter:
            step.desc.advance_counter = optimizer.step_count.data_ptr()
        self.graph = torch.cuda.CUDAGraph()
        self._workspace = step._ws   # the replays write into this allocation
        step._captured = True
        try:
            with torch.cuda.graph(self.graph):
                step.run(params, mvp, target, 0, inv_mvp)
                if optimizer is not None:
                    if fused_counter:
                        optimizer.step_advanced()
                    else:
                        optimizer.step()
        finally:
            step.desc.seed_device = None
            step.desc.advance_seed = 0
            step.desc.advance_counter = None
        self.set_seed(seed)

    def set_seed(self, seed: int) -> None:
        """Seed of the next replay: all 64 bits, as FusedStep.run uses them
        (stored as the two's-complement int64 of the unsigned seed)."""
        v = int(seed) & (2**64 - 1)
        self.seed_dev.fill_(v - 2**64 if v >= 2**63 else v)

    def replay(self) -> torch.Tensor:
        self.graph.replay()
        return self.step.loss
