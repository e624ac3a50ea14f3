"""Optimisers.

VectorAdam is the reference's torch optimiser (python/diffsoup/optimize.py:7-68;
Ling, Sharp & Jacobson, NeurIPS 2022): Adam with the second moment pooled
over the last dimension.  FusedAdam applies torch.optim.Adam's update and
VectorAdam's to every parameter group of a training step in one HIP launch.

VectorAdam's update, as in the reference:
    g1 <- b1 g1 + (1 - b1) g
    g2 <- b2 g2 + (1 - b2) sum_last(g^2)
    p  <- p - lr * (g1 / (1 - b1^t)) / (sqrt(g2 / (1 - b2^t)) + 1e-8)
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


class VectorAdam(torch.optim.Optimizer):
    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999)):
        super().__init__(params, dict(lr=lr, betas=betas))

    @torch.no_grad()
    def step(self):
        for group in self.param_groups:
            lr = group["lr"]
            b1, b2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                grad = p.grad
                state = self.state[p]
                if len(state) == 0:
                    state["step"] = 0
                    state["g1"] = torch.zeros_like(p)
                    state["g2"] = torch.zeros_like(p[..., :1])
                state["step"] += 1
                g1, g2 = state["g1"], state["g2"]
                g1.mul_(b1).add_(grad, alpha=1 - b1)
                g2.mul_(b2).add_(grad.square().sum(dim=-1, keepdim=True), alpha=1 - b2)
                t = state["step"]
                m1 = g1 / (1 - b1**t)
                m2 = g2 / (1 - b2**t)
                p.sub_(m1 / (m2.sqrt() + 1e-8), alpha=lr)


class AdamGroup(ctypes.Structure):
    """ds_adam_group of include/diffsoup_hip.h."""
    _fields_ = [("param", ctypes.c_void_p), ("grad", ctypes.c_void_p), ("exp_avg", ctypes.c_void_p),
                ("exp_avg_sq", ctypes.c_void_p), ("n", ctypes.c_int64), ("vector_dim", ctypes.c_int),
                ("lr", ctypes.c_double), ("beta1", ctypes.c_double), ("beta2", ctypes.c_double),
                ("eps", ctypes.c_double), ("param_out", ctypes.c_void_p), ("param_bf16_out", ctypes.c_void_p),
                ("lr_mult", ctypes.c_void_p), ("decay_steps", ctypes.c_int64), ("decay_final", ctypes.c_double)]


_lib.register("ds_adam_step", [ctypes.POINTER(AdamGroup), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p])
_lib.register("ds_adam_step_advanced", [ctypes.POINTER(AdamGroup), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p])


class FusedAdam:
    """Adam and VectorAdam over several (parameter, gradient) pairs in one
    HIP launch (ds_adam_step).

    ``groups`` is a list of dicts with keys ``param`` and ``grad`` (fp32 GPU
    tensors of the same shape, contiguous), and optionally ``lr`` (1e-3),
    ``betas`` ((0.9, 0.999)), ``eps`` (1e-8) and ``vector_dim`` (0 for Adam;
    k > 0 pools the second moment over the last dimension of width k, as
    VectorAdam does) and ``param_out`` (None: update in place; else a tensor
    of the parameter's shape that receives the updated values while
    ``param`` stays as it is) and ``param_bf16_out`` (None, or a bfloat16
    tensor of the parameter's shape that also receives the updated values
    rounded to bf16).  Gradients are read from ``grad`` (for example
    the buffers a fused training step writes), not from ``param.grad``.  The step
    counter lives in device memory, so a captured graph advances it.

    Learning-rate schedules that a captured graph follows: ``lr_mult``
    (a one-element fp32 device tensor shared by every group, read when the
    update runs: write it with ``set_lr_mult`` between replays) and
    ``decay=(steps, final)``, the reference's exp_decay_mult
    (examples/utils.py:435-442): iteration i (= the device step counter + 1)
    uses lr * final ** (min(max(i, 1), steps) / steps), formed on the device.
    The base ``lr`` itself is baked into each launch (``set_lr`` needs a
    re-capture).
    """

    def __init__(self, groups, lr_mult=False, decay=None):
        if not 1 <= len(groups) <= 16:
            raise ValueError("FusedAdam: 1..16 parameter groups")
        self.groups = []
        for g in groups:
            p, gr = g["param"], g["grad"]
            k = int(g.get("vector_dim", 0))
            for t, name in ((p, "param"), (gr, "grad")):
                _lib.require_device(t, name)
                if t.dtype != torch.float32 or not t.is_contiguous():
                    raise TypeError(f"FusedAdam: {name} must be a contiguous fp32 tensor")
            if p.shape != gr.shape:
                raise ValueError("FusedAdam: param and grad shapes differ")
            if k and p.shape[-1] != k:
                raise ValueError("FusedAdam: vector_dim must equal the last dimension")
            po = g.get("param_out")
            if po is not None:
                _lib.require_device(po, "param_out")
                if po.dtype != torch.float32 or not po.is_contiguous() or po.shape != p.shape:
                    raise TypeError("FusedAdam: param_out must be a contiguous fp32 tensor of the parameter's shape")
            m = torch.zeros_like(p)
            v = torch.zeros(p.numel() // k if k else p.numel(), dtype=torch.float32, device=p.device)
            b1, b2 = g.get("betas", (0.9, 0.999))
            pb = g.get("param_bf16_out")
            if pb is not None:
                _lib.require_device(pb, "param_bf16_out")
                if pb.dtype != torch.bfloat16 or not pb.is_contiguous() or pb.shape != p.shape:
                    raise TypeError("FusedAdam: param_bf16_out must be a contiguous bfloat16 tensor of the "
                                    "parameter's shape")
            self.groups.append(dict(param=p, grad=gr, param_out=po, param_bf16_out=pb, exp_avg=m, exp_avg_sq=v,
                                    vector_dim=k,
                                    lr=float(g.get("lr", 1e-3)), betas=(float(b1), float(b2)),
                                    eps=float(g.get("eps", 1e-8))))
        self.device = self.groups[0]["param"].device
        self.step_count = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.lr_mult = torch.ones(1, dtype=torch.float32, device=self.device) if lr_mult else None
        if decay is not None:
            steps, final = int(decay[0]), float(decay[1])
            if steps < 1 or not final > 0:
                raise ValueError("FusedAdam: decay=(steps >= 1, final > 0)")
            self.decay = (steps, final)
        else:
            self.decay = None
        self._build()

    def _build(self):
        arr = (AdamGroup * len(self.groups))()
        for a, g in zip(arr, self.groups):
            a.param, a.grad = g["param"].data_ptr(), g["grad"].data_ptr()
            a.exp_avg, a.exp_avg_sq = g["exp_avg"].data_ptr(), g["exp_avg_sq"].data_ptr()
            a.n, a.vector_dim = g["param"].numel(), g["vector_dim"]
            a.lr, (a.beta1, a.beta2), a.eps = g["lr"], g["betas"], g["eps"]
            a.param_out = None if g["param_out"] is None else g["param_out"].data_ptr()
            a.param_bf16_out = None if g["param_bf16_out"] is None else g["param_bf16_out"].data_ptr()
            a.lr_mult = None if self.lr_mult is None else self.lr_mult.data_ptr()
            a.decay_steps, a.decay_final = self.decay if self.decay is not None else (0, 0.0)
        self._arr = arr

    def set_lr_mult(self, mult):
        """Store the learning-rate multiplier the next update reads (a fill
        kernel on the current stream; outside any captured graph)."""
        if self.lr_mult is None:
            raise RuntimeError("FusedAdam was built without lr_mult")
        self.lr_mult.fill_(float(mult))

    def set_lr(self, index, lr):
        self.groups[index]["lr"] = float(lr)
        self._build()

    def step(self):
        _lib.check(_lib.lib().ds_adam_step(self._arr, len(self.groups), _lib.ptr(self.step_count),
                                           _lib.stream_ptr(self.device)), "ds_adam_step")

    def step_advanced(self):
        """The update for a counter that an earlier kernel on the stream has
        already advanced (a fused step given ``advance_counter``): iteration
        ``step_count``, and the counter is left as it is."""
        _lib.check(_lib.lib().ds_adam_step_advanced(self._arr, len(self.groups), _lib.ptr(self.step_count),
                                                    _lib.stream_ptr(self.device)), "ds_adam_step_advanced")
