"""VectorAdam: Adam with the second moment pooled over the last dimension
(python/diffsoup/optimize.py:7-68; Ling, Sharp & Jacobson, NeurIPS 2022).

Same update as the reference:
    g1 <- b1 g1 + (1 - b1) g
    g2 <- b2 g2 + (1 - b2) sum_last(g^2)
    p  <- p - lr * (g1 / (1 - b1^t)) / (sqrt(g2 / (1 - b2^t)) + 1e-8)
"""
from __future__ import annotations

import torch


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
