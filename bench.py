"""Benchmark: fused forward + backward iterations per second of DiffSoup's
rasterizer training step on one MI355X.

Workload (the reference's headline configuration that fits one GPU): a
15K-triangle soup (12K on a unit sphere + torus, 3K floaters), opacity
logits at lattice level 3, multires features at levels 0-3 with 8 fp32
channels concatenated into a 32->32->32->3 pixel MLP, 8 views at 800x800 on
a radius-4 orbit, stochastic opacity masking with a new seed per iteration,
MSE to a uniform random target, edge and opacity gradients.  Inputs are
generated on the host once and are resident in HBM before timing starts.

Prints one JSON line.  ``value`` is fp32 it/s of the fused step; the bf16-MLP
variant, covered fragments per second, the roofline of the dominant kernel
and the CPU baseline are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fwd+bwd iters/s @15K tris, 8x800^2 views"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-bf16", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--stage", default="auto", help="kernel stage timed for the roofline")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--profile-driver", action="store_true",
                    help="run four fused steps and a 256 MiB calibration kernel (for rocprofv3 --pmc), print nothing")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from diffsoup_amd import fused, scenes

    dev = torch.device("cuda:0")
    soup = scenes.sphere_torus_soup(0)
    if args.profile_driver:
        profile_driver(soup, dev)
        return
    prm = fused.SoupParams.from_soup(soup, dev)
    mvp = torch.from_numpy(soup.mvp).to(dev)
    target = torch.from_numpy(soup.target).to(dev)
    dims = dict(T=soup.T, V=soup.verts.shape[0], alpha_level=soup.alpha_level, levels=soup.levels,
                channels=soup.feat.shape[2], hidden=soup.mlp_dims[1])

    def make(bf16):
        return fused.FusedStep(soup.resolutions, dims["T"], dims["V"], dims["alpha_level"], dims["levels"],
                               dims["channels"], dims["hidden"], bf16=bf16, device=dev)

    def timed(step, steps, warmup, stage=None):
        for i in range(warmup):
            step.run(prm, mvp, target, seed=1000 + i)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        kev = []
        if stage is not None:
            kev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        e0.record(stream)
        for i in range(steps):
            if kev:
                step.time_stage(stage, kev[i][0], kev[i][1])
            step.run(prm, mvp, target, seed=2000 + i)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_time(e1) / steps
        kms = float(np.mean([a.elapsed_time(b) for a, b in kev])) if kev else None
        step.time_stage(None)
        return ms, wall / steps * 1e3, kms

    st32 = make(False)
    ms32, wall32, _ = timed(st32, args.steps, args.warmup)
    stats = st32.stats().cpu().numpy()
    covered = int(stats[0])
    out = {
        "metric": METRIC,
        "value": round(1000.0 / ms32, 2),
        "unit": "it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "sphere+torus triangle soup, 15K tris, 8 views 800x800, levels 0-3 x 8ch concat, "
                        "MLP 32-32-32-3, stochastic mask, MSE, edge + opacity grads",
            "triangles": soup.T, "views": soup.B, "resolution": list(soup.resolutions[0]),
            "pixels_per_step": soup.pixels, "parallelism": "single GPU",
        },
        "wall_ms_per_step": round(wall32, 4),
        "covered_fragments_per_step": covered,
        "gfrag_per_s": round(covered * 1000.0 / ms32 / 1e9, 4),
    }
    if not args.no_bf16:
        st16 = make(True)
        ms16, _, _ = timed(st16, args.steps, args.warmup)
        out["bf16_mlp_it_s"] = round(1000.0 / ms16, 2)
        out["bf16_vs_fp32"] = round(ms32 / ms16, 3)
        del st16

    if not args.no_plain:
        out["plain_fp32_it_s"] = round(plain_its(soup, prm, mvp, target, args.plain_steps), 3)
        out["speedup_vs_plain_fp32"] = round(out["value"] / out["plain_fp32_it_s"], 2)
        if "bf16_mlp_it_s" in out:
            out["bf16_speedup_vs_plain_fp32"] = round(out["bf16_mlp_it_s"] / out["plain_fp32_it_s"], 2)

    # roofline of the dominant kernel, timed with events on the launch stream
    stage = "resolve" if args.stage == "auto" else args.stage
    _, _, kms = timed(st32, max(10, args.steps // 2), 2, stage=stage)
    alg = st32.algorithmic_bytes(stage)
    achieved = alg / (kms * 1e-3) / 1e9
    out["roofline"] = {"kernel": stage, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                       "avg_launch_ms": round(kms, 5), "algorithmic_bytes": int(alg)}

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(soup, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def profile_driver(soup, dev):
    """Counter-collection driver: four fp32 fused steps, then torch.neg on a
    256 MiB tensor, a kernel that reads and writes exactly 256 MiB and so
    calibrates the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    from diffsoup_amd import fused

    prm = fused.SoupParams.from_soup(soup, dev)
    mvp = torch.from_numpy(soup.mvp).to(dev)
    target = torch.from_numpy(soup.target).to(dev)
    st = fused.FusedStep(soup.resolutions, soup.T, soup.verts.shape[0], soup.alpha_level, soup.levels,
                         soup.feat.shape[2], soup.mlp_dims[1], device=dev)
    for i in range(4):
        st.run(prm, mvp, target, seed=100 + i)
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(2):
        torch.neg(a, out=b)
    torch.cuda.synchronize()


def plain_its(soup, prm, mvp, target, steps):
    """The plain per-view FP32 path (fragment lists, global-atomic depth test,
    torch MLP, one view per launch) on the same workload and inputs."""
    import torch

    from diffsoup_amd import fused, plain

    leaves = fused.SoupParams(prm.verts.clone().requires_grad_(True), prm.tri,
                              prm.alpha_logits.clone().requires_grad_(True), prm.feat.clone().requires_grad_(True),
                              *[t.clone() for t in (prm.w1, prm.b1, prm.w2, prm.b2, prm.w3, prm.b3)])
    layers = []
    for w, b in ((leaves.w1, leaves.b1), (leaves.w2, leaves.b2), (leaves.w3, leaves.b3)):
        lin = torch.nn.Linear(w.shape[1], w.shape[0]).to(w.device)
        with torch.no_grad():
            lin.weight.copy_(w)
            lin.bias.copy_(b)
        layers.append(lin)
    ps = plain.PlainStep(leaves, layers, soup.levels, soup.alpha_level)

    def one(i):
        for t in [leaves.verts, leaves.alpha_logits, leaves.feat] + [p for l in layers for p in l.parameters()]:
            t.grad = None
        ps.run(mvp, target, soup.resolutions, seed=3000 + i)

    for i in range(2):
        one(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        one(10 + i)
    torch.cuda.synchronize()
    return steps / (time.perf_counter() - t0)


def cpu_baseline(soup, budget_s):
    """The CPU oracle's full step (C raster / gather / edge / opacity, torch
    CPU MLP in fp32) on a bounded sample of iterations, one thread."""
    import torch

    from oracle import oracle

    torch.set_num_threads(1)
    n, t0 = 0, time.perf_counter()
    while n < 1 or (time.perf_counter() - t0 < budget_s and n < 20):
        oracle.reference_step(soup, 2000 + n, dtype=torch.float32)
        n += 1
    dt = (time.perf_counter() - t0) / n
    return {"value": round(1.0 / dt, 4), "unit": "it/s", "cores": 1, "kind": "port",
            "sample": f"{n} full iterations of the same workload (single-threaded C oracle + torch CPU fp32 MLP)"}


if __name__ == "__main__":
    main()
