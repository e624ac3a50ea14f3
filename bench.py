"""Benchmark: fused forward + backward iterations per second of DiffSoup's
rasterizer training step on one MI355X.

Workload (the reference's headline configuration that fits one GPU): a
15K-triangle soup (12K on a unit sphere + torus, 3K floaters), opacity
logits at lattice level 3, multires features at levels 0-3 with 8 fp32
channels concatenated into a 32->32->32->3 pixel MLP, 8 views at 800x800 on
a radius-4 orbit, stochastic opacity masking with a new seed per iteration,
MSE to a uniform random target, edge and opacity gradients.  Inputs are
generated on the host once and are resident in HBM before timing starts.

Prints one JSON line.  ``value`` is fp32 it/s of the fused step; beside it:
the bf16 variant (MLP on bf16 MFMA with fp32 accumulation, feature rows
stored in bf16, every gradient fp32) with its image error against a float64
evaluation, covered fragments per second, the roofline of the dominant
kernel, the captured iteration of the three larger workloads (50K
triangles over 32 uneven views, 100K triangles at 16 x 1920x1080 in bf16,
300K triangles over 64 views) and the CPU baseline on the host's cores at
the 1K-triangle 256x256 and the 15K-triangle 8 x 800x800 configurations.
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
MFMA_FP32_PEAK_TFLOPS = 157.3
MFMA_BF16_PEAK_TFLOPS = 2500.0
METRIC = "fwd+bwd iters/s @15K tris, 8x800^2 views"
PROFILE_MD = os.path.join(ROOT, "profiles", "fused_step_sphere_torus_15k_8x800.md")


def stage_kernel(stage, step):
    """Kernel behind a timed stage of ds_train_step, as named in the rocprofv3
    summary (template arguments: tile edge or threads per tile; shading
    levels, channels, hidden width and precision)."""
    tile = step.tile
    # the bf16-MLP kernel that reads bf16 feature rows is a separate instance
    rows = ",bf16rows" if step.bf16 and step.bf16_features else ""
    shade = f"k_shade<{step.nl},{step.C},{step.hidden},{'true' if step.bf16 else 'false'}{rows}>"
    return {"resolve": f"k_resolve<{tile},true>", "shade": shade,
            "edge": "k_edge_grad_sorted", "aux": f"k_opacity_aux<{tile}>"}[stage]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the three larger workloads in the headline line")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: the process's CPU share, capped by OMP_NUM_THREADS)")
    ap.add_argument("--no-bf16", action="store_true")
    ap.add_argument("--stage", default="shade", choices=["resolve", "shade", "edge", "aux"],
                    help="kernel stage reported as the roofline (the dominant one by default)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--profile-driver", action="store_true",
                    help="run four fused steps and a 256 MiB calibration kernel (for rocprofv3 --pmc), print nothing")
    ap.add_argument("--workload", default="soup15k", choices=["soup15k"] + sorted(WORKLOADS),
                    help="soup15k (default) is the headline measurement; the others print their own line")
    ap.add_argument("--summarize-sq", nargs=3, metavar=("PASS1_CSV", "PASS2_CSV", "OUT_MD"),
                    help="write the SQ counter summary of two --pmc passes of --profile-driver")
    ap.add_argument("--summarize-rocprof", nargs=4, metavar=("KERNEL_STATS_CSV", "FETCH_CSV", "WRITE_CSV", "OUT_MD"),
                    help="write the markdown summary of a kernel trace and two counter runs, then exit")
    ap.add_argument("--summarize-workload-traffic", nargs=4, metavar=("WORKLOAD", "FETCH_CSV", "WRITE_CSV", "OUT_MD"),
                    help="write the HBM-traffic summary of two --pmc runs of --profile-driver --workload WORKLOAD")
    ap.add_argument("--summarize-workload-trace", nargs=3, metavar=("WORKLOAD", "KERNEL_STATS_CSV", "OUT_MD"),
                    help="write the markdown summary of a kernel trace of `bench.py --workload WORKLOAD`")
    return ap.parse_args()


def summarize_workload_trace(workload, stats_csv, out_md):
    """Markdown table of a kernel trace of one larger workload's bench run
    (autotuning, the captured iterations and the per-stage timing runs)."""
    import csv

    rows = list(csv.DictReader(open(stats_csv)))
    d = os.path.basename(os.path.dirname(stats_csv))
    cmd = (f"cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --stats --output-format csv "
           f"-d $REPO/gpurun_out/{d} -o trace -- python3 $REPO/bench.py --workload {workload}")
    out = [f"# Kernel trace: `bench.py --workload {workload}`\n",
           f"Workload: {WORKLOADS[workload][2]}.  The run autotunes the tile, times the captured iteration and "
           "then each stage; every launch of every kernel is in the table.  `$REPO` is the repository root.\n",
           "```\n" + cmd + "\n```\n",
           "| kernel | calls | avg us | min us | max us | % time |", "|---|---:|---:|---:|---:|---:|"]
    for r in rows[:24]:
        out.append(f"| `{_kernel_key(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    out += ["", "Summary written by:",
            f"```\npython3 bench.py --summarize-workload-trace {workload} {os.path.relpath(stats_csv, ROOT)} "
            f"{os.path.relpath(out_md, ROOT)}\n```"]
    with open(out_md, "w") as fh:
        fh.write("\n".join(out) + "\n")


def main():
    args = parse()
    if args.summarize_rocprof:
        summarize_rocprof(*args.summarize_rocprof)
        return
    if args.summarize_sq:
        summarize_sq(*args.summarize_sq)
        return
    if args.summarize_workload_trace:
        summarize_workload_trace(*args.summarize_workload_trace)
        return
    if args.summarize_workload_traffic:
        summarize_workload_traffic(*args.summarize_workload_traffic)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from diffsoup_amd import fused, scenes

    dev = torch.device("cuda:0")
    if args.workload != "soup15k":
        if args.profile_driver:
            profile_driver_workload(args.workload, dev)
        else:
            other_workload(args, dev)
        return
    soup = scenes.sphere_torus_soup(0)
    if args.profile_driver:
        profile_driver(soup, dev)
        return
    prm = fused.SoupParams.from_soup(soup, dev)
    mvp = torch.from_numpy(soup.mvp).to(dev)
    target = torch.from_numpy(soup.target).to(dev)
    dims = dict(T=soup.T, V=soup.verts.shape[0], alpha_level=soup.alpha_level, levels=soup.levels,
                channels=soup.feat.shape[2], hidden=soup.mlp_dims[1])

    def make(bf16, bf16_features=None, loss="mse"):
        """bf16: the pixel MLP on bf16 MFMA; bf16_features (default: as bf16)
        also stores the features in bf16.  Accumulation, depth, barycentrics
        and every gradient stay fp32.  loss: "mse" or "l1_ssim" (the
        reference's 0.8 L1 + 0.2 (1 - SSIM) / 2)."""
        f16 = bf16 if bf16_features is None else bf16_features
        return fused.FusedStep(soup.resolutions, dims["T"], dims["V"], dims["alpha_level"], dims["levels"],
                               dims["channels"], dims["hidden"], bf16=bf16, bf16_features=f16, device=dev,
                               loss=loss)

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

    def timed_graph(bf16, steps, warmup, bf16_features=None, loss="mse", keep_params=False):
        """The captured training iteration: forward + backward + Adam/VectorAdam,
        one graph replay per step.  The optimiser reads and writes every
        parameter and moment each iteration but stores the updated parameters
        beside the inputs, so each iteration sees the configured scene (trained
        in place on a random target, the soup's triangles grow and the workload
        would drift with the step count)."""
        step = make(bf16, bf16_features, loss)
        params = fused.SoupParams.from_soup(soup, dev)
        updated = fused.SoupParams.from_soup(soup, dev)
        if step.bf16_features:
            params.with_bf16_features()
            updated.with_bf16_features()
        tune = step.autotune(params, mvp, target, tiles=(8, 16, 32), reps=5, cache=None)
        opt = fused.soup_optimizer(step, params, out=updated)
        cap = fused.CapturedStep(step, params, mvp, target, opt, seed=1000)
        for _ in range(warmup):
            cap.replay()
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            cap.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if keep_params:
            return e0.elapsed_time(e1) / steps, wall / steps * 1e3, step, tune, params
        return e0.elapsed_time(e1) / steps, wall / steps * 1e3, step, tune

    ms32, wall32, gstep, tune32 = timed_graph(False, args.steps, args.warmup)
    covered = int(gstep.stats().cpu().numpy()[0])
    tile32 = gstep.tile
    del gstep
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
                        "MLP 32-32-32-3, stochastic mask with a new seed per iteration, MSE, edge + opacity "
                        "grads, Adam (features, opacities, MLP) + VectorAdam (vertices) written out of place, "
                        "one HIP graph per iteration",
            "triangles": soup.T, "views": soup.B, "resolution": list(soup.resolutions[0]),
            "pixels_per_step": soup.pixels, "tile": tile32, "parallelism": "single GPU",
        },
        "autotune_ms_per_step": {str(k): (None if v is None else round(v, 4)) for k, v in tune32.items()},
        "wall_ms_per_step": round(wall32, 4),
        "covered_fragments_per_step": covered,
        "gfrag_per_s": round(covered * 1000.0 / ms32 / 1e9, 4),
    }
    if not args.no_bf16:
        # bf16: MLP on bf16 MFMA (fp32 accumulation), bf16 feature rows (the
        # four residual-blend channels read from the fp32 master), fp32
        # depth, barycentrics and gradients
        ms16, _, g16, _, p16 = timed_graph(True, args.steps, args.warmup, keep_params=True)
        out["bf16_it_s"] = round(1000.0 / ms16, 2)
        out["bf16_vs_fp32"] = round(ms32 / ms16, 3)
        out["bf16_image_err_vs_fp64"] = float(f"{image_err_vs_fp64(g16, p16):.3e}")
        del g16, p16
        msm, _, gm, _, pm = timed_graph(True, args.steps, args.warmup, bf16_features=False, keep_params=True)
        out["bf16_mlp_fp32_features_it_s"] = round(1000.0 / msm, 2)
        out["bf16_mlp_fp32_features_image_err_vs_fp64"] = float(f"{image_err_vs_fp64(gm, pm):.3e}")
        del gm, pm
    # the same iteration with the reference's own image loss, 0.8 L1 + 0.2 (1 - SSIM) / 2
    msl, _, gl, _ = timed_graph(False, args.steps, args.warmup, loss="l1_ssim")
    del gl
    out["l1_ssim_loss_it_s"] = round(1000.0 / msl, 2)

    # forward + backward only (no optimiser, fixed parameters, eager launches)
    st32 = make(False)
    st32.set_tile(tile32)
    msfb, _, _ = timed(st32, args.steps, args.warmup)
    out["fwd_bwd_only_eager_it_s"] = round(1000.0 / msfb, 2)

    if not args.no_plain:
        out["plain_fp32_it_s"] = round(plain_its(soup, prm, mvp, target, args.plain_steps), 3)
        out["speedup_vs_plain_fp32"] = round(out["value"] / out["plain_fp32_it_s"], 2)
        if "bf16_it_s" in out:
            out["bf16_speedup_vs_plain_fp32"] = round(out["bf16_it_s"] / out["plain_fp32_it_s"], 2)

    # per-stage kernel time with HIP events recorded on the launch stream,
    # and the roofline of each stage's kernel
    traffic = profiled_traffic()
    stages = {}
    for stage in ("resolve", "shade", "edge", "aux"):
        _, _, kms = timed(st32, max(10, args.steps // 2), 2, stage=stage)
        stages[stage] = roofline(st32, stage, kms, traffic)
    out["roofline"] = stages[args.stage]
    out["stages"] = {k: {"avg_launch_us": round(v["avg_launch_ms"] * 1e3, 1), "bound": v["bound"],
                         "frac": v["frac"]} for k, v in stages.items()}
    # the whole iteration against HBM: its compulsory bytes (forward, backward
    # and the optimiser's 28 B per parameter) at the measured rate
    it_bytes = st32.iteration_bytes(adam=True)
    it_gbs = it_bytes * out["value"] / 1e9
    out["iteration_roofline"] = {"bound": "hbm", "algorithmic_bytes": int(it_bytes), "achieved": round(it_gbs, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(it_gbs / HBM_PEAK_GBS, 4),
                                 "ceiling_it_s": round(HBM_PEAK_GBS * 1e9 / it_bytes, 1)}
    del st32

    if not args.no_workloads:
        # the larger workloads' captured iterations, measured in this run
        out["workloads"] = {}
        for name in WORKLOADS:
            torch.cuda.empty_cache()
            w = run_workload(name, args, dev)
            out["workloads"][name] = {k: w[k] for k in ("value", "unit", "ms_per_step", "dtype", "gfrag_per_s",
                                                        "iteration_roofline", "stages")}
            out["workloads"][name]["workload"] = w["config"]["workload"]
            out["workloads"][name]["tile"] = w["config"]["tile"]
            print(f"workload {name}: {w['value']} it/s", file=sys.stderr, flush=True)

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(soup, args.cpu_threads)
    print(json.dumps(out), flush=True)


def image_err_vs_fp64(step, params):
    """Largest |colour - float64 colour| over the covered pixels of the step's
    last output.  The float64 colour takes the features gathered from the
    fp32 master table on the step's own raster (ds_multires_color_fwd,
    levels concatenated; its fp32 rounding is ~1e-8 here), then the MLP,
    sigmoid and residual blend in float64 (torch on the GPU) with the fp32
    master weights."""
    import torch

    from diffsoup_amd.multires import multires_feature_gather

    cov = step.rast[:, 3] > 0
    x = multires_feature_gather(step.rast[cov].contiguous(), step.lo, step.hi, params.feat, mode="concat").double()
    h = x
    for i, (w, b) in enumerate(((params.w1, params.b1), (params.w2, params.b2), (params.w3, params.b3))):
        h = h @ w.double().T + b.double()
        h = torch.relu(h) if i < 2 else torch.sigmoid(h)
    r = x[:, 3:4]
    ref = (1.0 - r) * x[:, :3] + r * h
    return float((step.color[cov].double() - ref).abs().max().item())


WORKLOADS = {
    "uneven50k": ("batched_uneven_soup", False,
                  "50K-triangle unbounded scene, 32 views of mixed resolution (1297x840, 979x546, 1237x822, "
                  "800x800), levels 0-3 x 8 ch fp32, MLP 32-32-32-3 fp32"),
    "bf16_100k": ("bf16_large_soup", True,
                  "100K-triangle unbounded scene, 16 views 1920x1080, levels 0-4 x 16 ch bf16 features, "
                  "MLP 80-64-64-3 on bf16 MFMA with fp32 accumulation"),
    "stress300k": ("stress_soup", True,
                   "300K triangles (40% in stacked depth layers), 64 views of mixed resolution up to 1920x1080, "
                   "bf16 features and MLP as bf16_100k"),
    # the reference's own training iteration at its synthetic driver's size
    "reference15k_l0": ("reference_sphere_torus_soup", False,
                        "the reference's own iteration (examples/02_synthetic.py:354-388): 15K triangles, 4 views "
                        "800x800, 7-channel features and opacities at levels 0-0 accumulated + sigmoid, SH2 inputs, "
                        "MLP 16-16-16-3 fp32, L1 + SSIM, edge + opacity grads, Adam/VectorAdam with exp_decay_mult"),
    "reference15k_l5": ("reference_sphere_torus_soup", False,
                        "as reference15k_l0 after the driver's lift to levels 2-5 (561 samples per triangle at "
                        "level 5)"),
}

# scene-generator keyword arguments of the workloads that take any
WORKLOAD_ARGS = {"reference15k_l0": dict(levels=(0, 0)), "reference15k_l5": dict(levels=(2, 5))}


def other_workload(args, dev):
    """One of the larger workloads on its own: prints its JSON line."""
    print(json.dumps(run_workload(args.workload, args, dev)), flush=True)


def run_workload(name, args, dev):
    """One of the larger workloads: the captured training iteration
    (forward + backward + Adam/VectorAdam written out of place) with the tile
    chosen by the autotuner, per-stage kernel times and their rooflines, and
    the whole iteration against HBM."""
    import torch

    from diffsoup_amd import fused, scenes

    gen, bf16, desc = WORKLOADS[name]
    t0 = time.perf_counter()
    soup = getattr(scenes, gen)(0, **WORKLOAD_ARGS.get(name, {}))
    gen_s = time.perf_counter() - t0
    mvp = torch.from_numpy(soup.mvp).to(dev)
    target = torch.from_numpy(soup.target).to(dev)
    ref_iter = soup.accumulate
    inv = torch.from_numpy(soup.inv_mvp).to(dev) if soup.sh2 else None
    step = fused.FusedStep(soup.resolutions, soup.T, soup.verts.shape[0], soup.alpha_level, soup.levels,
                           soup.feat.shape[2], soup.mlp_dims[1], bf16=bf16, bf16_features=bf16, device=dev,
                           loss="l1_ssim" if ref_iter else "mse", accumulate=soup.accumulate,
                           alpha_lo=soup.alpha_lo, sh2=soup.sh2)
    params = fused.SoupParams.from_soup(soup, dev)
    updated = fused.SoupParams.from_soup(soup, dev)
    if bf16:
        params.with_bf16_features()
        updated.with_bf16_features()
    tune = step.autotune(params, mvp, target, tiles=(8, 16, 32), reps=2, cache=None, inv_mvp=inv)
    # the reference's schedule over its 10,000 steps (examples/02_synthetic.py:341-349)
    opt = fused.soup_optimizer(step, params, out=updated, decay=(10_000, 0.01) if ref_iter else None)
    cap = fused.CapturedStep(step, params, mvp, target, opt, seed=1000, inv_mvp=inv)
    steps, warmup = min(args.steps, 20), min(args.warmup, 3)
    for _ in range(warmup):
        cap.replay()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        cap.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    covered = int(step.stats()[0].item())
    # per-stage kernel times on the eager step (same buffers, fixed parameters)
    stages = {}
    for stage in ("resolve", "shade", "edge", "aux"):
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(3)]
        for i in range(3):
            step.time_stage(stage, *evs[i])
            step.run(params, mvp, target, seed=3000 + i, inv_mvp=inv)
        torch.cuda.synchronize()
        kms = sum(a.elapsed_time(b) for a, b in evs) / 3
        step.time_stage(None)
        if ref_iter and stage == "shade":
            # colour pass, L1 + SSIM kernels and adjoint pass: no single roof
            stages[stage] = {"avg_launch_ms": round(kms, 5), "bound": "colour pass + L1/SSIM + adjoint pass",
                             "achieved": None, "unit": None, "frac": None}
        else:
            stages[stage] = roofline(step, stage, kms, {})
    out = {"metric": f"fwd+bwd iters/s ({name})", "value": round(1000.0 / ms, 2), "unit": "it/s",
           "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
           "dtype": "bf16" if bf16 else "f32", "data": "synthetic",
           "config": {"workload": desc, "triangles": soup.T, "views": soup.B, "pixels_per_step": soup.pixels,
                      "tile": step.tile, "parallelism": "single GPU"},
           "autotune_ms_per_step": {str(k): (None if v is None else round(v, 4)) for k, v in tune.items()},
           "covered_fragments_per_step": covered, "gfrag_per_s": round(covered * 1000.0 / ms / 1e9, 4),
           "stages": {k: {"avg_launch_us": round(v["avg_launch_ms"] * 1e3, 1), "bound": v["bound"],
                          "achieved": v["achieved"], "unit": v["unit"], "frac": v["frac"]} for k, v in stages.items()},
           "scene_generation_s": round(gen_s, 1)}
    it_bytes = step.iteration_bytes(adam=True)
    it_gbs = it_bytes * out["value"] / 1e9
    out["iteration_roofline"] = {"bound": "hbm", "algorithmic_bytes": int(it_bytes), "achieved": round(it_gbs, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(it_gbs / HBM_PEAK_GBS, 4),
                                 "ceiling_it_s": round(HBM_PEAK_GBS * 1e9 / it_bytes, 1)}
    del cap, opt, step, params, updated
    return out


def roofline(step, stage, kms, traffic):
    """Roofline entry of one stage's kernel.  The raster kernels are compared
    with HBM bandwidth.  Shading is compared with the lower of its two roofs:
    the MFMA peak of its precision, or its arithmetic intensity (algorithmic
    flops over algorithmic bytes) times HBM bandwidth.  The fp32 MLP is above
    the fp32 ridge (157.3 / 8 = 19.7 flop/B), so MFMA bounds it; the bf16
    MLP is far below the bf16 ridge (312 flop/B), so HBM bounds it."""
    kernel = stage_kernel(stage, step)
    tr = traffic.get(kernel)
    if stage == "shade":
        flops = step.algorithmic_flops(stage)
        nbytes = step.algorithmic_bytes(stage)
        ai = flops / nbytes
        mfma = MFMA_BF16_PEAK_TFLOPS if step.bf16 else MFMA_FP32_PEAK_TFLOPS
        if mfma <= ai * HBM_PEAK_GBS / 1e3:
            achieved = flops / (kms * 1e-3) / 1e12
            peak = mfma
            r = {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s"}
        else:
            achieved = nbytes / (kms * 1e-3) / 1e9
            peak = HBM_PEAK_GBS
            r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": peak, "unit": "GB/s",
                 "achieved_tflops": round(flops / (kms * 1e-3) / 1e12, 2), "mfma_peak_tflops": mfma}
        r.update(algorithmic_flops=int(flops), algorithmic_bytes=int(nbytes), flop_per_byte=round(ai, 1))
    else:
        nbytes = step.algorithmic_bytes(stage)
        achieved = nbytes / (kms * 1e-3) / 1e9
        peak = HBM_PEAK_GBS
        r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": peak, "unit": "GB/s",
             "algorithmic_bytes": int(nbytes)}
    r["frac"] = round(achieved / peak, 4)
    r["traffic"] = None if tr is None else int(tr)
    r["kernel"] = kernel
    r["avg_launch_ms"] = round(kms, 5)
    if tr is not None:
        r["traffic_source"] = os.path.relpath(PROFILE_MD, ROOT) + " (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, calibrated)"
        r["traffic_gbs"] = round(tr / (kms * 1e-3) / 1e9, 1)
    return r


def profiled_traffic():
    """HBM bytes per launch of each kernel from the committed rocprofv3
    counter summary (its 'traffic MiB' column); empty if it is absent."""
    out = {}
    if not os.path.exists(PROFILE_MD):
        return out
    in_table = False
    for line in open(PROFILE_MD):
        if line.startswith("| kernel | fetch MiB"):
            in_table = True
            continue
        if in_table:
            cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
            if len(cells) != 4 or cells[0].startswith("---"):
                if not line.strip():
                    break
                continue
            out[cells[0]] = float(cells[3]) * 1024 * 1024
    return out


def _kernel_key(name):
    """Short kernel name: the k_* function plus its template arguments."""
    import re
    m = re.search(r"::(k_\w+)", name)
    if m:
        k = m.group(1)
        cfg = re.search(r"Cfg<([^>]*)>((?:,\s*(?:true|false))*)>", name)
        if cfg:
            # k_shade<Cfg<levels, channels, hidden, bf16 MLP, SH2>, bf16 rows, two-pass form, direct rows>
            p = [x.strip() for x in cfg.group(1).split(",")]
            flags = [x.strip() for x in cfg.group(2).split(",") if x.strip()]
            name_ = ",".join(p[:4]) + (",sh2" if len(p) > 4 and p[4] == "true" else "")
            for flag, label in zip(flags, ("bf16rows", "two-pass", "direct")):
                if flag == "true":
                    name_ += "," + label
            return f"{k}<{name_}>"
        t = re.search(r"k_\w+<([^>(]*)>", name)
        return k + ("<" + t.group(1).replace(" ", "") + ">" if t else "")
    if "neg" in name:
        return "calibration torch.neg 256 MiB"
    return re.sub(r"\(.*", "", name)[:50]


def summarize_rocprof(stats_csv, fetch_csv, write_csv, out_md):
    """Markdown summary of `rocprofv3 --kernel-trace --stats` (per-kernel
    average duration) and two `--pmc` runs of --profile-driver (FETCH_SIZE,
    WRITE_SIZE in KiB per dispatch), scaled by what each counter reads for
    torch.neg over 256 MiB, which moves exactly 256 MiB each way."""
    import collections
    import csv

    rows = list(csv.DictReader(open(stats_csv)))
    res = collections.defaultdict(dict)
    for ctr, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        agg = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            agg[_kernel_key(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        for k, v in agg.items():
            res[k][ctr] = sum(v) / len(v)
    cal = res["calibration torch.neg 256 MiB"]
    fs, ws = 262144.0 / cal["FETCH_SIZE"], 262144.0 / cal["WRITE_SIZE"]
    cmd_stats = ("cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --stats --output-format csv "
                 "-d $REPO/gpurun_out/prof_bench -o bench -- python3 $REPO/bench.py --steps 20 --warmup 3 "
                 "--no-cpu-baseline --no-plain --no-workloads")
    cmd_pmc = "\n".join(
        f"cd /tmp && export TMPDIR=/tmp && rocprofv3 --pmc {c} --output-format csv -d $REPO/gpurun_out/pmc_{c} "
        f"-o run -- python3 $REPO/bench.py --profile-driver" for c in ("FETCH_SIZE", "WRITE_SIZE"))
    cmd_sum = ("python3 bench.py --summarize-rocprof gpurun_out/prof_bench/bench_kernel_stats.csv "
               "gpurun_out/pmc_FETCH_SIZE/run_counter_collection.csv "
               "gpurun_out/pmc_WRITE_SIZE/run_counter_collection.csv " + os.path.relpath(PROFILE_MD, ROOT))
    out = ["# Fused training step: sphere+torus soup, 15K triangles, 8 views 800x800\n",
           "bench.py runs the fp32-MLP step (`k_shade<4,8,32,false>`), the bf16-MLP step with bf16 "
           "feature rows (`k_shade<4,8,32,true,bf16rows>`) and with fp32 features (`k_shade<4,8,32,true>`), "
           "and the fp32 step with the L1 + SSIM loss (two `k_shade<4,8,32,false,two-pass>` passes around "
           "`k_ssim_maps<3>`, `k_ssim_grad<3>` and `k_photo_loss`); every other kernel is shared.  `$REPO` is "
           "the repository root.\n",
           "Kernel trace:\n", "```\n" + cmd_stats + "\n```\n",
           "| kernel | calls | avg us | min us | max us | % time |", "|---|---:|---:|---:|---:|---:|"]
    for r in rows[:30]:
        out.append(f"| `{_kernel_key(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    out += ["\nHBM traffic per launch (two counter runs, one counter each):\n", "```\n" + cmd_pmc + "\n```\n",
            f"Calibration (torch.neg over 256 MiB reads and writes 262144 KiB): FETCH_SIZE "
            f"{cal['FETCH_SIZE']:.0f} KiB, WRITE_SIZE {cal['WRITE_SIZE']:.0f} KiB; scale factors fetch "
            f"x{fs:.3f}, write x{ws:.3f}.\n",
            "| kernel | fetch MiB | write MiB | traffic MiB |", "|---|---:|---:|---:|"]
    order = sorted(res, key=lambda k: -(res[k].get("FETCH_SIZE", 0) * fs + res[k].get("WRITE_SIZE", 0) * ws))
    for k in order[:20]:
        if k.startswith("calibration") or not k.startswith("k_"):
            continue
        f = res[k].get("FETCH_SIZE", 0) * fs / 1024
        w = res[k].get("WRITE_SIZE", 0) * ws / 1024
        out.append(f"| `{k}` | {f:.1f} | {w:.1f} | {f + w:.1f} |")
    out += ["", "Summary written by:", "```\n" + cmd_sum + "\n```"]
    with open(out_md, "w") as fh:
        fh.write("\n".join(out) + "\n")


SQ_PASS1 = ("SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_ANY SQ_ACTIVE_INST_VALU "
            "SQ_ACTIVE_INST_LDS SQ_ACTIVE_INST_VMEM")
SQ_PASS2 = ("SQ_WAVES SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_LDS_BANK_CONFLICT "
            "SQ_VALU_MFMA_BUSY_CYCLES SQ_INSTS_MFMA")


def summarize_sq(pass1_csv, pass2_csv, out_md):
    """Markdown table of where each kernel's waves spend their cycles: the
    fractions of wave-cycles spent waiting (SQ_WAIT_ANY, SQ_WAIT_INST_ANY)
    and issuing VALU, LDS and memory instructions, with instruction counts
    and LDS bank-conflict cycles per dispatch."""
    import collections
    import csv

    res = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in (pass1_csv, pass2_csv):
        for r in csv.DictReader(open(path)):
            res[_kernel_key(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    cmd = "\n".join(
        f"cd /tmp && export TMPDIR=/tmp && rocprofv3 --pmc {ctrs} --output-format csv -d $REPO/gpurun_out/sq{i} "
        f"-o run -- python3 $REPO/bench.py --profile-driver" for i, ctrs in ((1, SQ_PASS1), (2, SQ_PASS2)))
    out = ["# Where the step's waves spend their cycles: 15K-triangle soup, 8 views 800x800\n",
           "Two counter passes over `bench.py --profile-driver` (four fp32 MSE steps, two L1 + SSIM steps); "
           "`$REPO` is the repository root. Fractions are of SQ_WAVE_CYCLES (cycles summed over resident "
           "waves); counts are per dispatch.\n", "```\n" + cmd + "\n```\n",
           "| kernel | waves | waiting (any) | waiting (instruction) | issuing VALU | issuing LDS | VALU insts "
           "| LDS insts | LDS bank-conflict cycles | MFMA insts |",
           "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    rows = []
    for k, ctr in res.items():
        d = {c: sum(v) / len(v) for c, v in ctr.items()}
        wc = d.get("SQ_WAVE_CYCLES", 0.0)
        if not k.startswith("k_") or wc <= 0:
            continue
        rows.append((wc, k, d))
    for wc, k, d in sorted(rows, reverse=True):
        f = lambda c: d.get(c, 0.0) / wc
        out.append(f"| `{k}` | {d.get('SQ_WAVES', 0):.0f} | {f('SQ_WAIT_ANY'):.2f} | {f('SQ_WAIT_INST_ANY'):.2f} | "
                   f"{f('SQ_ACTIVE_INST_VALU'):.2f} | {f('SQ_ACTIVE_INST_LDS'):.2f} | {d.get('SQ_INSTS_VALU', 0):.3g} | "
                   f"{d.get('SQ_INSTS_LDS', 0):.3g} | {d.get('SQ_LDS_BANK_CONFLICT', 0):.3g} | "
                   f"{d.get('SQ_INSTS_MFMA', 0):.3g} |")
    out += ["", "Summary written by:", "```\npython3 bench.py --summarize-sq " + pass1_csv + " " + pass2_csv + " "
            + out_md + "\n```"]
    with open(out_md, "w") as fh:
        fh.write("\n".join(out) + "\n")


def profile_driver(soup, dev):
    """Counter-collection driver: four fp32 fused steps, two with the L1 +
    SSIM loss, then torch.neg on a 256 MiB tensor, a kernel that reads and
    writes exactly 256 MiB and so calibrates the FETCH_SIZE / WRITE_SIZE
    counters."""
    import torch

    from diffsoup_amd import fused

    prm = fused.SoupParams.from_soup(soup, dev)
    mvp = torch.from_numpy(soup.mvp).to(dev)
    target = torch.from_numpy(soup.target).to(dev)
    st = fused.FusedStep(soup.resolutions, soup.T, soup.verts.shape[0], soup.alpha_level, soup.levels,
                         soup.feat.shape[2], soup.mlp_dims[1], device=dev)
    for i in range(4):
        st.run(prm, mvp, target, seed=100 + i)
    del st
    sl = fused.FusedStep(soup.resolutions, soup.T, soup.verts.shape[0], soup.alpha_level, soup.levels,
                         soup.feat.shape[2], soup.mlp_dims[1], device=dev, loss="l1_ssim")
    for i in range(2):
        sl.run(prm, mvp, target, seed=200 + i)
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(2):
        torch.neg(a, out=b)
    torch.cuda.synchronize()


def profile_driver_workload(name, dev, tile=16, steps=2):
    """Counter-collection driver for one larger workload: ``steps`` eager
    fused steps at tile ``tile`` (the tile the autotuner picks for these
    workloads), then the 256 MiB torch.neg calibration kernel."""
    import torch

    from diffsoup_amd import fused, scenes

    gen, bf16, _ = WORKLOADS[name]
    soup = getattr(scenes, gen)(0, **WORKLOAD_ARGS.get(name, {}))
    mvp = torch.from_numpy(soup.mvp).to(dev)
    target = torch.from_numpy(soup.target).to(dev)
    inv = torch.from_numpy(soup.inv_mvp).to(dev) if soup.sh2 else None
    st = fused.FusedStep(soup.resolutions, soup.T, soup.verts.shape[0], soup.alpha_level, soup.levels,
                         soup.feat.shape[2], soup.mlp_dims[1], bf16=bf16, bf16_features=bf16, device=dev, tile=tile,
                         loss="l1_ssim" if soup.accumulate else "mse", accumulate=soup.accumulate,
                         alpha_lo=soup.alpha_lo, sh2=soup.sh2)
    prm = fused.SoupParams.from_soup(soup, dev)
    if bf16:
        prm.with_bf16_features()
    for i in range(steps):
        st.run(prm, mvp, target, seed=100 + i, inv_mvp=inv)
    torch.cuda.synchronize()
    del st, prm
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(2):
        torch.neg(a, out=b)
    torch.cuda.synchronize()


def summarize_workload_traffic(workload, fetch_csv, write_csv, out_md):
    """Markdown table of HBM traffic per launch of one larger workload's step
    kernels (two counter runs of ``--profile-driver --workload``), scaled by
    the torch.neg calibration as in the headline summary."""
    import collections
    import csv

    res = collections.defaultdict(dict)
    for ctr, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        agg = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            agg[_kernel_key(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        for k, v in agg.items():
            res[k][ctr] = sum(v) / len(v)
    cal = res["calibration torch.neg 256 MiB"]
    fs, ws = 262144.0 / cal["FETCH_SIZE"], 262144.0 / cal["WRITE_SIZE"]
    cmd = "\n".join(
        f"cd /tmp && export TMPDIR=/tmp && rocprofv3 --pmc {c} --output-format csv -d $REPO/gpurun_out/pmc_{workload}_{c} "
        f"-o run -- python3 $REPO/bench.py --workload {workload} --profile-driver" for c in ("FETCH_SIZE", "WRITE_SIZE"))
    out = [f"# HBM traffic per launch: `bench.py --workload {workload}`\n",
           f"Workload: {WORKLOADS[workload][2]}.  Two eager fused steps at tile 16, one counter per run; `$REPO` is the "
           "repository root.\n", "```\n" + cmd + "\n```\n",
           f"Calibration (torch.neg over 256 MiB reads and writes 262144 KiB): FETCH_SIZE {cal['FETCH_SIZE']:.0f} KiB, "
           f"WRITE_SIZE {cal['WRITE_SIZE']:.0f} KiB; scale factors fetch x{fs:.3f}, write x{ws:.3f}.\n",
           "| kernel | fetch MiB | write MiB | traffic MiB |", "|---|---:|---:|---:|"]
    order = sorted(res, key=lambda k: -(res[k].get("FETCH_SIZE", 0) * fs + res[k].get("WRITE_SIZE", 0) * ws))
    for k in order[:20]:
        if k.startswith("calibration") or not k.startswith("k_"):
            continue
        f = res[k].get("FETCH_SIZE", 0) * fs / 1024
        w = res[k].get("WRITE_SIZE", 0) * ws / 1024
        out.append(f"| `{k}` | {f:.1f} | {w:.1f} | {f + w:.1f} |")
    out += ["", "Summary written by:", "```\npython3 bench.py --summarize-workload-traffic " + workload + " " + fetch_csv +
            " " + write_csv + " " + out_md + "\n```"]
    with open(out_md, "w") as fh:
        fh.write("\n".join(out) + "\n")


def plain_its(soup, prm, mvp, target, steps):
    """The plain per-view FP32 path (fragment lists, global-atomic depth test,
    torch MLP, one view per launch) on the same workload and inputs, with the
    reference's optimisers (torch Adam, VectorAdam) after each backward."""
    import torch

    from diffsoup_amd import fused, plain
    from diffsoup_amd.optimize import VectorAdam

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
    opt_soup = torch.optim.Adam([{"params": [leaves.feat], "lr": 5e-2}, {"params": [leaves.alpha_logits], "lr": 5e-2}])
    opt_vert = VectorAdam([leaves.verts], lr=1e-3)
    opt_mlp = torch.optim.Adam([p for l in layers for p in l.parameters()], lr=1e-2)

    def one(i):
        for o in (opt_soup, opt_vert, opt_mlp):
            o.zero_grad(set_to_none=True)
        ps.run(mvp, target, soup.resolutions, seed=3000 + i)
        for o in (opt_soup, opt_vert, opt_mlp):
            o.step()

    for i in range(2):
        one(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        one(10 + i)
    torch.cuda.synchronize()
    return steps / (time.perf_counter() - t0)


def cpu_threads(requested=0):
    """Threads for the CPU baseline: the requested count, else the CPUs this
    process may run on, capped by OMP_NUM_THREADS when that is set."""
    if requested > 0:
        return requested
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_baseline(soup, threads=0):
    """The CPU restatement's full training iteration on the host's cores:
    the C raster / gather / edge / opacity kernels on OpenMP threads, the
    MLP, blend and MSE in torch CPU fp32, then torch Adam and VectorAdam
    updates of every parameter from the gradients it produced.  Median of 5
    iterations after one warm-up, at the 15K-triangle 8 x 800x800 workload
    (``value``) and at the 1K-triangle 256x256 single-view configuration."""
    import numpy as np
    import torch

    from diffsoup_amd import scenes
    from diffsoup_amd.optimize import VectorAdam
    from oracle import oracle

    n = cpu_threads(threads)
    torch.set_num_threads(n)
    oracle.set_threads(n)
    t = lambda a: torch.tensor(np.asarray(a, np.float32))

    def its(s, reps=5):
        feat, alog, verts = t(s.feat), t(s.alpha_logits), t(s.verts)
        mlp = [t(x) for wb in s.mlp for x in wb]
        opts = [torch.optim.Adam([{"params": [feat], "lr": 5e-2}, {"params": [alog], "lr": 5e-2}]),
                VectorAdam([verts], lr=1e-3), torch.optim.Adam(mlp, lr=1e-2)]
        order = [0, 2, 4, 1, 3, 5]   # grad_mlp holds dW1, dW2, dW3, db1, db2, db3

        def one(k):
            r = oracle.reference_step(s, 2000 + k, dtype=torch.float32)
            feat.grad, alog.grad, verts.grad = t(r["grad_feat"]), t(r["grad_alpha_logits"]), t(r["grad_verts"])
            g, o = r["grad_mlp"], 0
            for slot in order:
                size = mlp[slot].numel()
                mlp[slot].grad = t(g[o:o + size].reshape(mlp[slot].shape))
                o += size
            for opt in opts:
                opt.step()

        one(0)
        ts = []
        for k in range(reps):
            t0 = time.perf_counter()
            one(1 + k)
            ts.append(time.perf_counter() - t0)
        return 1.0 / float(np.median(ts))

    try:
        v2 = its(soup)
        v1 = its(scenes.small_random_soup(0))
    finally:
        oracle.set_threads(1)
    return {"value": round(v2, 4), "unit": "it/s", "cores": n, "kind": "port",
            "sample": "median of 5 full iterations after 1 warm-up of the same 15K-triangle 8 x 800x800 workload: "
                      "the C restatement's raster, gather, scatter, edge and opacity kernels on OpenMP threads, "
                      "torch CPU fp32 MLP and MSE, torch Adam + VectorAdam updates",
            "config1_1k_tris_256x256_it_s": round(v1, 3)}


if __name__ == "__main__":
    main()
