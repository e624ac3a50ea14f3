"""Build libdiffsoup_hip.so in-tree with hipcc for gfx950.

Usage: python -m diffsoup_amd.build_lib   (or called from __graft_entry__.build)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libdiffsoup_hip.so")
SOURCES = ["capi.hip", "raster.hip", "multires.hip", "encoding.hip", "edge_grad.hip", "shade.hip", "step.hip", "plain.hip",
           "scan.hip", "optim.hip", "visibility.hip", "ssim.hip", "remesh.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no multiply-add contraction: coverage, barycentrics and depth must round
    # exactly like the CPU restatement
    "-ffp-contract=off",
    # hardware fp32 atomic add on coarse-grained device memory
    "-munsafe-fp-atomics",
    # no packed fp32 (v_pk_fma_f32 / v_pk_mul_f32 / v_pk_add_f32) in the
    # device code.  The device pass honours the feature; the host pass does
    # not know it and warns "not a recognized feature", which is harmless
    # (tests/test_capi.py compiles a probe that would pack and checks).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", CSRC, "-I", os.path.join(os.path.dirname(HERE), "include")] + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
