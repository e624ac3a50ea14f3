"""The C-ABI library loads on a machine without a GPU, exports every symbol
include/diffsoup_hip.h declares, and rejects bad arguments on the host
without launching anything.  CPU only."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "diffsoup_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ds_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    from diffsoup_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libdiffsoup_hip.so not built (run __graft_entry__.build())")
    return _lib.lib()


def test_every_declared_symbol_is_exported(L):
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_host_queries(L):
    assert L.ds_version().decode().startswith("diffsoup")
    assert [L.ds_feats_at_level(l) for l in range(5)] == [3, 6, 15, 45, 153]
    from diffsoup_amd import _lib
    H, W = _lib.int_array([800, 600]), _lib.int_array([800, 1000])
    assert L.ds_raster_tile_count(2, H, W, 16) == 50 * 50 + 38 * 63
    n = L.ds_raster_workspace_size(2, H, W, 1000, 16, 1 << 16)
    assert n > 2 * 1000 * 64


def test_bad_arguments_fail_without_a_gpu(L):
    from diffsoup_amd import _lib
    H, W = _lib.int_array([64]), _lib.int_array([64])
    rc = L.ds_rasterize_fwd(0, H, W, 3, None, 1, None, 0, None, 1, 0, 16, None, None, 0, 0, None)
    assert rc == -1 and b"view count" in L.ds_last_error()
    rc = L.ds_rasterize_fwd(1, H, W, 3, None, 1, None, 0, None, 1, 0, 12, None, None, 0, 0, None)
    assert rc == -1 and b"tile size" in L.ds_last_error()
    rc = L.ds_multires_color_fwd(10, None, 2, 1, 1, 4, None, 0, None, None)
    assert rc == -1 and b"level range" in L.ds_last_error()
    Hbig = _lib.int_array([70000])
    assert L.ds_raster_tile_count(1, Hbig, W, 16) == -1
    from diffsoup_amd import losses  # noqa: F401  (declares the SSIM signatures)
    assert L.ds_ssim_tiles(1, H, W, 11, None) == 4          # 54 x 54 windows in 60 x 16 tiles
    assert L.ds_ssim_tiles(1, H, W, 7, None) == -1 and b"ssim_tiles" in L.ds_last_error()


def test_core_dropin_surface_and_type_checks(L):
    """The drop-in _core module offers every function the reference's Python
    layer binds (src/main.cpp:30-385) and, like nanobind, raises TypeError on
    a wrong dtype, rank or layout before any GPU work."""
    import torch

    from diffsoup_amd import _core
    ref_names = ["compute_triangle_rects", "compute_fragments", "depth_test", "filter_valid_fragments",
                 "backward_edge_grad", "backward_opacity_aux_loss", "encode_view_dir_sh2",
                 "multires_triangle_alpha", "backward_multires_triangle_alpha", "multires_triangle_color",
                 "backward_multires_triangle_color", "accumulate_to_level_forward", "accumulate_to_level_backward",
                 "split_triangle_soup", "split_triangle_soup_clip"]
    assert all(callable(getattr(_core, n)) for n in ref_names)
    assert _core.__version__ == "0.1.0"
    pos = torch.zeros(1, 3, 4, dtype=torch.float64)
    tri = torch.zeros(1, 3, dtype=torch.int32)
    with pytest.raises(TypeError, match="pos"):
        _core.compute_triangle_rects(8, 8, pos, tri, torch.zeros(1, 4, dtype=torch.int32),
                                     torch.zeros(1, dtype=torch.int32))
    with pytest.raises(TypeError, match="frag_attrs"):
        _core.multires_triangle_alpha(torch.zeros(4, 4, dtype=torch.float32).t(), 0, 0,
                                      torch.zeros(1, 3), torch.zeros(4))
    # a CPU tensor of the right type is refused loudly: there is no CPU path
    from diffsoup_amd._lib import DiffSoupError
    with pytest.raises(DiffSoupError, match="GPU tensor"):
        _core.encode_view_dir_sh2(torch.zeros(1, 2, 2, 4), torch.zeros(1, 4, 4), torch.zeros(1, 2, 2, 9))


def test_device_code_has_no_packed_fp32(tmp_path):
    """The library's hipcc flags keep packed fp32 (v_pk_fma_f32 and friends)
    out of the device code: a probe whose float2 arithmetic hipcc -O3 packs
    without the flag compiles to unpacked instructions with them."""
    import shutil
    import subprocess

    from diffsoup_amd import build_lib

    if not os.path.exists("/opt/rocm/bin/hipcc") and not shutil.which("hipcc"):
        pytest.skip("hipcc not present")
    src = tmp_path / "probe.hip"
    src.write_text("#include <hip/hip_runtime.h>\n"
                   "__global__ void k(float2* a, const float2* b, float s) {\n"
                   "  int i = threadIdx.x; float2 x = a[i], y = b[i];\n"
                   "  x.x = x.x * s + y.x; x.y = x.y * s + y.y; a[i] = x; }\n")
    packed = re.compile(r"v_pk_(fma|mul|add)_f32")

    def device_asm(flags, name):
        out = tmp_path / name
        flags = [f for f in flags if f not in ("-shared", "-fPIC")]
        subprocess.run([build_lib.hipcc()] + flags + ["--cuda-device-only", "-S", str(src), "-o", str(out)],
                       check=True, capture_output=True)
        return out.read_text()

    with_flag = device_asm(build_lib.HIPCC_FLAGS, "with.s")
    assert not packed.search(with_flag)
    i = build_lib.HIPCC_FLAGS.index("-packed-fp32-ops")
    without = build_lib.HIPCC_FLAGS[:i - 3] + build_lib.HIPCC_FLAGS[i + 1:]
    assert packed.search(device_asm(without, "without.s")), "the probe no longer packs: it checks nothing"
