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
