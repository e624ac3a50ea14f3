"""ds_rasterize_fwd_noalloc (workspace in a fixed per-GPU device array)
against ds_rasterize_fwd (caller workspace): bit-identical rasters on two
calls in a row through the C-ABI, the second with a smaller batch, so the
scratch left by the first call must not leak into the second; and a batch
beyond the fixed scratch is rejected on the host."""
import ctypes

import numpy as np
import pytest
import torch

from diffsoup_amd import _lib, scenes
from diffsoup_amd.multires import feats_at_level

pytestmark = pytest.mark.gpu

_SIG = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
        ctypes.c_void_p, ctypes.c_void_p]


def _lib_noalloc():
    L = _lib.lib()
    fn = L.ds_rasterize_fwd_noalloc
    fn.argtypes = _SIG
    fn.restype = ctypes.c_int
    return L, fn


def _raster_pair(dev, s, seed, tile):
    L, noalloc = _lib_noalloc()
    pos = torch.from_numpy(scenes.project_vertices_np(s.verts, s.mvp)).to(dev).contiguous()
    tri = torch.from_numpy(s.tri).to(dev)
    alpha = torch.from_numpy((1 / (1 + np.exp(-s.alpha_logits.astype(np.float64)))).astype(np.float32)).to(dev)
    B, V = pos.shape[0], pos.shape[1]
    H = _lib.int_array([h for h, _ in s.resolutions])
    W = _lib.int_array([w for _, w in s.resolutions])
    P = sum(h * w for h, w in s.resolutions)
    stream = _lib.stream_ptr(dev)
    # the reference entry point with its own workspace
    cap = max(4 * B * s.T, 1 << 16)
    nbytes = int(L.ds_raster_workspace_size(B, H, W, s.T, tile, cap))
    ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=dev)
    wptr = (ws.data_ptr() + 255) & ~255
    want = torch.full((P, 4), -7.0, device=dev)
    _lib.check(L.ds_rasterize_fwd(B, H, W, V, pos.data_ptr(), s.T, tri.data_ptr(), s.alpha_level, alpha.data_ptr(),
                                  1, seed, tile, want.data_ptr(), wptr, nbytes, cap, stream), "ds_rasterize_fwd")
    got = torch.full((P, 4), -7.0, device=dev)
    _lib.check(noalloc(B, H, W, V, pos.data_ptr(), s.T, tri.data_ptr(), s.alpha_level, alpha.data_ptr(), 1, seed,
                       tile, got.data_ptr(), stream), "ds_rasterize_fwd_noalloc")
    torch.cuda.synchronize()
    return got.cpu().numpy(), want.cpu().numpy()


def test_noalloc_matches_workspace_entry_point(dev):
    big = scenes.sphere_torus_soup(3, n_surface=6000, n_float=1500, views=4, res=(320, 352))
    small = scenes.small_random_soup(5, T=700, res=(128, 96))
    for name, s, seed, tile in (("4 views 320x352, 7.5K tris", big, 91, 16),
                                ("1 view 128x96, 700 tris", small, 92, 8)):
        print(f"noalloc vs workspace: {name}, tile {tile}", flush=True)
        got, want = _raster_pair(dev, s, seed, tile)
        assert (want[:, 3] > 0).any()
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32), err_msg=name)


def test_noalloc_rejects_batches_beyond_the_scratch(dev):
    L, noalloc = _lib_noalloc()
    B, T = 64, 20000   # 64 x 20000 x 64 B of records alone exceed 64 MiB
    H = _lib.int_array([64] * B)
    W = _lib.int_array([64] * B)
    rc = noalloc(B, H, W, 3, None, T, None, 0, None, 1, 0, 16, None, _lib.stream_ptr(dev))
    assert rc == -1
    assert b"scratch" in L.ds_last_error()
