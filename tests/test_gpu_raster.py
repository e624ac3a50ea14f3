"""GPU parity of the HIP rasterizer against the CPU oracle.

The raster output (b0, b1, depth, triangle id + 1) must be bit-identical to
the oracle's for the same inputs and seed, stochastic or deterministic, for
every tile size, for uneven views packed in one launch, and when the bin
capacity overflows."""
import numpy as np
import pytest
import torch

from diffsoup_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _sig(x):
    return (1.0 / (1.0 + np.exp(-x.astype(np.float64)))).astype(np.float32)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def small():
    s = scenes.small_random_soup(0)
    return s, scenes.project_vertices_np(s.verts, s.mvp), _sig(s.alpha_logits)


@pytest.fixture(scope="module")
def medium():
    s = scenes.sphere_torus_soup(0)
    return s, scenes.project_vertices_np(s.verts, s.mvp), _sig(s.alpha_logits)


def _gpu_rast(dev, pos, tri, level, alpha, res, stochastic, seed, tile=16):
    import diffsoup_amd as ds
    B = pos.shape[0]
    T = tri.shape[0]
    out = ds.rasterize_multires_triangle_alpha(
        res, torch.from_numpy(pos).to(dev), torch.from_numpy(tri).to(dev), level,
        torch.from_numpy(alpha).reshape(T, -1, 1).contiguous().to(dev), stochastic, seed=seed, tile=tile)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(B, res[0], res[1], 4)


@pytest.mark.parametrize("stochastic,seed", [(False, 0), (True, 1), (True, 99)])
@pytest.mark.parametrize("tile", [8, 16, 32])
def test_small_soup_bit_exact(dev, small, stochastic, seed, tile):
    s, pos, alpha = small
    print(f"small soup tile={tile} stochastic={stochastic} seed={seed}", flush=True)
    ref = O.rasterize(s.resolutions[0], pos, s.tri, 0, alpha, stochastic, seed)
    got = _gpu_rast(dev, pos, s.tri, 0, alpha, s.resolutions[0], stochastic, seed, tile)
    assert (ref[..., 3] > 0).sum() > 1000
    np.testing.assert_array_equal(_bits(got), _bits(ref))


@pytest.mark.parametrize("stochastic,seed", [(False, 0), (True, 7)])
def test_sphere_torus_8x800_bit_exact(dev, medium, stochastic, seed):
    s, pos, alpha = medium
    print(f"sphere-torus 8x800^2 stochastic={stochastic}", flush=True)
    ref = O.rasterize(s.resolutions[0], pos, s.tri, s.alpha_level, alpha, stochastic, seed)
    got = _gpu_rast(dev, pos, s.tri, s.alpha_level, alpha, s.resolutions[0], stochastic, seed)
    ids_equal = (got[..., 3] == ref[..., 3]).mean()
    assert ids_equal == 1.0, f"triangle ids differ on {(1 - ids_equal) * got[..., 3].size:.0f} pixels"
    np.testing.assert_array_equal(_bits(got), _bits(ref))


def test_covered_fragment_count(dev, medium):
    import diffsoup_amd as ds
    s, pos, alpha = medium
    T = s.T
    rast, st = ds.rasterize_views(s.resolutions, torch.from_numpy(pos).to(dev),
                                  torch.from_numpy(s.tri).to(dev), s.alpha_level,
                                  torch.from_numpy(alpha).to(dev), True, seed=3, stats=True)
    st = st.cpu().numpy()
    n_ref = O.covered_fragments(800, 800, pos, s.tri)
    assert st[0] == n_ref
    assert st[2] == 0 and st[3] == 0


def test_uneven_views_one_launch(dev):
    import diffsoup_amd as ds
    s = scenes.sphere_torus_soup(1, n_surface=3000, n_float=1000, views=4)
    res = [(840, 1297), (546, 979), (822, 1237), (800, 800)]
    mvp = scenes.orbit_cameras(4, 3.0, res)
    pos = scenes.project_vertices_np(s.verts, mvp)
    alpha = _sig(s.alpha_logits)
    rast = ds.rasterize_views(res, torch.from_numpy(pos).to(dev), torch.from_numpy(s.tri).to(dev),
                              s.alpha_level, torch.from_numpy(alpha).to(dev), True, seed=11)
    got = rast.cpu().numpy()
    off = 0
    for b, (H, W) in enumerate(res):
        # the hash takes the view index within the batch: thresholds use view b
        fp, fa = O.fragments(H, W, pos[b:b + 1], s.tri)
        fp[:, 0] = b
        al = O.multires_alpha(fa, s.alpha_level, alpha)
        th = O.thresholds(fp, fa, True, 11)
        fp[:, 0] = 0
        ref = np.zeros((1, H, W, 4), np.float32)
        O.lib().dso_depth_test(1, H, W, fp.shape[0], fp, fa, al, th, ref)
        np.testing.assert_array_equal(_bits(got[off:off + H * W]), _bits(ref.reshape(-1, 4)))
        off += H * W


def test_capacity_overflow_falls_back_exactly(dev, small):
    import diffsoup_amd as ds
    from diffsoup_amd import _lib
    s, pos, alpha = small
    H, W = s.resolutions[0]
    T = s.T
    L = _lib.lib()
    Ha, Wa = _lib.int_array([H]), _lib.int_array([W])
    cap = 100
    nbytes = L.ds_raster_workspace_size(1, Ha, Wa, T, 16, cap)
    ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=dev)
    base = (ws.data_ptr() + 255) & ~255
    rast = torch.empty(H * W, 4, device=dev)
    pos_t = torch.from_numpy(pos).to(dev)
    tri_t = torch.from_numpy(s.tri).to(dev)
    a_t = torch.from_numpy(alpha).to(dev)
    rc = L.ds_rasterize_fwd(1, Ha, Wa, pos.shape[1], _lib.ptr(pos_t), T, _lib.ptr(tri_t), 0, _lib.ptr(a_t), 1, 5,
                            16, _lib.ptr(rast), base, nbytes, cap, _lib.stream_ptr(dev))
    assert rc == 0
    torch.cuda.synchronize()
    off = base - ws.data_ptr()
    stats = ws[off:off + 64].view(torch.int64).cpu().numpy()
    assert stats[2] == 1 and stats[1] > cap
    ref = O.rasterize((H, W), pos, s.tri, 0, alpha, True, 5)
    np.testing.assert_array_equal(_bits(rast.cpu().numpy()), _bits(ref.reshape(-1, 4)))


def test_empty_and_culled_inputs(dev):
    import diffsoup_amd as ds
    # all triangles behind the camera (w <= 0) or outside: background only
    pos = np.zeros((2, 6, 4), np.float32)
    pos[..., 3] = -1.0
    tri = np.arange(6, dtype=np.int32).reshape(2, 3)
    alpha = np.ones((2, 3, 1), np.float32)
    out = ds.rasterize_multires_triangle_alpha((33, 17), torch.from_numpy(pos).to(dev), torch.from_numpy(tri).to(dev),
                                               0, torch.from_numpy(alpha).to(dev), True, seed=1)
    assert out.shape == (2, 33, 17, 4) and out.abs().sum().item() == 0
    # zero triangles
    out = ds.rasterize_multires_triangle_alpha((8, 8), torch.zeros(1, 0, 4, device=dev),
                                               torch.zeros(0, 3, dtype=torch.int32, device=dev), 0,
                                               torch.zeros(0, 3, 1, device=dev), True, seed=1)
    assert out.abs().sum().item() == 0
