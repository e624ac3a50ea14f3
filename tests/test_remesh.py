"""Soup splitting (python/diffsoup/remesh.py, src/remesh.cpp,
src/remesh_clip.cpp) on the host: the library's splitter against the
reference's own splitters (compiled from its sources into oracle/_ref, and
their outputs committed in tests/golden/soup_split.npz), and against
properties the geometry fixes: vertex and face counts, area conservation
per source face, children inside their source face, the edge bound after a
threshold split, prefix consistency of the split sequence, hand-computed
splits in world and screen space.  CPU only (the splitter is host code)."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "soup_split.npz")


def _same(got, ref):
    """Bit-identical four-tuples (float arrays compared as bit patterns)."""
    for g, r in zip(got, ref):
        g = g.numpy() if torch.is_tensor(g) else g
        if g.dtype == np.float32:
            g, r = g.view(np.uint32), np.ascontiguousarray(r, np.float32).view(np.uint32)
        np.testing.assert_array_equal(g, r)


def random_soup(n, seed, size=0.1):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.0, 1.0, (n, 1, 3))
    v = (c + rng.uniform(-size, size, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    return v, np.arange(3 * n, dtype=np.int32).reshape(n, 3)


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(GOLDEN))


@pytest.mark.parametrize("name", ["w_cap", "w_tau", "w_ties"])
def test_world_split_matches_reference_vectors(golden, name):
    import diffsoup_amd as ds
    g = golden
    ns, tau = g[f"{name}_args"].tolist()
    got = ds.split_triangle_soup(_t(g[f"{name}_in_verts"]), _t(g[f"{name}_in_faces"]), int(ns), float(tau))
    _same(got, [g[f"{name}_{k}"] for k in ("verts", "faces", "face_map", "flags")])


@pytest.mark.parametrize("name", ["c_cap", "c_tau", "c_wide"])
def test_clip_split_matches_reference_vectors(golden, name):
    import diffsoup_amd as ds
    g = golden
    ns, tau_ratio, aspect = g[f"{name}_args"].tolist()
    H = 256
    W = int(round(H * aspect))
    got = ds.split_triangle_soup_clip((H, W), _t(g[f"{name}_mvp"]), _t(g[f"{name}_in_verts"]),
                                      _t(g[f"{name}_in_faces"]), _t(g[f"{name}_valid"]), int(ns), float(tau_ratio))
    _same(got, [g[f"{name}_{k}"] for k in ("verts", "faces", "face_map", "flags")])


@pytest.mark.skipif(not O.ref_remesh_available(), reason="oracle/_ref not built (the reference's sources are absent)")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_matches_reference_build_on_random_soups(seed):
    """Live comparison with the compiled reference on fresh inputs, including
    shared vertices (a mesh, not a soup) and degenerate triangles."""
    import diffsoup_amd as ds
    rng = np.random.default_rng(seed)
    v, f = random_soup(150, seed)
    f = np.concatenate([f, rng.integers(0, v.shape[0], (40, 3)).astype(np.int32)])   # shared and repeated corners
    for ns, tau in [(1, 0.0), (97, 0.0), (-1, 0.05), (400, 0.02), (-1, -1.0)]:
        _same(ds.split_triangle_soup(_t(v), _t(f), ns, tau), O.ref_split_soup(v, f, ns, tau))
    from diffsoup_amd import scenes
    s = scenes.small_random_soup(seed)
    m = s.mvp[0].astype(np.float32)
    valid = (rng.uniform(size=s.T) < 0.5).astype(np.int32)
    for ns, tr, res in [(300, 0.0, (256, 256)), (-1, 0.03, (200, 300)), (2000, 0.01, (256, 128))]:
        got = ds.split_triangle_soup_clip(res, _t(m), _t(s.verts), _t(s.tri), _t(valid), ns, tr)
        _same(got, O.ref_split_soup_clip(m, s.verts, s.tri, valid, ns, tr, res[1] / res[0]))


def _areas(v, f):
    v = v.astype(np.float64)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)


def test_world_split_properties():
    import diffsoup_amd as ds
    v, f = random_soup(120, 5)
    ns = 333
    V2, F2, fmap, flag = [x.numpy() for x in ds.split_triangle_soup(_t(v), _t(f), ns)]
    # three vertices and one face per split; the input is a prefix
    assert V2.shape[0] == v.shape[0] + 3 * ns and F2.shape[0] == f.shape[0] + ns
    np.testing.assert_array_equal(V2[: v.shape[0]], v)
    # the children of each source face tile it: areas add up, corners lie in its plane and inside it
    src = _areas(v, f)
    summed = np.bincount(fmap, weights=_areas(V2, F2), minlength=f.shape[0])
    np.testing.assert_allclose(summed, src, rtol=1e-5, atol=1e-12)
    a, b, c = (v[f[fmap, k]].astype(np.float64) for k in range(3))
    n = np.cross(b - a, c - a)
    for k in range(3):
        p = V2[F2[:, k]].astype(np.float64)
        # barycentric coordinates of each child corner in its source face
        l1 = np.einsum("ij,ij->i", np.cross(p - a, c - a), n) / np.einsum("ij,ij->i", n, n)
        l2 = np.einsum("ij,ij->i", np.cross(b - a, p - a), n) / np.einsum("ij,ij->i", n, n)
        assert np.all(l1 >= -1e-5) and np.all(l2 >= -1e-5) and np.all(l1 + l2 <= 1 + 1e-5)
    # flags: exactly the source faces never touched
    touched = np.zeros(f.shape[0], bool)
    touched[fmap[f.shape[0]:]] = True
    np.testing.assert_array_equal(flag[: f.shape[0]], (~touched).astype(np.int32))
    assert not flag[f.shape[0]:].any()
    unchanged = flag[: f.shape[0]] == 1
    np.testing.assert_array_equal(F2[: f.shape[0]][unchanged], f[unchanged])


def test_world_threshold_bound_and_prefix_consistency():
    import diffsoup_amd as ds
    v, f = random_soup(60, 9)
    tau = 0.03
    V2, F2, _, _ = [x.numpy() for x in ds.split_triangle_soup_until(_t(v), _t(f), tau)]
    e = np.stack([V2[F2[:, k]] - V2[F2[:, (k + 1) % 3]] for k in range(3)]).astype(np.float32)
    assert np.all((e * e).sum(-1) < np.float32(tau) * np.float32(tau))
    # k splits are the first k of k + 1: the vertex arrays nest and only one face changes
    for k in (5, 40):
        a = [x.numpy() for x in ds.split_triangle_soup(_t(v), _t(f), k)]
        b = [x.numpy() for x in ds.split_triangle_soup(_t(v), _t(f), k + 1)]
        np.testing.assert_array_equal(b[0][: a[0].shape[0]], a[0])
        assert (b[1][: a[1].shape[0]] != a[1]).any(axis=1).sum() == 1
        np.testing.assert_array_equal(b[2][: a[2].shape[0]], a[2])


def test_world_hand_computed_split():
    """One right triangle: the hypotenuse (edge 1, v1 -> v2) is bisected; the
    face becomes (v1, m, v0') and (m', v2, v0) is appended, with vertices
    m, m', v0' appended in that order."""
    import diffsoup_amd as ds
    v = np.array([[0, 0, 0], [2, 0, 0], [0, 1, 0]], np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    V2, F2, fmap, flag = [x.numpy() for x in ds.split_triangle_soup(_t(v), _t(f), 1)]
    np.testing.assert_array_equal(V2, np.array([[0, 0, 0], [2, 0, 0], [0, 1, 0], [1, 0.5, 0], [1, 0.5, 0],
                                                [0, 0, 0]], np.float32))
    np.testing.assert_array_equal(F2, np.array([[1, 3, 5], [4, 2, 0]], np.int32))
    np.testing.assert_array_equal(fmap, [0, 0])
    np.testing.assert_array_equal(flag, [0, 0])
    # no work: zero budget, or unbounded without a threshold
    for ns, tau in [(0, 0.0), (-1, 0.0)]:
        V3, F3, _, flag3 = [x.numpy() for x in ds.split_triangle_soup(_t(v), _t(f), ns, tau)]
        np.testing.assert_array_equal(V3, v)
        np.testing.assert_array_equal(flag3, [1])


def _perspective(fovy_deg=50.0, near=0.1, far=10.0, eye_z=3.0):
    t = 1.0 / np.tan(np.radians(fovy_deg) / 2)
    proj = np.array([[t, 0, 0, 0], [0, t, 0, 0], [0, 0, -(far + near) / (far - near), -2 * far * near / (far - near)],
                     [0, 0, -1, 0]], np.float64)
    view = np.eye(4)
    view[2, 3] = -eye_z
    return (proj @ view).astype(np.float32)


def test_clip_split_point_is_the_screen_midpoint():
    """A slanted edge seen in perspective: the split point projects onto the
    NDC midpoint of the projected ends (not the world midpoint), and its
    screen length is halved."""
    import diffsoup_amd as ds
    m = _perspective()
    v = np.array([[-0.8, -0.3, -1.0], [0.6, 0.4, 1.2], [0.0, 0.5, 0.0]], np.float32)
    f = np.array([[0, 1, 2]], np.int32)
    V2, F2, _, _ = [x.numpy() for x in ds.split_triangle_soup_clip((100, 100), _t(m), _t(v), _t(f),
                                                                    _t(np.ones(1, np.int32)), 1)]

    def ndc(p):
        c = m.astype(np.float64) @ np.append(p.astype(np.float64), 1.0)
        return c[:3] / c[3]
    lens = [np.linalg.norm((ndc(v[a]) - ndc(v[b]))[:2]) for a, b in ((0, 1), (1, 2), (2, 0))]
    e = int(np.argmax(lens))
    a, b = e, (e + 1) % 3
    mid = 0.5 * (ndc(v[a]) + ndc(v[b]))
    np.testing.assert_allclose(ndc(V2[3])[:2], mid[:2], atol=2e-6)
    # the point lies on the world segment a-b but away from its world midpoint
    d = V2[3].astype(np.float64) - v[a]
    seg = v[b].astype(np.float64) - v[a]
    s = d @ seg / (seg @ seg)
    np.testing.assert_allclose(d, s * seg, atol=1e-5)
    assert abs(s - 0.5) > 0.02


def test_clip_split_skips_offscreen_and_ineligible_faces():
    import diffsoup_amd as ds
    m = _perspective()
    v = np.array([[-0.5, -0.5, 0], [0.5, -0.5, 0], [0, 0.5, 0],        # on screen, eligible
                  [-0.5, -0.5, 0.1], [0.5, -0.5, 0.1], [0, 0.5, 0.1],  # on screen, not eligible
                  [-9, 0, 0], [0.5, 0, 0], [0, 0.3, 0]], np.float32)   # one corner far outside the view
    f = np.arange(9, dtype=np.int32).reshape(3, 3)
    valid = np.array([1, 0, 1], np.int32)
    V2, F2, fmap, flag = [x.numpy() for x in ds.split_triangle_soup_clip((64, 64), _t(m), _t(v), _t(f), _t(valid),
                                                                          -1, 0.05)]
    assert flag[1] == 1 and np.array_equal(F2[1], f[1])                  # not eligible: untouched
    # the face with a corner off screen only ever splits its on-screen edge
    # (7, 8): the two edges to vertex 6 are never queued, so every child
    # keeps a corner at vertex 6's position
    kids = np.nonzero(fmap == 2)[0]
    assert kids.size > 1
    assert all(np.any(np.all(V2[F2[k]] == v[6], axis=1)) for k in kids)
    assert (fmap == 0).sum() > 1


def test_split_rejects_bad_faces():
    import diffsoup_amd as ds
    from diffsoup_amd._lib import DiffSoupError
    v = np.zeros((3, 3), np.float32)
    with pytest.raises(DiffSoupError, match="outside"):
        ds.split_triangle_soup(_t(v), _t(np.array([[0, 1, 3]], np.int32)), 1)


def test_expand_and_remove_unreferenced():
    import diffsoup_amd as ds
    src = torch.arange(12.0).reshape(4, 3)
    np.testing.assert_array_equal(ds.expand_by_index(src, torch.tensor([3, 0, 0])).numpy(),
                                  src.numpy()[[3, 0, 0]])
    with pytest.raises(ValueError):
        ds.expand_by_index(src, torch.tensor([4]))
    v = torch.arange(18.0).reshape(6, 3)
    f = torch.tensor([[4, 1, 5], [5, 1, 4]], dtype=torch.int32)
    v2, f2 = ds.remove_unreferenced_vertices_from_soup(v, f)
    np.testing.assert_array_equal(v2.numpy(), v.numpy()[[1, 4, 5]])
    np.testing.assert_array_equal(f2.numpy(), [[1, 0, 2], [2, 0, 1]])
    assert f2.dtype == torch.int32


def test_core_dropin_splitters_match_reference_vectors(golden):
    """The drop-in _core entry points, NumPy in and out, as the reference's
    remesh.py calls them (python/diffsoup/remesh.py:37-39, 123-131)."""
    from diffsoup_amd import _core
    g = golden
    ns, tau = g["w_cap_args"].tolist()
    got = _core.split_triangle_soup(g["w_cap_in_verts"], g["w_cap_in_faces"], int(ns), float(tau))
    _same(got, [g[f"w_cap_{k}"] for k in ("verts", "faces", "face_map", "flags")])
    ns, tau_ratio, aspect = g["c_wide_args"].tolist()
    got = _core.split_triangle_soup_clip(g["c_wide_mvp"], g["c_wide_in_verts"], g["c_wide_in_faces"],
                                         g["c_wide_valid"], int(ns), float(tau_ratio), float(aspect))
    _same(got, [g[f"c_wide_{k}"] for k in ("verts", "faces", "face_map", "flags")])
    with pytest.raises(TypeError):
        _core.split_triangle_soup(g["w_cap_in_verts"].astype(np.float64), g["w_cap_in_faces"], 1, 0.0)


@pytest.mark.skipif(not O.ref_remesh_available(), reason="oracle/_ref not built (the reference's sources are absent)")
def test_edge_cases_match_reference_build():
    """Empty soups, vertices without faces, degenerate triangles (repeated
    corners: zero-length edges are still queued in world space and never in
    screen space), nothing eligible, a zero and a one-split budget."""
    import diffsoup_amd as ds
    empty_v, empty_f = np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32)
    _same(ds.split_triangle_soup(_t(empty_v), _t(empty_f), 5), O.ref_split_soup(empty_v, empty_f, 5))
    lone = np.random.default_rng(4).uniform(size=(7, 3)).astype(np.float32)
    _same(ds.split_triangle_soup(_t(lone), _t(empty_f), 5, 0.1), O.ref_split_soup(lone, empty_f, 5, 0.1))
    v, f = random_soup(20, 11)
    f[3] = [f[3][0], f[3][0], f[3][0]]          # a point
    f[4] = [f[4][0], f[4][1], f[4][1]]          # a segment
    for ns, tau in [(0, 0.0), (1, 0.0), (60, 0.0), (-1, 0.08)]:
        _same(ds.split_triangle_soup(_t(v), _t(f), ns, tau), O.ref_split_soup(v, f, ns, tau))
    from diffsoup_amd import scenes
    s = scenes.small_random_soup(3)
    m = s.mvp[0].astype(np.float32)
    none = np.zeros(s.T, np.int32)
    _same(ds.split_triangle_soup_clip((256, 256), _t(m), _t(s.verts), _t(s.tri), _t(none), -1, 0.01),
          O.ref_split_soup_clip(m, s.verts, s.tri, none, -1, 0.01, 1.0))
    tri = s.tri.copy()
    tri[:50, 1] = tri[:50, 0]                    # degenerate on screen too
    ok = np.ones(s.T, np.int32)
    _same(ds.split_triangle_soup_clip((256, 256), _t(m), _t(s.verts), _t(tri), _t(ok), 400, 0.0),
          O.ref_split_soup_clip(m, s.verts, tri, ok, 400, 0.0, 1.0))
