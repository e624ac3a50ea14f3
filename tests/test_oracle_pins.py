"""Pins for the CPU oracle (oracle/ds_oracle.c): every check here is fixed by
something other than the oracle's own formulas - published known-answer
vectors, closed forms, brute force in float64, adjoint identities, or the
statistics a correct mask must have.  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O


# --------------------------------------------------------------------------
# Mask hash
# --------------------------------------------------------------------------

def test_splitmix64_known_answers():
    # Published SplitMix64 output sequences (Steele, Lea & Flood 2014; the
    # seed-1234567 vector is the one used by common reimplementations' tests).
    assert O.splitmix64(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert O.splitmix64(1234567, 5) == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                        4593380528125082431, 16408922859458223821]


def test_mask_threshold_is_uniform_on_24_bits():
    rng = np.random.default_rng(0)
    us = np.array([O.mask_threshold(7, int(b), int(t), int(y), int(x))
                   for b, t, y, x in rng.integers(0, [8, 100000, 1080, 1920], size=(20000, 4))])
    assert us.min() >= 0.0 and us.max() < 1.0
    # every value is a multiple of 2^-24
    assert np.all(np.floor(us * 2**24) == us * 2**24)
    # mean and variance of U[0,1): 1/2, 1/12 (4-sigma bounds)
    n = len(us)
    assert abs(us.mean() - 0.5) < 4 * math.sqrt(1 / 12 / n)
    assert abs(us.var() - 1 / 12) < 0.005
    # P(keep) = P(u <= alpha) = alpha, binomial 4-sigma
    for a in (0.1, 0.5, 0.9):
        k = (us <= a).mean()
        assert abs(k - a) < 4 * math.sqrt(a * (1 - a) / n)


def test_mask_threshold_depends_on_every_coordinate():
    base = O.mask_threshold(3, 1, 2, 3, 4)
    for args in [(4, 1, 2, 3, 4), (3, 2, 2, 3, 4), (3, 1, 3, 3, 4), (3, 1, 2, 4, 4), (3, 1, 2, 3, 5)]:
        assert O.mask_threshold(*args) != base


# --------------------------------------------------------------------------
# Lattice
# --------------------------------------------------------------------------

def test_feats_at_level_closed_form():
    for L in range(0, 9):
        R = 1 << L
        assert O.feats_at_level(L) == (R + 1) * (R + 2) // 2
    assert [O.feats_at_level(L) for L in range(5)] == [3, 6, 15, 45, 153]


def _interp(b0, b1, lo, hi, feat):
    out = np.zeros(feat.shape[1], np.float32)
    O.lib().dso_interp(b0, b1, lo, hi, feat.shape[1], np.ascontiguousarray(feat, np.float32), out)
    return out


def _lattice_points(L):
    R = 1 << L
    pts = []
    for n in range(R + 1):
        for y in range(n + 1):
            pts.append((n - y, y))   # index n(n+1)/2 + y
    return pts


@pytest.mark.parametrize("L", [0, 1, 2, 3, 4])
def test_lattice_vertex_identity_and_linear_precision(L):
    R = 1 << L
    pts = _lattice_points(L)
    assert len(pts) == O.feats_at_level(L)
    # at level 0 index 0,1,2 are the triangle corners (1,0)=vertex0, (0,1)=vertex1 ... check by
    # value: a feature equal to the sample's own (x, y) must be returned at every lattice vertex
    feat = np.array([[x / R, y / R, 1.0] for x, y in pts], np.float32)
    for x, y in pts:
        out = _interp(x / R, y / R, L, L, feat)
        np.testing.assert_allclose(out, [x / R, y / R, 1.0], atol=1e-6)
    # affine data a + b*u + c*v is reproduced at arbitrary points (linear precision)
    rng = np.random.default_rng(L)
    coef = rng.normal(size=(3, 4))
    feat = np.array([coef[0] + coef[1] * (x / R) + coef[2] * (y / R) for x, y in pts], np.float32)
    for _ in range(200):
        u, v = rng.uniform(0, 1, 2)
        if u + v > 1:
            u, v = 1 - u, 1 - v
        u, v = np.float32(u), np.float32(v)
        out = _interp(float(u), float(v), L, L, feat)
        np.testing.assert_allclose(out, coef[0] + coef[1] * float(u) + coef[2] * float(v), atol=2e-5)


def test_level_sum_and_offsets():
    rng = np.random.default_rng(1)
    S = O.feats_in_levels(1, 3)
    feat = rng.normal(size=(S, 2)).astype(np.float32)
    u, v = 0.3, 0.45
    off = 0
    tot = np.zeros(2, np.float64)
    for L in (1, 2, 3):
        tot += _interp(u, v, L, L, feat[off:off + O.feats_at_level(L)])
        off += O.feats_at_level(L)
    np.testing.assert_allclose(_interp(u, v, 1, 3, feat), tot, atol=1e-6)


# --------------------------------------------------------------------------
# Coverage, rects, brute force
# --------------------------------------------------------------------------

def test_intersect_recovers_planted_barycentrics():
    rng = np.random.default_rng(2)
    for _ in range(200):
        P = rng.uniform(-1, 1, size=(3, 4)).astype(np.float32)
        P[:, 3] = rng.uniform(0.5, 3.0, size=3)
        b = rng.dirichlet([1, 1, 1])
        # perspective-correct: the screen point of barycentrics b in clip space
        clip = (b[:, None] * P.astype(np.float64)).sum(0)
        nx, ny = clip[0] / clip[3], clip[1] / clip[3]
        ok, (b0, b1, z, w) = O.intersect(np.float32(ny), np.float32(nx), P[0], P[1], P[2])
        if min(b) > 1e-3:
            assert ok
        np.testing.assert_allclose([b0, b1], b[:2], atol=2e-4)
        np.testing.assert_allclose(z / w, clip[2] / clip[3], atol=2e-4)


def _clip_from_ndc(pts_ndc, w=1.0, z=0.0):
    return np.array([[x * w, y * w, z * w, w] for x, y in pts_ndc], np.float32)


def test_rects_closed_form():
    H, W = 10, 20
    # NDC box [-0.5,0.5]x[-0.2,0.6]: x pixels floor(0.25*20)=5 .. floor(0.75*20)=15 -> [5,16)
    tri = np.array([[0, 1, 2]], np.int32)
    P = _clip_from_ndc([(-0.5, -0.2), (0.5, -0.2), (0.0, 0.6)])[None]
    rects, counts, total = O.triangle_rects(H, W, P, tri)
    # y: floor(0.4*10)=4 .. floor(0.8*10)=8 -> [4,9)
    assert rects.tolist() == [[4, 5, 5, 11]] and total == 55
    # w <= 0 on one vertex culls
    P2 = P.copy()
    P2[0, 1, 3] = -0.1
    assert O.triangle_rects(H, W, P2, tri)[2] == 0
    # all vertices right of the frustum (shared outcode bit) culls
    P3 = _clip_from_ndc([(1.5, 0), (2, 0.2), (1.2, -0.3)])[None]
    assert O.triangle_rects(H, W, P3, tri)[2] == 0
    # partially outside: clamped to the image
    P4 = _clip_from_ndc([(-3, -3), (3, -3), (0, 3)])[None]
    assert O.triangle_rects(H, W, P4, tri)[0].tolist() == [[0, H, 0, W]]


def _brute_force(H, W, pos, tri, alpha):
    """float64 per-pixel x per-triangle rasterizer with a 0.5 threshold: the
    winning triangle id (or -1) and whether the decision is unambiguous."""
    B = pos.shape[0]
    ids = -np.ones((B, H, W), np.int64)
    clear = np.ones((B, H, W), bool)
    P64 = pos.astype(np.float64)
    for b in range(B):
        for y in range(H):
            for x in range(W):
                nx = -1 + 2 * (x + 0.5) / W
                ny = -1 + 2 * (y + 0.5) / H
                best, bestd, second = -1, np.inf, np.inf
                for t in range(tri.shape[0]):
                    p = P64[b, tri[t]]
                    q = p[:, :2] - np.array([nx, ny]) * p[:, 3:4]
                    A0 = q[1, 0] * q[2, 1] - q[1, 1] * q[2, 0]
                    A1 = q[2, 0] * q[0, 1] - q[2, 1] * q[0, 0]
                    A2 = q[0, 0] * q[1, 1] - q[0, 1] * q[1, 0]
                    A = A0 + A1 + A2
                    if A == 0:
                        continue
                    bb = np.array([A0, A1, A2]) / A
                    if np.min(np.abs(bb)) < 1e-4 and np.min(bb) > -1e-4:
                        clear[b, y, x] = False
                    if np.min(bb) < 0:
                        continue
                    d = (bb @ p[:, 2]) / (bb @ p[:, 3])
                    if not (-1 < d <= 1) or alpha[t] < 0.5:
                        continue
                    if d < bestd:
                        second, bestd, best = bestd, d, t
                    elif d < second:
                        second = d
                if second - bestd < 1e-5:
                    clear[b, y, x] = clear[b, y, x] and best == -1
                ids[b, y, x] = best
    return ids, clear


def test_oracle_matches_float64_brute_force():
    rng = np.random.default_rng(3)
    H, W, T = 24, 32, 12
    verts = rng.uniform(-0.8, 0.8, size=(3 * T, 3)).astype(np.float32)
    verts[:, 2] = rng.uniform(-0.5, 0.5, size=3 * T)
    pos = np.concatenate([verts, np.ones((3 * T, 1), np.float32)], -1)[None]
    pos[..., 3] = rng.uniform(0.8, 1.2, size=3 * T)
    pos[..., :3] *= pos[..., 3:4]
    tri = np.arange(3 * T, dtype=np.int32).reshape(T, 3)
    alpha_src = rng.uniform(0, 1, size=(T, 3)).astype(np.float32)
    alpha_src[:, :] = alpha_src[:, :1]  # constant per triangle
    rast = O.rasterize((H, W), pos, tri, 0, alpha_src, stochastic=False)
    ids, clear = _brute_force(H, W, pos, tri, alpha_src[:, 0])
    got = rast[..., 3].astype(np.int64) - 1
    assert clear.mean() > 0.8
    np.testing.assert_array_equal(got[clear], ids[clear])
    assert (ids >= 0).sum() > 50


# --------------------------------------------------------------------------
# Adjoints and SH
# --------------------------------------------------------------------------

def test_color_gather_adjoint():
    rng = np.random.default_rng(4)
    T, lo, hi, C = 5, 0, 3, 3
    S = O.feats_in_levels(lo, hi)
    rast = np.zeros((40, 4), np.float32)
    u = rng.uniform(0, 1, size=(40, 2))
    u = np.where(u.sum(1, keepdims=True) > 1, 1 - u, u)
    rast[:, :2] = u
    rast[:, 3] = rng.integers(0, T + 1, size=40)  # some background
    for concat in (False, True):
        x = rng.normal(size=(T, S, C)).astype(np.float32)
        fx = O.multires_color(rast, lo, hi, x, concat).astype(np.float64)
        y = rng.normal(size=fx.shape).astype(np.float32)
        aty = O.multires_color_bwd(rast, lo, hi, x.shape, y, concat).astype(np.float64)
        assert abs((fx * y).sum() - (x * aty).sum()) < 1e-4 * (1 + abs((fx * y).sum()))
        assert np.all(fx[rast[:, 3] == 0] == 0)


def test_concat_slice_equals_single_level():
    rng = np.random.default_rng(5)
    T, C = 3, 2
    S = O.feats_in_levels(1, 3)
    x = rng.normal(size=(T, S, C)).astype(np.float32)
    rast = np.array([[0.2, 0.3, 0.0, 1.0], [0.6, 0.1, 0.0, 3.0]], np.float32)
    cat = O.multires_color(rast, 1, 3, x, True)
    off = 0
    for k, L in enumerate((1, 2, 3)):
        s = O.feats_at_level(L)
        one = O.multires_color(rast, L, L, np.ascontiguousarray(x[:, off:off + s]), False)
        np.testing.assert_array_equal(cat[:, k * C:(k + 1) * C], one)
        off += s


def test_accumulate_identity_and_adjoint():
    rng = np.random.default_rng(6)
    T, C = 4, 3
    # a single level accumulated onto itself is the identity
    x = rng.normal(size=(T, O.feats_at_level(3), C)).astype(np.float32)
    np.testing.assert_allclose(O.accumulate_fwd(3, 3, 3, x), x, atol=1e-6)
    # adjoint identity for a multi-level stack
    S = O.feats_in_levels(0, 3)
    x = rng.normal(size=(T, S, C)).astype(np.float32)
    fx = O.accumulate_fwd(0, 3, 2, x).astype(np.float64)
    y = rng.normal(size=fx.shape).astype(np.float32)
    aty = O.accumulate_bwd(0, 3, 2, x.shape, y).astype(np.float64)
    assert abs((fx * y).sum() - (x * aty).sum()) < 1e-4 * (1 + abs((fx * y).sum()))
    # accumulated value at a target vertex equals the level-summed gather there
    rast = np.array([[0.5, 0.25, 0, 1]], np.float32)   # (x,y)=(2,1) on level 2
    g = O.multires_color(rast, 0, 3, x, False)[0]
    k = (3 * 4) // 2 + 1
    np.testing.assert_allclose(O.accumulate_fwd(0, 3, 2, x)[0, k], g, atol=1e-5)


def test_sh2_orthonormal_basis_and_axis_direction():
    # basis orthonormality over the sphere pins the constants
    rng = np.random.default_rng(7)
    d = rng.normal(size=(200000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x, y, z = d.T
    c0, c1 = 0.28209479177387814, 0.4886025119029199
    Y = np.stack([np.full_like(x, c0), -c1 * y, c1 * z, -c1 * x, 1.0925484305920792 * x * y,
                  -1.0925484305920792 * y * z, 0.31539156525252005 * (2 * z * z - x * x - y * y),
                  -1.0925484305920792 * x * z, 0.5462742152960396 * (x * x - y * y)], 1)
    G = (Y.T @ Y) * (4 * math.pi / len(x))
    np.testing.assert_allclose(G, np.eye(9), atol=0.03)
    # the oracle's direction for a 1x1 image is the optical axis, pointing at the camera
    from diffsoup_amd import scenes
    mvp = scenes.orbit_cameras(1, 3.0, [(1, 1)], elevation_deg=30.0, azimuth0=0.7)[0].astype(np.float64)
    inv = np.linalg.inv(mvp).astype(np.float32)[None]
    rast = np.array([[[[0.2, 0.2, 0.0, 1.0]]]], np.float32)
    enc = O.sh2(rast, inv)[0, 0, 0]
    eye = 3.0 * np.array([math.cos(math.radians(30)) * math.sin(0.7), math.sin(math.radians(30)),
                          math.cos(math.radians(30)) * math.cos(0.7)])
    X, Yv, Z = eye / np.linalg.norm(eye)
    np.testing.assert_allclose(enc[[1, 2, 3]], [-c1 * Yv, c1 * Z, -c1 * X], atol=1e-5)
    assert O.sh2(np.zeros((1, 1, 1, 4), np.float32), inv).sum() == 0


# --------------------------------------------------------------------------
# Opacity auxiliary gradient: unbiasedness by enumerating the masks
# --------------------------------------------------------------------------

def _depth_test_with(fp, fa, alpha, keep_by_tri, B, H, W):
    """Depth test with thresholds forcing each triangle's fragments kept
    (threshold 0) or dropped (threshold 2 > any opacity)."""
    thr = np.where(keep_by_tri[fa[:, 3].astype(int) - 1], 0.0, 2.0).astype(np.float32)
    rast = np.zeros((B, H, W, 4), np.float32)
    O.lib().dso_depth_test(B, H, W, fp.shape[0], np.ascontiguousarray(fp, np.int32),
                           np.ascontiguousarray(fa, np.float32), np.ascontiguousarray(alpha, np.float32), thr, rast)
    return rast


def test_opacity_aux_estimator_is_unbiased():
    """Two overlapping triangles with constant opacities a0 (front) and a1
    (back), flat colours and a white background.  Over the four keep/drop
    outcomes, E[L] = sum_o P(o) L(o) with L the per-pixel L1 summed over
    pixels is multilinear in (a0, a1), and the score-function estimator of
    python/diffsoup/rasterize.py:139-196 must average to dE[L]/da exactly.
    A per-triangle constant opacity spreads over its 3 level-0 lattice values
    with weights summing to 1, so the estimator summed over a triangle's
    lattice values is compared."""
    H, W = 12, 16
    pos = np.array([[[-0.8, -0.7, -0.2, 1.0], [0.6, -0.5, -0.2, 1.0], [-0.2, 0.8, -0.2, 1.0],
                     [-0.4, -0.9, 0.3, 1.0], [0.9, 0.2, 0.3, 1.0], [-0.7, 0.6, 0.3, 1.0]]], np.float32)
    tri = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    rng = np.random.default_rng(5)
    target = rng.uniform(0, 1, size=(1, H, W, 3)).astype(np.float32)
    flat = np.array([[0.9, 0.1, 0.2], [0.1, 0.7, 0.3]], np.float32)
    fp, fa = O.fragments(H, W, pos, tri)
    assert fp.shape[0] > 50
    for a0, a1 in ((0.3, 0.6), (0.8, 0.25)):
        src = np.array([[a0] * 3, [a1] * 3], np.float32)
        alpha = O.multires_alpha(fa, 0, src)

        def outcome(k0, k1):
            rast = _depth_test_with(fp, fa, alpha, np.array([k0, k1], bool), 1, H, W)
            ids = rast[..., 3].astype(int)
            color = np.where(ids[..., None] > 0, flat[np.maximum(ids - 1, 0)], 1.0).astype(np.float32)
            L = float(np.abs(color.astype(np.float64) - target).sum())
            est = O.opacity_aux_grad(color, target, rast, pos, tri, 0, src).astype(np.float64).sum(axis=1)
            return L, est

        outs = {(k0, k1): outcome(k0, k1) for k0 in (0, 1) for k1 in (0, 1)}

        def expected_loss(b0, b1):
            p = lambda k, a: a if k else 1.0 - a
            return sum(p(k0, b0) * p(k1, b1) * outs[(k0, k1)][0] for (k0, k1) in outs)

        mean_est = sum((a0 if k0 else 1 - a0) * (a1 if k1 else 1 - a1) * outs[(k0, k1)][1] for (k0, k1) in outs)
        h = 1e-3   # E[L] is linear in each opacity, so central differences are exact up to rounding
        d0 = (expected_loss(a0 + h, a1) - expected_loss(a0 - h, a1)) / (2 * h)
        d1 = (expected_loss(a0, a1 + h) - expected_loss(a0, a1 - h)) / (2 * h)
        np.testing.assert_allclose(mean_est, [d0, d1], rtol=1e-4)
        assert abs(d0) > 1.0 and abs(d1) > 1.0


# --------------------------------------------------------------------------
# Edge gradients: the estimator against finite differences of the image
# --------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["silhouette", "occlusion", "intersection", "interior"])
def test_edge_grad_against_finite_differences(name):
    """The oracle's edge-gradient estimator against float64 central
    differences of L = sum g . colour, rendered by an independent
    point-sampling renderer (tests/fd_render.py): across silhouettes,
    occlusion edges and intersection lines it matches dL/dpos to the pixel
    discretisation, and inside a triangle it is exactly twice dL/dpos (each
    same-triangle pixel pair is visited from both pixels).  A wrong sign,
    axis, image scale (W/2 vs H/2), dropped perspective (w) or depth term or
    a mis-attributed occlusion edge fails."""
    import fd_render as F
    sc = F.scene(name)
    rast = O.rasterize((sc["H"], sc["W"]), sc["pos"][None], sc["tri"], 0,
                       np.ones((sc["tri"].shape[0], 3), np.float32), stochastic=False)
    ids, _ = F.render(sc["pos"], sc["tri"], sc["H"], sc["W"])
    np.testing.assert_array_equal(rast[0, ..., 3].astype(int) - 1, ids)
    color, g = F.image(sc)
    est = O.edge_grad(color, g, rast, sc["pos"][None], sc["tri"])[0].astype(np.float64)
    want = F.expected(sc)
    assert np.abs(want[:, [0, 1, 3]]).max(0).min() > 0.05 * np.abs(want).max()   # x, y and w terms all matter
    if name == "intersection":
        assert np.abs(want[:, 2]).min() > 0.1 * np.abs(want).max()   # depth moves the intersection line
    else:
        assert np.all(want[:, 2] == 0) and np.all(est[:, 2] == 0)    # depth moves no pixel here
    np.testing.assert_allclose(est, want, rtol=0, atol=F.TOLERANCE[name] * np.abs(want).max())


@pytest.mark.parametrize("levels", [(0, 0), (1, 3)])
def test_reference_composition_gradients_match_central_differences(levels):
    """The oracle's reference-composition step (sigmoid of accumulated
    multires features and opacities, single-level gather, SH2 inputs, the
    16 -> 16 -> 16 -> 3 ColorMLP, MSE) against float64 central differences of
    its own loss: feature-source and MLP gradients (the raster does not
    depend on either, so the loss is smooth in them)."""
    import dataclasses

    from diffsoup_amd import scenes
    s = scenes.reference_training_soup(4, T=400, views=1, res=(48, 56), levels=levels, alpha_levels=levels)
    r = O.reference_step(s, 9)
    cov = r["covered"]
    assert cov.sum() > 100
    rng = np.random.default_rng(1)
    tris = np.unique(r["rast"][cov, 3].astype(int) - 1)
    eps = 1e-3

    def loss_with(**kw):
        return O.reference_step(dataclasses.replace(s, **kw), 9, pos=r["pos"], alpha=r["alpha"])["loss"]

    for _ in range(5):
        t, k, c = int(rng.choice(tris)), int(rng.integers(0, s.feat.shape[1])), int(rng.integers(0, 7))
        fp, fm = s.feat.copy(), s.feat.copy()
        fp[t, k, c] += eps
        fm[t, k, c] -= eps
        fd = (loss_with(feat=fp) - loss_with(feat=fm)) / (2 * eps)
        got = r["grad_feat"][t, k, c]
        assert abs(got - fd) <= 5e-3 * abs(fd) + 2e-9, (t, k, c, got, fd)
    for _ in range(4):   # W1 entries, including the SH2 columns 7..15
        i, j = int(rng.integers(0, 16)), int(rng.integers(0, 16))
        mp = [(w.copy(), b.copy()) for w, b in s.mlp]
        mm = [(w.copy(), b.copy()) for w, b in s.mlp]
        mp[0][0][i, j] += eps
        mm[0][0][i, j] -= eps
        fd = (loss_with(mlp=mp) - loss_with(mlp=mm)) / (2 * eps)
        assert abs(r["grad_mlp"][16 * i + j] - fd) <= 5e-3 * abs(fd) + 2e-9, (i, j, r["grad_mlp"][16 * i + j], fd)
