"""Synthetic triangle soups and cameras for the benchmark and test workloads.

Everything is generated on the CPU with NumPy's PCG64 from an integer seed,
so a workload is fully determined by (name, seed).  Cameras are pinhole
look-at cameras on a circular orbit around the origin with an OpenGL-style
projection; ``mvp`` is row-major and clip = mvp @ [x, y, z, 1], exactly what
the reference's project_vertices consumes (examples/utils.py:412-426).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def perspective(fovy_deg: float, aspect: float, near: float, far: float) -> np.ndarray:
    f = 1.0 / math.tan(math.radians(fovy_deg) / 2.0)
    m = np.zeros((4, 4), np.float64)
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = (far + near) / (near - far)
    m[2, 3] = 2.0 * far * near / (near - far)
    m[3, 2] = -1.0
    return m


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> np.ndarray:
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def orbit_cameras(n: int, radius: float, resolutions, fovy=50.0, near=0.1, far=10.0,
                  elevation_deg=20.0, azimuth0=0.0):
    """n views at azimuths azimuth0 + 2 pi k / n.  Returns mvp [n,4,4] f32."""
    mvps = []
    el = math.radians(elevation_deg)
    for k in range(n):
        az = azimuth0 + 2.0 * math.pi * k / n
        eye = radius * np.array([math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)])
        H, W = resolutions[k]
        mvps.append(perspective(fovy, W / H, near, far) @ look_at(eye))
    return np.stack(mvps).astype(np.float32)


def _tangent_frame(n):
    n = n / np.linalg.norm(n, axis=-1, keepdims=True)
    a = np.where(np.abs(n[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    t1 = np.cross(n, a)
    t1 /= np.linalg.norm(t1, axis=-1, keepdims=True)
    t2 = np.cross(n, t1)
    return t1, t2


def _planar_triangles(rng, centres, normals, edge, jitter):
    """Equilateral triangles of the given edge (a scalar or one per centre,
    shape [N,1]) in the tangent planes, with a random in-plane rotation and
    per-vertex uniform jitter."""
    t1, t2 = _tangent_frame(normals)
    th = rng.uniform(0, 2 * math.pi, size=(len(centres), 1))
    rad = edge / math.sqrt(3.0)
    verts = []
    for k in range(3):
        ang = th + 2.0 * math.pi * k / 3.0
        v = centres + rad * (np.cos(ang) * t1 + np.sin(ang) * t2)
        v = v + rng.uniform(-jitter, jitter, size=v.shape)
        verts.append(v)
    return np.stack(verts, axis=1)  # [N,3,3]


@dataclass
class Soup:
    verts: np.ndarray            # [V,3] f32 (soup: 3 unique vertices per triangle)
    tri: np.ndarray              # [T,3] i32
    alpha_logits: np.ndarray     # [T, S_alpha] f32
    alpha_level: int
    feat: np.ndarray             # [T, S_tot, C] f32
    levels: tuple                # (lo, hi)
    mvp: np.ndarray              # [B,4,4] f32
    resolutions: list            # [(H, W)] per view
    target: np.ndarray           # [P,3] f32 packed over views
    mlp_dims: tuple              # (Din, Dh, Dout)
    mlp: list = field(default_factory=list)   # [(W [out,in], b [out])] f32
    feature_mode: str = "concat"
    name: str = ""
    # the reference's own composition (examples/02_synthetic.py:354-364):
    # feat / alpha_logits are multires sources over levels / (alpha_lo,
    # alpha_level), accumulated to the top level and passed through a sigmoid
    accumulate: bool = False
    alpha_lo: int = 0
    sh2: bool = False                          # SH2 of the view direction after the features
    inv_mvp: np.ndarray | None = None          # [B,4,4] f32 (sh2)

    @property
    def T(self):
        return self.tri.shape[0]

    @property
    def B(self):
        return self.mvp.shape[0]

    @property
    def pixels(self):
        return sum(h * w for h, w in self.resolutions)


def _soup_from_tris(tris):
    T = tris.shape[0]
    verts = tris.reshape(-1, 3).astype(np.float32)
    tri = np.arange(3 * T, dtype=np.int32).reshape(T, 3)
    return verts, tri


def _xavier(rng, fan_out, fan_in):
    a = math.sqrt(6.0 / (fan_in + fan_out))
    return rng.uniform(-a, a, size=(fan_out, fan_in)).astype(np.float32)


def _kaiming(rng, fan_out, fan_in):
    a = math.sqrt(6.0 / fan_in)   # kaiming_uniform_, relu gain sqrt(2): bound = sqrt(2)*sqrt(3/fan_in)
    return rng.uniform(-a, a, size=(fan_out, fan_in)).astype(np.float32)


def _mlp(rng, din, dh, dout, init="xavier"):
    w1 = _xavier(rng, dh, din) if init == "xavier" else _kaiming(rng, dh, din)
    w2 = _xavier(rng, dh, dh) if init == "xavier" else _kaiming(rng, dh, dh)
    w3 = _xavier(rng, dout, dh)
    z = lambda n: np.zeros(n, np.float32)
    return [(w1, z(dh)), (w2, z(dh)), (w3, z(dout))]


def feats_at_level(level):
    return 3 if level == 0 else ((1 << (level - 1)) + 1) * ((1 << level) + 1)


def small_random_soup(seed=0, T=1000, res=(256, 256)) -> Soup:
    """1,000 triangles, vertices within 0.1 of uniform centres in [-1,1]^3;
    logits N(0,1) at level 0; level-0 features, 8 channels ~N(0,0.1); MLP
    8->16->16->3 Xavier; one 256x256 view on a radius-3 orbit, fovy 50;
    MSE target U[0,1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(-1, 1, size=(T, 1, 3))
    tris = centres + rng.uniform(-0.1, 0.1, size=(T, 3, 3))
    verts, tri = _soup_from_tris(tris)
    logits = rng.normal(0, 1, size=(T, feats_at_level(0))).astype(np.float32)
    feat = rng.normal(0, 0.1, size=(T, feats_at_level(0), 8)).astype(np.float32)
    mvp = orbit_cameras(1, 3.0, [res], fovy=50.0, near=0.1, far=10.0, elevation_deg=0.0)
    target = rng.uniform(0, 1, size=(res[0] * res[1], 3)).astype(np.float32)
    return Soup(verts, tri, logits, 0, feat, (0, 0), mvp, [res], target, (8, 16, 3),
                _mlp(rng, 8, 16, 3), "concat", "random_soup_1k")


def sphere_torus_soup(seed=0, n_surface=12000, n_float=3000, views=8, res=(800, 800),
                      levels=(0, 3), C=8, alpha_level=3, radius=4.0, dh=32) -> Soup:
    """12K triangles on a unit sphere + torus (R=0.7, r=0.25; edge 0.02,
    per-vertex jitter +-0.005) and 3K floaters (edge 0.03) in [-1.2,1.2]^3;
    logits N(0,1) at alpha_level; levels 0..3 (1,2,4,8 subdivisions), C=8
    ~N(0,0.1) concatenated per level into the MLP (32->32->32->3); views on a
    radius-4 orbit at fovy 50; MSE target U[0,1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a_s, a_t = 4 * math.pi, 4 * math.pi**2 * 0.7 * 0.25
    n_s = int(round(n_surface * a_s / (a_s + a_t)))
    n_t = n_surface - n_s
    # sphere
    p = rng.normal(size=(n_s, 3))
    p /= np.linalg.norm(p, axis=-1, keepdims=True)
    tri_s = _planar_triangles(rng, p, p, 0.02, 0.005)
    # torus around the y axis, area-uniform by rejection on the tube angle
    u = rng.uniform(0, 2 * math.pi, size=4 * n_t)
    v = rng.uniform(0, 2 * math.pi, size=4 * n_t)
    keep = rng.uniform(0, 1, size=4 * n_t) < (0.7 + 0.25 * np.cos(v)) / 0.95
    u, v = u[keep][:n_t], v[keep][:n_t]
    R, r = 0.7, 0.25
    c = np.stack([(R + r * np.cos(v)) * np.cos(u), r * np.sin(v), (R + r * np.cos(v)) * np.sin(u)], -1)
    nrm = np.stack([np.cos(v) * np.cos(u), np.sin(v), np.cos(v) * np.sin(u)], -1)
    tri_t = _planar_triangles(rng, c, nrm, 0.02, 0.005)
    # floaters
    cf = rng.uniform(-1.2, 1.2, size=(n_float, 3))
    nf = rng.normal(size=(n_float, 3))
    tri_f = _planar_triangles(rng, cf, nf, 0.03, 0.0)
    tris = np.concatenate([tri_s, tri_t, tri_f], 0)
    verts, tri = _soup_from_tris(tris)
    T = tri.shape[0]
    lo, hi = levels
    S_tot = sum(feats_at_level(l) for l in range(lo, hi + 1))
    logits = rng.normal(0, 1, size=(T, feats_at_level(alpha_level))).astype(np.float32)
    feat = rng.normal(0, 0.1, size=(T, S_tot, C)).astype(np.float32)
    resolutions = [tuple(res)] * views
    mvp = orbit_cameras(views, radius, resolutions, fovy=50.0, near=0.1, far=10.0)
    P = sum(h * w for h, w in resolutions)
    target = rng.uniform(0, 1, size=(P, 3)).astype(np.float32)
    din = (hi - lo + 1) * C
    return Soup(verts, tri, logits, alpha_level, feat, levels, mvp, resolutions, target, (din, dh, 3),
                _mlp(rng, din, dh, 3), "concat", f"sphere_torus_{T // 1000}k")


MIXED_RESOLUTIONS = [(840, 1297), (546, 979), (822, 1237), (800, 800)]   # (H, W)


def _features_and_mlp(rng, T, levels, C, alpha_level, dh, feat_std=0.1):
    lo, hi = levels
    S_tot = sum(feats_at_level(l) for l in range(lo, hi + 1))
    logits = rng.normal(0, 1, size=(T, feats_at_level(alpha_level))).astype(np.float32)
    feat = rng.normal(0, feat_std, size=(T, S_tot, C)).astype(np.float32)
    din = (hi - lo + 1) * C
    return logits, feat, din, _mlp(rng, din, dh, 3)


def unbounded_soup(seed=0, T=50_000, views=32, resolutions=None, levels=(0, 3), C=8, alpha_level=3, dh=32,
                   layered=0.0, res_scale=1.0, radius=4.0, edge_scale=1.0, name=None) -> Soup:
    """An unbounded scene: 70% of the triangles within radius 2 of the origin
    (edge 0.01), 30% at radii log-uniform in [2, 100] with edge proportional
    to the radius (0.005 r); near/far 0.1/100.  ``layered`` > 0 turns that
    fraction of the triangles into copies of others pushed outward by 1-3%
    of their radius, stacking depth layers for depth-test contention.  Views
    on a radius-4 orbit at fovy 50, each with a resolution drawn by the seed
    from ``resolutions`` (default: 1297x840, 979x546, 1237x822, 800x800),
    scaled by ``res_scale`` (scaled-down tests pair it with ``edge_scale`` to
    keep the pixels per triangle); features N(0, 0.1) per level and channel,
    MLP (levels x C) -> dh -> dh -> 3 (Xavier); MSE target U[0,1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n_lay = int(round(T * layered))
    n_base = T - n_lay
    n_in = int(round(0.7 * n_base))
    n_out = n_base - n_in
    c_in = rng.normal(size=(n_in, 3))
    c_in *= (2.0 * rng.uniform(0, 1, size=(n_in, 1)) ** (1 / 3)) / np.linalg.norm(c_in, axis=-1, keepdims=True)
    tri_in = _planar_triangles(rng, c_in, rng.normal(size=(n_in, 3)), 0.01 * edge_scale, 0.0)
    d = rng.normal(size=(n_out, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rad = np.exp(rng.uniform(math.log(2.0), math.log(100.0), size=(n_out, 1)))
    c_out = d * rad
    tri_out = _planar_triangles(rng, c_out, rng.normal(size=(n_out, 3)), 0.005 * edge_scale * rad, 0.0)
    tris = np.concatenate([tri_in, tri_out], 0)
    if n_lay:
        src = rng.integers(0, n_base, size=n_lay)
        cen = tris[src].mean(axis=1, keepdims=True)
        push = 1.0 + rng.uniform(0.01, 0.03, size=(n_lay, 1, 1))
        tris = np.concatenate([tris, tris[src] + cen * (push - 1.0)], 0)
    verts, tri = _soup_from_tris(tris)
    T = tri.shape[0]
    pool = resolutions or MIXED_RESOLUTIONS
    pick = rng.integers(0, len(pool), size=views)
    res = [(max(8, int(round(pool[k][0] * res_scale))), max(8, int(round(pool[k][1] * res_scale)))) for k in pick]
    logits, feat, din, mlp = _features_and_mlp(rng, T, levels, C, alpha_level, dh)
    mvp = orbit_cameras(views, radius, res, fovy=50.0, near=0.1, far=100.0)
    P = sum(h * w for h, w in res)
    target = rng.uniform(0, 1, size=(P, 3)).astype(np.float32)
    return Soup(verts, tri, logits, alpha_level, feat, levels, mvp, res, target, (din, dh, 3), mlp, "concat",
                name or f"unbounded_{T // 1000}k")


def batched_uneven_soup(seed=0, **kw) -> Soup:
    """50K-triangle unbounded scene, 32 views of mixed resolution, features
    and MLP as the 15K-triangle soup (levels 0-3 x 8 channels, 32-32-32-3)."""
    return unbounded_soup(seed, **{"T": 50_000, "views": 32, **kw})


def bf16_large_soup(seed=0, **kw) -> Soup:
    """100K-triangle unbounded scene, 16 views at 1920x1080, levels 0-4 with
    16 channels (stored in bf16 by the step), MLP 80->64->64->3."""
    args = {"T": 100_000, "views": 16, "resolutions": [(1080, 1920)], "levels": (0, 4), "C": 16,
            "alpha_level": 4, "dh": 64}
    args.update(kw)
    return unbounded_soup(seed, **args)


def stress_soup(seed=0, **kw) -> Soup:
    """300K triangles from the unbounded generator, 40% of them in stacked
    depth layers, 64 views of mixed resolution up to 1920x1080, features and
    MLP as the 100K-triangle bf16 scene."""
    args = {"T": 300_000, "views": 64, "resolutions": MIXED_RESOLUTIONS + [(1080, 1920)], "levels": (0, 4),
            "C": 16, "alpha_level": 4, "dh": 64, "layered": 0.4}
    args.update(kw)
    return unbounded_soup(seed, **args)


def reference_training_soup(seed=0, T=3000, views=2, res=(192, 224), levels=(0, 0), alpha_levels=(0, 0),
                            C=7, dh=16, src_std=0.5, elevation_deg=20.0) -> Soup:
    """The reference's training configuration (examples/02_synthetic.py:
    288-306): feat_dim 7 features and opacities as multires sources over
    ``levels`` / ``alpha_levels`` (the driver starts at levels (0, 0) and
    lifts to (2, 5)), accumulated to the top level and passed through a
    sigmoid, the 9 SH2 values of the view direction after the features, and
    ColorMLP(16, 16, n_layers=2) (kaiming hidden layers, xavier last, zero
    biases).  The sources are drawn N(0, src_std) instead of the driver's
    zeros so that every term is exercised; triangles as in small_random_soup
    (vertices within 0.1 of uniform centres in [-1,1]^3), views on a
    radius-3 orbit at fovy 50, L1 + SSIM target U[0,1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(-1, 1, size=(T, 1, 3))
    tris = centres + rng.uniform(-0.1, 0.1, size=(T, 3, 3))
    verts, tri = _soup_from_tris(tris)
    lo, hi = levels
    alo, ahi = alpha_levels
    S = sum(feats_at_level(l) for l in range(lo, hi + 1))
    Sa = sum(feats_at_level(l) for l in range(alo, ahi + 1))
    logits = rng.normal(0, 1, size=(T, Sa)).astype(np.float32)
    feat = rng.normal(0, src_std, size=(T, S, C)).astype(np.float32)
    mvp = orbit_cameras(views, 3.0, [res] * views, fovy=50.0, near=0.1, far=10.0, elevation_deg=elevation_deg)
    inv_mvp = np.linalg.inv(mvp.astype(np.float64)).astype(np.float32)
    target = rng.uniform(0, 1, size=(views * res[0] * res[1], 3)).astype(np.float32)
    return Soup(verts, tri, logits, ahi, feat, (lo, hi), mvp, [res] * views, target, (C + 9, dh, 3),
                _mlp(rng, C + 9, dh, 3, init="kaiming"), "accumulate", f"reference_training_{T}",
                accumulate=True, alpha_lo=alo, sh2=True, inv_mvp=inv_mvp)


def reference_sphere_torus_soup(seed=0, views=4, res=(800, 800), levels=(0, 0), src_std=0.5) -> Soup:
    """The reference's own training iteration at the size of its synthetic
    driver (examples/02_synthetic.py: 15K primitives, batch of 4 views at
    800x800): the sphere + torus + floaters geometry of sphere_torus_soup,
    with 7-channel feature and opacity sources over ``levels`` (accumulated
    and passed through a sigmoid), SH2 inputs and ColorMLP(16, 16,
    n_layers=2)."""
    import dataclasses

    g = sphere_torus_soup(seed, views=views, res=res)
    r = reference_training_soup(seed + 1, T=g.T, views=views, res=res, levels=levels, alpha_levels=levels,
                                src_std=src_std)
    inv_mvp = np.linalg.inv(g.mvp.astype(np.float64)).astype(np.float32)
    return dataclasses.replace(r, verts=g.verts, tri=g.tri, mvp=g.mvp, inv_mvp=inv_mvp, target=g.target,
                               name=f"reference_sphere_torus_{g.T}_levels{levels[0]}-{levels[1]}")


def project_vertices_np(verts, mvp):
    """clip[b, v] = mvp[b] @ [verts[v], 1] in float32 (examples/utils.py:412-426)."""
    vh = np.concatenate([verts, np.ones((verts.shape[0], 1), np.float32)], -1)
    return np.einsum("bij,nj->bni", mvp.astype(np.float64), vh.astype(np.float64)).astype(np.float32)
