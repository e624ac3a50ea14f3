"""A float64 point-sampling renderer and central finite differences of a
linear image loss, for checking the edge-gradient estimator
(src/cuda/rasterize.cu:653-903) against the derivative of the image it
differentiates.

The renderer is independent of the oracle: for every pixel centre it solves
the clip-space barycentrics of each triangle from homogeneous 2-D edge
functions, keeps the nearest triangle whose depth z/w lies in (-1, 1], and
shades it with a flat colour per triangle or with per-vertex colours
interpolated by those barycentrics.  The loss is L = sum_p g_p . c_p for a
fixed weight image g, so dL/dcolour = g is the estimator's colour gradient.
"""
from __future__ import annotations

import numpy as np

BACKGROUND = np.array([0.1, 0.5, 0.3])


def render(pos, tri, H, W):
    """pos [V, 4] clip space -> (triangle id [H, W], -1 for background;
    barycentrics [H, W, 3])."""
    pos = np.asarray(pos, np.float64)
    ys, xs = np.mgrid[0:H, 0:W]
    nx = -1 + 2 * (xs + 0.5) / W
    ny = -1 + 2 * (ys + 0.5) / H
    ids = np.full((H, W), -1)
    bary = np.zeros((H, W, 3))
    nearest = np.full((H, W), np.inf)
    for t in range(tri.shape[0]):
        p = pos[tri[t]]
        qx = p[:, 0, None, None] - nx[None] * p[:, 3, None, None]
        qy = p[:, 1, None, None] - ny[None] * p[:, 3, None, None]
        A = np.stack([qx[1] * qy[2] - qy[1] * qx[2], qx[2] * qy[0] - qy[2] * qx[0], qx[0] * qy[1] - qy[0] * qx[1]])
        b = A / A.sum(0)
        d = (b * p[:, 2, None, None]).sum(0) / (b * p[:, 3, None, None]).sum(0)
        hit = (b.min(0) >= 0) & (d > -1) & (d <= 1) & (d < nearest)
        ids[hit] = t
        nearest[hit] = d[hit]
        bary[hit] = np.moveaxis(b, 0, -1)[hit]
    return ids, bary


def shade(ids, bary, tri, flat=None, vertex_colors=None):
    """[H, W, 3] image: flat[t] or sum_k b_k vertex_colors[tri[t, k]] where
    triangle t is seen, BACKGROUND elsewhere."""
    H, W = ids.shape
    c = np.broadcast_to(BACKGROUND, (H, W, 3)).copy()
    seen = ids >= 0
    if vertex_colors is None:
        c[seen] = flat[ids[seen]]
    else:
        c[seen] = (bary[seen][:, :, None] * vertex_colors[tri[ids[seen]]]).sum(1)
    return c


def loss(pos, tri, H, W, g, flat=None, vertex_colors=None):
    ids, bary = render(pos, tri, H, W)
    return float((g * shade(ids, bary, tri, flat, vertex_colors)).sum())


def central_differences(pos, tri, H, W, g, steps, flat=None, vertex_colors=None):
    """dL/dpos [V, 4] by central differences, averaged over the step sizes
    ``steps`` (clip-space units).  Averaging over steps of several pixels
    smooths the pixel-count noise of a moving silhouette."""
    pos = np.asarray(pos, np.float64)
    out = np.zeros_like(pos)
    for h in steps:
        for v in range(pos.shape[0]):
            for j in range(4):
                q = pos.copy()
                q[v, j] += h
                up = loss(q, tri, H, W, g, flat, vertex_colors)
                q[v, j] -= 2 * h
                down = loss(q, tri, H, W, g, flat, vertex_colors)
                out[v, j] += (up - down) / (2 * h)
    return out / len(steps)


def _clip(ndc, w):
    ndc = np.asarray(ndc, np.float64)
    w = np.asarray(w, np.float64)
    return np.concatenate([ndc * w[:, None], w[:, None]], 1).astype(np.float32)


def _weights(H, W):
    ys, xs = np.mgrid[0:H, 0:W]
    return np.stack([1 + 0.3 * np.sin(xs / 17.0), -0.5 + 0.2 * np.cos(ys / 11.0), 2.0 + 0 * xs], -1)


def _edge_distance(pos, H, W):
    """Distance in pixels from each pixel centre to the nearest edge line of
    the single triangle pos[0:3]."""
    ndc = pos[:3, :2].astype(np.float64) / pos[:3, 3:4]
    px = np.stack([(ndc[:, 0] + 1) * W / 2 - 0.5, (ndc[:, 1] + 1) * H / 2 - 0.5], 1)
    ys, xs = np.mgrid[0:H, 0:W]
    d = np.full((H, W), np.inf)
    for a, b in ((0, 1), (1, 2), (2, 0)):
        e = px[b] - px[a]
        n = np.array([-e[1], e[0]]) / np.hypot(*e)
        d = np.minimum(d, np.abs((xs - px[a, 0]) * n[0] + (ys - px[a, 1]) * n[1]))
    return d


def scene(name, H=96, W=128):
    """Scenes with every vertex at a different w, so the estimator's w terms
    matter, and H != W, so its x and y scales differ.

    silhouette: one flat triangle over the background.
    occlusion:  a flat triangle partly in front of another.
    intersection: two flat triangles that cross in depth, so part of the
                visible boundary is their intersection line, which moves with
                every coordinate of their vertices, depth included.
    interior:   one triangle with per-vertex colours and a weight image that
                vanishes within 3 pixels of the triangle's edges, so that only
                same-triangle pixel pairs carry weight and L is smooth in the
                vertex positions for steps well under 3 pixels.
    Returns a dict with pos [V, 4] float32, tri, g, flat / vertex_colors and
    the central-difference steps."""
    tri1 = np.array([[0, 1, 2]], np.int32)
    ndc1 = [[-0.7, -0.6, 0.1], [0.75, -0.4, 0.2], [-0.1, 0.8, 0.3]]
    w1 = [0.9, 1.1, 1.3]
    pixel_steps = [k * 2.0 / W for k in (2, 3, 4, 5, 6, 7, 8)]
    g = _weights(H, W)
    if name == "silhouette":
        return dict(H=H, W=W, pos=_clip(ndc1, w1), tri=tri1, g=g, flat=np.array([[0.9, 0.2, 0.4]]),
                    vertex_colors=None, steps=pixel_steps)
    if name == "occlusion":
        ndc = ndc1 + [[-0.3, -0.9, 0.4], [0.9, 0.3, 0.5], [-0.8, 0.5, 0.3]]
        ndc = [[x, y, z - 0.4] if i < 3 else [x, y, z] for i, (x, y, z) in enumerate(ndc)]
        return dict(H=H, W=W, pos=_clip(ndc, w1 + [1.0, 1.2, 0.8]), tri=np.array([[0, 1, 2], [3, 4, 5]], np.int32),
                    g=g, flat=np.array([[0.9, 0.2, 0.4], [0.3, 0.8, 0.6]]), vertex_colors=None, steps=pixel_steps)
    if name == "intersection":
        ndc = [[-0.8, -0.7, -0.5], [0.8, -0.6, 0.6], [0.0, 0.85, 0.05],
               [-0.6, 0.7, 0.1], [0.7, 0.6, 0.1], [0.1, -0.85, 0.1]]
        return dict(H=H, W=W, pos=_clip(ndc, w1 + [1.0, 1.2, 0.8]), tri=np.array([[0, 1, 2], [3, 4, 5]], np.int32),
                    g=g, flat=np.array([[0.9, 0.2, 0.4], [0.3, 0.8, 0.6]]), vertex_colors=None, steps=pixel_steps)
    if name == "interior":
        pos = _clip(ndc1, w1)
        ids, _ = render(pos, tri1, H, W)
        ramp = np.clip((_edge_distance(pos, H, W) - 3.0) / 8.0, 0, 1) * (ids >= 0)
        return dict(H=H, W=W, pos=pos, tri=tri1, g=g * ramp[..., None], flat=None,
                    vertex_colors=np.array([[0.9, 0.1, 0.3], [0.2, 0.8, 0.5], [0.4, 0.3, 0.95]]), steps=[1e-3])
    raise KeyError(name)


# Relative to the largest |dL/dpos| of the scene.  Silhouette steps count
# pixel centres, so the agreement there is limited by pixel discretisation
# (measured 1%, 2.7% and 1.3% at 96 x 128, halving at twice the
# resolution); the interior is smooth (measured 2e-4).
TOLERANCE = {"silhouette": 0.03, "occlusion": 0.045, "intersection": 0.035, "interior": 2e-3}

# The estimator visits each same-triangle pixel pair from both of its pixels
# (src/cuda/rasterize.cu:694-745 runs for every covered pixel and each of its
# four neighbours), so inside a triangle it returns twice the derivative of
# the image; a silhouette pair is visited only from its covered side.
INTERIOR_FACTOR = 2.0


def expected(sc):
    """The estimator's expected value for a scene: the finite differences,
    times INTERIOR_FACTOR for the interior scene."""
    fd = central_differences(sc["pos"], sc["tri"], sc["H"], sc["W"], sc["g"], sc["steps"], sc["flat"],
                             sc["vertex_colors"])
    return fd * (INTERIOR_FACTOR if sc["vertex_colors"] is not None else 1.0)


def image(sc):
    """(colour [1, H, W, 3] float32, weight [1, H, W, 3] float32) for the
    estimator, rendered at the scene's positions."""
    ids, bary = render(sc["pos"], sc["tri"], sc["H"], sc["W"])
    c = shade(ids, bary, sc["tri"], sc["flat"], sc["vertex_colors"])
    return c[None].astype(np.float32), sc["g"][None].astype(np.float32)
