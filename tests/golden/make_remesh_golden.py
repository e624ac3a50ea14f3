"""Writes tests/golden/soup_split.npz from the reference's own soup splitters
(src/remesh.cpp, src/remesh_clip.cpp), compiled from the reference's sources
into oracle/_ref by ``make -C oracle ref``.

    python tests/golden/make_remesh_golden.py        (from the repository root)

Cases (inputs and expected outputs only):

* ``w_cap``: a 200-triangle random soup, 300 world-space bisections.
* ``w_tau``: the same soup split until every edge is below 0.06.
* ``w_ties``: 40 copies of one isosceles triangle plus 40 copies of its
  mirror image, 150 bisections: every queued length ties with dozens of
  others, so the order of the outputs is the heap's tie order.
* ``c_cap`` / ``c_tau``: the first 400 triangles of the 1,000-triangle
  random soup seen through its 256x256 camera, a random two thirds of the
  faces eligible, 500 screen-space bisections / split until below 0.03 image
  heights.
* ``c_wide``: the same soup through a 2:1 viewport (aspect 2), until 0.04.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "soup_split.npz")


def random_soup(n, seed, size=0.1):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.0, 1.0, (n, 1, 3))
    v = (c + rng.uniform(-size, size, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    return v, np.arange(3 * n, dtype=np.int32).reshape(n, 3)


def tie_soup():
    a = np.array([[0, 0, 0], [1, 0, 0], [0.5, 0.75, 0]], np.float32)
    b = a * np.array([-1, 1, 1], np.float32) + np.array([3, 0, 0], np.float32)
    v = np.concatenate([np.tile(a, (40, 1)), np.tile(b, (40, 1))]).astype(np.float32)
    return v, np.arange(v.shape[0], dtype=np.int32).reshape(-1, 3)


def main():
    from diffsoup_amd import scenes
    from oracle import oracle as O

    O.build()
    if not O.ref_remesh_available():
        raise SystemExit("the reference's sources are not present: oracle/_ref cannot be built")
    out = {}

    def world(name, v, f, ns, tau):
        r = O.ref_split_soup(v, f, ns, tau)
        out.update({f"{name}_in_verts": v, f"{name}_in_faces": f,
                    f"{name}_args": np.array([ns, tau], np.float64)})
        for k, a in zip(("verts", "faces", "face_map", "flags"), r):
            out[f"{name}_{k}"] = a

    def clip(name, m, v, f, valid, ns, tau_ratio, aspect):
        r = O.ref_split_soup_clip(m, v, f, valid, ns, tau_ratio, aspect)
        out.update({f"{name}_in_verts": v, f"{name}_in_faces": f, f"{name}_mvp": m, f"{name}_valid": valid,
                    f"{name}_args": np.array([ns, tau_ratio, aspect], np.float64)})
        for k, a in zip(("verts", "faces", "face_map", "flags"), r):
            out[f"{name}_{k}"] = a

    v, f = random_soup(200, 7)
    world("w_cap", v, f, 300, 0.0)
    world("w_tau", v, f, -1, 0.06)
    tv, tf = tie_soup()
    world("w_ties", tv, tf, 150, 0.0)

    s = scenes.small_random_soup(0)
    m = s.mvp[0].astype(np.float32)
    cv, cf = s.verts[:1200].copy(), s.tri[:400].copy()
    valid = (np.random.default_rng(11).uniform(size=400) < 2 / 3).astype(np.int32)
    clip("c_cap", m, cv, cf, valid, 500, 0.0, 1.0)
    clip("c_tau", m, cv, cf, valid, -1, 0.03, 1.0)
    clip("c_wide", m, cv, cf, valid, -1, 0.04, 2.0)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: " + ", ".join(f"{k} {out[k + '_faces'].shape[0]} faces"
                                       for k in ("w_cap", "w_tau", "w_ties", "c_cap", "c_tau", "c_wide")))


if __name__ == "__main__":
    main()
