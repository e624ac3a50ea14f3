"""CPU oracle for DiffSoup's rasterizer hot path (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker or the timed CPU
baseline, never as the thing measured or shipped.  The product path
(``diffsoup_amd``) must fail loudly when its HIP library is missing rather
than fall back to anything here.

The numeric kernels live in ``ds_oracle.c`` (plain C, strict IEEE fp32, no
contraction), wrapped here with ctypes over NumPy arrays.  The compositions
below restate the reference's Python layer:

* ``rasterize`` - python/diffsoup/multires.py:52-102 with
  python/diffsoup/rasterize.py:40-126 (rects -> fragments -> compaction ->
  opacity -> threshold -> depth test), with the counter-based hash threshold
  in place of torch.rand and the compaction order fixed to
  (view, triangle, rectangle position).
* ``opacity_aux_grad`` - python/diffsoup/rasterize.py:148-196.
* ``shade_reference`` - ColorMLP (python/diffsoup/multires.py:109-207) and the
  training-step composition of examples/02_synthetic.py:354-380, evaluated in
  float64 with torch on the CPU.

Pinning: see the oracle tests (tests/test_oracle_*.py) - closed forms,
brute force, the SplitMix64 known-answer vector and FP64 finite differences.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libds_oracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_c = ctypes


REF_ROOT = "/root/reference"
_REF_REMESH = os.path.join(_HERE, "_ref", "libref_remesh.so")
_ref_remesh = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.isdir(os.path.join(REF_ROOT, "src")):
        # the reference's own soup splitters, where its sources are present
        subprocess.run(["make", "-s", "-C", _HERE, "ref", f"REF={REF_ROOT}"], check=True)
    return _LIB_PATH


def ref_remesh_available() -> bool:
    return os.path.exists(_REF_REMESH)


def _ref_remesh_lib():
    global _ref_remesh
    if _ref_remesh is None:
        if not ref_remesh_available():
            raise FileNotFoundError(f"{_REF_REMESH} is not built (oracle.build() where the reference is present)")
        L = ctypes.CDLL(_REF_REMESH)
        ip = ctypes.POINTER(ctypes.c_int)
        tail = [_c.c_int, _c.c_int, _f32p, _i32p, _i32p, _i32p, ip, ip]
        L.ref_split_soup.argtypes = [_f32p, _c.c_int, _i32p, _c.c_int, _c.c_int, _c.c_float] + tail
        L.ref_split_soup_clip.argtypes = ([_f32p, _f32p, _c.c_int, _i32p, _c.c_int, _i32p, _c.c_int, _c.c_float,
                                           _c.c_float] + tail)
        _ref_remesh = L
    return _ref_remesh


def _ref_split(call, nv, nt, budget):
    """Run one reference splitter through the driver, growing the output
    buffers until they hold the result."""
    cap_v, cap_t = nv + 3 * max(budget, 0) + 3, nt + max(budget, 0) + 1
    while True:
        ov, of = np.zeros((cap_v, 3), np.float32), np.zeros((cap_t, 3), np.int32)
        om, ofl = np.zeros(cap_t, np.int32), np.zeros(cap_t, np.int32)
        got_v, got_t = ctypes.c_int(), ctypes.c_int()
        rc = call(cap_v, cap_t, ov, of, om, ofl, ctypes.byref(got_v), ctypes.byref(got_t))
        if rc == 0:
            n, m = got_v.value, got_t.value
            return ov[:n].copy(), of[:m].copy(), om[:m].copy(), ofl[:m].copy()
        cap_v, cap_t = got_v.value, got_t.value


def ref_split_soup(verts, faces, num_splits, tau=0.0):
    """The reference's TriangleSoupSplitter (src/remesh.cpp), compiled from its
    own sources: (verts, faces, face_map, same_as_original) as numpy arrays."""
    L = _ref_remesh_lib()
    v, f = _f32(verts).reshape(-1, 3), _i32(faces).reshape(-1, 3)
    return _ref_split(lambda *o: L.ref_split_soup(v, v.shape[0], f, f.shape[0], int(num_splits), float(tau), *o),
                      v.shape[0], f.shape[0], int(num_splits))


def ref_split_soup_clip(mvp, verts, faces, valid, num_splits, tau_ratio, aspect_wh):
    """The reference's TriangleSoupSplitterClip (src/remesh_clip.cpp)."""
    L = _ref_remesh_lib()
    m, v, f = _f32(mvp).reshape(4, 4), _f32(verts).reshape(-1, 3), _i32(faces).reshape(-1, 3)
    ok = _i32(valid).reshape(-1)
    return _ref_split(lambda *o: L.ref_split_soup_clip(m, v, v.shape[0], f, f.shape[0], ok, int(num_splits),
                                                       float(tau_ratio), float(aspect_wh), *o),
                      v.shape[0], f.shape[0], int(num_splits))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.dso_splitmix64_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.dso_splitmix64_next.restype = ctypes.c_uint64
        L.dso_mask_threshold.argtypes = [_c.c_uint64, _c.c_uint32, _c.c_uint32, _c.c_uint32, _c.c_uint32]
        L.dso_mask_threshold.restype = _c.c_float
        L.dso_triangle_rects.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _f32p, _c.c_int, _i32p, _i32p, _i64p]
        L.dso_triangle_rects.restype = _c.c_int64
        L.dso_fragments.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _f32p, _c.c_int, _i32p, _i32p, _i32p, _f32p]
        L.dso_fragments.restype = _c.c_int64
        L.dso_intersect.argtypes = [_c.c_float, _c.c_float, _f32p, _f32p, _f32p] + [ctypes.POINTER(_c.c_float)] * 4
        L.dso_intersect.restype = _c.c_int
        L.dso_feats_at_level.argtypes = [_c.c_uint32]
        L.dso_feats_at_level.restype = _c.c_uint32
        L.dso_interp.argtypes = [_c.c_float, _c.c_float, _c.c_int, _c.c_int, _c.c_int, _f32p, _f32p]
        L.dso_multires_alpha.argtypes = [_c.c_int64, _f32p, _c.c_int, _c.c_int, _f32p, _f32p]
        L.dso_mask_thresholds.argtypes = [_c.c_int64, _i32p, _f32p, _c.c_int, _c.c_uint64, _f32p]
        L.dso_depth_test.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int64, _i32p, _f32p, _f32p, _f32p, _f32p]
        L.dso_multires_color.argtypes = [_c.c_int64, _f32p, _c.c_int, _c.c_int, _c.c_int, _f32p, _c.c_int, _f32p]
        L.dso_multires_color_bwd.argtypes = [_c.c_int64, _f32p, _c.c_int, _c.c_int, _c.c_int, _f32p, _c.c_int, _f32p]
        L.dso_accumulate_fwd.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _f32p, _f32p]
        L.dso_accumulate_bwd.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _f32p, _f32p]
        L.dso_sh2.argtypes = [_c.c_int, _c.c_int, _c.c_int, _f32p, _f32p, _f32p]
        L.dso_edge_grad.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _f32p, _f32p, _f32p, _c.c_int, _f32p, _f32p, _c.c_int, _i32p]
        L.dso_opacity_aux.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _f32p, _f32p, _f32p, _c.c_int64,
                                      _i32p, _f32p, _f32p, _c.c_int, _c.c_int, _f32p]
        L.dso_set_threads.argtypes = [_c.c_int]
        L.dso_get_threads.restype = _c.c_int
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """OpenMP threads of the C restatement (1, the default, keeps every loop
    in its written order and every result deterministic; more is for the
    timed CPU baseline, where scatters become atomic adds)."""
    lib().dso_set_threads(int(n))


def get_threads() -> int:
    return int(lib().dso_get_threads())


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def feats_at_level(level: int) -> int:
    return int(lib().dso_feats_at_level(level))


def feats_in_levels(lo: int, hi: int) -> int:
    return sum(feats_at_level(l) for l in range(lo, hi + 1))


def splitmix64(seed: int, n: int) -> list[int]:
    st = ctypes.c_uint64(seed)
    return [int(lib().dso_splitmix64_next(ctypes.byref(st))) for _ in range(n)]


def mask_threshold(seed: int, view: int, tri: int, y: int, x: int) -> float:
    return float(lib().dso_mask_threshold(seed, view, tri, y, x))


def intersect(ndc_y, ndc_x, p0, p1, p2):
    out = [ctypes.c_float() for _ in range(4)]
    ok = lib().dso_intersect(ndc_y, ndc_x, _f32(p0), _f32(p1), _f32(p2), *[ctypes.byref(o) for o in out])
    return bool(ok), tuple(o.value for o in out)


def triangle_rects(H, W, pos, tri):
    """Reference rects (src/cuda/rasterize.cu:43-137): [B*T,4] (h0,h_len,w0,w_len), counts."""
    pos = _f32(pos)
    tri = _i32(tri)
    B, V, _ = pos.shape
    T = tri.shape[0]
    rects = np.zeros((B * T, 4), np.int32)
    counts = np.zeros(B * T, np.int64)
    total = lib().dso_triangle_rects(H, W, B, V, pos, T, tri, rects, counts)
    return rects, counts, int(total)


def fragments(H, W, pos, tri):
    """Valid fragments in canonical order: frag_pix [N,3] (view,y,x), frag_attrs [N,4]."""
    pos = _f32(pos)
    tri = _i32(tri)
    B, V, _ = pos.shape
    T = tri.shape[0]
    rects, _, total = triangle_rects(H, W, pos, tri)
    fp = np.zeros((max(total, 1), 3), np.int32)
    fa = np.zeros((max(total, 1), 4), np.float32)
    n = lib().dso_fragments(H, W, B, V, pos, T, tri, rects, fp, fa)
    return fp[:n].copy(), fa[:n].copy()


def multires_alpha(frag_attrs, level, alpha_src):
    alpha_src = _f32(alpha_src).reshape(alpha_src.shape[0], -1)
    N = frag_attrs.shape[0]
    out = np.zeros(max(N, 1), np.float32)
    lib().dso_multires_alpha(N, _f32(frag_attrs), level, alpha_src.shape[1], alpha_src, out)
    return out[:N]


def thresholds(frag_pix, frag_attrs, stochastic, seed, view_base=0):
    """Mask thresholds; the hash sees view index frag_pix[:,0] + view_base."""
    N = frag_pix.shape[0]
    out = np.zeros(max(N, 1), np.float32)
    fp = _i32(frag_pix)
    if view_base:
        fp = fp.copy()
        fp[:, 0] += view_base
    lib().dso_mask_thresholds(N, fp, _f32(frag_attrs), int(bool(stochastic)), seed, out)
    return out[:N]


def rasterize(resolution, pos, tri, level, alpha_src, stochastic=True, seed=0, return_frags=False, view_base=0):
    """rast [B,H,W,4] for one resolution shared by all views; view b of this
    call draws the mask of view view_base + b."""
    H, W = resolution
    pos = _f32(pos)
    B = pos.shape[0]
    fp, fa = fragments(H, W, pos, tri)
    alpha = multires_alpha(fa, level, alpha_src)
    thr = thresholds(fp, fa, stochastic, seed, view_base)
    rast = np.zeros((B, H, W, 4), np.float32)
    lib().dso_depth_test(B, H, W, fp.shape[0], _i32(fp) if fp.size else np.zeros((1, 3), np.int32), _f32(fa) if fa.size else np.zeros((1, 4), np.float32),
                         alpha if alpha.size else np.zeros(1, np.float32), thr if thr.size else np.zeros(1, np.float32), rast)
    if return_frags:
        return rast, (fp, fa, alpha, thr)
    return rast


def multires_color(rast, lo, hi, feat, concat=False):
    rast = _f32(rast)
    feat = _f32(feat)
    C = feat.shape[2]
    P = rast.size // 4
    width = (hi - lo + 1) * C if concat else C
    out = np.zeros((P, width), np.float32)
    lib().dso_multires_color(P, rast, lo, hi, C, feat, int(concat), out)
    return out.reshape(rast.shape[:-1] + (width,))


def multires_color_bwd(rast, lo, hi, feat_shape, grad_out, concat=False):
    rast = _f32(rast)
    g = np.zeros(feat_shape, np.float32)
    P = rast.size // 4
    lib().dso_multires_color_bwd(P, rast, lo, hi, feat_shape[2], g, int(concat), _f32(grad_out))
    return g


def accumulate_fwd(lo, hi, target, feat):
    feat = _f32(feat)
    T, _, C = feat.shape
    out = np.zeros((T, feats_at_level(target), C), np.float32)
    lib().dso_accumulate_fwd(T, lo, hi, target, C, feat, out)
    return out


def accumulate_bwd(lo, hi, target, feat_shape, grad_out):
    g = np.zeros(feat_shape, np.float32)
    T, _, C = feat_shape
    lib().dso_accumulate_bwd(T, lo, hi, target, C, g, _f32(grad_out))
    return g


def sh2(rast, inv_mvp):
    rast = _f32(rast)
    B, H, W, _ = rast.shape
    out = np.zeros((B, H, W, 9), np.float32)
    lib().dso_sh2(B, H, W, rast, _f32(inv_mvp), out)
    return out


def edge_grad(color, grad_color, rast, pos, tri):
    color = _f32(color)
    B, H, W, C = color.shape
    pos = _f32(pos)
    V = pos.shape[1]
    tri = _i32(tri)
    gp = np.zeros_like(pos)
    lib().dso_edge_grad(B, H, W, C, color, _f32(grad_color), _f32(rast), V, pos, gp, tri.shape[0], tri)
    return gp


def opacity_aux_grad(color, target, rast, pos, tri, level, alpha_src):
    """Unscaled grad w.r.t. alpha_src [T,S] (python/diffsoup/rasterize.py:162-181)."""
    color = _f32(color)
    B, H, W, C = color.shape
    alpha_src = _f32(alpha_src).reshape(alpha_src.shape[0], -1)
    fp, fa = fragments(H, W, pos, tri)
    alpha = multires_alpha(fa, level, alpha_src)
    g = np.zeros_like(alpha_src)
    N = fp.shape[0]
    if N:
        lib().dso_opacity_aux(B, H, W, C, color, _f32(target), _f32(rast), N, _i32(fp), _f32(fa), alpha, level,
                              alpha_src.shape[1], g)
    return g


def covered_fragments(H, W, pos, tri) -> int:
    return int(fragments(H, W, pos, tri)[0].shape[0])


# ---------------------------------------------------------------------------
# Shading reference in float64 (torch on the CPU): ColorMLP + composite + MSE.
# ---------------------------------------------------------------------------

def shade_reference(x, covered, weights, target, background=1.0):
    """x [P, Din] features, covered [P] bool, weights = [(W1,b1),(W2,b2),(W3,b3)]
    (torch layout, W [out,in]); returns color [P,3] and loss (MSE mean over
    P*3) in float64, plus the MLP as a torch function for autograd checks.

    ColorMLP.forward (python/diffsoup/multires.py:185-207): covered pixels get
    (1 - x[:,3]) * x[:,:3] + x[:,3] * sigmoid(mlp(x)); the driver composites
    a white background (examples/02_synthetic.py:367-368).
    """
    import torch

    xt = torch.as_tensor(np.asarray(x, np.float64))
    cov = torch.as_tensor(np.asarray(covered, bool))
    h = xt
    Ws = [(torch.as_tensor(np.asarray(w, np.float64)), torch.as_tensor(np.asarray(b, np.float64))) for w, b in weights]
    for i, (w, b) in enumerate(Ws):
        h = h @ w.T + b
        h = torch.relu(h) if i < len(Ws) - 1 else torch.sigmoid(h)
    res = xt[:, 3:4]
    blended = (1.0 - res) * xt[:, :3] + res * h
    color = torch.where(cov[:, None], blended, torch.full_like(blended, background))
    tgt = torch.as_tensor(np.asarray(target, np.float64))
    loss = ((color - tgt) ** 2).mean()
    return color.numpy(), float(loss)


# ---------------------------------------------------------------------------
# Whole training step on the CPU (the timed CPU baseline and the fused-step
# checker): raster -> concatenated gather -> MLP + blend + MSE (torch, CPU) ->
# MLP adjoint -> feature scatter -> edge gradients -> opacity gradient.
# Composition: examples/02_synthetic.py:354-388 with an MSE loss.
# ---------------------------------------------------------------------------

def project(verts, mvp):
    """clip = mvp @ [v, 1] evaluated in float32 in the order m0 x + m1 y + m2 z + m3."""
    v = np.asarray(verts, np.float32)
    m = np.asarray(mvp, np.float32)
    out = np.empty((m.shape[0], v.shape[0], 4), np.float32)
    for i in range(4):
        out[..., i] = (m[:, None, i, 0] * v[None, :, 0] + m[:, None, i, 1] * v[None, :, 1]
                       + m[:, None, i, 2] * v[None, :, 2] + m[:, None, i, 3])
    return out


def reference_step(soup, seed, dtype=None, pos=None, alpha=None, color_in=None, grad_color_in=None, loss="mse"):
    """Full forward + backward of the fused step (views of any resolution).

    Returns a dict with rast, color, grad_color, loss and the gradients of the
    MLP (torch layout), features, world vertices and opacity logits.  With
    color_in / grad_color_in given, the edge and opacity gradients use those
    images (to isolate kernels from upstream rounding).  loss is "mse" or
    "l1_ssim": 0.8 mean|target - color| + 0.2 (1 - SSIM) / 2 with SSIM the
    mean over views of ssim_fp64 per view (examples/02_synthetic.py:376-380)."""
    import torch

    dtype = dtype or torch.float64
    s = soup
    B = s.B
    acc = getattr(s, "accumulate", False)
    sh = getattr(s, "sh2", False)
    res = [tuple(r) for r in s.resolutions]
    off = np.cumsum([0] + [h * w for h, w in res])
    lo, hi = s.levels
    if pos is None:
        pos = project(s.verts, s.mvp)
    if acc:
        # the reference's composition: sigmoid(accumulate_to_level(...)) of the
        # multires sources, rounded once from double like the step
        # (examples/02_synthetic.py:355-356)
        a_src = s.alpha_logits.reshape(s.T, -1, 1)
        a_acc = accumulate_fwd(s.alpha_lo, s.alpha_level, s.alpha_level, a_src).reshape(s.T, -1)
        if alpha is None:
            alpha = (1.0 / (1.0 + np.exp(-a_acc.astype(np.float64)))).astype(np.float32)
        f_acc = accumulate_fwd(lo, hi, hi, s.feat)
        feat_tab = (1.0 / (1.0 + np.exp(-f_acc.astype(np.float64)))).astype(np.float32)
    elif alpha is None:
        alpha = (1.0 / (1.0 + np.exp(-s.alpha_logits.astype(np.float64)))).astype(np.float32)
    # views one at a time (any mix of resolutions), packed like the step's buffers
    rast = np.concatenate([rasterize((H, W), pos[b:b + 1], s.tri, s.alpha_level, alpha, True, seed,
                                     view_base=b).reshape(-1, 4) for b, (H, W) in enumerate(res)])
    P = rast.shape[0]
    covered = rast[:, 3] > 0
    rast_cov = np.ascontiguousarray(rast[covered])   # background gathers are zero and carry no gradient
    if acc:
        X = multires_color(rast_cov, hi, hi, feat_tab, False).reshape(rast_cov.shape[0], -1)
    else:
        X = multires_color(rast_cov, lo, hi, s.feat, True).reshape(rast_cov.shape[0], -1)
    DF = X.shape[1]
    if sh:
        # encode_view_dir_sh2 per view (src/cuda/encoding.cu:89-130), concatenated after the features
        enc = np.concatenate([sh2(rast[off[b]:off[b + 1]].reshape(1, H, W, 4), s.inv_mvp[b:b + 1]).reshape(-1, 9)
                              for b, (H, W) in enumerate(res)])
        X = np.concatenate([X, enc[covered]], axis=1)
    idx = torch.from_numpy(np.nonzero(covered)[0])
    Xt = torch.tensor(X, dtype=dtype, requires_grad=True)
    Ws = [(torch.tensor(w, dtype=dtype, requires_grad=True), torch.tensor(b, dtype=dtype, requires_grad=True))
          for w, b in s.mlp]
    h = Xt
    for i, (w, b) in enumerate(Ws):
        h = h @ w.T + b
        h = torch.relu(h) if i < len(Ws) - 1 else torch.sigmoid(h)
    r3 = Xt[:, 3:4]
    blend = (1 - r3) * Xt[:, :3] + r3 * h
    tgt = torch.from_numpy(np.ascontiguousarray(s.target)).to(dtype)
    color = torch.ones(P, 3, dtype=dtype).index_put((idx,), blend)
    gcol = None
    if loss == "mse":
        loss = ((color - tgt) ** 2).mean()
    else:
        views = [ssim_fp64(tgt[off[b]:off[b + 1]].reshape(1, H, W, 3).permute(0, 3, 1, 2),
                           color[off[b]:off[b + 1]].reshape(1, H, W, 3).permute(0, 3, 1, 2))
                 for b, (H, W) in enumerate(res)]
        loss = 0.8 * (tgt - color).abs().mean() + 0.2 * 0.5 * (1 - torch.stack(views).mean())
        gcol = torch.autograd.grad(loss, color, retain_graph=True)[0]
    loss.backward()
    dXf = np.ascontiguousarray(Xt.grad.numpy()[:, :DF].astype(np.float32))
    if acc:
        g_acc = multires_color_bwd(rast_cov, hi, hi, feat_tab.shape, dXf, False)
        grad_feat = accumulate_bwd(lo, hi, hi, s.feat.shape, g_acc * (feat_tab * (1.0 - feat_tab)))
    else:
        grad_feat = multires_color_bwd(rast_cov, lo, hi, s.feat.shape, dXf, True)
    color32 = color.detach().numpy().astype(np.float32)
    if gcol is not None:
        grad_color32 = gcol.numpy().astype(np.float32)
    else:
        grad_color32 = (2.0 * (color.detach() - tgt) / color.numel()).numpy().astype(np.float32)
    ci = color32 if color_in is None else color_in
    gi = grad_color32 if grad_color_in is None else grad_color_in
    gpos = np.zeros((B, pos.shape[1], 4), np.float32)
    ga = np.zeros((s.T, feats_at_level(s.alpha_level)), np.float32)
    for b, (H, W) in enumerate(res):
        sl = slice(off[b], off[b + 1])
        rb = rast[sl].reshape(1, H, W, 4)
        gpos[b:b + 1] = edge_grad(ci[sl].reshape(1, H, W, 3), gi[sl].reshape(1, H, W, 3), rb, pos[b:b + 1], s.tri)
        ga += opacity_aux_grad(ci[sl].reshape(1, H, W, 3), s.target[sl].reshape(1, H, W, 3), rb, pos[b:b + 1],
                               s.tri, s.alpha_level, alpha)
    gverts = np.einsum("bij,bvi->vj", s.mvp.astype(np.float64), gpos.astype(np.float64))[:, :3]
    galog = ga / (3 * P) * alpha * (1 - alpha)
    if acc:
        galog = accumulate_bwd(s.alpha_lo, s.alpha_level, s.alpha_level, (s.T, s.alpha_logits.shape[1], 1),
                               galog.reshape(s.T, -1, 1)).reshape(s.T, -1)
    gmlp = np.concatenate([Ws[0][0].grad.numpy().ravel(), Ws[1][0].grad.numpy().ravel(),
                           Ws[2][0].grad.numpy().ravel(), Ws[0][1].grad.numpy(), Ws[1][1].grad.numpy(),
                           Ws[2][1].grad.numpy()])
    return dict(rast=rast, color=color.detach().numpy(), grad_color=grad_color32, loss=float(loss.detach()),
                grad_feat=grad_feat, grad_verts=gverts, grad_alpha_logits=galog, grad_mlp=gmlp,
                covered=covered, pos=pos, alpha=alpha)


# ---------------------------------------------------------------------------
# Photometric loss (examples/02_synthetic.py:376-380).  SSIM comes from the
# reference's dependency pytorch_msssim (unpinned in requirements.txt; the
# algorithm below restates the published pytorch_msssim 1.0.0 functions
# _fspecial_gauss_1d, gaussian_filter, _ssim and ssim), evaluated in float64
# with torch on the CPU; gradients come from torch autograd.
# ---------------------------------------------------------------------------

def gaussian_window_1d(size=11, sigma=1.5):
    """_fspecial_gauss_1d: exp(-c^2 / (2 sigma^2)) over c = -size//2 .. size//2,
    normalised, in float32 (as there)."""
    import torch
    coords = torch.arange(size, dtype=torch.float32) - size // 2
    g = torch.exp(-(coords ** 2) / (2 * sigma ** 2))
    return g / g.sum()


def gaussian_filter_valid(x, win):
    """gaussian_filter: the 1-D window along H then W with no padding (each
    output is one full window), skipping a side shorter than the window;
    x (N, C, H, W) float64."""
    import torch.nn.functional as F
    C = x.shape[1]
    w = win.to(x.dtype).reshape(1, 1, 1, -1).repeat(C, 1, 1, 1)
    out = x
    if x.shape[2] >= w.shape[-1]:
        out = F.conv2d(out, w.transpose(2, 3), groups=C)
    if x.shape[3] >= w.shape[-1]:
        out = F.conv2d(out, w, groups=C)
    return out


def ssim_map_fp64(X, Y, data_range=1.0, win=None, K=(0.01, 0.03)):
    """_ssim's per-window map (N, C, H', W') in float64."""
    import torch
    win = gaussian_window_1d() if win is None else win
    X, Y = X.to(torch.float64), Y.to(torch.float64)
    c1, c2 = (K[0] * data_range) ** 2, (K[1] * data_range) ** 2
    mu1, mu2 = gaussian_filter_valid(X, win), gaussian_filter_valid(Y, win)
    s11 = gaussian_filter_valid(X * X, win) - mu1 * mu1
    s22 = gaussian_filter_valid(Y * Y, win) - mu2 * mu2
    s12 = gaussian_filter_valid(X * Y, win) - mu1 * mu2
    cs = (2 * s12 + c2) / (s11 + s22 + c2)
    return ((2 * mu1 * mu2 + c1) / (mu1 * mu1 + mu2 * mu2 + c1)) * cs


def ssim_fp64(X, Y, data_range=1.0, size_average=True, win=None, K=(0.01, 0.03), nonnegative_ssim=False):
    """pytorch_msssim.ssim on (N, C, H, W) images, float64."""
    import torch
    per_channel = torch.flatten(ssim_map_fp64(X, Y, data_range, win, K), 2).mean(-1)
    if nonnegative_ssim:
        per_channel = torch.relu(per_channel)
    return per_channel.mean() if size_average else per_channel.mean(1)


def l1_ssim_loss_fp64(color, target, l1_weight=0.8, ssim_weight=0.2):
    """The reference's image loss on [B, H, W, C] images and its gradient with
    respect to color (float64, torch autograd): (loss, grad)."""
    import torch
    c = torch.as_tensor(np.asarray(color), dtype=torch.float64).clone().requires_grad_(True)
    t = torch.as_tensor(np.asarray(target), dtype=torch.float64)
    loss = l1_weight * (t - c).abs().mean() + ssim_weight * 0.5 * (
        1 - ssim_fp64(t.permute(0, 3, 1, 2), c.permute(0, 3, 1, 2)))
    loss.backward()
    return float(loss), c.grad.numpy()
