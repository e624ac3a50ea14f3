"""On-disk formats either side of training, so a model trained with this
library opens in the reference's viewers and a reference checkpoint loads here.

* ``save_checkpoint`` / ``load_checkpoint``: the ``final_params.pt``
  dictionary of the reference's training scripts
  (examples/02_synthetic.py:478-492, examples/01_mip360.py:357-371): keys
  feat_acc, alpha_acc, V, F, color_mlp (a ColorMLP state dict), Rmin, Rmax,
  feat_dim, H, W, steps, losses, seed, and optional K, flip_z, up.  Loading
  uses torch's weights-only unpickler: such a file holds tensors, numbers,
  strings, lists and an ordered dict and nothing that runs.
* ``write_ply``: binary little-endian PLY with float xyz vertices and
  uchar-counted int32 triangles (examples/06_export_web.py:116-135).
* ``pack_face_color_lut``: per-texel features and opacity as the [H, 4096, 8]
  float32 table the viewers sample (examples/05_benchmark_fps.py:56-79);
  ``lut_rgba8`` splits it into the two RGBA8 textures of the web viewer.
* ``save_mlp_json``: the three layers as flat row-major arrays
  (examples/06_export_web.py:143-160).
* ``export_web``: the whole web-viewer asset directory of
  examples/06_export_web.py:178-296 (mesh.ply, lut0.png, lut1.png,
  mlp_weights.json, meta.json, models.json).
"""
from __future__ import annotations

import json
import math
import os
import struct
import zlib
from collections import OrderedDict

import numpy as np
import torch

from .multires import feats_at_level

CHECKPOINT_KEYS = ("feat_acc", "alpha_acc", "V", "F", "color_mlp", "Rmin", "Rmax", "feat_dim", "H", "W", "steps",
                   "losses", "seed")


def save_checkpoint(path, feat_acc, alpha_acc, verts, faces, color_mlp, Rmin, Rmax, H, W, steps=0, losses=(),
                    seed=0, **extra):
    """Write ``final_params.pt`` with the reference's keys.  ``color_mlp`` is a
    ColorMLP (or its state dict); ``extra`` adds optional keys (K, flip_z, up)."""
    state = color_mlp.state_dict() if hasattr(color_mlp, "state_dict") else color_mlp
    ckpt = {
        "feat_acc": feat_acc.detach().cpu(),
        "alpha_acc": alpha_acc.detach().cpu(),
        "V": verts.detach().cpu(),
        "F": faces.detach().cpu(),
        "color_mlp": OrderedDict((k, v.detach().cpu()) for k, v in state.items()),
        "Rmin": int(Rmin),
        "Rmax": int(Rmax),
        "feat_dim": int(feat_acc.shape[-1]),
        "H": int(H),
        "W": int(W),
        "steps": int(steps),
        "losses": [float(x) for x in losses],
        "seed": int(seed),
    }
    for k, v in extra.items():
        ckpt[k] = v.detach().cpu() if torch.is_tensor(v) else v
    torch.save(ckpt, path)


def load_checkpoint(path) -> dict:
    """Read a ``final_params.pt`` (the reference's or ours) with the
    weights-only loader; checks the required keys."""
    ckpt = torch.load(path, map_location="cpu", weights_only=True)
    missing = [k for k in ("feat_acc", "alpha_acc", "V", "F", "color_mlp", "Rmax") if k not in ckpt]
    if missing:
        raise KeyError(f"checkpoint {path} lacks {missing}")
    return ckpt


def detect_up(ckpt: dict):
    """World up direction from checkpoint metadata (examples/06_export_web.py:
    103-110): an explicit "up", else (0, -1, 0) for COLMAP scenes (flip_z),
    else (0, 0, 1)."""
    if "up" in ckpt:
        u = ckpt["up"]
        return (float(u[0]), float(u[1]), float(u[2]))
    if "flip_z" in ckpt:
        return (0.0, -1.0, 0.0)
    return (0.0, 0.0, 1.0)


def write_ply(path, verts, faces) -> None:
    """Binary little-endian PLY: float x, y, z per vertex; per face a uchar 3
    and three int32 indices."""
    v = np.ascontiguousarray(np.asarray(verts, np.float32).reshape(-1, 3))
    f = np.asarray(faces, np.int32).reshape(-1, 3)
    header = ("ply\nformat binary_little_endian 1.0\n"
              f"element vertex {v.shape[0]}\nproperty float x\nproperty float y\nproperty float z\n"
              f"element face {f.shape[0]}\nproperty list uchar int vertex_indices\nend_header\n")
    rec = np.empty(f.shape[0], dtype=np.dtype([("n", "u1"), ("idx", "<i4", (3,))]))
    rec["n"] = 3
    rec["idx"] = f
    with open(path, "wb") as fh:
        fh.write(header.encode("ascii"))
        fh.write(v.astype("<f4").tobytes())
        fh.write(rec.tobytes())


def read_ply(path):
    """(verts [N,3] float32, faces [M,3] int32) of a PLY written by write_ply."""
    with open(path, "rb") as fh:
        data = fh.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:end].decode("ascii").split("\n")
    nv = int(next(l for l in head if l.startswith("element vertex")).split()[-1])
    nf = int(next(l for l in head if l.startswith("element face")).split()[-1])
    v = np.frombuffer(data, "<f4", 3 * nv, end).reshape(nv, 3).astype(np.float32)
    rec = np.frombuffer(data, np.dtype([("n", "u1"), ("idx", "<i4", (3,))]), nf, end + 12 * nv)
    if nf and not (rec["n"] == 3).all():
        raise ValueError("only triangle faces are supported")
    return v, rec["idx"].astype(np.int32)


def pack_face_color_lut(feat_acc, alpha_acc, num_faces: int, level: int, width: int = 4096) -> np.ndarray:
    """Per-texel features and opacity of ``num_faces`` faces at ``level`` as a
    row-major [ceil(N / W), W, feat_dim + 1] float32 table, N = faces x texels
    per face, W = min(width, N), zero padded."""
    S = feats_at_level(level)
    N = num_faces * S
    fa = np.asarray(feat_acc, np.float32)
    aa = np.asarray(alpha_acc, np.float32)
    fa = fa.reshape(-1, fa.shape[-1])
    aa = aa.reshape(-1, aa.shape[-1])
    if fa.shape[0] < N or aa.shape[0] < N:
        raise ValueError(f"feature tables hold {fa.shape[0]} / {aa.shape[0]} texels, {N} needed")
    flat = np.concatenate([fa[:N], aa[:N]], axis=-1)
    tex_w = min(width, N)
    tex_h = math.ceil(N / tex_w) if N else 0
    out = np.zeros((tex_h * tex_w, flat.shape[1]), np.float32)
    out[:N] = flat
    return out.reshape(tex_h, tex_w, flat.shape[1])


def lut_rgba8(lut: np.ndarray):
    """The two RGBA8 textures of an 8-channel LUT: channels 0-3 and 4-7,
    scaled by 255, clipped and truncated."""
    if lut.shape[-1] != 8:
        raise ValueError("the web viewer's LUT has 8 channels (7 features + opacity)")
    return tuple((lut[..., k:k + 4] * 255).clip(0, 255).astype(np.uint8) for k in (0, 4))


def write_png_rgba8(path, rgba: np.ndarray) -> None:
    """8-bit RGBA PNG (filter 0 on every row, zlib-compressed)."""
    a = np.ascontiguousarray(rgba, np.uint8)
    h, w = a.shape[:2]
    raw = np.concatenate([np.zeros((h, 1), np.uint8), a.reshape(h, 4 * w)], axis=1).tobytes()

    def chunk(tag, body):
        return struct.pack(">I", len(body)) + tag + body + struct.pack(">I", zlib.crc32(tag + body) & 0xFFFFFFFF)
    with open(path, "wb") as fh:
        fh.write(b"\x89PNG\r\n\x1a\n")
        fh.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)))
        fh.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        fh.write(chunk(b"IEND", b""))


def mlp_weights(state_dict):
    """(W1, b1, W2, b2, W3, b3) float32 from a ColorMLP state dict, in layer
    order (examples/06_export_web.py:80-100)."""
    weights = [v for k, v in state_dict.items() if "weight" in k]
    biases = [v for k, v in state_dict.items() if "bias" in k]
    if len(weights) < 3 or len(biases) < 3:
        raise ValueError(f"expected >= 3 linear layers, got keys {list(state_dict)}")
    as_np = lambda t: t.detach().cpu().numpy().astype(np.float32)
    return (as_np(weights[0]), as_np(biases[0]), as_np(weights[1]), as_np(biases[1]), as_np(weights[2]),
            as_np(biases[2]))


def save_mlp_json(path, W1, b1, W2, b2, W3, b3) -> None:
    data = {k: np.asarray(v, np.float32).ravel().tolist()
            for k, v in (("W1", W1), ("b1", b1), ("W2", W2), ("b2", b2), ("W3", W3), ("b3", b3))}
    with open(path, "w") as fh:
        json.dump(data, fh)


def export_web(ckpt, output_dir, scene: str, up=None, background=None) -> str:
    """Write the web viewer's assets for one checkpoint (a path or a loaded
    dict) into output_dir/scene and list the scene in output_dir/models.json."""
    if not isinstance(ckpt, dict):
        ckpt = load_checkpoint(ckpt)
    scene_dir = os.path.join(output_dir, scene)
    os.makedirs(scene_dir, exist_ok=True)
    verts = ckpt["V"].numpy().astype(np.float32)
    faces = ckpt["F"].numpy().astype(np.int32)
    level = int(ckpt["Rmax"])
    write_ply(os.path.join(scene_dir, "mesh.ply"), verts, faces)
    lut = pack_face_color_lut(ckpt["feat_acc"].numpy(), ckpt["alpha_acc"].numpy(), faces.shape[0], level)
    lut0, lut1 = lut_rgba8(lut)
    write_png_rgba8(os.path.join(scene_dir, "lut0.png"), lut0)
    write_png_rgba8(os.path.join(scene_dir, "lut1.png"), lut1)
    save_mlp_json(os.path.join(scene_dir, "mlp_weights.json"), *mlp_weights(ckpt["color_mlp"]))
    up = tuple(up) if up is not None else detect_up(ckpt)
    if background is None:
        background = [0.0, 0.0, 0.0] if "flip_z" in ckpt else [1.0, 1.0, 1.0]
    meta = {"up": list(up), "level": level, "background": list(background), "num_faces": int(faces.shape[0]),
            "num_verts": int(verts.shape[0])}
    with open(os.path.join(scene_dir, "meta.json"), "w") as fh:
        json.dump(meta, fh, indent=2)
    models_path = os.path.join(output_dir, "models.json")
    models = []
    if os.path.exists(models_path):
        with open(models_path) as fh:
            models = json.load(fh)
    if scene not in models:
        models.append(scene)
        models.sort()
    with open(models_path, "w") as fh:
        json.dump(models, fh, indent=2)
    return scene_dir
