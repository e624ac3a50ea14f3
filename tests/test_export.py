"""On-disk formats (examples/02_synthetic.py:478-492, examples/05_benchmark_fps.py:
47-102, examples/06_export_web.py:116-296): the checkpoint's keys and a
weights-only round trip into ColorMLP, the PLY byte layout, the LUT packing
against its closed form, PNGs decoded by PIL and by hand, the MLP JSON and the
web asset directory.  CPU only."""
import json
import os
import struct
import zlib

import numpy as np
import pytest
import torch

from diffsoup_amd import export as E
from diffsoup_amd.multires import ColorMLP


def _model(F=120, level=2, feat_dim=7, seed=0):
    g = torch.Generator().manual_seed(seed)
    S = E.feats_at_level(level)
    V = torch.rand(3 * F, 3, generator=g)
    Fa = torch.arange(3 * F, dtype=torch.int32).reshape(F, 3)
    feat = torch.rand(F, S, feat_dim, generator=g)
    alpha = torch.rand(F, S, 1, generator=g)
    torch.manual_seed(seed)
    mlp = ColorMLP(feat_dim + 9, 3, hidden_dim=16, n_layers=2)
    return V, Fa, feat, alpha, mlp


def test_checkpoint_round_trip(tmp_path):
    V, F, feat, alpha, mlp = _model()
    p = tmp_path / "final_params.pt"
    E.save_checkpoint(p, feat, alpha, V, F, mlp, 0, 2, 64, 80, steps=10, losses=[0.5, 0.25], seed=3, up=(0, 1, 0))
    ck = E.load_checkpoint(p)
    assert set(E.CHECKPOINT_KEYS) <= set(ck)
    assert ck["Rmin"] == 0 and ck["Rmax"] == 2 and ck["feat_dim"] == 7 and ck["H"] == 64 and ck["W"] == 80
    assert ck["losses"] == [0.5, 0.25] and ck["seed"] == 3 and E.detect_up(ck) == (0.0, 1.0, 0.0)
    assert torch.equal(ck["feat_acc"], feat) and torch.equal(ck["F"], F)
    assert list(ck["color_mlp"]) == ["mlp.0.weight", "mlp.0.bias", "mlp.2.weight", "mlp.2.bias", "mlp.4.weight",
                                     "mlp.4.bias"]
    m2 = ColorMLP(16, 3, hidden_dim=16, n_layers=2)
    m2.load_state_dict(ck["color_mlp"])
    x = torch.rand(5, 16)
    assert torch.equal(m2.mlp(x), mlp.mlp(x))
    del ck["up"]
    assert E.detect_up(ck) == (0.0, 0.0, 1.0)
    ck["flip_z"] = True
    assert E.detect_up(ck) == (0.0, -1.0, 0.0)


def test_ply_layout(tmp_path):
    V, F, *_ = _model(F=7)
    p = tmp_path / "mesh.ply"
    E.write_ply(p, V.numpy(), F.numpy())
    data = open(p, "rb").read()
    header = (b"ply\nformat binary_little_endian 1.0\nelement vertex 21\nproperty float x\nproperty float y\n"
              b"property float z\nelement face 7\nproperty list uchar int vertex_indices\nend_header\n")
    assert data.startswith(header)
    assert len(data) == len(header) + 21 * 12 + 7 * 13
    body = data[len(header):]
    assert body[:12] == struct.pack("<3f", *V[0].tolist())
    assert body[21 * 12:21 * 12 + 13] == b"\x03" + struct.pack("<3i", 0, 1, 2)
    v2, f2 = E.read_ply(p)
    np.testing.assert_array_equal(v2, V.numpy())
    np.testing.assert_array_equal(f2, F.numpy())


@pytest.mark.parametrize("F,level", [(10, 0), (300, 3), (91, 2)])
def test_lut_packing(F, level):
    _, _, feat, alpha, _ = _model(F=F, level=level)
    lut = E.pack_face_color_lut(feat.numpy(), alpha.numpy(), F, level)
    N = F * E.feats_at_level(level)
    W = min(4096, N)
    assert lut.shape == (-(-N // W), W, 8) and lut.dtype == np.float32
    flat = lut.reshape(-1, 8)
    np.testing.assert_array_equal(flat[:N, :7], feat.numpy().reshape(-1, 7))
    np.testing.assert_array_equal(flat[:N, 7], alpha.numpy().reshape(-1))
    assert not flat[N:].any()


def test_rgba8_and_png(tmp_path):
    lut = np.array([[[0.0, 0.5, 0.999, 1.2, -0.1, 1.0, 0.25, 0.7]]], np.float32)
    a, b = E.lut_rgba8(lut)
    np.testing.assert_array_equal(a[0, 0], [0, 127, 254, 255])
    np.testing.assert_array_equal(b[0, 0], [0, 255, 63, 178])
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (13, 37, 4), dtype=np.uint8)
    p = tmp_path / "lut0.png"
    E.write_png_rgba8(p, img)
    from PIL import Image
    np.testing.assert_array_equal(np.asarray(Image.open(p).convert("RGBA")), img)
    # decoded by hand: the IDAT stream is one filter byte (0) and the RGBA row per scanline
    data = open(p, "rb").read()
    i = data.index(b"IDAT")
    n = struct.unpack(">I", data[i - 4:i])[0]
    raw = np.frombuffer(zlib.decompress(data[i + 4:i + 4 + n]), np.uint8).reshape(13, 1 + 37 * 4)
    assert not raw[:, 0].any()
    np.testing.assert_array_equal(raw[:, 1:].reshape(13, 37, 4), img)


def test_web_export(tmp_path):
    V, F, feat, alpha, mlp = _model(F=50, level=1)
    ck = tmp_path / "final_params.pt"
    E.save_checkpoint(ck, feat, alpha, V, F, mlp, 0, 1, 32, 32)
    out = tmp_path / "web"
    E.export_web(str(ck), str(out), "b_scene")
    d = E.export_web(str(ck), str(out), "a_scene", background=[0.2, 0.2, 0.2])
    assert sorted(os.listdir(d)) == ["lut0.png", "lut1.png", "mesh.ply", "meta.json", "mlp_weights.json"]
    assert json.load(open(out / "models.json")) == ["a_scene", "b_scene"]
    meta = json.load(open(os.path.join(d, "meta.json")))
    assert meta == {"up": [0.0, 0.0, 1.0], "level": 1, "background": [0.2, 0.2, 0.2], "num_faces": 50,
                    "num_verts": 150}
    w = json.load(open(os.path.join(d, "mlp_weights.json")))
    assert list(w) == ["W1", "b1", "W2", "b2", "W3", "b3"]
    lin = mlp.linears()
    np.testing.assert_array_equal(np.array(w["W1"], np.float32), lin[0].weight.detach().numpy().ravel())
    np.testing.assert_array_equal(np.array(w["b3"], np.float32), lin[2].bias.detach().numpy())
    from PIL import Image
    lut = E.pack_face_color_lut(feat.numpy(), alpha.numpy(), 50, 1)
    np.testing.assert_array_equal(np.asarray(Image.open(os.path.join(d, "lut1.png"))), E.lut_rgba8(lut)[1])
