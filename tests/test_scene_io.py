"""On-disk scenes (msplat_amd.scene_io): PLY files of the 3D Gaussian-splatting
layout, written by hand here (struct-packed bytes) and by save_ply."""

import struct

import numpy as np
import pytest
import torch

from msplat_amd import scene_io


def _write(path, props, rows, fmt="binary_little_endian", extra_header=()):
    head = ["ply", f"format {fmt} 1.0", *extra_header, f"element vertex {len(rows)}"]
    head += [f"property {t} {name}" for t, name in props]
    head += ["end_header"]
    codes = {"float": "f", "double": "d", "uchar": "B", "int": "i"}
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        for r in rows:
            f.write(struct.pack("<" + "".join(codes[t] for t, _ in props), *r))


BASE = [("float", "x"), ("float", "y"), ("float", "z"), ("float", "opacity"), ("float", "scale_0"),
        ("float", "scale_1"), ("float", "scale_2"), ("float", "rot_0"), ("float", "rot_1"), ("float", "rot_2"),
        ("float", "rot_3")]


def test_hand_written_file_channel_major_sh(tmp_path):
    # degree 1, three channels: f_rest_{c*3 + k-1} holds coefficient k of channel c
    props = BASE + [("float", f"f_dc_{c}") for c in range(3)] + [("float", f"f_rest_{j}") for j in range(9)]
    props.insert(3, ("uchar", "red"))  # unrelated properties are skipped
    rows = []
    for g in range(2):
        base = [1.0 + g, 2.0, 3.0, 77, 0.0, 0.0, 0.5, -0.5, 2.0, 0.0, 0.0, 0.0]
        rows.append(base + [10 * g + c for c in range(3)] + [100 * g + j for j in range(9)])
    p = tmp_path / "hand.ply"
    _write(p, props, rows, extra_header=["comment made by hand"])
    s = scene_io.load_ply(str(p), activate=False)
    assert s["sh_degree"] == 1 and tuple(s["shs"].shape) == (2, 3, 4)
    for g in range(2):
        for c in range(3):
            assert s["shs"][g, c, 0] == 10 * g + c
            for k in range(1, 4):
                assert s["shs"][g, c, k] == 100 * g + c * 3 + (k - 1)
    assert torch.equal(s["xyz"], torch.tensor([[1.0, 2.0, 3.0], [2.0, 2.0, 3.0]]))
    a = scene_io.load_ply(str(p))  # activations: exp, sigmoid, normalised quaternion
    assert torch.allclose(a["scale"][0], torch.exp(torch.tensor([0.0, 0.5, -0.5])))
    assert torch.allclose(a["opacity"][0], torch.tensor([0.5]))
    assert torch.equal(a["rotate"][0], torch.tensor([1.0, 0.0, 0.0, 0.0]))


def test_double_properties_and_features(tmp_path):
    props = [(("double" if t == "float" else t), n) for t, n in BASE] + [("float", "feat_1"), ("float", "feat_0")]
    rows = [[0.5, 0.25, 0.125, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 9.0, 8.0]]
    p = tmp_path / "d.ply"
    _write(p, props, rows)
    s = scene_io.load_ply(str(p), activate=False)
    assert s["xyz"].dtype == torch.float32
    assert torch.equal(s["feature"], torch.tensor([[8.0, 9.0]]))  # feat_0, feat_1 in index order
    assert "shs" not in s


def test_round_trip_raw_bitexact(tmp_path):
    rng = np.random.default_rng(0)
    n = 5000
    t = lambda *s: torch.from_numpy(rng.normal(size=s).astype(np.float32))
    xyz, scale, rot, op, shs, feat = t(n, 3), t(n, 3), t(n, 4), t(n, 1), t(n, 3, 16), t(n, 5)
    p = tmp_path / "raw.ply"
    scene_io.save_ply(str(p), xyz, scale, rot, op, shs=shs, feature=feat, activated=False)
    s = scene_io.load_ply(str(p), activate=False)
    for k, v in (("xyz", xyz), ("scale", scale), ("rotate", rot), ("opacity", op), ("shs", shs),
                 ("feature", feat)):
        assert torch.equal(s[k], v), k
    assert s["sh_degree"] == 3


def test_round_trip_activated(tmp_path):
    rng = np.random.default_rng(1)
    n = 3000
    xyz = torch.from_numpy(rng.uniform(-3, 3, (n, 3)).astype(np.float32))
    scale = torch.from_numpy(np.exp(rng.uniform(-5, -2, (n, 3))).astype(np.float32))
    rot = torch.nn.functional.normalize(torch.from_numpy(rng.normal(size=(n, 4)).astype(np.float32)), dim=1)
    op = torch.from_numpy(rng.uniform(0.01, 0.99, (n, 1)).astype(np.float32))
    feat = torch.from_numpy(rng.uniform(0, 1, (n, 3)).astype(np.float32))
    p = tmp_path / "act.ply"
    scene_io.save_ply(str(p), xyz, scale, rot, op, feature=feat)
    s = scene_io.load_ply(str(p))
    assert torch.equal(s["xyz"], xyz) and torch.equal(s["feature"], feat)
    assert float(((s["scale"] - scale).abs() / scale).max()) < 1e-6
    assert float((s["opacity"] - op).abs().max()) < 1e-6
    assert float((s["rotate"] - rot).abs().max()) < 1e-6


def test_rejects_bad_files(tmp_path):
    p = tmp_path / "a.ply"
    _write(p, BASE, [[0.0] * 11], fmt="ascii")
    with pytest.raises(ValueError, match="binary_little_endian"):
        scene_io.load_ply(str(p))
    _write(p, BASE[:-1], [[0.0] * 10])
    with pytest.raises(ValueError, match="rot_3"):
        scene_io.load_ply(str(p))
    _write(p, BASE, [[0.0] * 11] * 3)
    data = p.read_bytes()
    p.write_bytes(data[:-8])
    with pytest.raises(ValueError, match="do not fit"):
        scene_io.load_ply(str(p))
    p.write_bytes(b"plx\n")
    with pytest.raises(ValueError, match="not a PLY"):
        scene_io.load_ply(str(p))
    props = BASE + [("float", "f_dc_0"), ("float", "f_rest_0")]
    _write(p, props, [[0.0] * 13])
    with pytest.raises(ValueError, match="square"):
        scene_io.load_ply(str(p))
