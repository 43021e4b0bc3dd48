"""The window finite differences (tests/band.py) against the oracle's own
fp64 adjoint on a small SH + depth scene: the per-Gaussian directional
derivatives agree, which pins the FD machinery the large-scene GPU
gradient checks rely on (the adjoint itself is pinned by whole-image fp64
FD in test_oracle_blend / test_oracle_geometry / test_oracle_sh)."""

import numpy as np
import torch

from msplat_amd import scenes
from tests.band import WindowFD, sh_scene_backward, sh_scene_forward


def test_window_fd_matches_oracle_adjoint():
    W, H, n = 96, 80, 600
    gen = torch.Generator().manual_seed(41)
    g = {k: v.numpy().astype(np.float64) for k, v in scenes.make_gaussians(gen, n, 1.0).items()}
    cam = scenes.make_camera(gen, W, H, 4.0, 1.171875 * W)
    intr, extr, campos = cam.intr.numpy(), cam.extr.numpy(), cam.eye.numpy().astype(np.float64)
    shs = torch.randn(n, 3, 16, generator=gen).numpy().astype(np.float64) * 0.3
    target = torch.rand(4, H, W, generator=gen).numpy().astype(np.float64)
    args = (g["xyz"], g["scale"], g["rotate"], g["opacity"], shs, intr, extr, campos, W, H, 0.0)
    out, st = sh_scene_forward(*args, depth_channel=True)
    grads = sh_scene_backward(*args, st, 2.0 * (out - target), depth_channel=True)
    fd = WindowFD(st["geo"], g["opacity"], st["feat"], st["idx"], st["rng"], W, H, target, 0.0, intr, extr,
                  campos=campos, shs=shs, depth_channel=True)
    rng = np.random.default_rng(5)
    live = np.nonzero((st["geo"]["radius"] > 0) & (np.abs(grads["dopacity"]) > 1e-6))[0]
    checked = 0
    for gi in rng.choice(live, 30, replace=False):
        p = fd.params(gi, g["xyz"], g["scale"], g["rotate"], g["opacity"])
        h = np.r_[np.full(3, 1e-6), 1e-6 * g["scale"][gi], np.full(4, 1e-6), 1e-6, np.full(3, 1e-6)]
        v = h * rng.normal(size=p.shape)
        an = (grads["dxyz"][gi] @ v[0:3] + grads["dscale"][gi] @ v[3:6] + grads["drotate"][gi] @ v[6:10]
              + grads["dopacity"][gi] * v[10] + grads["dshs"][gi, :, 0] @ v[11:14])
        d1, d4 = fd.directional(gi, p, v)
        if abs(d1 - d4) > 1e-4 * max(abs(d1), 1e-12):  # a blending threshold crossed within the step
            continue
        checked += 1
        assert abs(d1 - an) <= 1e-5 * max(abs(an), 1e-9), (gi, d1, an)
    assert checked >= 25
