"""Seeded synthetic scenes and cameras for the benchmark workloads.

All tensors are drawn on the CPU with ``torch.Generator`` so a seed gives the
same scene on any machine; callers move them to the GPU.  Distributions:
xyz ~ U[-a, a]^3, scale = exp(U(-5, -2)) per axis, rotate = normalize(N(0,1)^4)
as (w,x,y,z), opacity = sigmoid(N(0,1)) [N,1].  Cameras are OpenCV look-at
views of the origin (x right, y down, z forward).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch


@dataclass
class Camera:
    W: int
    H: int
    intr: torch.Tensor  # [4] fx, fy, cx, cy
    extr: torch.Tensor  # [3,4] world-to-camera [R|t]
    eye: torch.Tensor   # [3] camera centre

    def to(self, device):
        return Camera(self.W, self.H, self.intr.to(device), self.extr.to(device), self.eye.to(device))


def look_at(eye: torch.Tensor, target=(0.0, 0.0, 0.0)) -> torch.Tensor:
    """World-to-camera [R|t] (OpenCV) looking from ``eye`` at ``target`` with up = +z."""
    eye = eye.double()
    z = torch.tensor(target, dtype=torch.float64) - eye
    z = z / z.norm()
    up = torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64)
    if torch.cross(z, up, dim=0).norm() < 1e-6:
        up = torch.tensor([0.0, 1.0, 0.0], dtype=torch.float64)
    x = torch.cross(z, up, dim=0)
    x = x / x.norm()
    y = torch.cross(z, x, dim=0)
    R = torch.stack([x, y, z], 0)
    t = -R @ eye
    return torch.cat([R, t[:, None]], 1).float()


def random_direction(gen: torch.Generator) -> torch.Tensor:
    v = torch.randn(3, generator=gen, dtype=torch.float64)
    return v / v.norm()


def make_camera(gen: torch.Generator, W: int, H: int, dist: float, fx: float, fy: float | None = None,
                cx: float | None = None, cy: float | None = None) -> Camera:
    eye = random_direction(gen) * dist
    intr = torch.tensor([fx, fy if fy is not None else fx, W / 2 if cx is None else cx,
                         H / 2 if cy is None else cy], dtype=torch.float32)
    return Camera(W, H, intr, look_at(eye), eye.float())


def make_gaussians(gen: torch.Generator, n: int, half_extent: float) -> dict:
    xyz = (torch.rand(n, 3, generator=gen) * 2 - 1) * half_extent
    scale = torch.exp(torch.rand(n, 3, generator=gen) * 3 - 5)
    q = torch.randn(n, 4, generator=gen)
    rotate = q / q.norm(dim=1, keepdim=True)
    opacity = torch.sigmoid(torch.randn(n, 1, generator=gen))
    return dict(xyz=xyz, scale=scale, rotate=rotate, opacity=opacity)


def view_dirs(xyz: torch.Tensor, eye: torch.Tensor) -> torch.Tensor:
    d = xyz - eye[None]
    return d / d.norm(dim=1, keepdim=True)


def config_small(seed: int = 0, n: int = 10_000, W: int = 256, H: int = 256, C: int = 3) -> dict:
    """Plain forward workload: N=10k, 256x256, intr [300,300,128,128], eye at distance 4, feats U(0,1)."""
    gen = torch.Generator().manual_seed(seed)
    g = make_gaussians(gen, n, 1.0)
    cam = make_camera(gen, W, H, 4.0, 300.0 * W / 256, cx=W / 2, cy=H / 2)
    g["feature"] = torch.rand(n, C, generator=gen)
    return dict(gaussians=g, camera=cam, bg=0.0)


def config_sh3_1080p(seed: int = 1, n: int = 1_000_000, W: int = 1920, H: int = 1080) -> dict:
    """fwd+bwd at scale: N Gaussians in [-3,3]^3, 1920x1080, fx = fy = 1.171875 W (the
    256-px workload's field of view), SH degree 3 [N,3,16] ~ N(0, 0.3) plus a depth
    channel (C = 4), L2 loss against a U(0,1) target."""
    gen = torch.Generator().manual_seed(seed)
    g = make_gaussians(gen, n, 3.0)
    cam = make_camera(gen, W, H, 4.0, 1.171875 * W)
    g["shs"] = torch.randn(n, 3, 16, generator=gen) * 0.3
    target = torch.rand(4, H, W, generator=gen)
    return dict(gaussians=g, camera=cam, bg=0.0, target=target)


def config_sh10_wide(seed: int = 2, n: int = 500_000, W: int = 800, H: int = 800, extra: int = 32) -> dict:
    """High-order SH + wide features: SH degree 10 [N,3,121] with sigma_l = 0.1 / max(l, 1)
    per band, plus `extra` N(0,1) channels (C = 3 + extra), 800x800."""
    gen = torch.Generator().manual_seed(seed)
    g = make_gaussians(gen, n, 3.0)
    cam = make_camera(gen, W, H, 4.0, 1.171875 * W)
    sigma = torch.tensor([0.1 / max(l, 1) for l in range(11) for _ in range(2 * l + 1)])
    g["shs"] = torch.randn(n, 3, 121, generator=gen) * sigma
    g["extra"] = torch.randn(n, extra, generator=gen)
    return dict(gaussians=g, camera=cam, bg=0.0)


def config_multiview(seed: int = 3, n: int = 2_000_000, views: int = 16) -> dict:
    """Batched views: B cameras on a sphere of radius 5, W ~ U{400..1920}, H ~ U{300..1080},
    fx = fy ~ U(0.8, 1.5) * W, C = 3."""
    gen = torch.Generator().manual_seed(seed)
    g = make_gaussians(gen, n, 3.0)
    cams = []
    for _ in range(views):
        W = int(torch.randint(400, 1921, (1,), generator=gen))
        H = int(torch.randint(300, 1081, (1,), generator=gen))
        f = float(torch.empty(1).uniform_(0.8, 1.5, generator=gen)) * W
        cams.append(make_camera(gen, W, H, 5.0, f))
    g["feature"] = torch.rand(n, 3, generator=gen)
    return dict(gaussians=g, cameras=cams, bg=0.0)


def fov_focal(W: int, fov_deg: float) -> float:
    return W / (2 * math.tan(math.radians(fov_deg) / 2))


SERVING_SIZES = ((640, 480), (1280, 720), (1920, 1080))


def config_serving(seed: int = 4, counts=(100_000, 500_000, 1_500_000, 3_000_000)) -> dict:
    """Resident scenes for serving: distributions as the 1080p workload (xyz in [-3,3]^3), C = 3."""
    gen = torch.Generator().manual_seed(seed)
    out = {}
    for i, n in enumerate(counts):
        g = make_gaussians(gen, n, 3.0)
        g["feature"] = torch.rand(n, 3, generator=gen)
        out[f"scene{i}_{n}"] = g
    return out


def serving_camera(rng, W: int, H: int):
    """A random look-at camera at distance 4..5, fx = fy = 1.171875 W (rng: random.Random)."""
    gen = torch.Generator().manual_seed(rng.randrange(1 << 30))
    return make_camera(gen, W, H, 4.0 + rng.random(), 1.171875 * W)
