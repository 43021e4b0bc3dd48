"""Shared helpers for the test suite (inputs, conversions, comparisons)."""

from __future__ import annotations

import numpy as np
import torch

from msplat_amd import scenes


def small_scene(seed=0, n=2000, W=128, H=96, C=3, half_extent=1.0, dist=4.0, fx=None):
    """A seeded scene small enough for the CPU oracle (seconds)."""
    gen = torch.Generator().manual_seed(seed)
    g = scenes.make_gaussians(gen, n, half_extent)
    cam = scenes.make_camera(gen, W, H, dist, fx if fx is not None else 1.171875 * W)
    g["feature"] = torch.rand(n, C, generator=gen)
    return g, cam


def np_(t):
    return t.detach().cpu().numpy()


def dev(t, device="cuda"):
    return t.to(device)


def rel_err(a, b, floor=1e-12):
    """max |a-b| / max(|b|) over the whole array (scale-aware relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), floor)) if a.size else 0.0


def assert_bitexact(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if a.dtype.kind == "f":
        same = (a == b) | (np.isnan(a) & np.isnan(b))
    else:
        same = a == b
    bad = np.argwhere(~same)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first at {bad[0].tolist()}: {a[tuple(bad[0])]} vs " \
                          f"{b[tuple(bad[0])]}"
