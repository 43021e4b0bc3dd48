"""Pin the sort oracle (CPU only): the reference's known-answer test, brute
force with Python's stable sorted() over the signed 64-bit keys, invariants."""

import json
import os
import struct

import numpy as np
import pytest
import torch

from oracle import sort as osort
from tests.conftest import GOLDEN


def test_known_answer():
    kat = json.load(open(os.path.join(GOLDEN, "sort_kat.json")))
    uv = np.array(kat["uv"], np.float32)
    idx, rng, keys = osort.sort_gaussian(uv, np.array(kat["depth"], np.float32), np.array(kat["radius"]),
                                         kat["W"], kat["H"])
    assert idx.tolist() == kat["idx_sorted"]
    assert rng.tolist() == kat["tile_range"]
    _, _, area = osort.gaussian_keys(uv, np.array(kat["depth"], np.float32), np.array(kat["radius"]), kat["W"],
                                     kat["H"])
    assert area.tolist() == kat["tiles"]


def brute(uv, depth, radius, W, H):
    """Keys by explicit loops, signed 64-bit, then Python's stable sort."""
    gx, gy = (W + 15) // 16, (H + 15) // 16
    items = []
    for g in range(len(uv)):
        r = int(radius[g])
        if r <= 0:
            continue
        x0 = min(gx, max(0, int((np.float32(uv[g, 0]) - np.float32(r)) / np.float32(16))))
        y0 = min(gy, max(0, int((np.float32(uv[g, 1]) - np.float32(r)) / np.float32(16))))
        x1 = min(gx, max(0, int((np.float32(uv[g, 0]) + np.float32(r) + np.float32(15)) / np.float32(16))))
        y1 = min(gy, max(0, int((np.float32(uv[g, 1]) + np.float32(r) + np.float32(15)) / np.float32(16))))
        dbits = struct.unpack("<i", struct.pack("<f", float(depth[g])))[0]  # int32 bit pattern
        for y in range(y0, y1):
            for x in range(x0, x1):
                key = ((y * gx + x) << 32) | (dbits & 0xFFFFFFFF)
                if dbits < 0:
                    key |= 0xFFFFFFFF << 32  # sign extension
                if key >= 1 << 63:
                    key -= 1 << 64
                items.append((key, g))
    items = sorted(items, key=lambda kv: kv[0])  # stable
    keys = [k for k, _ in items]
    ids = [g for _, g in items]
    rng = np.zeros((gx * gy, 2), np.int32)
    for i, k in enumerate(keys):
        t = k >> 32
        if t < 0:
            continue
        if i == 0 or (keys[i - 1] >> 32) != t:
            rng[t, 0] = i
        if i == len(keys) - 1 or (keys[i + 1] >> 32) != t:
            rng[t, 1] = i + 1
    return np.array(keys, np.int64), np.array(ids, np.int32), rng


@pytest.mark.parametrize("seed", range(6))
def test_brute_force(seed):
    rng = np.random.default_rng(seed)
    W, H = int(rng.integers(20, 90)), int(rng.integers(20, 70))
    n = 60
    uv = rng.uniform(-20, max(W, H) + 20, (n, 2)).astype(np.float32)
    depth = rng.choice(np.float32([-1.5, -0.25, 0.5, 1.0, 2.0]), n)  # ties and negative depths
    radius = rng.integers(-1, 30, n)
    bk, bi, br = brute(uv, depth, radius, W, H)
    idx, trange, keys = osort.sort_gaussian(uv, depth, radius, W, H)
    assert keys.tolist() == bk.tolist()
    assert idx.tolist() == bi.tolist()
    assert np.array_equal(trange, br)


def test_invariants():
    rng = np.random.default_rng(9)
    W, H = 300, 200
    n = 2000
    uv = rng.uniform(-50, 350, (n, 2)).astype(np.float32)
    depth = rng.uniform(-1, 5, n).astype(np.float32)
    radius = rng.integers(0, 80, n)
    idx, trange, keys = osort.sort_gaussian(uv, depth, radius, W, H)
    _, _, area = osort.gaussian_keys(uv, depth, radius, W, H)
    assert np.all(np.diff(keys) >= 0)
    assert np.array_equal(np.bincount(idx, minlength=n), area)
    neg = int((keys < 0).sum())
    assert neg == int(area[depth < 0].sum())
    live = trange[trange[:, 1] > trange[:, 0]]
    # ranges tile [neg, K) exactly, in tile order
    assert live[0, 0] == neg and live[-1, 1] == len(keys)
    assert np.all(live[1:, 0] == live[:-1, 1])
    for t in range(len(trange)):
        s, e = trange[t]
        assert np.all((keys[s:e] >> 32) == t)
        assert np.all(np.diff(depth[idx[s:e]]) >= 0)  # front to back within a tile


@pytest.mark.parametrize("cell", [1, 2, 3, 8])
def test_cell_lists_filter_to_tile_lists(cell):
    """A tile's list is its cell's list filtered by tile rectangle, for every tile."""
    rng_ = np.random.default_rng(cell)
    n, W, H = 300, 200, 136
    uv = rng_.uniform(-40, 240, (n, 2)).astype(np.float32)
    depth = rng_.uniform(-1, 8, n).astype(np.float32)
    depth[:5] = depth[5:10]  # ties in depth
    depth[10] = -0.0         # sign bit set: no tile
    radius = rng_.integers(-2, 60, n)
    idx, trng, _ = osort.sort_gaussian(uv, depth, radius, W, H)
    rect, ids, crng = osort.cell_lists(uv, depth, radius, W, H, cell)
    gx, gy = osort.grid(W, H)
    ncx = (gx + cell - 1) // cell
    assert 10 not in ids.tolist()
    for t in range(gx * gy):
        tx, ty = t % gx, t // gx
        c = (ty // cell) * ncx + tx // cell
        s, e = crng[c]
        walk = [g for g in ids[s:e] if rect[g, 0] <= tx < rect[g, 2] and rect[g, 1] <= ty < rect[g, 3]]
        a, b = trng[t]
        assert walk == idx[a:b].tolist(), f"tile {t}"


def test_cell_lists_single_cell_is_depth_sort():
    """With one cell covering the grid the list is the stable depth sort of the live Gaussians."""
    rng_ = np.random.default_rng(7)
    n, W, H = 200, 64, 48
    uv = rng_.uniform(0, 1, (n, 2)).astype(np.float32) * np.float32([64, 48])  # inside the image
    depth = rng_.choice(np.float32([0.5, 1.0, 2.0, 3.0]), n)
    radius = rng_.integers(1, 20, n)
    _, ids, crng = osort.cell_lists(uv, depth, radius, W, H, 8)
    assert crng.shape == (1, 2) and tuple(crng[0]) == (0, n)
    expect = sorted(range(n), key=lambda g: (float(depth[g]), g))
    assert ids.tolist() == expect


@pytest.mark.parametrize("seed,W,H,rows", [(3, 200, 136, (3, 6)), (4, 97, 61, (0, 1)), (5, 160, 112, (5, 7))])
def test_band_lists_equal_full_sort_on_the_band(seed, W, H, rows):
    """band_lists gives every tile of the band its full-sort list (ids and
    order: ties, negative depths and boxes crossing the band included) and
    every other tile nothing."""
    from msplat_amd import scenes
    from oracle import pipeline
    gen = torch.Generator().manual_seed(seed)
    g = scenes.make_gaussians(gen, 4000, 1.5)
    cam = scenes.make_camera(gen, W, H, 4.0, 1.2 * W)
    geo = pipeline.stage_geometry(g["xyz"].numpy(), g["scale"].numpy(), g["rotate"].numpy(), cam.intr.numpy(),
                                  cam.extr.numpy(), W, H)
    depth = geo["depth"].copy()
    depth[::50] = depth[1::50][: len(depth[::50])]  # ties
    depth[::37] = -depth[::37]  # negative depths: keyed into tile -1
    assert (depth < 0).any()
    idx, rng, _ = osort.sort_gaussian(geo["uv"], depth, geo["radius"], W, H)
    bi, br = osort.band_lists(geo["uv"], depth, geo["radius"], W, H, *rows)
    gx = (W + 15) // 16
    for t in range(rng.shape[0]):
        band = bi[br[t, 0]:br[t, 1]]
        if rows[0] <= t // gx < rows[1]:
            assert np.array_equal(idx[rng[t, 0]:rng[t, 1]], band), t
        else:
            assert len(band) == 0
