"""Tile keys, stable key sort and tile ranges on the CPU (test infrastructure only).

Restates msplat/sort_gaussian.py:42-52 and msplat/src/sort_gaussian.cu:17-71:
keys are ``(tile << 32) | int64(int32 depth bits)`` (sign-extended, so
negative depths land in "tile -1"), sorted as signed int64 by a stable sort,
and tile_range[t] = [first, last+1) of tile t (untouched tiles stay (0,0)).
"""

from __future__ import annotations

import numpy as np

from .geometry import TILE, tile_rect


def grid(W, H):
    return (W + TILE - 1) // TILE, (H + TILE - 1) // TILE


def gaussian_keys(uv, depth, radius, W, H):
    """sort_gaussian.cu:17-43: keys [K] int64 and ids [K] int32 in Gaussian order.

    The per-Gaussian count is the tile box area (what ewa_project reports as
    ``tiles``); Gaussians with radius <= 0 write nothing.
    """
    gx, gy = grid(W, H)
    uv = np.asarray(uv, np.float32)
    r = np.asarray(radius).reshape(-1).astype(np.int64)
    d = np.asarray(depth, np.float32).reshape(-1)
    x0, y0, x1, y1 = tile_rect(uv[:, 0], uv[:, 1], r, gx, gy, np.float32)
    live = r > 0
    area = np.where(live, (x1 - x0) * (y1 - y0), 0)
    dbits = d.view(np.int32).astype(np.int64)  # sign-extended
    area = area.astype(np.int64)
    k = int(area.sum())
    if k == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int32), area
    # Gaussian g owns slots [start_g, start_g + area_g), its tiles row-major
    ids = np.repeat(np.arange(len(area), dtype=np.int64), area)
    start = np.cumsum(area) - area
    t = np.arange(k, dtype=np.int64) - start[ids]
    w = np.maximum(x1 - x0, 1)[ids]
    tiles = (y0[ids] + t // w) * gx + (x0[ids] + t % w)
    keys = (tiles.astype(np.int64) << 32) | dbits[ids]
    return keys, ids.astype(np.int32), area


def sort_keys(keys, ids):
    """torch.sort + gather of sort_gaussian.py:49-50, stable (ties keep Gaussian order)."""
    perm = np.argsort(keys, kind="stable")
    return keys[perm], ids[perm]


def tile_range(key_sorted, W, H):
    """sort_gaussian.cu:45-71 restated without the out-of-bounds write for tile -1."""
    gx, gy = grid(W, H)
    T = gx * gy
    rng = np.zeros((T, 2), np.int32)
    k = len(key_sorted)
    if k == 0:
        return rng
    tiles = (key_sorted >> 32).astype(np.int64)
    starts = np.r_[0, np.nonzero(tiles[1:] != tiles[:-1])[0] + 1]
    ends = np.r_[starts[1:], k]
    for s, e in zip(starts, ends):
        t = tiles[s]
        if 0 <= t < T:
            rng[t] = (s, e)
    return rng


def sort_gaussian(uv, depth, radius, W, H):
    """Full sort stage: (idx_sorted [K] int32, tile_range [T,2] int32, key_sorted [K] int64)."""
    keys, ids, _ = gaussian_keys(uv, depth, radius, W, H)
    ks, idx = sort_keys(keys, ids)
    return idx, tile_range(ks, W, H), ks


def cell_lists(uv, depth, radius, W, H, cell):
    """Cell lists of the rasterizer (no reference counterpart; they replace the
    sorted tile lists inside rasterization, msplat/__init__.py:83-91).

    Every Gaussian with radius > 0, a non-negative depth bit pattern (sign bit
    clear: the reference key of a negative depth sorts into no tile) and a
    non-empty tile rectangle is listed in each cell of ``cell x cell`` tiles its
    rectangle touches, cells row-major, each cell's list ordered by (depth bits
    as unsigned, id).  Returns (rect [n,4] int64 = x0, y0, x1, y1; cell_ids
    [count] int32; cell_range [cells,2] int32).
    """
    gx, gy = grid(W, H)
    ncx, ncy = (gx + cell - 1) // cell, (gy + cell - 1) // cell
    uv = np.asarray(uv, np.float32)
    r = np.asarray(radius).reshape(-1).astype(np.int64)
    dbits = np.asarray(depth, np.float32).reshape(-1).view(np.uint32).astype(np.int64)
    x0, y0, x1, y1 = tile_rect(uv[:, 0], uv[:, 1], r, gx, gy, np.float32)
    live = (r > 0) & (dbits >> 31 == 0) & ((x1 - x0) * (y1 - y0) > 0)
    cx0, cy0 = x0 // cell, y0 // cell
    cx1, cy1 = (x1 + cell - 1) // cell, (y1 + cell - 1) // cell
    cells, ids = [], []
    for g in np.nonzero(live)[0]:
        for cy in range(cy0[g], cy1[g]):
            for cx in range(cx0[g], cx1[g]):
                cells.append(cy * ncx + cx)
                ids.append(g)
    cells = np.asarray(cells, np.int64)
    ids = np.asarray(ids, np.int64)
    order = np.lexsort((ids, dbits[ids] if len(ids) else ids, cells))
    cells, ids = cells[order], ids[order]
    rng = np.zeros((ncx * ncy, 2), np.int32)
    for c in np.unique(cells):
        w = np.nonzero(cells == c)[0]
        rng[c] = (w[0], w[-1] + 1)
    rect = np.stack([x0, y0, x1, y1], 1).astype(np.int64)
    rect[~live] = 0
    return rect, ids.astype(np.int32), rng


def band_lists(uv, depth, radius, W, H, row0, row1):
    """sort_gaussian (sort_gaussian.cu:17-71, sort_gaussian.py:42-52)
    restricted to the tiles of tile rows [row0, row1): every such tile gets
    the list (ids and order) the full sort gives it, every other tile an
    empty range.  Only the Gaussians whose tile box meets the band are keyed,
    so a band of a large frame costs a fraction of the full sort."""
    gx, gy = grid(W, H)
    uv = np.asarray(uv, np.float32)
    r = np.asarray(radius).reshape(-1).astype(np.int64)
    _, y0, _, y1 = tile_rect(uv[:, 0], uv[:, 1], r, gx, gy, np.float32)
    sub = np.nonzero((r > 0) & (y0 < row1) & (y1 > row0))[0]
    keys, ids, _ = gaussian_keys(uv[sub], np.asarray(depth, np.float32).reshape(-1)[sub], r[sub], W, H)
    tiles = keys >> 32  # -1 for negative depths (sign extension): in no band
    keep = (tiles >= row0 * gx) & (tiles < row1 * gx)
    ks, idx = sort_keys(keys[keep], sub[ids[keep]].astype(np.int32))  # sub ascending: ties keep Gaussian order
    return idx, tile_range(ks, W, H)
