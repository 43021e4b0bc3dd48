"""Tile binning and depth sort: the drop-in for msplat.sort_gaussian (msplat/sort_gaussian.py:8-52).

Besides ``sort_gaussian`` this module exposes the reference's two step-API
kernels, ``compute_gaussian_key`` (msplat/src/sort_gaussian.cu:74-113) and
``compute_tile_gaussian_range`` (:115-142), for callers that drive the sort
themselves.
"""

from __future__ import annotations

import torch

from . import _lib, binning

TILE = 16


def tile_grid(W: int, H: int) -> tuple[int, int]:
    return (W + TILE - 1) // TILE, (H + TILE - 1) // TILE


def sort_gaussian(uv, depth, W: int, H: int, radius, tiles):
    """Return ``(idx_sorted [K] int32, tile_range [T,2] int32)``.

    Intersections are ordered by the reference's signed 64-bit key
    ``tile << 32 | depth bits`` (stable): per tile, front to back, ties by
    Gaussian index; Gaussians with negative depth go to "tile -1" and appear
    first in idx_sorted but in no tile range.  ``tiles`` must equal the
    number of tiles of each Gaussian's radius box (as ewa_project returns).
    """
    idx_sorted, tile_range, _ = sort_gaussian_keys(uv, depth, W, H, radius, tiles, want_keys=False)
    return idx_sorted, tile_range


def sort_gaussian_keys(uv, depth, W: int, H: int, radius, tiles=None, want_keys: bool = True):
    """sort_gaussian plus the sorted 64-bit keys (for checking and debugging)."""
    u = _lib.f32("uv", uv)
    d = _lib.f32("depth", depth).view(-1)
    r = _lib.i32("radius", radius).view(-1)
    t = _lib.i32("tiles", tiles).view(-1) if tiles is not None else None
    n = u.shape[0]
    dev = u.device
    gx, gy = tile_grid(int(W), int(H))
    with torch.cuda.device(dev):
        stream = _lib.stream_ptr()
        prep_bytes = int(_lib.query("msplat_sort_prepare_workspace_bytes", n, int(W), int(H)))
        prep = torch.empty(prep_bytes, dtype=torch.uint8, device=dev)
        count = torch.empty(1, dtype=torch.int64, device=dev)
        _lib.call("msplat_sort_prepare", _lib.ptr(u), _lib.ptr(d), _lib.ptr(r), _lib.ptr(t), n, int(W), int(H),
                  _lib.ptr(prep), prep_bytes, _lib.ptr(count), stream)
        k = int(count.item())  # the first host synchronisation of the sort
        if k == -2:
            binning.check_waits(dev)  # clears the device's wait state
            raise RuntimeError(binning.WAIT_EXPIRED)
        if k < 0:
            raise RuntimeError("sort_gaussian: tiles disagree with the tile boxes of (uv, radius)")
        idx_sorted = torch.empty(k, dtype=torch.int32, device=dev)
        tile_range = torch.empty((gx * gy, 2), dtype=torch.int32, device=dev)
        keys = torch.empty(k, dtype=torch.int64, device=dev) if want_keys else None
        emit_bytes = int(_lib.query("msplat_sort_emit_workspace_bytes", n, k, int(W), int(H)))
        ws = torch.empty(max(emit_bytes, 1), dtype=torch.uint8, device=dev)
        _lib.call("msplat_sort_emit", _lib.ptr(u), _lib.ptr(d), _lib.ptr(r), n, int(W), int(H), _lib.ptr(prep), k,
                  _lib.ptr(ws), emit_bytes, _lib.ptr(idx_sorted), _lib.ptr(tile_range), _lib.ptr(keys), stream)
        # the second (as the reference's two .item() calls): the class sort's
        # bounded waits must have ended in time, or the lists are undefined
        binning.check_waits(dev)
    return idx_sorted, tile_range, keys


def compute_gaussian_key(uv, depth, W: int, H: int, radius, tiles_cumsum):
    """Reference-order keys: ``(key [K] int64, idx [K] int32)`` from the inclusive cumsum of tiles."""
    u = _lib.f32("uv", uv)
    off = _lib.i32("tiles", tiles_cumsum).view(-1)
    n = u.shape[0]
    k = int(off[-1].item()) if n > 0 else 0
    keys = torch.empty(k, dtype=torch.int64, device=u.device)
    idx = torch.empty(k, dtype=torch.int32, device=u.device)
    with torch.cuda.device(u.device):
        _lib.call("msplat_gaussian_keys", _lib.ptr(u), _lib.ptr(_lib.f32("depth", depth).view(-1)),
                  _lib.ptr(_lib.i32("radius", radius).view(-1)), _lib.ptr(off), n, int(W), int(H), k,
                  _lib.ptr(keys), _lib.ptr(idx), _lib.stream_ptr())
    return keys, idx


def compute_tile_gaussian_range(W: int, H: int, tiles_cumsum, key_sorted):
    """Per-tile [start, end) of sorted keys; ``tiles_cumsum`` is unused beyond the reference signature."""
    del tiles_cumsum
    ks = key_sorted.contiguous()
    _lib.check_device("key_sorted", ks)
    gx, gy = tile_grid(int(W), int(H))
    tile_range = torch.empty((gx * gy, 2), dtype=torch.int32, device=ks.device)
    with torch.cuda.device(ks.device):
        _lib.call("msplat_tile_range_from_keys", _lib.ptr(ks), ks.numel(), int(W), int(H), _lib.ptr(tile_range),
                  _lib.stream_ptr())
    return tile_range
