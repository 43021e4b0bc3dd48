"""Expired look-back waits never pass silently.

Every look-back wait of the sorts and the cell binning is bounded; one that
expires leaves the lists undefined.  msplat_debug_force_expiry makes every
wait give up at once (without reading its predecessors), so these tests drive
each detection path while every kernel still runs to its end and writes only
inside its outputs:
  * an expiry in the counting pass makes the host's count read raise;
  * one in the depth sort or the emission (after the count was read) makes
    the blending forward write NaN, leaves the gradients at zero, and makes
    the next binning raise; after that the device is clean again and results
    are bit-exact with the ones before;
  * the sort_gaussian API raises at either of its two host reads.
"""

import pytest
import torch

from msplat_amd import _lib, binning
from msplat_amd.binning import CellLists
from msplat_amd.rasterization import preprocess_forward
from msplat_amd.sort_gaussian import sort_gaussian_keys, tile_grid
from tests.helpers import small_scene

pytestmark = pytest.mark.gpu
DEV = "cuda"
W, H, CELL = 256, 192, 8


def _force(on):
    _lib.call("msplat_debug_force_expiry", int(on), _lib.stream_ptr())


@pytest.fixture(scope="module")
def scene():
    # 20 000 Gaussians: more than one workgroup in every pass, so forced waits expire
    g, cam = small_scene(21, n=20000, W=W, H=H, half_extent=2.0, C=3)
    d = {k: v.to(DEV) for k, v in g.items()}
    uv, depth, conic, radius, tiles = preprocess_forward(d["xyz"], d["scale"], d["rotate"], cam.intr.to(DEV),
                                                         cam.extr.to(DEV), W, H)
    binning.check_waits()  # start clean
    ref = binning.bin_gaussians(uv, depth, W, H, radius, CELL)
    out = binning.rasterize_forward(uv, conic, d["opacity"], d["feature"], ref, 0.1, W, H)
    return d, uv, depth, conic, radius, tiles, ref, out


def _assert_clean_again(scene):
    d, uv, depth, conic, radius, tiles, ref, ref_out = scene
    b = binning.bin_gaussians(uv, depth, W, H, radius, CELL)
    assert torch.equal(b.cell_ids, ref.cell_ids) and torch.equal(b.cell_range, ref.cell_range)
    out = binning.rasterize_forward(uv, conic, d["opacity"], d["feature"], b, 0.1, W, H)
    for a, r in zip(out, ref_out):
        assert torch.equal(a, r)
    binning.check_waits()


def test_expiry_in_counting_pass_raises_at_count_read(scene):
    d, uv, depth, conic, radius, tiles, ref, ref_out = scene
    print("forced expiry: counting pass", flush=True)
    _force(1)
    try:
        with pytest.raises(RuntimeError, match="look-back wait"):
            binning.bin_gaussians(uv, depth, W, H, radius, CELL)
    finally:
        _force(0)
    torch.cuda.synchronize()
    _assert_clean_again(scene)


def test_expiry_after_count_poisons_and_raises_next(scene):
    d, uv, depth, conic, radius, tiles, ref, ref_out = scene
    print("forced expiry: depth sort and emission", flush=True)
    n = uv.shape[0]
    rect = torch.empty((n, 2), dtype=torch.int32, device=DEV)
    prep_bytes = int(_lib.query("msplat_bin_prepare_workspace_bytes", n, W, H, CELL))
    prep = torch.empty(prep_bytes, dtype=torch.uint8, device=DEV)
    count = torch.empty(2, dtype=torch.int64, device=DEV)
    Wa, Ha = (_lib.I32 * 1)(W), (_lib.I32 * 1)(H)
    args = (_lib.ptr(uv), _lib.ptr(depth.reshape(-1)), _lib.ptr(radius), n, 1, Wa, Ha, None, None, CELL,
            _lib.ptr(rect), _lib.ptr(prep), prep_bytes, _lib.ptr(count), _lib.stream_ptr())
    _lib.call("msplat_bin_views_count", *args)  # not forced: the count is exact
    _force(1)
    try:
        _lib.call("msplat_bin_views_order", *args)
        k = int(count[0].item())
        assert k == ref.cell_ids.numel() and int(count[1].item()) == 0
        gx, gy = tile_grid(W, H)
        ncells = ((gx + CELL - 1) // CELL) * ((gy + CELL - 1) // CELL)
        cell_ids = torch.empty(k, dtype=torch.int32, device=DEV)
        cell_range = torch.empty((ncells, 2), dtype=torch.int32, device=DEV)
        emit_bytes = int(_lib.query("msplat_bin_emit_workspace_bytes", n, k, W, H, CELL))
        ws = torch.empty(max(emit_bytes, 1), dtype=torch.uint8, device=DEV)
        _lib.call("msplat_bin_emit", _lib.ptr(uv), _lib.ptr(depth.reshape(-1)), _lib.ptr(radius), n, W, H, CELL,
                  _lib.ptr(prep), k, None, _lib.ptr(ws), emit_bytes, _lib.ptr(cell_ids), _lib.ptr(cell_range),
                  _lib.stream_ptr())
    finally:
        _force(0)
    bins = CellLists(rect, cell_ids, cell_range, CELL)
    out, T, nc, walk = binning.rasterize_forward(uv, conic, d["opacity"], d["feature"], bins, 0.1, W, H)
    assert bool(torch.isnan(out).all()) and bool(torch.isnan(T).all()) and int(nc.abs().sum()) == 0
    gout = torch.ones_like(out)
    grads = binning.rasterize_backward(uv, conic, d["opacity"], d["feature"], bins, 0.1, W, H, T, nc, walk, gout)
    for g in grads:
        assert int(torch.count_nonzero(g)) == 0
    with pytest.raises(RuntimeError, match="look-back wait"):  # the next binning reads the device's wait state
        binning.bin_gaussians(uv, depth, W, H, radius, CELL)
    torch.cuda.synchronize()
    _assert_clean_again(scene)


def test_sort_gaussian_raises_at_either_read(scene):
    d, uv, depth, conic, radius, tiles, ref, ref_out = scene
    print("forced expiry: sort_gaussian", flush=True)
    good = sort_gaussian_keys(uv, depth, W, H, radius, tiles)
    _force(1)
    try:  # depth sort: the count read reports it
        with pytest.raises(RuntimeError, match="look-back wait"):
            sort_gaussian_keys(uv, depth, W, H, radius, tiles)
    finally:
        _force(0)
    torch.cuda.synchronize()
    # the class sort only: step 1 clean, step 2 forced, the second read raises
    n = uv.shape[0]
    prep_bytes = int(_lib.query("msplat_sort_prepare_workspace_bytes", n, W, H))
    prep = torch.empty(prep_bytes, dtype=torch.uint8, device=DEV)
    count = torch.empty(1, dtype=torch.int64, device=DEV)
    _lib.call("msplat_sort_prepare", _lib.ptr(uv), _lib.ptr(depth.reshape(-1)), _lib.ptr(radius), _lib.ptr(tiles),
              n, W, H, _lib.ptr(prep), prep_bytes, _lib.ptr(count), _lib.stream_ptr())
    k = int(count.item())
    assert k == good[0].numel() and k > 4096  # several class-sort tiles
    gx, gy = tile_grid(W, H)
    idx = torch.empty(k, dtype=torch.int32, device=DEV)
    rng = torch.empty((gx * gy, 2), dtype=torch.int32, device=DEV)
    keys = torch.empty(k, dtype=torch.int64, device=DEV)
    emit_bytes = int(_lib.query("msplat_sort_emit_workspace_bytes", n, k, W, H))
    ws = torch.empty(emit_bytes, dtype=torch.uint8, device=DEV)
    _force(1)
    try:
        _lib.call("msplat_sort_emit", _lib.ptr(uv), _lib.ptr(depth.reshape(-1)), _lib.ptr(radius), n, W, H,
                  _lib.ptr(prep), k, _lib.ptr(ws), emit_bytes, _lib.ptr(idx), _lib.ptr(rng), _lib.ptr(keys),
                  _lib.stream_ptr())
    finally:
        _force(0)
    with pytest.raises(RuntimeError, match="look-back wait"):
        binning.check_waits()
    binning.check_waits()  # cleared by the raise
    again = sort_gaussian_keys(uv, depth, W, H, radius, tiles)
    for a, b in zip(again, good):
        assert torch.equal(a, b)
