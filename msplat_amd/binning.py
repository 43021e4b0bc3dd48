"""Cell binning and blending over cell lists: the rasterizer's replacement for
the sort_gaussian + alpha_blending pair inside msplat.rasterization
(msplat/__init__.py:83-91).

The reference materialises every (tile, Gaussian) pair of the tile
rectangles, sorts them by ``tile << 32 | depth bits`` and blends each tile's
list front to back.  Most of those pairs lie behind pixels that saturate
early and are never read.  Here Gaussians are binned into cells of
``cell x cell`` tiles, each cell's list in (depth bits, id) order, and a tile
walks its cell's list keeping the Gaussians whose tile rectangle covers it.
That filtered sequence is exactly the tile's reference list, so rendered
images, final_T and ncontrib are identical to sort_gaussian + alpha_blending
(and gradients equal up to float summation order).
"""

from __future__ import annotations

import dataclasses
import threading

import torch

from . import _lib
from .sort_gaussian import tile_grid


@dataclasses.dataclass
class CellLists:
    rect: torch.Tensor        # [n, 2] int32: packed tile rectangle per Gaussian
    cell_ids: torch.Tensor    # [capacity] int32: Gaussian ids, cell-major, (depth, id) order within a cell
    cell_range: torch.Tensor  # [cells, 2] int32: [start, end) per cell
    cell: int                 # cell edge in tiles
    count: torch.Tensor | None = None  # [2] int64 on the device: the true number of entries, wait state

    def overflowed(self) -> bool:
        """True if the entries did not fit the capacity (synchronises)."""
        return self.count is not None and int(self.count[0].item()) > self.cell_ids.numel()

    def check(self) -> None:
        """Raise RuntimeError if a bounded look-back wait of this binning (or of
        an earlier one on the device) expired, which leaves the lists undefined
        (synchronises; clears the device's wait state)."""
        check_waits(self.cell_range.device)


WAIT_EXPIRED = ("a bounded look-back wait of a sort or cell binning on this device expired: the lists it wrote "
                "were undefined and the images rendered from them were written as NaN")


def check_waits(device=None, clear: bool = True) -> None:
    """Raise RuntimeError if the device's expired-wait word is set (synchronises);
    the word is cleared first when ``clear``."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    with torch.cuda.device(dev):
        out = torch.empty(1, dtype=torch.int64, device=dev)
        _lib.call("msplat_wait_status", _lib.ptr(out), int(bool(clear)), _lib.stream_ptr())
        if int(out.item()) != 0:
            raise RuntimeError(WAIT_EXPIRED)


def _raise_expired() -> None:
    _lib.call("msplat_wait_status", None, 1, _lib.stream_ptr())  # cleared: the next binning starts afresh
    raise RuntimeError(WAIT_EXPIRED)


# Capacity plan of a graph capture in progress (msplat_amd.graphs): binning
# calls without an explicit capacity record their counts (warm-up) or take a
# preset capacity (capture).
_plan = None


def set_plan(plan):
    global _plan
    prev, _plan = _plan, plan
    return prev


def _planned(capacity):
    if capacity is not None or _plan is None:
        return capacity
    return _plan.capacity()


def _observe(count: torch.Tensor, capacity, k: int):
    if _plan is not None:
        _plan.observe(k if capacity is None else -1, count)


_host = threading.local()


def _host_count():
    """A pinned host word and an event per thread for the entry count's copy
    (kept: creating and destroying an event per call costs more than the wait)."""
    buf = getattr(_host, "count", None)
    if buf is None:
        buf = _host.count = (torch.empty(2, dtype=torch.int64, pin_memory=True), torch.cuda.Event())
    return buf


def prepare_views(u, d, r, n: int, B: int, W, H, F, N, cell: int, rect, prep, prep_bytes: int, count,
                  capacity, queued=None) -> int:
    """Step 1 of the binning on the current stream: the counting pass, then the
    depth order.  Without a capacity, the entry count is copied to the host
    between the two and read once the sort is queued behind it (the GPU sorts
    while the count travels); returns that count, or the capacity.  ``count``
    is a [2] int64 device tensor: the entries, then the device's wait state
    (non-zero: an earlier binning's look-back wait expired; raises).
    ``queued()``, when given, enqueues further independent work on the stream
    before the host waits: the GPU runs it after the sort instead of idling
    while the host issues the emission."""
    args = (_lib.ptr(u), _lib.ptr(d), _lib.ptr(r), n, B, W, H, F, N, int(cell), _lib.ptr(rect), _lib.ptr(prep),
            prep_bytes, _lib.ptr(count), _lib.stream_ptr())
    _lib.call("msplat_bin_views_count", *args)
    if capacity is None:
        host, ready = _host_count()
        host.copy_(count, non_blocking=True)
        ready.record()
    _lib.call("msplat_bin_views_order", *args)
    if queued is not None:
        queued()
    if capacity is not None:
        return int(capacity)
    # the only host synchronisation of the binning: poll the event (a blocking
    # event wait can sleep far longer than the copy takes when the GPU is idle)
    while not ready.query():
        pass
    k, expired = int(host[0]), int(host[1])
    if k < 0 or expired:
        _raise_expired()
    return k


def cell_grid(W: int, H: int, cell: int) -> tuple[int, int]:
    gx, gy = tile_grid(W, H)
    return (gx + cell - 1) // cell, (gy + cell - 1) // cell


def bin_gaussians(uv, depth, W: int, H: int, radius, cell: int = 8, capacity: int | None = None,
                  queued=None) -> CellLists:
    """Cell lists of one view.  With ``capacity`` the entry count stays on the
    device (no host synchronisation: the call can be captured in a graph);
    ``overflowed()`` tells afterwards whether the capacity was too small.
    ``queued``: see prepare_views."""
    u = _lib.f32("uv", uv)
    d = _lib.f32("depth", depth).view(-1)
    r = _lib.i32("radius", radius).view(-1)
    n = u.shape[0]
    dev = u.device
    W, H, cell = int(W), int(H), int(cell)
    capacity = _planned(capacity)
    ncx, ncy = cell_grid(W, H, cell)
    with torch.cuda.device(dev):
        stream = _lib.stream_ptr()
        rect = torch.empty((n, 2), dtype=torch.int32, device=dev)
        prep_bytes = int(_lib.query("msplat_bin_prepare_workspace_bytes", n, W, H, cell))
        prep = torch.empty(max(prep_bytes, 1), dtype=torch.uint8, device=dev)
        count = torch.empty(2, dtype=torch.int64, device=dev)
        Wa, Ha = (_lib.I32 * 1)(W), (_lib.I32 * 1)(H)
        k = prepare_views(u, d, r, n, 1, Wa, Ha, None, None, cell, rect, prep, prep_bytes, count, capacity, queued)
        _observe(count, capacity, k)
        cell_ids = torch.empty(k, dtype=torch.int32, device=dev)
        cell_range = torch.empty((ncx * ncy, 2), dtype=torch.int32, device=dev)
        emit_bytes = int(_lib.query("msplat_bin_emit_workspace_bytes", n, k, W, H, cell))
        ws = torch.empty(max(emit_bytes, 1), dtype=torch.uint8, device=dev)
        _lib.call("msplat_bin_emit", _lib.ptr(u), _lib.ptr(d), _lib.ptr(r), n, W, H, cell, _lib.ptr(prep), k,
                  None if capacity is None else _lib.ptr(count), _lib.ptr(ws), emit_bytes, _lib.ptr(cell_ids),
                  _lib.ptr(cell_range), stream)
    return CellLists(rect, cell_ids, cell_range, cell, count)


def rasterize_forward(uv, conic, opacity, feature, bins: CellLists, bg: float, W: int, H: int, ppl: int = 0,
                      cull: bool = True, target=None):
    """``(rendered [C,H,W], final_T [H,W], ncontrib [H,W] int32, walk [T,2] int32)``,
    and with ``target`` [C,H,W] also the fused loss ``sum((rendered - target)^2)``
    (a 0-d tensor) as a fifth element.

    ``cull`` skips, per tile, Gaussians that give no pixel of it alpha >= 1/255
    (rendered and final_T are unchanged; ncontrib counts culled-list
    positions).  The adjoint must be called with the same ``cull``."""
    u = _lib.f32("uv", uv)
    f = _lib.f32("feature", feature)
    n, C = f.shape
    dev = u.device
    W, H = int(W), int(H)
    gx, gy = tile_grid(W, H)
    out = torch.empty((C, H, W), dtype=torch.float32, device=dev)
    final_T = torch.empty((H, W), dtype=torch.float32, device=dev)
    ncontrib = torch.empty((H, W), dtype=torch.int32, device=dev)
    walk = torch.empty((gx * gy, 2), dtype=torch.int32, device=dev)
    common = (_lib.ptr(u), _lib.ptr(_lib.f32("conic", conic)), _lib.ptr(_lib.f32("opacity", opacity)), _lib.ptr(f),
              n, C, _lib.ptr(bins.rect), _lib.ptr(bins.cell_ids), _lib.ptr(bins.cell_range), bins.cell, float(bg),
              W, H, int(ppl), int(bool(cull)))
    with torch.cuda.device(dev):
        if target is None:
            _lib.call("msplat_rasterize_forward", *common, _lib.ptr(out), _lib.ptr(final_T), _lib.ptr(ncontrib),
                      _lib.ptr(walk), _lib.stream_ptr())
            return out, final_T, ncontrib, walk
        t = _lib.f32("target", target)
        if tuple(t.shape) != (C, H, W):
            raise RuntimeError(f"rasterize_forward: target shape {tuple(t.shape)} != {(C, H, W)}")
        nb = int(_lib.query("msplat_rasterize_l2_workspace_bytes", W, H, C))
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
        loss = torch.empty((), dtype=torch.float32, device=dev)
        _lib.call("msplat_rasterize_forward_l2", *common, _lib.ptr(t), _lib.ptr(ws), nb, _lib.ptr(out),
                  _lib.ptr(final_T), _lib.ptr(ncontrib), _lib.ptr(walk), _lib.ptr(loss), _lib.stream_ptr())
    return out, final_T, ncontrib, walk, loss


def rasterize_backward(uv, conic, opacity, feature, bins: CellLists, bg: float, W: int, H: int, final_T, ncontrib,
                       walk, dL_drendered, ppl: int = 0, cull: bool = True, l2=None):
    """``(dL_duv [P,2], dL_dconic [P,3], dL_dopacity [P,1], dL_dfeature [P,C])``.

    ``l2 = (rendered, target, dL_dloss)``: the fused loss's adjoint, where
    dL/drendered = 2 (rendered - target) dL_dloss, plus ``dL_drendered`` when
    that is not None."""
    u = _lib.f32("uv", uv)
    f = _lib.f32("feature", feature)
    n, C = f.shape
    dev = u.device
    duv = torch.empty((n, 2), dtype=torch.float32, device=dev)
    dconic = torch.empty((n, 3), dtype=torch.float32, device=dev)
    dopac = torch.empty((n, 1), dtype=torch.float32, device=dev)
    dfeat = torch.empty((n, C), dtype=torch.float32, device=dev)
    common = (_lib.ptr(u), _lib.ptr(_lib.f32("conic", conic)), _lib.ptr(_lib.f32("opacity", opacity)), _lib.ptr(f),
              n, C, _lib.ptr(bins.rect), _lib.ptr(bins.cell_ids), _lib.ptr(bins.cell_range), bins.cell, float(bg),
              int(W), int(H), int(ppl), int(bool(cull)), _lib.ptr(final_T), _lib.ptr(ncontrib), _lib.ptr(walk))
    g = None if dL_drendered is None else _lib.f32("dL_drendered", dL_drendered)
    with torch.cuda.device(dev):
        if l2 is None:
            _lib.call("msplat_rasterize_backward", *common, _lib.ptr(g), _lib.ptr(duv), _lib.ptr(dconic),
                      _lib.ptr(dopac), _lib.ptr(dfeat), _lib.stream_ptr())
        else:
            rendered, target, dL_dloss = l2
            _lib.call("msplat_rasterize_backward_l2", *common, _lib.ptr(g), _lib.ptr(_lib.f32("rendered", rendered)),
                      _lib.ptr(_lib.f32("target", target)), _lib.ptr(_lib.f32("dL_dloss", dL_dloss.reshape(1))),
                      _lib.ptr(duv), _lib.ptr(dconic), _lib.ptr(dopac), _lib.ptr(dfeat), _lib.stream_ptr())
    return duv, dconic, dopac, dfeat
