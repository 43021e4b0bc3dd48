"""Many cameras in one launch: the rasterizer for a batch of views.

``rasterization_views`` renders the same Gaussians through B <= 64 cameras of
(possibly) different image sizes with one launch per stage -- projection /
covariance / EWA for all views, one cell binning over every (view, Gaussian)
pair, one blending forward and one adjoint over the global tile space of all
views -- and differentiates with respect to the Gaussians and every camera.

Each view's image equals ``msplat_amd.rasterization`` of that camera (the
same device code walks the same per-view lists); gradients of the shared
parameters are the sums over views.  The reference renders one camera per
call (msplat/__init__.py:22-93); this module has no reference counterpart.
"""

from __future__ import annotations

import torch

from . import _lib, binning, tuning
from .binning import CellLists

MAX_VIEWS = 64


def _sizes(sizes):
    if not 1 <= len(sizes) <= MAX_VIEWS:
        raise RuntimeError(f"rasterization_views: {len(sizes)} views (1..{MAX_VIEWS} supported)")
    W = (_lib.I32 * len(sizes))(*[int(w) for w, _ in sizes])
    H = (_lib.I32 * len(sizes))(*[int(h) for _, h in sizes])
    return W, H


def _extr12(extrs):
    e = _lib.f32("extrs", extrs)
    B = e.shape[0]
    return e.reshape(B, -1)[:, :12].contiguous() if e.reshape(B, -1).shape[1] != 12 else e.reshape(B, 12)


def _ranges(first, count, B):
    """Host arrays of per-view Gaussian ranges (None, None = all views draw all Gaussians)."""
    if first is None and count is None:
        return None, None
    if len(first) != B or len(count) != B:
        raise RuntimeError("rasterization_views: one (first, count) range per view")
    return (_lib.I64 * B)(*[int(f) for f in first]), (_lib.I64 * B)(*[int(c) for c in count])


def rows_of(n: int, B: int, count=None) -> int:
    return B * n if count is None else int(sum(int(c) for c in count))


def tiles_of(sizes) -> int:
    return sum(((w + 15) // 16) * ((h + 15) // 16) for w, h in sizes)


def cells_of(sizes, cell: int) -> int:
    return sum((((w + 15) // 16 + cell - 1) // cell) * (((h + 15) // 16 + cell - 1) // cell) for w, h in sizes)


def preprocess_views_forward(xyz, scale, rotate, intrs, extrs, sizes, nearest: float = 0.0, extent: float = 1.3,
                             first=None, count=None):
    """Per-view ``(uv [R,2], depth [R,1], conic [R,3], radius [R], tiles [R])``, R rows (B*n by default)."""
    x = _lib.f32("xyz", xyz)
    n, B = x.shape[0], len(sizes)
    dev = x.device
    W, H = _sizes(sizes)
    F, N = _ranges(first, count, B)
    R = rows_of(n, B, count)
    uv = torch.empty((R, 2), dtype=torch.float32, device=dev)
    depth = torch.empty((R, 1), dtype=torch.float32, device=dev)
    conic = torch.empty((R, 3), dtype=torch.float32, device=dev)
    radius = torch.empty(R, dtype=torch.int32, device=dev)
    tiles = torch.empty(R, dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        _lib.call("msplat_preprocess_views_forward", _lib.ptr(x), _lib.ptr(_lib.f32("scale", scale)),
                  _lib.ptr(_lib.f32("rotate", rotate)), n, B, _lib.ptr(_lib.f32("intrs", intrs).reshape(B, 4)),
                  _lib.ptr(_extr12(extrs)), W, H, F, N, float(nearest), float(extent), _lib.ptr(uv),
                  _lib.ptr(depth), _lib.ptr(conic), _lib.ptr(radius), _lib.ptr(tiles), _lib.stream_ptr())
    return uv, depth, conic, radius, tiles


def preprocess_views_backward(xyz, scale, rotate, intrs, extrs, sizes, depth, radius, dL_duv, dL_dconic,
                              want_cam: bool = False):
    """``(dxyz, dscale, drotate, dintrs [B,4] | None, dextrs [B,12] | None)`` summed over views."""
    x = _lib.f32("xyz", xyz)
    n, B = x.shape[0], len(sizes)
    dev = x.device
    dxyz = torch.empty((n, 3), dtype=torch.float32, device=dev)
    dscale = torch.empty((n, 3), dtype=torch.float32, device=dev)
    drot = torch.empty((n, 4), dtype=torch.float32, device=dev)
    dintr = torch.empty((B, 4), dtype=torch.float32, device=dev) if want_cam else None
    dextr = torch.empty((B, 12), dtype=torch.float32, device=dev) if want_cam else None
    with torch.cuda.device(dev):
        _lib.call("msplat_preprocess_views_backward", _lib.ptr(x), _lib.ptr(_lib.f32("scale", scale)),
                  _lib.ptr(_lib.f32("rotate", rotate)), n, B, _lib.ptr(_lib.f32("intrs", intrs).reshape(B, 4)),
                  _lib.ptr(_extr12(extrs)), _lib.ptr(_lib.f32("depth", depth)), _lib.ptr(_lib.i32("radius", radius)),
                  _lib.ptr(_lib.f32("dL_duv", dL_duv)), None, _lib.ptr(_lib.f32("dL_dconic", dL_dconic)),
                  _lib.ptr(dxyz), _lib.ptr(dscale), _lib.ptr(drot), _lib.ptr(dintr), _lib.ptr(dextr),
                  _lib.stream_ptr())
    return dxyz, dscale, drot, dintr, dextr


def bin_views(uv, depth, radius, n: int, sizes, cell: int = 8, capacity: int | None = None, first=None,
              count=None) -> CellLists:
    """Cell lists over the global cell space of all views (entries are rows b*n + g).

    With ``capacity`` the entry count stays on the device (no host
    synchronisation, so the call can be captured in a graph)."""
    u = _lib.f32("uv", uv)
    d = _lib.f32("depth", depth).view(-1)
    r = _lib.i32("radius", radius).view(-1)
    B = len(sizes)
    dev = u.device
    W, H = _sizes(sizes)
    F, N = _ranges(first, count, B)
    R = rows_of(n, B, count)
    capacity = binning._planned(capacity)
    with torch.cuda.device(dev):
        stream = _lib.stream_ptr()
        rect = torch.empty((R, 2), dtype=torch.int32, device=dev)
        prep_bytes = int(_lib.query("msplat_bin_views_prepare_workspace_bytes", R))
        prep = torch.empty(max(prep_bytes, 1), dtype=torch.uint8, device=dev)
        cnt = torch.empty(2, dtype=torch.int64, device=dev)
        k = binning.prepare_views(u, d, r, n, B, W, H, F, N, cell, rect, prep, prep_bytes, cnt, capacity)
        binning._observe(cnt, capacity, k)
        cell_ids = torch.empty(k, dtype=torch.int32, device=dev)
        cell_range = torch.empty((cells_of(sizes, cell), 2), dtype=torch.int32, device=dev)
        emit_bytes = int(_lib.query("msplat_bin_views_emit_workspace_bytes", k))
        ws = torch.empty(max(emit_bytes, 1), dtype=torch.uint8, device=dev)
        _lib.call("msplat_bin_views_emit", _lib.ptr(u), _lib.ptr(d), _lib.ptr(r), n, B, W, H, F, N, int(cell),
                  _lib.ptr(prep), k, None if capacity is None else _lib.ptr(cnt), _lib.ptr(ws), emit_bytes,
                  _lib.ptr(cell_ids), _lib.ptr(cell_range), stream)
    return CellLists(rect, cell_ids, cell_range, int(cell), cnt)


def rasterize_views_forward(uv, conic, opacity, feature, n: int, sizes, bins: CellLists, bg: float,
                            ppl: int = 0, cull: bool = True, first=None, count=None):
    """Concatenated ``(rendered, final_T, ncontrib, walk)`` of all views."""
    f = _lib.f32("feature", feature)
    B = len(sizes)
    per_view = f.dim() == 3
    C = f.shape[-1]
    dev = f.device
    W, H = _sizes(sizes)
    F, N = _ranges(first, count, B)
    P = sum(w * h for w, h in sizes)
    out = torch.empty(C * P, dtype=torch.float32, device=dev)
    final_T = torch.empty(P, dtype=torch.float32, device=dev)
    ncontrib = torch.empty(P, dtype=torch.int32, device=dev)
    walk = torch.empty((tiles_of(sizes), 2), dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        _lib.call("msplat_rasterize_views_forward", _lib.ptr(_lib.f32("uv", uv)), _lib.ptr(_lib.f32("conic", conic)),
                  _lib.ptr(_lib.f32("opacity", opacity)), _lib.ptr(f), n, C, int(per_view), B, W, H, F, N,
                  _lib.ptr(bins.rect), _lib.ptr(bins.cell_ids), _lib.ptr(bins.cell_range), bins.cell, float(bg),
                  int(ppl), int(bool(cull)), _lib.ptr(out), _lib.ptr(final_T), _lib.ptr(ncontrib), _lib.ptr(walk),
                  _lib.stream_ptr())
    return out, final_T, ncontrib, walk


def rasterize_views_backward(uv, conic, opacity, feature, n: int, sizes, bins: CellLists, bg: float, final_T,
                             ncontrib, walk, dL_drendered, ppl: int = 0, cull: bool = True):
    """``(dL_duv [B*n,2], dL_dconic [B*n,3], dL_dopacity [n,1], dL_dfeature like feature)``."""
    f = _lib.f32("feature", feature)
    B = len(sizes)
    per_view = f.dim() == 3
    C = f.shape[-1]
    dev = f.device
    W, H = _sizes(sizes)
    duv = torch.empty((B * n, 2), dtype=torch.float32, device=dev)
    dconic = torch.empty((B * n, 3), dtype=torch.float32, device=dev)
    dopac = torch.empty((n, 1), dtype=torch.float32, device=dev)
    dfeat = torch.empty(f.shape, dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        _lib.call("msplat_rasterize_views_backward", _lib.ptr(_lib.f32("uv", uv)),
                  _lib.ptr(_lib.f32("conic", conic)), _lib.ptr(_lib.f32("opacity", opacity)), _lib.ptr(f), n, C,
                  int(per_view), B, W, H, _lib.ptr(bins.rect), _lib.ptr(bins.cell_ids), _lib.ptr(bins.cell_range),
                  bins.cell, float(bg), int(ppl), int(bool(cull)), _lib.ptr(final_T), _lib.ptr(ncontrib),
                  _lib.ptr(walk), _lib.ptr(_lib.f32("dL_drendered", dL_drendered)), _lib.ptr(duv),
                  _lib.ptr(dconic), _lib.ptr(dopac), _lib.ptr(dfeat), _lib.stream_ptr())
    return duv, dconic, dopac, dfeat


def split_views(flat, C: int, sizes):
    """Concatenated [C*P] buffer -> list of [C, H, W] views."""
    parts = torch.split(flat, [C * w * h for w, h in sizes])
    return [p.view(C, h, w) for p, (w, h) in zip(parts, sizes)]


def rasterization_views(xyz, scale, rotate, opacity, feature, intrs, extrs, sizes, bg: float,
                        views_per_launch: int | None = None):
    """Render B cameras in one launch per stage; returns a list of ``[C, H_b, W_b]`` images.

    feature: [P, C] shared by the views, or [B, P, C] per view (e.g. SH colours
    evaluated per camera).  intrs [B, 4]; extrs [B, 3, 4] or [B, 4, 4];
    sizes: B pairs (W, H).  Differentiable with respect to every tensor input.
    ``views_per_launch`` (default: the autotuner's choice for these sizes, else
    all) splits the batch into consecutive groups of that many views, one
    launch per stage each; images do not depend on the grouping.
    """
    sizes = [(int(w), int(h)) for w, h in sizes]
    C = feature.shape[-1]
    B = len(sizes)
    v = views_per_launch if views_per_launch is not None else tuning.views_per_launch_for(sizes, C)
    if v is None or v >= B:
        flat = _RasterizeViews.apply(xyz, scale, rotate, opacity, feature, intrs, extrs, tuple(sizes), bg)
        return split_views(flat, C, sizes)
    if v < 1:
        raise RuntimeError(f"rasterization_views: views_per_launch = {v}")
    out = []
    for b0 in range(0, B, v):
        sl = slice(b0, min(B, b0 + v))
        feat = feature[sl] if feature.dim() == 3 else feature
        flat = _RasterizeViews.apply(xyz, scale, rotate, opacity, feat, intrs[sl], extrs[sl], tuple(sizes[sl]), bg)
        out += split_views(flat, C, sizes[sl])
    return out


class _RasterizeViews(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, scale, rotate, opacity, feature, intrs, extrs, sizes, bg):
        n = xyz.shape[0]
        cfg = tuning.config_for_views(sizes, feature.shape[-1])
        uv, depth, conic, radius, _ = preprocess_views_forward(xyz, scale, rotate, intrs, extrs, sizes)
        bins = bin_views(uv, depth, radius, n, sizes, cfg.cell)
        feat = feature.reshape(len(sizes), n, -1) if feature.dim() == 3 else feature
        out, final_T, ncontrib, walk = rasterize_views_forward(uv, conic, opacity, feat, n, sizes, bins, bg,
                                                               cfg.ppl_fwd, cfg.cull)
        ctx.sizes, ctx.bg, ctx.cell, ctx.ppl_bwd, ctx.cull = sizes, bg, cfg.cell, cfg.ppl_bwd, cfg.cull
        ctx.feature_shape = feature.shape
        ctx.save_for_backward(xyz, scale, rotate, opacity, feat, intrs, extrs, uv, depth, conic, radius, bins.rect,
                              bins.cell_ids, bins.cell_range, final_T, ncontrib, walk)
        return out

    @staticmethod
    def backward(ctx, dL_dout):
        (xyz, scale, rotate, opacity, feat, intrs, extrs, uv, depth, conic, radius, rect, cell_ids, cell_range,
         final_T, ncontrib, walk) = ctx.saved_tensors
        n = xyz.shape[0]
        bins = CellLists(rect, cell_ids, cell_range, ctx.cell)
        duv, dconic, dopac, dfeat = rasterize_views_backward(uv, conic, opacity, feat, n, ctx.sizes, bins, ctx.bg,
                                                             final_T, ncontrib, walk, dL_dout.contiguous(),
                                                             ctx.ppl_bwd, ctx.cull)
        want_cam = intrs.requires_grad or extrs.requires_grad
        dxyz, dscale, drot, dintr, dextr = preprocess_views_backward(xyz, scale, rotate, intrs, extrs, ctx.sizes,
                                                                     depth, radius, duv, dconic, want_cam)
        dextr_full = None
        if dextr is not None and extrs.requires_grad:
            B = len(ctx.sizes)
            g = torch.zeros((B, extrs[0].numel()), dtype=torch.float32, device=dextr.device)
            g[:, :12] = dextr
            dextr_full = g.view(extrs.shape).to(extrs.dtype)
        return (dxyz, dscale, drot, dopac.view_as(opacity), dfeat.view(ctx.feature_shape),
                dintr.view_as(intrs) if (dintr is not None and intrs.requires_grad) else None, dextr_full, None,
                None)
