"""Rasterization of SH colours, fused: the view-dependent colour of every
Gaussian (``compute_sh_points``: degree-L SH of the direction from the camera
centre), optional extra channels and optionally the camera-space depth, all
rendered in one pass.

Equals ``rasterization(xyz, ..., cat([compute_sh_points(shs, xyz, campos),
extra, project_point(xyz, ...)[1]], 1), ...)``, the composition a training
step of the reference builds from separate ops (msplat/compute_sh.py,
msplat/project_point.py, msplat/__init__.py:22-93), with three differences
in how it runs:

* the SH kernel writes its channels straight into the feature rows (no
  ``cat``) and evaluates only the Gaussians the projection kept (no others
  enter a cell list); it is queued behind the binning's depth sort, so the
  GPU runs it while the host waits for the entry count (beside the sort's
  latency-bound kernels, a bandwidth-bound kernel over every Gaussian slowed
  them more than it saved); the SH adjoint runs after the geometry adjoint
  and adds its xyz gradient into the geometry's (beside it, on a side
  stream, both contending for bandwidth, plus a separate addition, the pair
  took longer); only a campos gradient, the negated sum of the SH part
  alone, keeps the side-stream form;
* the depth channel is the fused preprocessing's own depth (no second
  projection), and its gradient enters the geometry adjoint directly;
* one autograd node instead of four.

The reference has no such op; the composition's results are reproduced:
images bit for bit, gradients to float summation order.
"""

from __future__ import annotations

import torch

from . import _lib, tuning
from .binning import CellLists, bin_gaussians, rasterize_backward, rasterize_forward
from .project_point import camera_grad_shape
from .rasterization import preprocess_backward, preprocess_forward

_side_streams: dict[int, torch.cuda.Stream] = {}


def _side_stream(dev: torch.device) -> torch.cuda.Stream:
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    if idx not in _side_streams:
        _side_streams[idx] = torch.cuda.Stream(device=idx)  # default priority
    return _side_streams[idx]


def rasterization_sh(xyz, scale, rotate, opacity, shs, intr, extr, campos, W: int, H: int, bg: float, extra=None,
                     depth_channel: bool = False, ndc=None):
    """Render ``[Csh + E + depth_channel, H, W]``: SH colours of shs [P, Csh, D]
    seen from campos [3], then extra [P, E] (optional), then the depth.
    Differentiable with respect to xyz, scale, rotate, opacity, shs, extra,
    intr, extr and campos."""
    return _RasterizeSH.apply(xyz, scale, rotate, opacity, shs, intr, extr, campos, int(W), int(H), float(bg),
                              extra, bool(depth_channel), ndc)


def rasterization_sh_l2(xyz, scale, rotate, opacity, shs, intr, extr, campos, W: int, H: int, bg: float, target,
                        extra=None, depth_channel: bool = False, ndc=None):
    """``(image, loss)``: ``rasterization_sh`` and ``l2_loss(image, target)`` in
    one op.  The loss is summed inside the blending forward and its gradient
    formed inside the blending adjoint (no loss or gradient image launches);
    ``image`` stays differentiable for other loss terms.  Results: the image
    as rasterization_sh's, the loss to fp32 rounding of the same sum,
    gradients equal to the composition's."""
    return _RasterizeSHL2.apply(xyz, scale, rotate, opacity, shs, intr, extr, campos, int(W), int(H), float(bg),
                                extra, bool(depth_channel), ndc, target)


def _forward(ctx, xyz, scale, rotate, opacity, shs, intr, extr, campos, W, H, bg, extra, depth_channel, ndc,
             target=None):
    s = _lib.f32("shs", shs)
    n, Csh, D = s.shape
    x = _lib.f32("xyz", xyz)
    cp = _lib.f32("campos", campos).reshape(3)
    E = 0 if extra is None else int(extra.shape[1])
    C = Csh + E + (1 if depth_channel else 0)
    dev = x.device
    feat = torch.empty((n, C), dtype=torch.float32, device=dev)
    cfg = tuning.config_for(W, H, C)
    vis = torch.empty(n, dtype=torch.uint8, device=dev)  # radius > 0, written by the preprocessing
    with torch.cuda.device(dev):
        uv, depth, conic, radius, _ = preprocess_forward(xyz, scale, rotate, intr, extr, W, H,
                                                         depth_col=feat[:, C - 1] if depth_channel else None,
                                                         visible=vis)
        if E:
            feat[:, Csh:Csh + E].copy_(extra)
        # SH colours into feature columns [0, Csh) for the projected Gaussians
        # only (the others are in no cell list: their rows get zeros and their
        # coefficients are not read), queued behind the binning's depth sort

        def sh_colours():
            _lib.call("msplat_compute_sh_points_forward", _lib.ptr(s), _lib.ptr(x), _lib.ptr(cp), _lib.ptr(vis), n,
                      Csh, D, _lib.ptr(feat), C, _lib.stream_ptr())

        bins = bin_gaussians(uv, depth, W, H, radius, cfg.cell, queued=sh_colours)
        res = rasterize_forward(uv, conic, opacity, feat, bins, bg, W, H, cfg.ppl_fwd, cfg.cull, target=target)
    out, final_T, ncontrib, walk = res[:4]
    ctx.W, ctx.H, ctx.bg, ctx.cell, ctx.ppl_bwd, ctx.cull = W, H, bg, cfg.cell, cfg.ppl_bwd, cfg.cull
    ctx.dims = (Csh, D, E, C, depth_channel)
    ctx.campos_shape = campos.shape
    ctx.has_ndc = ndc is not None
    ctx.l2 = target is not None
    l2_saved = (out, _lib.f32("target", target)) if ctx.l2 else ()
    ctx.save_for_backward(xyz, scale, rotate, opacity, s, intr, extr, cp, feat, uv, depth, conic, radius,
                          bins.rect, bins.cell_ids, bins.cell_range, final_T, ncontrib, walk, *l2_saved)
    return out if target is None else (out, res[4])


def _backward(ctx, dL_dout, dL_dloss=None):
    saved = ctx.saved_tensors
    (xyz, scale, rotate, opacity, s, intr, extr, cp, feat, uv, depth, conic, radius, rect, cell_ids, cell_range,
     final_T, ncontrib, walk) = saved[:19]
    l2 = (saved[19], saved[20], dL_dloss) if (ctx.l2 and dL_dloss is not None) else None
    Csh, D, E, C, depth_channel = ctx.dims
    n = xyz.shape[0]
    dev = xyz.device
    main = torch.cuda.current_stream(dev)
    side = _side_stream(dev)
    bins = CellLists(rect, cell_ids, cell_range, ctx.cell)
    if l2 is None and dL_dout is None:  # neither output reaches the loss
        dL_dout = torch.zeros((C, ctx.H, ctx.W), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        duv, dconic, dopac, dfeat = rasterize_backward(uv, conic, opacity, feat, bins, ctx.bg, ctx.W, ctx.H,
                                                       final_T, ncontrib, walk, dL_dout, ctx.ppl_bwd, ctx.cull,
                                                       l2=l2)
        x = _lib.f32("xyz", xyz)
        dshs = torch.empty((n, Csh, D), dtype=torch.float32, device=dev)
        if Csh > 0 and not ctx.needs_input_grad[7]:
            # the geometry adjoint, then the SH adjoint (feature columns [0, Csh))
            # adding its xyz gradient into the geometry's: no gradient array or
            # addition of its own
            dxyz, dscale, drot, dintr, dextr = preprocess_backward(
                xyz, scale, rotate, intr, extr, depth, radius, duv, dconic,
                dL_ddepth=dfeat[:, C - 1] if depth_channel else None,
                want_intr=intr.requires_grad, want_extr=extr.requires_grad)
            _lib.call("msplat_compute_sh_points_backward_add", _lib.ptr(s), _lib.ptr(x), _lib.ptr(cp), None,
                      _lib.ptr(dfeat), C, n, Csh, D, _lib.ptr(dshs), _lib.ptr(dxyz), main.cuda_stream)
            dcampos = None
        else:
            dxyz, dscale, drot, dintr, dextr, dcampos = _geometry_and_sh_adjoints(
                ctx, main, side, xyz, scale, rotate, intr, extr, depth, radius, duv, dconic, dfeat, s, x, cp, dshs)
    dextra = dfeat[:, Csh:Csh + E] if (E and ctx.needs_input_grad[11]) else None
    dndc = None
    if ctx.has_ndc:
        dndc = duv * torch.tensor([0.5 * ctx.W, 0.5 * ctx.H], dtype=duv.dtype, device=duv.device)[None]
    return (dxyz, dscale, drot, dopac.view_as(opacity), dshs, dintr,
            camera_grad_shape(extr, dextr) if dextr is not None else None, dcampos, None, None, None,
            dextra, None, dndc)


def _geometry_and_sh_adjoints(ctx, main, side, xyz, scale, rotate, intr, extr, depth, radius, duv, dconic, dfeat, s,
                              x, cp, dshs):
    """The SH adjoint on a side stream beside the geometry adjoint, with its own
    xyz gradient (the campos gradient is its negated sum)."""
    Csh, D, E, C, depth_channel = ctx.dims
    n = x.shape[0]
    dev = x.device
    dxyz_sh = torch.empty((n, 3), dtype=torch.float32, device=dev)
    ready = torch.cuda.Event()
    ready.record(main)
    side.wait_event(ready)
    with torch.cuda.stream(side):
        _lib.call("msplat_compute_sh_points_backward", _lib.ptr(s), _lib.ptr(x), _lib.ptr(cp), None,
                  _lib.ptr(dfeat), C, n, Csh, D, _lib.ptr(dshs), _lib.ptr(dxyz_sh), side.cuda_stream)
        sh_done = torch.cuda.Event()
        sh_done.record(side)
    for t in (s, x, cp, dfeat, dshs, dxyz_sh):
        t.record_stream(side)
    dxyz, dscale, drot, dintr, dextr = preprocess_backward(
        xyz, scale, rotate, intr, extr, depth, radius, duv, dconic,
        dL_ddepth=dfeat[:, C - 1] if depth_channel else None,
        want_intr=intr.requires_grad, want_extr=extr.requires_grad)
    main.wait_event(sh_done)
    dxyz.add_(dxyz_sh)
    dcampos = (-dxyz_sh.sum(0)).view(ctx.campos_shape) if ctx.needs_input_grad[7] else None
    return dxyz, dscale, drot, dintr, dextr, dcampos


class _RasterizeSH(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, scale, rotate, opacity, shs, intr, extr, campos, W, H, bg, extra, depth_channel, ndc):
        return _forward(ctx, xyz, scale, rotate, opacity, shs, intr, extr, campos, W, H, bg, extra, depth_channel,
                        ndc)

    @staticmethod
    def backward(ctx, dL_dout):
        return _backward(ctx, dL_dout)


class _RasterizeSHL2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, scale, rotate, opacity, shs, intr, extr, campos, W, H, bg, extra, depth_channel, ndc,
                target):
        ctx.set_materialize_grads(False)  # an unused image passes no gradient image
        out, loss = _forward(ctx, xyz, scale, rotate, opacity, shs, intr, extr, campos, W, H, bg, extra,
                             depth_channel, ndc, target)
        ctx.target_grad = ctx.needs_input_grad[14]
        return out, loss

    @staticmethod
    def backward(ctx, dL_dout, dL_dloss):
        grads = _backward(ctx, dL_dout, dL_dloss)
        dtarget = None
        if ctx.target_grad and dL_dloss is not None:  # d/dtarget of sum (out - t)^2 = -2 (out - t) dL
            out, t = ctx.saved_tensors[19], ctx.saved_tensors[20]
            dtarget = (t - out) * (2.0 * dL_dloss)
        return (*grads, dtarget)
