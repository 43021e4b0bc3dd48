"""The full pipeline: the drop-in for msplat.rasterization (msplat/__init__.py:22-93).

The reference chains five autograd ops (project_point, compute_cov3d with
visible = depth != 0, ewa_project, sort_gaussian, alpha_blending).  Here the
three per-Gaussian stages run as one fused kernel in each direction, the
sort + blend pair runs over cell lists (msplat_amd.binning: no (tile,
Gaussian) pairs are materialised), and the whole chain is a single autograd
node with a hand-written adjoint.  Results equal the chain's: the fused
kernels share the per-stage device code and the cell walk visits each tile's
reference list in order.
"""

from __future__ import annotations

import torch

from . import _lib
from . import tuning
from .binning import CellLists, bin_gaussians, rasterize_backward, rasterize_forward
from .project_point import camera_grad_shape


def _column(t, n: int):
    """(pointer, stride) of a float32 device vector of n elements, or None when
    it is not one (a copy is then needed)."""
    if (t.dtype == torch.float32 and t.is_cuda and t.dim() in (1, 2) and t.shape[0] == n
            and (t.dim() == 1 or t.shape[1] == 1) and (n <= 1 or t.stride(0) >= 1)):
        return t.data_ptr(), max(int(t.stride(0)), 1)
    return None


def preprocess_forward(xyz, scale, rotate, intr, extr, W: int, H: int, nearest: float = 0.0,
                       extent: float = 1.3, depth_col=None, visible=None):
    """Fused projection + covariance + EWA: ``(uv, depth, conic, radius, tiles)``.
    ``depth_col`` (a float32 device column, e.g. a feature channel) also
    receives the depths; ``visible`` (a uint8 [n] device tensor) receives
    radius > 0."""
    x = _lib.f32("xyz", xyz)
    n = x.shape[0]
    dev = x.device
    uv = torch.empty((n, 2), dtype=torch.float32, device=dev)
    depth = torch.empty((n, 1), dtype=torch.float32, device=dev)
    conic = torch.empty((n, 3), dtype=torch.float32, device=dev)
    radius = torch.empty(n, dtype=torch.int32, device=dev)
    tiles = torch.empty(n, dtype=torch.int32, device=dev)
    col, ld = None, 1
    if depth_col is not None:
        c = _column(depth_col, n)
        if c is None:
            raise RuntimeError("preprocess_forward: depth_col must be a float32 device column of n elements")
        col, ld = c
    if visible is not None and (visible.dtype != torch.uint8 or not visible.is_cuda or visible.numel() != n
                                or not visible.is_contiguous()):
        raise RuntimeError("preprocess_forward: visible must be a contiguous uint8 device tensor of n elements")
    with torch.cuda.device(dev):
        _lib.call("msplat_preprocess_forward_col", _lib.ptr(x), _lib.ptr(_lib.f32("scale", scale)),
                  _lib.ptr(_lib.f32("rotate", rotate)), n, _lib.ptr(_lib.f32("intr", intr)),
                  _lib.ptr(_lib.f32("extr", extr)), int(W), int(H), float(nearest), float(extent), _lib.ptr(uv),
                  _lib.ptr(depth), _lib.ptr(conic), _lib.ptr(radius), _lib.ptr(tiles), col, ld, _lib.ptr(visible),
                  _lib.stream_ptr())
    return uv, depth, conic, radius, tiles


def preprocess_backward(xyz, scale, rotate, intr, extr, depth, radius, dL_duv, dL_dconic, dL_ddepth=None,
                        want_intr: bool = False, want_extr: bool = False):
    """Adjoint of preprocess_forward: ``(dxyz, dscale, drotate, dintr|None, dextr12|None)``."""
    x = _lib.f32("xyz", xyz)
    n = x.shape[0]
    dev = x.device
    dxyz = torch.empty((n, 3), dtype=torch.float32, device=dev)
    dscale = torch.empty((n, 3), dtype=torch.float32, device=dev)
    drot = torch.empty((n, 4), dtype=torch.float32, device=dev)
    dintr = torch.empty(4, dtype=torch.float32, device=dev) if want_intr else None
    dextr = torch.empty(12, dtype=torch.float32, device=dev) if want_extr else None
    g_ptr, g_ld = None, 1
    if dL_ddepth is not None:  # a strided column (a feature channel's gradient) is read in place
        c = _column(dL_ddepth.detach(), n)
        g_d = None
        if c is None:
            g_d = _lib.f32("dL_ddepth", dL_ddepth)
            c = (g_d.data_ptr(), 1)
        g_ptr, g_ld = c
    with torch.cuda.device(dev):
        _lib.call("msplat_preprocess_backward_ld", _lib.ptr(x), _lib.ptr(_lib.f32("scale", scale)),
                  _lib.ptr(_lib.f32("rotate", rotate)), n, _lib.ptr(_lib.f32("intr", intr)),
                  _lib.ptr(_lib.f32("extr", extr)), _lib.ptr(_lib.f32("depth", depth)),
                  _lib.ptr(_lib.i32("radius", radius)), _lib.ptr(_lib.f32("dL_duv", dL_duv)), g_ptr, g_ld,
                  _lib.ptr(_lib.f32("dL_dconic", dL_dconic)), _lib.ptr(dxyz), _lib.ptr(dscale), _lib.ptr(drot),
                  _lib.ptr(dintr), _lib.ptr(dextr), _lib.stream_ptr())
    return dxyz, dscale, drot, dintr, dextr


def rasterization(xyz, scale, rotate, opacity, feature, intr, extr, W: int, H: int, bg: float, ndc=None):
    """Render ``feature`` [P,C] of 3D Gaussians to a ``[C, H, W]`` image.

    xyz/scale [P,3], rotate [P,4] quaternions (w,x,y,z), opacity [P,1],
    intr = [fx, fy, cx, cy], extr = world-to-camera [R|t] ([3,4] or [4,4]).
    ``ndc`` [P,2] only collects ``dL/duv * (W/2, H/2)``.
    """
    return _Rasterize.apply(xyz, scale, rotate, opacity, feature, intr, extr, W, H, bg, ndc)


class _Rasterize(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, scale, rotate, opacity, feature, intr, extr, W, H, bg, ndc):
        cfg = tuning.config_for(W, H, feature.shape[1])
        uv, depth, conic, radius, tiles = preprocess_forward(xyz, scale, rotate, intr, extr, W, H)
        bins = bin_gaussians(uv, depth, W, H, radius, cfg.cell)
        out, final_T, ncontrib, walk = rasterize_forward(uv, conic, opacity, feature, bins, bg, W, H, cfg.ppl_fwd,
                                                         cfg.cull)
        ctx.W, ctx.H, ctx.bg, ctx.cell, ctx.ppl_bwd, ctx.cull = W, H, bg, cfg.cell, cfg.ppl_bwd, cfg.cull
        ctx.has_ndc = ndc is not None
        ctx.save_for_backward(xyz, scale, rotate, opacity, feature, intr, extr, uv, depth, conic, radius,
                              bins.rect, bins.cell_ids, bins.cell_range, final_T, ncontrib, walk)
        return out

    @staticmethod
    def backward(ctx, dL_dout):
        (xyz, scale, rotate, opacity, feature, intr, extr, uv, depth, conic, radius, rect, cell_ids, cell_range,
         final_T, ncontrib, walk) = ctx.saved_tensors
        bins = CellLists(rect, cell_ids, cell_range, ctx.cell)
        duv, dconic, dopac, dfeat = rasterize_backward(uv, conic, opacity, feature, bins, ctx.bg, ctx.W, ctx.H,
                                                       final_T, ncontrib, walk, dL_dout, ctx.ppl_bwd, ctx.cull)
        need_geom = ctx.needs_input_grad[0] or ctx.needs_input_grad[1] or ctx.needs_input_grad[2] \
            or intr.requires_grad or extr.requires_grad
        dxyz = dscale = drot = dintr = dextr = None
        if need_geom:
            dxyz, dscale, drot, dintr, dextr = preprocess_backward(
                xyz, scale, rotate, intr, extr, depth, radius, duv, dconic,
                want_intr=intr.requires_grad, want_extr=extr.requires_grad)
        dndc = None
        if ctx.has_ndc:
            dndc = duv * torch.tensor([0.5 * ctx.W, 0.5 * ctx.H], dtype=duv.dtype, device=duv.device)[None]
        return (dxyz, dscale, drot, dopac.view_as(opacity), dfeat,
                dintr, camera_grad_shape(extr, dextr) if dextr is not None else None,
                None, None, None, dndc)
