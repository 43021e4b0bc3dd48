"""Tile-based alpha blending: the drop-in for msplat.alpha_blending (msplat/alpha_blending.py:7-135)."""

from __future__ import annotations

import torch

from . import _lib


def blend_forward(uv, conic, opacity, feature, idx_sorted, tile_range, bg: float, W: int, H: int):
    """Raw forward: ``(rendered [C,H,W], final_T [H,W], ncontrib [H,W] int32)``."""
    u = _lib.f32("uv", uv)
    f = _lib.f32("feature", feature)
    n, C = f.shape
    dev = u.device
    out = torch.empty((C, H, W), dtype=torch.float32, device=dev)
    final_T = torch.empty((H, W), dtype=torch.float32, device=dev)
    ncontrib = torch.empty((H, W), dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        _lib.call("msplat_alpha_blending_forward", _lib.ptr(u), _lib.ptr(_lib.f32("conic", conic)),
                  _lib.ptr(_lib.f32("opacity", opacity)), _lib.ptr(f), n, C,
                  _lib.ptr(_lib.i32("idx_sorted", idx_sorted)), _lib.ptr(_lib.i32("tile_range", tile_range)),
                  float(bg), int(W), int(H), _lib.ptr(out), _lib.ptr(final_T), _lib.ptr(ncontrib),
                  _lib.stream_ptr())
    return out, final_T, ncontrib


def blend_backward(uv, conic, opacity, feature, idx_sorted, tile_range, bg: float, W: int, H: int, final_T,
                   ncontrib, dL_drendered):
    """Raw adjoint: ``(dL_duv [P,2], dL_dconic [P,3], dL_dopacity [P,1], dL_dfeature [P,C])``."""
    u = _lib.f32("uv", uv)
    f = _lib.f32("feature", feature)
    n, C = f.shape
    dev = u.device
    duv = torch.empty((n, 2), dtype=torch.float32, device=dev)
    dconic = torch.empty((n, 3), dtype=torch.float32, device=dev)
    dopac = torch.empty((n, 1), dtype=torch.float32, device=dev)
    dfeat = torch.empty((n, C), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        _lib.call("msplat_alpha_blending_backward", _lib.ptr(u), _lib.ptr(_lib.f32("conic", conic)),
                  _lib.ptr(_lib.f32("opacity", opacity)), _lib.ptr(f), n, C,
                  _lib.ptr(_lib.i32("idx_sorted", idx_sorted)), _lib.ptr(_lib.i32("tile_range", tile_range)),
                  float(bg), int(W), int(H), _lib.ptr(_lib.f32("final_T", final_T)),
                  _lib.ptr(_lib.i32("ncontrib", ncontrib)), _lib.ptr(_lib.f32("dL_drendered", dL_drendered)),
                  _lib.ptr(duv), _lib.ptr(dconic), _lib.ptr(dopac), _lib.ptr(dfeat), _lib.stream_ptr())
    return duv, dconic, dopac, dfeat


def alpha_blending(uv, conic, opacity, feature, idx_sorted, tile_range, bg: float, W: int, H: int, ndc=None):
    """Blend sorted 2D Gaussians front to back per 16x16 tile; returns ``[C, H, W]``.

    ``ndc`` only receives a gradient, ``dL/duv * (W/2, H/2)`` (used for
    densification statistics).
    """
    return _AlphaBlending.apply(uv, conic, opacity, feature, idx_sorted, tile_range, bg, W, H, ndc)


class _AlphaBlending(torch.autograd.Function):
    @staticmethod
    def forward(ctx, uv, conic, opacity, feature, idx_sorted, tile_range, bg, W, H, ndc):
        out, final_T, ncontrib = blend_forward(uv, conic, opacity, feature, idx_sorted, tile_range, bg, W, H)
        ctx.W, ctx.H, ctx.bg = W, H, bg
        ctx.has_ndc = ndc is not None
        ctx.save_for_backward(uv, conic, opacity, feature, idx_sorted, tile_range, final_T, ncontrib)
        return out

    @staticmethod
    def backward(ctx, dL_drendered):
        uv, conic, opacity, feature, idx_sorted, tile_range, final_T, ncontrib = ctx.saved_tensors
        duv, dconic, dopac, dfeat = blend_backward(uv, conic, opacity, feature, idx_sorted, tile_range, ctx.bg,
                                                   ctx.W, ctx.H, final_T, ncontrib, dL_drendered)
        dndc = None
        if ctx.has_ndc:
            dndc = duv * torch.tensor([0.5 * ctx.W, 0.5 * ctx.H], dtype=duv.dtype, device=duv.device)[None]
        return (duv.to(uv.dtype), dconic.to(conic.dtype), dopac.view_as(opacity).to(opacity.dtype),
                dfeat.to(feature.dtype), None, None, None, None, None, dndc)
