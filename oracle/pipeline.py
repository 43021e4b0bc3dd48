"""The whole rasterization chain on the CPU (test infrastructure only).

msplat/__init__.py:71-91 composed from the oracle stages, forward and the
adjoint chain autograd would run (blend -> EWA -> covariance, plus the
projection branch with dL/ddepth = 0).
"""

from __future__ import annotations

import numpy as np

from . import blend, geometry, sort


def stage_geometry(xyz, scale, rotate, intr, extr, W, H, dtype=np.float32):
    """Projection, covariance and EWA (msplat/__init__.py:71-82) with the fp32
    discrete decisions (culling, radius, tile boxes) and, for dtype=np.float64,
    the continuous quantities recomputed in fp64.  Returns a dict of uv,
    depth, visible, cov3d, conic, radius, tiles."""
    uv, depth = geometry.project_forward(xyz, intr, extr, W, H, dtype=np.float32)
    visible = depth != 0
    cov3d = geometry.cov3d_forward(scale, rotate, visible, dtype=np.float32)
    conic, radius, tiles = geometry.ewa_forward(xyz, cov3d, intr, extr, uv, W, H, visible, dtype=np.float32)
    out = dict(uv=uv, depth=depth, visible=visible, cov3d=cov3d, conic=conic, radius=radius, tiles=tiles)
    if np.dtype(dtype) != np.float32:
        uv64, d64 = geometry.project_forward(xyz, intr, extr, W, H, nearest=0.0, extent=0.0, dtype=dtype)
        out["uv"] = np.where(visible[:, None], uv64, 0.0)
        out["depth"] = np.where(visible, d64, 0.0)
        out["cov3d"] = geometry.cov3d_forward(scale, rotate, visible, dtype=dtype)
        craw = geometry.ewa_forward(xyz, out["cov3d"], intr, extr, out["uv"], W, H, visible, dtype=dtype, raw=True)
        out["conic"] = np.where((radius > 0)[:, None], craw, 0.0)
        out["uv32"], out["depth32"] = uv, depth
    return out


def stage_geometry_backward(xyz, scale, rotate, intr, extr, geo, duv, dconic, ddepth=None, dtype=np.float32):
    """Adjoint of stage_geometry: (dxyz, dscale, drotate, dintr, dextr), the
    chain autograd runs through ewa_project, compute_cov3d and project_point
    (ddepth: the depth's own gradient, e.g. from a depth feature channel)."""
    dxyz_e, dcov, dintr_e, dextr_e = geometry.ewa_backward(xyz, geo["cov3d"], intr, extr, geo["radius"], dconic,
                                                           dtype=dtype)
    dscale, drot = geometry.cov3d_backward(scale, rotate, dcov, geo["visible"], dtype=dtype)
    dxyz_p, dintr_p, dextr_p = geometry.project_backward(xyz, intr, extr, geo["depth"], duv, ddepth, dtype=dtype)
    return dxyz_p + dxyz_e, dscale, drot, dintr_p + dintr_e, dextr_p + dextr_e


def rasterize_forward(xyz, scale, rotate, opacity, feature, intr, extr, W, H, bg, dtype=np.float32):
    """Returns the rendered image [C,H,W] and a dict of intermediates.

    The discrete decisions (culling, radius, tile boxes, keys) are always the
    fp32 ones, as the GPU takes them; with dtype=np.float64 the continuous
    quantities (uv, depth, covariances, conics, blending) are then recomputed
    in fp64, so the comparison measures arithmetic error only.
    """
    uv, depth = geometry.project_forward(xyz, intr, extr, W, H, dtype=np.float32)
    visible = depth != 0
    cov3d = geometry.cov3d_forward(scale, rotate, visible, dtype=np.float32)
    conic, radius, tiles = geometry.ewa_forward(xyz, cov3d, intr, extr, uv, W, H, visible, dtype=np.float32)
    idx, rng, keys = sort.sort_gaussian(uv, depth, radius, W, H)
    if np.dtype(dtype) != np.float32:
        uv64, d64 = geometry.project_forward(xyz, intr, extr, W, H, nearest=0.0, extent=0.0, dtype=dtype)
        uv = np.where(visible[:, None], uv64, 0.0)
        depth = np.where(visible, d64, 0.0)
        cov3d = geometry.cov3d_forward(scale, rotate, visible, dtype=dtype)
        craw = geometry.ewa_forward(xyz, cov3d, intr, extr, uv, W, H, visible, dtype=dtype, raw=True)
        conic = np.where((radius > 0)[:, None], craw, 0.0)
    out, fT, nc, evaluated = blend.blend_forward(uv, conic, opacity, feature, idx, rng, bg, W, H, dtype=dtype)
    inter = dict(uv=uv, depth=depth, visible=visible, cov3d=cov3d, conic=conic, radius=radius, tiles=tiles,
                 idx_sorted=idx, tile_range=rng, key_sorted=keys, final_T=fT, ncontrib=nc, evaluated=evaluated)
    return out, inter


def rasterize_backward(xyz, scale, rotate, opacity, feature, intr, extr, W, H, bg, inter, dout, dtype=np.float32):
    """Adjoint of rasterize_forward: dict of dxyz, dscale, drotate, dopacity, dfeature, dintr, dextr, duv."""
    duv, dconic, dopac, dfeat = blend.blend_backward(inter["uv"], inter["conic"], opacity, feature,
                                                     inter["idx_sorted"], inter["tile_range"], bg, W, H,
                                                     inter["final_T"], inter["ncontrib"], dout, dtype=dtype)
    dxyz_e, dcov, dintr_e, dextr_e = geometry.ewa_backward(xyz, inter["cov3d"], intr, extr, inter["radius"], dconic,
                                                           dtype=dtype)
    dscale, drot = geometry.cov3d_backward(scale, rotate, dcov, inter["visible"], dtype=dtype)
    dxyz_p, dintr_p, dextr_p = geometry.project_backward(xyz, intr, extr, inter["depth"], duv, None, dtype=dtype)
    return dict(dxyz=dxyz_p + dxyz_e, dscale=dscale, drotate=drot, dopacity=dopac, dfeature=dfeat,
                dintr=dintr_p + dintr_e, dextr=dextr_p + dextr_e, duv=duv, dconic=dconic)
