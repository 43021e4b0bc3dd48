"""Spherical harmonics, degree 0..10: the drop-in for msplat.compute_sh (msplat/compute_sh.py:8-64)."""

from __future__ import annotations

import torch

from . import _lib


def compute_sh(shs, view_dirs, visible=None):
    """Evaluate SH: ``value[p, c] = sum_k Y_k(view_dirs[p]) * shs[p, c, k]``.

    shs is [P, C, D] with D = (deg+1)^2, deg <= 10; view_dirs [P, 3] are
    expected to be unit vectors (the basis is the reference's polynomial form,
    so non-unit inputs give the same values and gradients as the reference).
    No +0.5 offset and no clamp.  Invisible rows give 0.
    """
    if visible is None:
        visible = torch.ones_like(shs[:, 0, 0], dtype=torch.bool)
    return _ComputeSH.apply(shs, view_dirs, visible)


class _ComputeSH(torch.autograd.Function):
    @staticmethod
    def forward(ctx, shs, view_dirs, visible):
        s = _lib.f32("shs", shs)
        n, C, D = s.shape
        value = torch.empty((n, C), dtype=torch.float32, device=s.device)
        with torch.cuda.device(s.device):
            _lib.call("msplat_compute_sh_forward", _lib.ptr(s), _lib.ptr(_lib.f32("view_dirs", view_dirs)),
                      _lib.ptr(_lib.u8("visible", visible)), n, C, D, _lib.ptr(value), _lib.stream_ptr())
        ctx.save_for_backward(shs, view_dirs, visible)
        return value

    @staticmethod
    def backward(ctx, dL_dvalue):
        shs, view_dirs, visible = ctx.saved_tensors
        s = _lib.f32("shs", shs)
        n, C, D = s.shape
        dshs = torch.empty((n, C, D), dtype=torch.float32, device=s.device)
        ddirs = torch.empty((n, 3), dtype=torch.float32, device=s.device)
        with torch.cuda.device(s.device):
            _lib.call("msplat_compute_sh_backward", _lib.ptr(s), _lib.ptr(_lib.f32("view_dirs", view_dirs)),
                      _lib.ptr(_lib.u8("visible", visible)), _lib.ptr(_lib.f32("dL_dvalue", dL_dvalue)), n, C,
                      D, _lib.ptr(dshs), _lib.ptr(ddirs), _lib.stream_ptr())
        return dshs.to(shs.dtype), ddirs.to(view_dirs.dtype), None
