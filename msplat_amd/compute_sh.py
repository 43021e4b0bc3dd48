"""Spherical harmonics, degree 0..10: the drop-in for msplat.compute_sh (msplat/compute_sh.py:8-64)."""

from __future__ import annotations

import torch

from . import _lib


def compute_sh(shs, view_dirs, visible=None):
    """Evaluate SH: ``value[p, c] = sum_k Y_k(view_dirs[p]) * shs[p, c, k]``.

    shs is [P, C, D] with D = (deg+1)^2, deg <= 10; view_dirs [P, 3] are
    expected to be unit vectors (the basis is the reference's polynomial form,
    so non-unit inputs give the same values and gradients as the reference).
    No +0.5 offset and no clamp.  Invisible rows give 0 (``visible=None``:
    every row is visible).
    """
    return _ComputeSH.apply(shs, view_dirs, visible)


def compute_sh_points(shs, xyz, campos, visible=None):
    """SH of the view directions from one camera centre, fused into one kernel.

    Equals ``compute_sh(shs, (xyz - campos) / |xyz - campos|)``, the view
    direction computed in-kernel as ``v / sqrt((vx*vx + vy*vy) + vz*vz)``
    (the reference leaves the directions to user code: compute_sh.py:8-33).
    The adjoint returns gradients for shs, xyz (through the normalisation)
    and campos [3] (minus the sum of the xyz gradients).
    """
    return _ComputeSHPoints.apply(shs, xyz, campos, visible)


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


class _ComputeSHPoints(torch.autograd.Function):
    @staticmethod
    def forward(ctx, shs, xyz, campos, visible):
        s = _lib.f32("shs", shs)
        n, C, D = s.shape
        x = _lib.f32("xyz", xyz)
        if x.shape != (n, 3):
            raise RuntimeError(f"compute_sh_points: xyz must be [{n}, 3], got {tuple(x.shape)}")
        c = _lib.f32("campos", campos).reshape(3)
        value = torch.empty((n, C), dtype=torch.float32, device=s.device)
        with torch.cuda.device(s.device):
            _lib.call("msplat_compute_sh_points_forward", _lib.ptr(s), _lib.ptr(x), _lib.ptr(c),
                      _lib.ptr(_lib.u8("visible", visible)), n, C, D, _lib.ptr(value), C, _lib.stream_ptr())
        ctx.save_for_backward(shs, xyz, campos, visible)
        return value

    @staticmethod
    def backward(ctx, dL_dvalue):
        shs, xyz, campos, visible = ctx.saved_tensors
        s = _lib.f32("shs", shs)
        n, C, D = s.shape
        dshs = torch.empty((n, C, D), dtype=torch.float32, device=s.device)
        dxyz = torch.empty((n, 3), dtype=torch.float32, device=s.device)
        with torch.cuda.device(s.device):
            _lib.call("msplat_compute_sh_points_backward", _lib.ptr(s), _lib.ptr(_lib.f32("xyz", xyz)),
                      _lib.ptr(_lib.f32("campos", campos).reshape(3)), _lib.ptr(_lib.u8("visible", visible)),
                      _lib.ptr(_lib.f32("dL_dvalue", dL_dvalue)), C, n, C, D, _lib.ptr(dshs), _lib.ptr(dxyz),
                      _lib.stream_ptr())
        dcampos = None
        if ctx.needs_input_grad[2]:
            dcampos = (-dxyz.sum(0)).view(campos.shape).to(campos.dtype)
        return dshs.to(shs.dtype), dxyz.to(xyz.dtype), dcampos, None
