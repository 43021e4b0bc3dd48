"""EWA splatting: the drop-in for msplat.ewa_project (msplat/ewa_project.py:8-94)."""

from __future__ import annotations

import torch

from . import _lib
from .project_point import camera_grad_shape


def ewa_project(xyz, cov3d, intr, extr, uv, W: int, H: int, visible=None):
    """Project 3D covariances to screen conics.

    Returns ``(conic [P,3], radius [P] int32, tiles [P] int32)``: conic is the
    inverse of the 2D covariance (+0.3 on the diagonal), radius = ceil(3
    sqrt(lambda_max)), tiles = number of 16x16 tiles the radius box touches.
    """
    if visible is None:
        visible = torch.ones_like(uv[:, 0], dtype=torch.bool)
    return _EWAProject.apply(xyz, cov3d, intr, extr, uv, W, H, visible)


class _EWAProject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, cov3d, intr, extr, uv, W, H, visible):
        x = _lib.f32("xyz", xyz)
        n = x.shape[0]
        dev = x.device
        conic = torch.empty((n, 3), dtype=torch.float32, device=dev)
        radius = torch.empty(n, dtype=torch.int32, device=dev)
        tiles = torch.empty(n, dtype=torch.int32, device=dev)
        with torch.cuda.device(dev):
            _lib.call("msplat_ewa_project_forward", _lib.ptr(x), _lib.ptr(_lib.f32("cov3d", cov3d)),
                      _lib.ptr(_lib.f32("intr", intr)), _lib.ptr(_lib.f32("extr", extr)),
                      _lib.ptr(_lib.f32("uv", uv)), _lib.ptr(_lib.u8("visible", visible)), n, int(W), int(H),
                      _lib.ptr(conic), _lib.ptr(radius), _lib.ptr(tiles), _lib.stream_ptr())
        ctx.save_for_backward(xyz, cov3d, intr, extr, radius)
        ctx.mark_non_differentiable(radius, tiles)
        return conic, radius, tiles

    @staticmethod
    def backward(ctx, dL_dconic, dL_dradius, dL_dtiles):
        xyz, cov3d, intr, extr, radius = ctx.saved_tensors
        x = _lib.f32("xyz", xyz)
        n = x.shape[0]
        dev = x.device
        dxyz = torch.empty((n, 3), dtype=torch.float32, device=dev)
        dcov = torch.empty((n, 6), dtype=torch.float32, device=dev)
        dintr = torch.empty(4, dtype=torch.float32, device=dev) if intr.requires_grad else None
        dextr = torch.empty(12, dtype=torch.float32, device=dev) if extr.requires_grad else None
        with torch.cuda.device(dev):
            _lib.call("msplat_ewa_project_backward", _lib.ptr(x), _lib.ptr(_lib.f32("cov3d", cov3d)),
                      _lib.ptr(_lib.f32("intr", intr)), _lib.ptr(_lib.f32("extr", extr)),
                      _lib.ptr(_lib.i32("radius", radius)), _lib.ptr(_lib.f32("dL_dconic", dL_dconic)), n,
                      _lib.ptr(dxyz), _lib.ptr(dcov), _lib.ptr(dintr), _lib.ptr(dextr), _lib.stream_ptr())
        return (dxyz.to(xyz.dtype), dcov.to(cov3d.dtype), dintr,
                camera_grad_shape(extr, dextr) if dextr is not None else None, None, None, None, None)
