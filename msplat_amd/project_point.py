"""Pinhole projection: the drop-in for msplat.project_point (msplat/project_point.py:8-98)."""

from __future__ import annotations

import torch

from . import _lib


def camera_grad_shape(extr: torch.Tensor, flat12: torch.Tensor) -> torch.Tensor:
    """Return the 12 extrinsic partials shaped like ``extr``.

    The kernels read the first 12 floats of extr as a row-major [R|t]
    (msplat/src/project_point.cu:28-30), so a [4,4] extrinsic gets those
    partials in its first three rows and zeros in the last.  (The reference
    always returns [3,4], which autograd rejects for a [4,4] input.)
    """
    if extr.numel() == 12:
        return flat12.view(extr.shape).to(extr.dtype)
    g = torch.zeros(extr.numel(), dtype=torch.float32, device=flat12.device)
    g[:12] = flat12
    return g.view(extr.shape).to(extr.dtype)


def project_point(xyz, intr, extr, W: int, H: int, nearest: float = 0.0, extent: float = 1.3):
    """Project 3D points to pixels; returns ``(uv [P,2], depth [P,1])``.

    u = fx * tx / (tz + 1e-7) + cx - 0.5 (same for v); points with
    ``nearest > 0 and tz <= nearest`` or uv outside the ``extent``-scaled
    image get uv = depth = 0.
    """
    return _ProjectPoint.apply(xyz, intr, extr, W, H, nearest, extent)


class _ProjectPoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, intr, extr, W, H, nearest, extent):
        x = _lib.f32("xyz", xyz)
        i = _lib.f32("intr", intr)
        e = _lib.f32("extr", extr)
        n = x.shape[0]
        uv = torch.empty((n, 2), dtype=torch.float32, device=x.device)
        depth = torch.empty((n, 1), dtype=torch.float32, device=x.device)
        with torch.cuda.device(x.device):
            _lib.call("msplat_project_point_forward", _lib.ptr(x), n, _lib.ptr(i), _lib.ptr(e), int(W), int(H),
                      float(nearest), float(extent), _lib.ptr(uv), _lib.ptr(depth), _lib.stream_ptr())
        ctx.save_for_backward(xyz, intr, extr, depth)
        return uv, depth

    @staticmethod
    def backward(ctx, dL_duv, dL_ddepth):
        xyz, intr, extr, depth = ctx.saved_tensors
        x = _lib.f32("xyz", xyz)
        n = x.shape[0]
        dev = x.device
        dxyz = torch.empty((n, 3), dtype=torch.float32, device=dev)
        dintr = torch.empty(4, dtype=torch.float32, device=dev) if intr.requires_grad else None
        dextr = torch.empty(12, dtype=torch.float32, device=dev) if extr.requires_grad else None
        g_uv = _lib.f32("dL_duv", dL_duv)
        g_d = _lib.f32("dL_ddepth", dL_ddepth) if dL_ddepth is not None else None
        with torch.cuda.device(dev):
            _lib.call("msplat_project_point_backward", _lib.ptr(x), n, _lib.ptr(_lib.f32("intr", intr)),
                      _lib.ptr(_lib.f32("extr", extr)), _lib.ptr(_lib.f32("depth", depth)), _lib.ptr(g_uv),
                      _lib.ptr(g_d), _lib.ptr(dxyz), _lib.ptr(dintr), _lib.ptr(dextr), _lib.stream_ptr())
        return (dxyz.to(xyz.dtype), dintr, camera_grad_shape(extr, dextr) if dextr is not None else None,
                None, None, None, None)
