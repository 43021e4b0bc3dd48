"""3D covariance: the drop-in for msplat.compute_cov3d (msplat/compute_cov3d.py:7-64)."""

from __future__ import annotations

import torch

from . import _lib


def compute_cov3d(scales, uquats, visible=None):
    """Sigma = R S^2 R^T as the upper triangle [P,6] = (00,01,02,11,12,22).

    ``uquats`` is (w,x,y,z) and is used as given (not normalised); rows with
    ``visible == False`` stay zero.
    """
    if visible is None:
        visible = torch.ones_like(scales[:, 0], dtype=torch.bool)
    return _ComputeCov3D.apply(scales, uquats, visible)


class _ComputeCov3D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, scales, uquats, visible):
        s = _lib.f32("scales", scales)
        q = _lib.f32("uquats", uquats)
        v = _lib.u8("visible", visible)
        n = s.shape[0]
        cov = torch.empty((n, 6), dtype=torch.float32, device=s.device)
        with torch.cuda.device(s.device):
            _lib.call("msplat_compute_cov3d_forward", _lib.ptr(s), _lib.ptr(q), _lib.ptr(v), n, _lib.ptr(cov),
                      _lib.stream_ptr())
        ctx.save_for_backward(scales, uquats, visible)
        return cov

    @staticmethod
    def backward(ctx, dL_dcov3d):
        scales, uquats, visible = ctx.saved_tensors
        s = _lib.f32("scales", scales)
        n = s.shape[0]
        ds = torch.empty((n, 3), dtype=torch.float32, device=s.device)
        dq = torch.empty((n, 4), dtype=torch.float32, device=s.device)
        with torch.cuda.device(s.device):
            _lib.call("msplat_compute_cov3d_backward", _lib.ptr(s), _lib.ptr(_lib.f32("uquats", uquats)),
                      _lib.ptr(_lib.u8("visible", visible)), _lib.ptr(_lib.f32("dL_dcov3d", dL_dcov3d)), n,
                      _lib.ptr(ds), _lib.ptr(dq), _lib.stream_ptr())
        return ds.to(scales.dtype), dq.to(uquats.dtype), None
