"""Squared-error image loss, fused: ``l2_loss(x, target) = sum((x - target)^2)``.

One reduction launch forward and one elementwise launch backward, where the
torch expression ``((x - target) ** 2).sum()`` runs three kernels forward
and four backward.  The reference computes its training loss in torch
(tutorials/gs_2d.py:73-90); this is the rasterizer's training-step loss next
to the hot path.  The sum is independent of scheduling (fixed per-workgroup
trees, partials added in order in fp64).
"""

from __future__ import annotations

import torch

from . import _lib


def l2_loss(x, target):
    """sum((x - target)^2) over all elements, differentiable with respect to x (and target)."""
    return _L2Loss.apply(x, target)


class _L2Loss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, target):
        a = _lib.f32("x", x)
        t = _lib.f32("target", target)
        if a.shape != t.shape:
            raise RuntimeError(f"l2_loss: shapes {tuple(a.shape)} and {tuple(t.shape)} differ")
        out = torch.empty((), dtype=torch.float32, device=a.device)
        with torch.cuda.device(a.device):
            nb = int(_lib.query("msplat_sq_diff_sum_workspace_bytes"))
            ws = torch.empty(nb, dtype=torch.uint8, device=a.device)
            _lib.call("msplat_sq_diff_sum", _lib.ptr(a), _lib.ptr(t), a.numel(), _lib.ptr(ws), nb, _lib.ptr(out),
                      _lib.stream_ptr())
        ctx.save_for_backward(a, t)
        return out

    @staticmethod
    def backward(ctx, dL):
        a, t = ctx.saved_tensors
        g = _lib.f32("dL", dL.reshape(1))
        gx = gt = None
        with torch.cuda.device(a.device):
            gx = torch.empty_like(a)
            _lib.call("msplat_sq_diff_grad", _lib.ptr(a), _lib.ptr(t), a.numel(), _lib.ptr(g), _lib.ptr(gx),
                      _lib.stream_ptr())
        if ctx.needs_input_grad[1]:
            gt = -gx
        return (gx if ctx.needs_input_grad[0] else None), gt
