"""The training step next to the rasterizer: Adam, and densification.

The reference trains its Gaussians with ``torch.optim.Adam``, one learning
rate per attribute tensor (tutorials/gs_3d.py:165-172, tutorials/gs_2d.py:
32-36), and every 256 iterations prunes Gaussians of opacity below 0.01 and
refills their slots with copies of the largest Gaussians, keeping the count
fixed (tutorials/gs_3d.py:193-225).  The rasterizer also hands every
Gaussian's screen-space position gradient back through ``ndc``
(msplat/alpha_blending.py:107-110): the statistic densification schemes
accumulate.

* ``Adam``: a drop-in for ``torch.optim.Adam`` (no weight decay, amsgrad or
  maximize) whose ``step()`` updates every parameter in one HIP launch per
  8 tensors.
* ``DensifyStats``: the dL/dndc consumer: per-Gaussian sums of |dL/dndc| and
  of the views that saw the Gaussian.
* ``prune_and_clone``: the tutorial's fixed-count prune and refill, on the
  device (mask, stable descending order of the scores, row copies).
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .binning import WAIT_EXPIRED


def _param(name: str, t: torch.Tensor) -> torch.Tensor:
    _lib.check_device(name, t)
    if t.dtype != torch.float32 or not t.is_contiguous():
        raise RuntimeError(f"{name}: the kernels update fp32 contiguous tensors in place "
                           f"(got {t.dtype}, contiguous={t.is_contiguous()})")
    return t


class Adam(torch.optim.Optimizer):
    """``torch.optim.Adam(params, lr, betas, eps)`` with the update in one HIP
    launch per 8 parameter tensors (grouped by device).

    Same parameter groups, state keys (``step``, ``exp_avg``, ``exp_avg_sq``)
    and update as torch's Adam without weight decay; parameters without a
    gradient are skipped.  Parameters must be fp32 contiguous HIP tensors."""

    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8, weight_decay: float = 0.0,
                 amsgrad: bool = False, maximize: bool = False):
        if weight_decay != 0.0 or amsgrad or maximize:
            raise RuntimeError("msplat_amd.train.Adam: weight_decay, amsgrad and maximize are not supported")
        if not 0.0 <= lr:
            raise ValueError(f"Invalid learning rate: {lr}")
        if not 0.0 <= eps:
            raise ValueError(f"Invalid epsilon value: {eps}")
        if not (0.0 <= betas[0] < 1.0 and 0.0 <= betas[1] < 1.0):
            raise ValueError(f"Invalid beta parameters: {betas}")
        super().__init__(params, dict(lr=lr, betas=tuple(betas), eps=eps))

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        # one launch per (device, betas, eps) and 8 tensors
        batches: dict = {}
        for group in self.param_groups:
            beta1, beta2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                _param("param", p)
                g = p.grad
                if g.is_sparse:
                    raise RuntimeError("msplat_amd.train.Adam: sparse gradients are not supported")
                g = _lib.f32("grad", g)
                st = self.state[p]
                if len(st) == 0:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p, memory_format=torch.contiguous_format)
                    st["exp_avg_sq"] = torch.zeros_like(p, memory_format=torch.contiguous_format)
                st["step"] += 1
                key = (p.device, float(beta1), float(beta2), float(group["eps"]))
                batches.setdefault(key, []).append((p, g, st["exp_avg"], st["exp_avg_sq"], float(group["lr"]),
                                                    int(st["step"])))
        for (dev, beta1, beta2, eps), items in batches.items():
            with torch.cuda.device(dev):
                k = len(items)
                arr = lambda vals: (ctypes.c_void_p * k)(*vals)
                _lib.call("msplat_adam_step", k, arr([_lib.ptr(t[0]) for t in items]),
                          arr([_lib.ptr(t[1]) for t in items]), arr([_lib.ptr(t[2]) for t in items]),
                          arr([_lib.ptr(t[3]) for t in items]), (_lib.I64 * k)(*[t[0].numel() for t in items]),
                          (ctypes.c_float * k)(*[t[4] for t in items]), (_lib.I64 * k)(*[t[5] for t in items]),
                          beta1, beta2, eps, _lib.stream_ptr())
        return loss

    def reset_state(self) -> None:
        """Forget every moment and step count (what re-creating the optimizer
        after densification does, tutorials/gs_3d.py:216-223)."""
        self.state.clear()


class DensifyStats:
    """Per-Gaussian densification statistics from dL/dndc.

    ``add(ndc_grad, radius)`` adds |dL/dndc| to ``accum`` and 1 to ``denom``
    for every Gaussian the view saw (radius > 0; without radii, a non-zero
    gradient); ``mean()`` is accum / denom (0 where never seen)."""

    def __init__(self, n: int, device="cuda"):
        self.accum = torch.zeros(n, dtype=torch.float32, device=device)
        self.denom = torch.zeros(n, dtype=torch.float32, device=device)

    def add(self, ndc_grad: torch.Tensor, radius: torch.Tensor | None = None) -> None:
        d = _lib.f32("ndc_grad", ndc_grad)
        n = self.accum.numel()
        if tuple(d.shape) != (n, 2):
            raise RuntimeError(f"DensifyStats.add: ndc_grad shape {tuple(d.shape)} != {(n, 2)}")
        r = None
        if radius is not None:
            r = _lib.i32("radius", radius).view(-1)
            if r.numel() != n:
                raise RuntimeError(f"DensifyStats.add: {r.numel()} radii for {n} Gaussians")
        with torch.cuda.device(d.device):
            _lib.call("msplat_densify_accumulate", _lib.ptr(d), _lib.ptr(r), n, _lib.ptr(self.accum),
                      _lib.ptr(self.denom), _lib.stream_ptr())

    def mean(self) -> torch.Tensor:
        return torch.where(self.denom > 0, self.accum / self.denom.clamp_min(1.0), torch.zeros_like(self.accum))

    def reset(self) -> None:
        self.accum.zero_()
        self.denom.zero_()


def prune_and_clone(arrays, opacity: torch.Tensor, threshold: float = 0.01, score: torch.Tensor | None = None,
                    scale: torch.Tensor | None = None, sync: bool = True):
    """tutorials/gs_3d.py:193-225 on the device: the Gaussians with
    ``opacity < threshold`` (activated opacity, [N] or [N,1]) receive, in
    ascending index order, the rows of the Gaussians of largest score in
    descending order (stable), in every tensor of ``arrays`` (row-major fp32
    [N, ...], updated in place).  ``score`` defaults to the largest entry of
    each row of ``scale`` [N,3] (the raw scale parameters, as the tutorial
    ranks them).  A pruned Gaussian that is also a donor keeps its row.

    Returns the number of pruned Gaussians (one host synchronisation), or,
    with ``sync=False``, a 0-d int64 device tensor holding it."""
    op = _lib.f32("opacity", opacity).reshape(-1)
    n = op.numel()
    if score is None and scale is None:
        raise RuntimeError("prune_and_clone: give a score or the scale parameters")
    sc = None if score is None else _lib.f32("score", score).reshape(-1)
    sl = None if scale is None or sc is not None else _lib.f32("scale", scale)
    if sc is not None and sc.numel() != n:
        raise RuntimeError(f"prune_and_clone: {sc.numel()} scores for {n} Gaussians")
    if sl is not None and tuple(sl.shape) != (n, 3):
        raise RuntimeError(f"prune_and_clone: scale shape {tuple(sl.shape)} != {(n, 3)}")
    arrays = list(arrays)
    if len(arrays) > 8:
        raise RuntimeError("prune_and_clone: at most 8 arrays per call")
    widths = []
    for k, a in enumerate(arrays):
        _param(f"arrays[{k}]", a)
        if a.shape[0] != n:
            raise RuntimeError(f"prune_and_clone: arrays[{k}] has {a.shape[0]} rows for {n} Gaussians")
        widths.append(a.numel() // n if n else 1)
    dev = op.device
    count = torch.empty((), dtype=torch.int64, device=dev)
    with torch.cuda.device(dev):
        nb = int(_lib.query("msplat_prune_clone_workspace_bytes", n))
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
        k = len(arrays)
        _lib.call("msplat_prune_clone", _lib.ptr(op), float(threshold), _lib.ptr(sc), _lib.ptr(sl), n, k,
                  (ctypes.c_void_p * max(k, 1))(*[_lib.ptr(a) for a in arrays]),
                  (ctypes.c_int * max(k, 1))(*widths), _lib.ptr(ws), nb, _lib.ptr(count), _lib.stream_ptr())
    if not sync:
        return count
    m = int(count.item())
    if m < 0:
        _lib.call("msplat_wait_status", None, 1, _lib.stream_ptr())
        raise RuntimeError(WAIT_EXPIRED)
    return m
