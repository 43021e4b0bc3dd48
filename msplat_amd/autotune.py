"""Autotuner for the rasterizer's launch parameters (cell size, pixels per lane).

For one scene, camera and channel count it times, on the GPU, the binning for
every candidate cell size and the blending forward and adjoint for every
pixels-per-lane, and records the fastest combination for that (W, H, C) in
``msplat_amd.tuning`` (which ``rasterization`` consults).  Every candidate
produces the same image, so the choice is purely a speed one; the report
keeps all timings.
"""

from __future__ import annotations

import itertools

import torch

from . import tuning
from .binning import bin_gaussians, rasterize_backward, rasterize_forward
from .rasterization import preprocess_forward


def _time(fn, reps: int) -> float:
    """Median milliseconds of ``fn`` over ``reps`` runs (after one warm-up)."""
    fn()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def autotune(xyz, scale, rotate, opacity, feature, intr, extr, W: int, H: int, bg: float = 0.0,
             cells=(4, 8, 16), ppls=(1, 2, 4), reps: int = 5, record: bool = True, cull: bool = True):
    """Return ``(best BlendConfig, report)``; ``report`` maps each candidate to its timings (ms)."""
    with torch.no_grad():
        uv, depth, conic, radius, _ = preprocess_forward(xyz, scale, rotate, intr, extr, W, H)
        feature = feature.detach().float().contiguous()
        opacity = opacity.detach().float().contiguous()
        C = feature.shape[1]
        dout = torch.rand((C, H, W), device=uv.device, generator=torch.Generator(uv.device).manual_seed(0))
        report = {}
        best = None
        for cell in cells:
            t_bin = _time(lambda: bin_gaussians(uv, depth, W, H, radius, cell), reps)
            bins = bin_gaussians(uv, depth, W, H, radius, cell)
            fwd, bwd = {}, {}
            for p in ppls:
                fwd[p] = _time(lambda: rasterize_forward(uv, conic, opacity, feature, bins, bg, W, H, p, cull),
                               reps)
            _, final_T, ncontrib, walk = rasterize_forward(uv, conic, opacity, feature, bins, bg, W, H, 0, cull)
            for p in ppls:
                bwd[p] = _time(lambda: rasterize_backward(uv, conic, opacity, feature, bins, bg, W, H, final_T,
                                                          ncontrib, walk, dout, p, cull), reps)
            for pf, pb in itertools.product(ppls, ppls):
                total = t_bin + fwd[pf] + bwd[pb]
                report[(cell, pf, pb)] = dict(bin=t_bin, fwd=fwd[pf], bwd=bwd[pb], total=total)
                if best is None or total < report[best]["total"]:
                    best = (cell, pf, pb)
    cfg = tuning.BlendConfig(cell=best[0], ppl_fwd=best[1], ppl_bwd=best[2], cull=cull)
    if record:
        tuning.record(W, H, C, cfg)
    return cfg, report
