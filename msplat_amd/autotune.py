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


def _views_step(args, intrs, extrs, sizes, bg, v):
    from . import views
    from .loss import l2_loss

    imgs = views.rasterization_views(*args, intrs, extrs, sizes, bg, views_per_launch=v)
    loss = sum(l2_loss(im, torch.zeros_like(im)) for im in imgs)
    return torch.autograd.grad(loss, [a for a in args if a.requires_grad])


def autotune_views(xyz, scale, rotate, opacity, feature, intrs, extrs, sizes, bg: float = 0.0,
                   groups=(1, 2, 4, 8, 16, 32, 64), cells=(8, 16, 32), ppls=((2, 4), (4, 4)), reps: int = 3,
                   record: bool = True):
    """The batch schedule of a many-view training step: how many views share
    a launch, and the cell size and pixels per lane of those launches.

    Times a forward + backward (sum of squares of every image) over all the
    views for each launch configuration with every view in one launch, then,
    with the fastest, for each number of views per launch.  Returns
    ``(views_per_launch, BlendConfig, report)`` and records both in
    ``msplat_amd.tuning`` for these sizes (and for every group's sizes)."""
    sizes = [(int(w), int(h)) for w, h in sizes]
    B = len(sizes)
    C = feature.shape[-1]
    args = [t.detach().requires_grad_() for t in (xyz, scale, rotate, opacity, feature)]
    report: dict = {}
    best_cfg, best_t = None, None
    for cell, (pf, pb) in itertools.product(cells, ppls):
        cfg = tuning.BlendConfig(cell=cell, ppl_fwd=pf, ppl_bwd=pb)
        with tuning.override(cfg):
            t = _time(lambda: _views_step(args, intrs, extrs, sizes, bg, B), reps)
        report[("launch", cell, pf, pb)] = t
        if best_t is None or t < best_t:
            best_cfg, best_t = cfg, t
    best_v, best_vt = B, best_t
    for v in sorted({g for g in groups if g < B}):
        with tuning.override(best_cfg):
            t = _time(lambda: _views_step(args, intrs, extrs, sizes, bg, v), reps)
        report[("views_per_launch", v)] = t
        if t < best_vt:
            best_v, best_vt = v, t
    report[("views_per_launch", B)] = best_t
    if record:
        tuning.record_views_per_launch(sizes, C, None if best_v >= B else best_v)
        for b0 in range(0, B, best_v):
            tuning.record_views(sizes[b0:b0 + best_v], C, best_cfg)
    return best_v, best_cfg, report


def autotune_serving(server, make_request, rate: float, candidates=(4, 8, 16, 32), duration: float = 1.0,
                     seed: int = 0):
    """The batch size of a render server: the ``max_views`` (requests per
    launch) with the lowest p99 latency under Poisson arrivals at ``rate``
    (ties: the higher throughput).  Sets it on ``server`` and returns
    ``(max_views, report)`` with each candidate's statistics."""
    from . import serving

    report = {}
    prev = server.max_views
    best = None
    try:
        for v in candidates:
            server.set_max_views(v)
            st = serving.run_poisson(server, make_request, rate=rate, duration=duration, seed=seed)
            report[v] = st
            key = (st["p99_ms"], -st["throughput_req_s"])
            if best is None or key < best[0]:
                best = (key, v)
    finally:
        server.set_max_views(prev)
    server.set_max_views(best[1])
    return best[1], report
