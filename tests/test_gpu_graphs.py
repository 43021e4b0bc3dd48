"""HIP-graph capture of a training step: every replay (with eager work and
synchronisations in between) equals the eager step, for one view and for a
many-view batch; inputs updated in place are picked up by the replay."""

import pytest
import torch

import msplat_amd as ms
from msplat_amd import graphs, scenes, views
from tests.helpers import np_, rel_err

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="HIP graph capture disabled: the first capture of the single-view step "
                                      "crashed in the HIP runtime at capture end (segfault in capture_end); "
                                      "cause not identified")]
DEV = "cuda"


def _params(seed, n):
    gen = torch.Generator().manual_seed(seed)
    g = scenes.make_gaussians(gen, n, 1.5)
    g["feature"] = torch.rand(n, 3, generator=gen)
    return {k: v.to(DEV).requires_grad_() for k, v in g.items()}, gen


def _check(step, g, eager, reps=3):
    for r in range(reps):
        torch.randn(1000, device=DEV).sum().item()  # eager work and a synchronisation between replays
        out = g()
        torch.cuda.synchronize()
        assert not g.overflowed()
        assert torch.equal(out[0], eager[0]), f"replay {r}: image"
        for a, b in zip(out[1:], eager[1:]):
            assert rel_err(np_(a), np_(b)) < 1e-5, f"replay {r}: gradient"


def test_graphed_single_view_step():
    P, gen = _params(0, 4000)
    cam = scenes.make_camera(gen, 144, 96, 4.0, 1.171875 * 144).to(DEV)
    target = torch.rand(3, 96, 144, generator=gen).to(DEV)
    keys = ("xyz", "scale", "rotate", "opacity", "feature")

    def step():
        out = ms.rasterization(P["xyz"], P["scale"], P["rotate"], P["opacity"], P["feature"], cam.intr, cam.extr,
                               144, 96, 0.0)
        loss = ((out - target) ** 2).sum()
        return (out,) + torch.autograd.grad(loss, [P[k] for k in keys])

    eager = step()
    g = graphs.GraphedStep(step)
    assert len(g.capacities) == 1
    _check(step, g, eager)
    with torch.no_grad():  # in-place input update: the replay must see it
        P["xyz"].add_(0.01)
    eager2 = step()
    _check(step, g, eager2, reps=2)


def test_graphed_views_step():
    P, gen = _params(1, 3000)
    sizes = [(128, 96), (64, 48), (96, 80)]
    cams = [scenes.make_camera(gen, w, h, 4.0, 1.171875 * w).to(DEV) for w, h in sizes]
    intrs = torch.stack([c.intr for c in cams])
    extrs = torch.stack([c.extr for c in cams]).requires_grad_()
    keys = ("xyz", "scale", "rotate", "opacity", "feature")

    def step():
        imgs = views.rasterization_views(P["xyz"], P["scale"], P["rotate"], P["opacity"], P["feature"], intrs,
                                         extrs, sizes, 0.0)
        loss = sum((im ** 2).sum() for im in imgs)
        return (torch.cat([im.reshape(-1) for im in imgs]),) + torch.autograd.grad(loss, [P[k] for k in keys] +
                                                                                  [extrs])

    eager = step()
    g = graphs.GraphedStep(step)
    _check(step, g, eager)
