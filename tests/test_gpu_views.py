"""Many views in one launch: every stage per view equals the single-view path
(bit-exact preprocessing, bins and images), shared-parameter gradients equal
the sums of the single-view gradients, and per-view camera gradients match."""

import numpy as np
import pytest
import torch

import msplat_amd as ms
from msplat_amd import binning, scenes, views
from msplat_amd.rasterization import preprocess_forward
from oracle import sort as osort
from tests.helpers import np_, rel_err

pytestmark = pytest.mark.gpu
DEV = "cuda"
SIZES = [(160, 120), (96, 64), (200, 136)]


def _scene(seed=0, n=3000, C=3, sizes=SIZES):
    gen = torch.Generator().manual_seed(seed)
    g = scenes.make_gaussians(gen, n, 1.5)
    g["feature"] = torch.rand(n, C, generator=gen)
    cams = [scenes.make_camera(gen, w, h, 4.0, 1.171875 * w) for w, h in sizes]
    return {k: v.to(DEV) for k, v in g.items()}, [c.to(DEV) for c in cams]


def _stack(cams):
    return torch.stack([c.intr for c in cams]), torch.stack([c.extr for c in cams])


def test_preprocess_views_bitexact():
    g, cams = _scene(1)
    intrs, extrs = _stack(cams)
    out = views.preprocess_views_forward(g["xyz"], g["scale"], g["rotate"], intrs, extrs, SIZES)
    n = g["xyz"].shape[0]
    for b, (c, (w, h)) in enumerate(zip(cams, SIZES)):
        ref = preprocess_forward(g["xyz"], g["scale"], g["rotate"], c.intr, c.extr, w, h)
        for a, r in zip(out, ref):
            assert torch.equal(a[b * n:(b + 1) * n].reshape(r.shape), r), f"view {b}"


@pytest.mark.parametrize("cell", [1, 4, 8])
def test_bin_views_bitexact(cell):
    g, cams = _scene(2)
    intrs, extrs = _stack(cams)
    uv, depth, conic, radius, _ = views.preprocess_views_forward(g["xyz"], g["scale"], g["rotate"], intrs, extrs,
                                                                 SIZES)
    n = g["xyz"].shape[0]
    b_all = views.bin_views(uv, depth, radius, n, SIZES, cell)
    ids, rng = np_(b_all.cell_ids), np_(b_all.cell_range)
    c0 = 0
    for b, (w, h) in enumerate(SIZES):
        sl = slice(b * n, (b + 1) * n)
        _, oids, orng = osort.cell_lists(np_(uv[sl]), np_(depth[sl]), np_(radius[sl]), w, h, cell)
        for c in range(orng.shape[0]):
            s, e = rng[c0 + c]
            os_, oe = orng[c]
            assert (ids[s:e] - b * n).tolist() == oids[os_:oe].tolist(), f"view {b} cell {c}"
        c0 += orng.shape[0]
    assert c0 == rng.shape[0]


@pytest.mark.parametrize("per_view", [False, True])
def test_views_equal_single(per_view):
    g, cams = _scene(3)
    n, C = g["xyz"].shape[0], 3
    B = len(SIZES)
    intrs, extrs = _stack(cams)
    feats = torch.rand(B, n, C, generator=torch.Generator().manual_seed(9)).to(DEV) if per_view else g["feature"]

    def leaves():
        return [t.detach().clone().requires_grad_() for t in (g["xyz"], g["scale"], g["rotate"], g["opacity"],
                                                               feats, intrs, extrs)]

    la = leaves()
    imgs = views.rasterization_views(*la, SIZES, 0.1)
    gouts = [torch.randn(im.shape, generator=torch.Generator().manual_seed(b)).to(DEV) for b, im in enumerate(imgs)]
    sum((im * go).sum() for im, go in zip(imgs, gouts)).backward()

    lb = leaves()
    for b, (c, (w, h)) in enumerate(zip(cams, SIZES)):
        fb = lb[4][b] if per_view else lb[4]
        img = ms.rasterization(lb[0], lb[1], lb[2], lb[3], fb, lb[5][b], lb[6][b], w, h, 0.1)
        assert torch.equal(img, imgs[b]), f"view {b} image"
        (img * gouts[b]).sum().backward()
    for a, r, w in zip(la, lb, ("xyz", "scale", "rotate", "opacity", "feature", "intr", "extr")):
        assert rel_err(np_(a.grad), np_(r.grad)) < 1e-4, w


def test_views_extr_4x4():
    g, cams = _scene(4, sizes=SIZES[:2])
    intrs = torch.stack([c.intr for c in cams])
    e4 = torch.zeros(2, 4, 4, device=DEV)
    e4[:, :3] = torch.stack([c.extr for c in cams])
    e4[:, 3, 3] = 1.0
    e4.requires_grad_()
    imgs = views.rasterization_views(g["xyz"], g["scale"], g["rotate"], g["opacity"], g["feature"], intrs, e4,
                                     SIZES[:2], 0.0)
    sum(im.sum() for im in imgs).backward()
    assert e4.grad.shape == (2, 4, 4) and bool((e4.grad[:, 3] == 0).all())


def test_views_limits():
    g, cams = _scene(5, n=100, sizes=[(32, 32)])
    with pytest.raises(RuntimeError):
        views.rasterization_views(g["xyz"], g["scale"], g["rotate"], g["opacity"], g["feature"],
                                  cams[0].intr[None].repeat(65, 1), cams[0].extr[None].repeat(65, 1, 1),
                                  [(32, 32)] * 65, 0.0)
