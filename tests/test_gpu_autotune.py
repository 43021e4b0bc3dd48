"""The autotuner: every choice it can make renders the same images.

* autotune(): the recorded (cell, pixels per lane) renders the image of the
  default configuration bit for bit, gradients within 1e-5;
* autotune_views(): the recorded views per launch and launch configuration
  render every view bit for bit as one launch of all views with the
  defaults, gradients within 1e-4 (sums over views in another order);
* autotune_serving(): picks one of its candidates and the server then still
  renders every request as a single-view call does."""

import random

import pytest
import torch

import msplat_amd as ms
from msplat_amd import autotune, scenes, serving, tuning, views
from tests.helpers import np_, rel_err, small_scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture(autouse=True)
def _clean_table():
    tuning.clear()
    yield
    tuning.clear()


def _grads(fn, leaves):
    out = fn()
    imgs = out if isinstance(out, list) else [out]
    loss = sum((im * im).sum() for im in imgs)
    return [im.detach() for im in imgs], torch.autograd.grad(loss, leaves)


def test_autotune_choice_renders_the_default_image():
    g, cam = small_scene(31, n=8000, W=320, H=208, half_extent=1.5, C=4)
    leaves = [g[k].to(DEV).requires_grad_() for k in ("xyz", "scale", "rotate", "opacity", "feature")]
    intr, extr = cam.intr.to(DEV), cam.extr.to(DEV)
    run = lambda: ms.rasterization(*leaves, intr, extr, cam.W, cam.H, 0.1)
    with tuning.override(tuning.DEFAULT):
        ref_img, ref_g = _grads(run, leaves)
    best, rep = autotune.autotune(*leaves, intr, extr, cam.W, cam.H, 0.1, cells=(4, 8, 16), ppls=(1, 2, 4), reps=2)
    print(f"autotune chose {best} of {len(rep)} candidates", flush=True)
    assert tuning.config_for(cam.W, cam.H, 4) == best
    img, gr = _grads(run, leaves)
    assert torch.equal(img[0], ref_img[0])
    for a, b in zip(gr, ref_g):
        assert rel_err(np_(a), np_(b)) < 1e-5


def test_autotune_views_schedule_renders_the_same_views():
    gen = torch.Generator().manual_seed(5)
    n = 6000
    g = scenes.make_gaussians(gen, n, 1.5)
    g["feature"] = torch.rand(n, 3, generator=gen)
    sizes = [(96, 64), (160, 120), (64, 48), (128, 96), (200, 150), (80, 80)]
    cams = [scenes.make_camera(gen, w, h, 4.0, 1.171875 * w) for w, h in sizes]
    intrs = torch.stack([c.intr for c in cams]).to(DEV)
    extrs = torch.stack([c.extr for c in cams]).to(DEV).requires_grad_()
    leaves = [g[k].to(DEV).requires_grad_() for k in ("xyz", "scale", "rotate", "opacity", "feature")]
    with tuning.override(tuning.DEFAULT):
        ref_img, ref_g = _grads(lambda: views.rasterization_views(*leaves, intrs, extrs, sizes, 0.0,
                                                                  views_per_launch=len(sizes)), leaves + [extrs])
    v, cfg, rep = autotune.autotune_views(*leaves, intrs, extrs, sizes, 0.0, groups=(1, 2, 4), reps=2)
    print(f"autotune_views chose {v} views per launch, {cfg}", flush=True)
    assert 1 <= v <= len(sizes)
    img, gr = _grads(lambda: views.rasterization_views(*leaves, intrs, extrs, sizes, 0.0), leaves + [extrs])
    for a, b in zip(img, ref_img):
        assert torch.equal(a, b)
    for a, b in zip(gr, ref_g):
        assert rel_err(np_(a), np_(b)) < 1e-4
    for vv in (1, 2, 4):  # every grouping the tuner tried renders the same views
        with tuning.override(cfg):
            imgs = views.rasterization_views(*leaves, intrs, extrs, sizes, 0.0, views_per_launch=vv)
        for a, b in zip(imgs, ref_img):
            assert torch.equal(a.detach(), b)


def test_autotune_serving_keeps_results():
    gen = torch.Generator().manual_seed(0)
    st = serving.SceneStore(DEV, channels=3)
    for name, n in (("a", 3000), ("b", 5000)):
        g = scenes.make_gaussians(gen, n, 1.0)
        st.add(name, g["xyz"], g["scale"], g["rotate"], g["opacity"], torch.rand(n, 3, generator=gen))

    def make(rng):
        W, H = rng.choice([(96, 64), (128, 96)])
        cam = scenes.make_camera(torch.Generator().manual_seed(rng.randrange(1 << 30)), W, H, 4.0, 1.171875 * W)
        return serving.Request(rng.choice(["a", "b"]), cam.intr, cam.extr, W, H)

    server = serving.RenderServer(st, hbm_budget=4 << 30, max_views=4)
    best, rep = autotune.autotune_serving(server, make, rate=500.0, candidates=(2, 8), duration=0.3)
    print(f"autotune_serving chose max_views={best}: {rep}", flush=True)
    assert best in (2, 8) and server.max_views == best
    rng = random.Random(9)
    reqs = [make(rng) for _ in range(10)]
    for r in reqs:
        server.submit(r)
    server.drain()
    for r in reqs:
        s = st.scenes[r.scene]
        sl = slice(s.first, s.first + s.count)
        with torch.no_grad():
            alone = ms.rasterization(st.xyz[sl], st.scale[sl], st.rotate[sl], st.opacity[sl], st.feature[sl],
                                     r.intr.to(DEV), r.extr.to(DEV), r.W, r.H, r.bg)
        assert torch.equal(r.image, alone)
