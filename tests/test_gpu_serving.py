"""The render server: every request's image is the same alone (a single-view
rasterization call) and in the mix (batched with other scenes, sizes and
arrival times); admission respects the memory budget; the Poisson driver
reports consistent latency statistics."""

import random

import pytest
import torch

import msplat_amd as ms
from msplat_amd import scenes, serving

pytestmark = pytest.mark.gpu
DEV = "cuda"
SIZES = [(96, 64), (128, 96), (160, 120)]


def _store():
    gen = torch.Generator().manual_seed(0)
    st = serving.SceneStore(DEV, channels=3)
    for name, n, ext in (("a", 3000, 1.0), ("b", 6000, 2.0), ("c", 1500, 0.5)):
        g = scenes.make_gaussians(gen, n, ext)
        st.add(name, g["xyz"], g["scale"], g["rotate"], g["opacity"], torch.rand(n, 3, generator=gen))
    return st


def _request(rng, st):
    name = rng.choice(sorted(st.scenes))
    W, H = rng.choice(SIZES)
    gen = torch.Generator().manual_seed(rng.randrange(1 << 30))
    cam = scenes.make_camera(gen, W, H, 4.0, 1.171875 * W)
    return serving.Request(name, cam.intr, cam.extr, W, H, bg=0.0)


def _alone(st, r):
    s = st.scenes[r.scene]
    sl = slice(s.first, s.first + s.count)
    with torch.no_grad():
        return ms.rasterization(st.xyz[sl], st.scale[sl], st.rotate[sl], st.opacity[sl], st.feature[sl],
                                r.intr.to(DEV), r.extr.to(DEV), r.W, r.H, r.bg)


@pytest.mark.parametrize("budget_mb", [4096, 8])
def test_same_alone_and_in_mix(budget_mb):
    st = _store()
    rng = random.Random(budget_mb)
    server = serving.RenderServer(st, hbm_budget=budget_mb << 20, max_views=6, streams=3)
    reqs = [_request(rng, st) for _ in range(14)]
    for i, r in enumerate(reqs):  # staggered arrivals: some join while others are on the GPU
        server.submit(r)
        if i % 3 == 2:
            server.step()
    server.drain()
    assert len(server.done) == len(reqs) and server.used == 0
    for r in reqs:
        assert torch.equal(r.image, _alone(st, r)), f"request on scene {r.scene} {r.W}x{r.H} (batch {r.batch})"
    batches = {r.batch for r in reqs}
    assert len(batches) >= (2 if budget_mb == 4096 else 5)


def test_poisson_stats():
    st = _store()
    server = serving.RenderServer(st, hbm_budget=2 << 30, max_views=8)
    stats = serving.run_poisson(server, lambda rng: _request(rng, st), rate=300.0, duration=0.5, seed=1, warmup=4)
    assert stats["requests"] > 0
    assert 0 < stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"]


def test_capacity_overflow_renders_again():
    """A capacity estimate far too small truncates the first batches' lists; the
    server sees the overflow when they retire, raises the estimate and renders
    them again, so every image still equals the single-view rendering."""
    st = _store()
    rng = random.Random(7)
    server = serving.RenderServer(st, hbm_budget=4096 << 20, max_views=6, streams=2, entries_per_row=1e-4,
                                  min_capacity=16)
    reqs = [_request(rng, st) for _ in range(9)]
    for r in reqs:
        server.submit(r)
    server.drain()
    assert server.rerenders >= 1 and server.entries_per_row > 1e-4
    assert len(server.done) == len(reqs) and server.used == 0
    for r in reqs:
        assert torch.equal(r.image, _alone(st, r)), f"request on scene {r.scene} {r.W}x{r.H} (batch {r.batch})"


def test_scene_loaded_from_ply(tmp_path):
    """A scene written to disk (scene_io's PLY layout) and loaded into a second
    store renders every request exactly as the in-memory scene does."""
    from msplat_amd import scene_io

    st = _store()
    disk = serving.SceneStore(DEV, channels=3)
    for name in sorted(st.scenes):
        s = st.scenes[name]
        sl = slice(s.first, s.first + s.count)
        path = str(tmp_path / f"{name}.ply")
        scene_io.save_ply(path, st.xyz[sl], st.scale[sl], st.rotate[sl], st.opacity[sl], feature=st.feature[sl],
                          activated=False)
        disk.add_ply(name, path, activate=False)
    for k in ("xyz", "scale", "rotate", "opacity", "feature"):
        assert torch.equal(getattr(disk, k), getattr(st, k)), k
    rng = random.Random(3)
    server = serving.RenderServer(disk, hbm_budget=4096 << 20, max_views=4)
    reqs = [_request(rng, disk) for _ in range(6)]
    for r in reqs:
        server.submit(r)
    server.drain()
    for r in reqs:
        assert torch.equal(r.image, _alone(st, r))
