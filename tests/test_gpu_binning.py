"""Cell binning and blending over cell lists (the rasterizer's replacement for
sort_gaussian + alpha_blending inside rasterization): bins bit-exact with the
oracle; images, final_T and ncontrib bit-identical to alpha_blending over the
reference's sorted tile lists for every cell size and pixels-per-lane; gradients
equal up to float summation order."""

import numpy as np
import pytest
import torch

from msplat_amd import binning
from msplat_amd.alpha_blending import blend_backward, blend_forward
from msplat_amd.rasterization import preprocess_forward
from msplat_amd.sort_gaussian import sort_gaussian
from oracle import sort as osort
from tests.helpers import np_, rel_err, small_scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _projected(seed, n, W, H, ext=2.0, C=4):
    g, cam = small_scene(seed, n=n, W=W, H=H, half_extent=ext, C=C)
    d = {k: v.to(DEV) for k, v in g.items()}
    uv, depth, conic, radius, tiles = preprocess_forward(d["xyz"], d["scale"], d["rotate"], cam.intr.to(DEV),
                                                         cam.extr.to(DEV), W, H)
    return d, uv, depth, conic, radius, tiles


def _check_bins(uv, depth, radius, W, H, cell):
    b = binning.bin_gaussians(uv, depth, W, H, radius, cell)
    rect, ids, rng = osort.cell_lists(np_(uv), np_(depth), np_(radius), W, H, cell)
    r = np_(b.rect).view(np.uint32).astype(np.int64)
    got = np.stack([r[:, 0] & 0xffff, r[:, 0] >> 16, r[:, 1] & 0xffff, r[:, 1] >> 16], 1)
    live = rect[:, 2] > rect[:, 0]
    assert np.array_equal(got[live], rect[live]), "tile rectangles"
    assert np.array_equal(np_(b.cell_ids), ids), "cell lists"
    assert np.array_equal(np_(b.cell_range), rng), "cell ranges"
    return b


@pytest.mark.parametrize("seed,n,W,H,cell", [(0, 3000, 128, 96, 1), (1, 5000, 200, 120, 3),
                                              (2, 4000, 256, 256, 8), (3, 2000, 100, 70, 64),
                                              (5, 6000, 640, 480, 1)])
def test_bins_bitexact(seed, n, W, H, cell):
    """Up to 256 cells the emission writes the lists directly (per-cell look-back);
    640 x 480 with cells of one tile has 1200 cells: emission + class sort."""
    _, uv, depth, _, radius, _ = _projected(seed, n, W, H)
    _check_bins(uv, depth, radius, W, H, cell)


@pytest.mark.parametrize("W,H,cell", [(128, 96, 1), (256, 192, 1), (240, 240, 1)])
def test_bins_bitexact_large_boxes(W, H, cell):
    """The emission marks rows of more than 16 cells by cell (the union of the
    wave's large boxes, a 64-bit word per 64 cells): a scene whose first
    Gaussians are large enough to cover dozens of cells, on grids of 48, 192
    and 225 cells (one, three and four words), bit-exact with the oracle."""
    g, cam = small_scene(11, n=3000, W=W, H=H, half_extent=1.5)
    g["scale"][:400] = torch.exp(torch.empty(400, 3).uniform_(-2.0, -0.8, generator=torch.Generator().manual_seed(3)))
    d = {k: v.to(DEV) for k, v in g.items()}
    uv, depth, conic, radius, tiles = preprocess_forward(d["xyz"], d["scale"], d["rotate"], cam.intr.to(DEV),
                                                         cam.extr.to(DEV), W, H)
    rect, _, _ = osort.cell_lists(np_(uv), np_(depth), np_(radius), W, H, cell)
    area = (rect[:, 2] - rect[:, 0]) * (rect[:, 3] - rect[:, 1])
    gx, gy = (W + 15) // 16, (H + 15) // 16
    big = (area > 16) & (np_(radius).reshape(-1) > 0)
    last_cell = (rect[:, 3] - 1) * gx + rect[:, 2] - 1  # cell = tile with cells of one tile
    beyond = int((big & (last_cell >= 64)).sum())
    print(f"{W}x{H} cell {cell}: {int(big.sum())} rows of more than 16 cells, {beyond} reaching past the "
          f"first 64 cells", flush=True)
    assert big.sum() >= 200 and (gx * gy <= 64 or beyond >= 200)
    _check_bins(uv, depth, radius, W, H, cell)


def test_bins_bitexact_many_tiles():
    """2.2 M Gaussians: the depth sort's capacity spans 538 tiles of 4096 keys;
    its length is on the device, so it still takes the one-launch passes with
    look-back (up to 640 tiles)."""
    _, uv, depth, _, radius, _ = _projected(7, 2_200_000, 640, 480)
    _check_bins(uv, depth, radius, 640, 480, 4)


def test_bins_bitexact_multi_launch_sorts():
    """2.8 M Gaussians on 1200 cells: the depth sort's capacity (684 tiles) and
    the class sort's millions of entries both exceed what the look-back passes
    take, so both sorts run as histogram, scan and scatter launches per pass.
    A capacity below the count writes only slots below it."""
    _, uv, depth, _, radius, _ = _projected(8, 2_800_000, 640, 480)
    b = _check_bins(uv, depth, radius, 640, 480, 1)
    k = b.cell_ids.numel()
    print(f"2.8 M Gaussians, cell 1: {k} entries", flush=True)
    assert k > 320 * 4096
    short = binning.bin_gaussians(uv, depth, 640, 480, radius, 1, capacity=k // 3)
    assert short.overflowed() and int(short.count[0].item()) == k
    assert int(short.cell_range.max()) <= k // 3


def test_bins_ties_and_negative_depth():
    gen = np.random.default_rng(11)
    n, W, H = 3000, 160, 112
    uv = torch.from_numpy(gen.uniform(-30, 190, (n, 2)).astype(np.float32)).to(DEV)
    depth = gen.choice(np.float32([-1.0, -0.0, 0.0, 0.5, 1.0, 2.0]), n)
    depth = torch.from_numpy(depth[:, None]).to(DEV)
    radius = torch.from_numpy(gen.integers(-3, 50, n).astype(np.int32)).to(DEV)
    for cell in (1, 2, 5):
        _check_bins(uv, depth, radius, W, H, cell)


@pytest.mark.parametrize("cell", [1, 8])
def test_bins_capacity_mode(cell):
    """With a capacity the count stays on the device: a capacity above the count
    gives the exact lists (entries past the count unused); one below it writes
    only `capacity` entries and overflowed() reports it."""
    _, uv, depth, _, radius, _ = _projected(4, 4000, 192, 128)
    exact = binning.bin_gaussians(uv, depth, 192, 128, radius, cell)
    k = exact.cell_ids.numel()
    assert k > 100 and not exact.overflowed()
    roomy = binning.bin_gaussians(uv, depth, 192, 128, radius, cell, capacity=k + 777)
    assert not roomy.overflowed() and int(roomy.count[0].item()) == k
    assert torch.equal(roomy.cell_ids[:k], exact.cell_ids) and torch.equal(roomy.cell_range, exact.cell_range)
    short = binning.bin_gaussians(uv, depth, 192, 128, radius, cell, capacity=k // 3)
    assert short.overflowed() and int(short.count[0].item()) == k
    assert int(short.cell_range.max()) <= k // 3


def test_bins_empty():
    uv = torch.zeros(0, 2, device=DEV)
    b = binning.bin_gaussians(uv, torch.zeros(0, 1, device=DEV), 64, 64, torch.zeros(0, dtype=torch.int32,
                                                                                     device=DEV), 4)
    assert b.cell_ids.numel() == 0 and int(b.cell_range.abs().sum()) == 0


@pytest.mark.parametrize("C", [1, 3, 4, 8, 35])
def test_rasterize_equals_tile_lists(C):
    W, H = 176, 120
    d, uv, depth, conic, radius, tiles = _projected(4 + C, 6000, W, H, ext=2.0, C=C)
    op, feat = d["opacity"], d["feature"]
    idx, trng = sort_gaussian(uv, depth, W, H, radius, tiles)
    ref_out, ref_T, ref_nc = blend_forward(uv, conic, op, feat, idx, trng, 0.2, W, H)
    gout = torch.randn(ref_out.shape, generator=torch.Generator().manual_seed(C)).to(DEV)
    ref_g = blend_backward(uv, conic, op, feat, idx, trng, 0.2, W, H, ref_T, ref_nc, gout)
    for cell in (1, 8, 16):
        b = binning.bin_gaussians(uv, depth, W, H, radius, cell)
        for ppl in ((1, 2, 4) if cell != 16 else (4,)):  # cell 16, 4 pixels per lane: the benchmarked launch
            for cull in (False, True):
                print(f"C={C} cell={cell} ppl={ppl} cull={cull}", flush=True)
                out, T, nc, walk = binning.rasterize_forward(uv, conic, op, feat, b, 0.2, W, H, ppl, cull)
                assert torch.equal(out, ref_out) and torch.equal(T, ref_T)
                if cull:  # culled Gaussians only shift list positions
                    assert bool((nc <= ref_nc).all())
                else:
                    assert torch.equal(nc, ref_nc)
                g = binning.rasterize_backward(uv, conic, op, feat, b, 0.2, W, H, T, nc, walk, gout, ppl, cull)
                for a, r, w in zip(g, ref_g, ("duv", "dconic", "dopacity", "dfeature")):
                    assert rel_err(np_(a), np_(r)) < 1e-5, f"{w} cell={cell} ppl={ppl} cull={cull}"


def test_walk_record():
    """walk[t] = (list index after the last contributor, its tile-list position)."""
    W, H = 128, 96
    d, uv, depth, conic, radius, tiles = _projected(9, 4000, W, H)
    b = binning.bin_gaussians(uv, depth, W, H, radius, 4)
    _, _, nc, walk = binning.rasterize_forward(uv, conic, d["opacity"], d["feature"], b, 0.0, W, H, cull=False)
    gx, gy = (W + 15) // 16, (H + 15) // 16
    ncp = np.zeros((gy * 16, gx * 16), np.int64)
    ncp[:H, :W] = np_(nc)
    tile_last = ncp.reshape(gy, 16, gx, 16).max(axis=(1, 3)).reshape(-1)
    walk = np_(walk)
    assert np.array_equal(walk[:, 1], tile_last)
    rect = np_(b.rect).view(np.uint32).astype(np.int64)
    ids, rng = np_(b.cell_ids), np_(b.cell_range)
    ncx = (gx + 3) // 4
    for t in range(gx * gy):
        tx, ty = t % gx, t // gx
        s, e = rng[(ty // 4) * ncx + tx // 4]
        covers = [(rect[g, 0] & 0xffff) <= tx < (rect[g, 1] & 0xffff) and (rect[g, 0] >> 16) <= ty < (rect[g, 1] >> 16)
                  for g in ids[s:e]]
        pos = np.nonzero(covers)[0] + s
        if tile_last[t] == 0:
            assert walk[t, 0] == 0
        else:
            assert walk[t, 0] == pos[tile_last[t] - 1] + 1


@pytest.mark.parametrize("C", [3, 4])
def test_grouped_adjoint_long_lists(C):
    """A dense, faint scene: tiles walk hundreds of cell-list positions before
    saturating, beyond the 256 positions the grouped adjoint (2 x 2 tiles per
    workgroup, cells of an even number of tiles, 4 pixels per lane) keeps in
    LDS, so its global-memory path for the positions past them runs."""
    W, H = 96, 80
    d, uv, depth, conic, radius, tiles = _projected(40 + C, 20000, W, H, ext=1.0, C=C)
    op = (d["opacity"] * 0.08).contiguous()
    feat = d["feature"]
    idx, trng = sort_gaussian(uv, depth, W, H, radius, tiles)
    ref_out, ref_T, ref_nc = blend_forward(uv, conic, op, feat, idx, trng, 0.0, W, H)
    assert int(ref_nc.max()) > 600, int(ref_nc.max())
    gout = torch.randn(ref_out.shape, generator=torch.Generator().manual_seed(C)).to(DEV)
    ref_g = blend_backward(uv, conic, op, feat, idx, trng, 0.0, W, H, ref_T, ref_nc, gout)
    for cell in (2, 16):
        b = binning.bin_gaussians(uv, depth, W, H, radius, cell)
        for cull in (False, True):
            print(f"long lists C={C} cell={cell} cull={cull}", flush=True)
            out, T, nc, walk = binning.rasterize_forward(uv, conic, op, feat, b, 0.0, W, H, 4, cull)
            assert torch.equal(out, ref_out) and torch.equal(T, ref_T)
            g = binning.rasterize_backward(uv, conic, op, feat, b, 0.0, W, H, T, nc, walk, gout, 4, cull)
            for a, r, w in zip(g, ref_g, ("duv", "dconic", "dopacity", "dfeature")):
                assert rel_err(np_(a), np_(r)) < 1e-5, f"{w} cell={cell} cull={cull}"
