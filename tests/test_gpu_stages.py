"""GPU parity of every stage against the CPU oracle, through the C ABI.

Bit-exact where the design promises it (per-Gaussian geometry and SH are
the same fp32 operation sequence on both sides; the sort is integer work);
tolerances are stated where arithmetic order legitimately differs (camera
gradients reduced in a different order, blending with a hardware exp).
"""

import numpy as np
import pytest
import torch

import msplat_amd as ms
from msplat_amd.rasterization import preprocess_backward, preprocess_forward
from msplat_amd.sort_gaussian import compute_gaussian_key, compute_tile_gaussian_range, sort_gaussian_keys
from oracle import blend as oblend
from oracle import geometry as og
from oracle import sh as osh
from oracle import sort as osort
from tests.helpers import assert_bitexact, np_, rel_err, small_scene

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _cam(cam):
    return cam.intr.to(DEV), cam.extr.to(DEV)


@pytest.mark.parametrize("seed,nearest", [(0, 0.0), (1, 0.0), (2, 0.5)])
def test_project_point_forward_bitexact(seed, nearest):
    g, cam = small_scene(seed, n=5000, half_extent=3.0)
    intr, extr = _cam(cam)
    uv, depth = ms.project_point(g["xyz"].to(DEV), intr, extr, cam.W, cam.H, nearest=nearest)
    ouv, od = og.project_forward(np_(g["xyz"]), np_(cam.intr), np_(cam.extr), cam.W, cam.H, nearest=nearest)
    assert_bitexact(np_(uv), ouv, "uv")
    assert_bitexact(np_(depth)[:, 0], od, "depth")
    assert (od == 0).any() and (od != 0).any()


def test_project_point_backward():
    g, cam = small_scene(3, n=4000, half_extent=2.0)
    xyz = g["xyz"].to(DEV).requires_grad_()
    intr = cam.intr.to(DEV).requires_grad_()
    extr = cam.extr.to(DEV).requires_grad_()
    uv, depth = ms.project_point(xyz, intr, extr, cam.W, cam.H)
    gen = torch.Generator().manual_seed(7)
    guv = torch.randn(uv.shape, generator=gen)
    gd = torch.randn(depth.shape, generator=gen)
    (uv * guv.to(DEV)).sum().add_((depth * gd.to(DEV)).sum()).backward()
    odxyz, odi, ode = og.project_backward(np_(g["xyz"]), np_(cam.intr), np_(cam.extr), np_(depth)[:, 0], np_(guv),
                                          np_(gd)[:, 0])
    assert_bitexact(np_(xyz.grad), odxyz, "dxyz")
    assert rel_err(np_(intr.grad), odi) < 1e-4
    assert rel_err(np_(extr.grad).reshape(-1), ode) < 1e-4


def test_cov3d_bitexact():
    g, _ = small_scene(4, n=3000)
    vis = torch.rand(3000, generator=torch.Generator().manual_seed(1)) > 0.2
    s = g["scale"].to(DEV).requires_grad_()
    q = (g["rotate"] * 1.3).to(DEV).requires_grad_()  # unnormalised on purpose
    cov = ms.compute_cov3d(s, q, vis.to(DEV))
    ocov = og.cov3d_forward(np_(g["scale"]), np_(q), np_(vis))
    assert_bitexact(np_(cov), ocov, "cov3d")
    gc = torch.randn(cov.shape, generator=torch.Generator().manual_seed(2))
    (cov * gc.to(DEV)).sum().backward()
    ods, odq = og.cov3d_backward(np_(g["scale"]), np_(q), np_(gc), np_(vis))
    assert_bitexact(np_(s.grad), ods, "dscale")
    assert_bitexact(np_(q.grad), odq, "dquat")


def _ewa_inputs(seed):
    g, cam = small_scene(seed, n=4000, half_extent=2.0)
    ouv, od = og.project_forward(np_(g["xyz"]), np_(cam.intr), np_(cam.extr), cam.W, cam.H)
    vis = od != 0
    ocov = og.cov3d_forward(np_(g["scale"]), np_(g["rotate"]), vis)
    return g, cam, ouv, od, vis, ocov


def test_ewa_forward_backward():
    g, cam, ouv, od, vis, ocov = _ewa_inputs(5)
    intr = cam.intr.to(DEV).requires_grad_()
    extr = cam.extr.to(DEV).requires_grad_()
    xyz = g["xyz"].to(DEV).requires_grad_()
    cov = torch.from_numpy(ocov).to(DEV).requires_grad_()
    conic, radius, tiles = ms.ewa_project(xyz, cov, intr, extr, torch.from_numpy(ouv).to(DEV), cam.W, cam.H,
                                          torch.from_numpy(vis).to(DEV))
    oconic, orad, otil = og.ewa_forward(np_(g["xyz"]), ocov, np_(cam.intr), np_(cam.extr), ouv, cam.W, cam.H, vis)
    assert_bitexact(np_(conic), oconic, "conic")
    assert_bitexact(np_(radius), orad, "radius")
    assert_bitexact(np_(tiles), otil, "tiles")
    assert (orad > 0).sum() > 100
    gc = torch.randn(conic.shape, generator=torch.Generator().manual_seed(3))
    (conic * gc.to(DEV)).sum().backward()
    odxyz, odcov, odi, ode = og.ewa_backward(np_(g["xyz"]), ocov, np_(cam.intr), np_(cam.extr), orad, np_(gc))
    assert_bitexact(np_(xyz.grad), odxyz, "dxyz")
    assert_bitexact(np_(cov.grad), odcov, "dcov3d")
    assert rel_err(np_(intr.grad), odi) < 1e-4
    assert rel_err(np_(extr.grad).reshape(-1), ode) < 1e-4


def test_fused_preprocess_equals_chain():
    g, cam = small_scene(6, n=5000, half_extent=3.0)
    intr, extr = _cam(cam)
    xyz, s, q = g["xyz"].to(DEV), g["scale"].to(DEV), g["rotate"].to(DEV)
    uv, depth, conic, radius, tiles = preprocess_forward(xyz, s, q, intr, extr, cam.W, cam.H)
    uv2, d2 = ms.project_point(xyz, intr, extr, cam.W, cam.H)
    vis = d2[:, 0] != 0
    cov = ms.compute_cov3d(s, q, vis)
    c2, r2, t2 = ms.ewa_project(xyz, cov, intr, extr, uv2, cam.W, cam.H, vis)
    for a, b, w in ((uv, uv2, "uv"), (depth, d2, "depth"), (conic, c2, "conic"), (radius, r2, "radius"),
                    (tiles, t2, "tiles")):
        assert_bitexact(np_(a), np_(b), w)
    gen = torch.Generator().manual_seed(9)
    guv = torch.randn(uv.shape, generator=gen).to(DEV)
    gco = torch.randn(conic.shape, generator=gen).to(DEV)
    dxyz, dscale, drot, _, _ = preprocess_backward(xyz, s, q, intr, extr, depth, radius, guv, gco)
    xyz_, s_, q_ = xyz.clone().requires_grad_(), s.clone().requires_grad_(), q.clone().requires_grad_()
    uv3, d3 = ms.project_point(xyz_, intr, extr, cam.W, cam.H)
    cov3 = ms.compute_cov3d(s_, q_, d3[:, 0] != 0)
    c3, _, _ = ms.ewa_project(xyz_, cov3, intr, extr, uv3.detach(), cam.W, cam.H, d3[:, 0] != 0)
    ((uv3 * guv).sum() + (c3 * gco).sum()).backward()
    assert_bitexact(np_(dxyz), np_(xyz_.grad), "dxyz")
    assert_bitexact(np_(dscale), np_(s_.grad), "dscale")
    assert_bitexact(np_(drot), np_(q_.grad), "drot")


@pytest.mark.parametrize("deg", list(range(11)))
def test_sh_bitexact(deg):
    D = (deg + 1) ** 2
    gen = torch.Generator().manual_seed(100 + deg)
    n, C = 777, 3
    dirs = torch.randn(n, 3, generator=gen)
    dirs = dirs / dirs.norm(dim=1, keepdim=True)
    shs = torch.randn(n, C, D, generator=gen)
    vis = torch.rand(n, generator=gen) > 0.1
    s = shs.to(DEV).requires_grad_()
    d = dirs.to(DEV).requires_grad_()
    val = ms.compute_sh(s, d, vis.to(DEV))
    assert_bitexact(np_(val), osh.sh_forward(np_(shs), np_(dirs), np_(vis)), "value")
    gv = torch.randn(n, C, generator=gen)
    (val * gv.to(DEV)).sum().backward()
    ods, odd = osh.sh_backward(np_(shs), np_(dirs), np_(gv), np_(vis))
    assert_bitexact(np_(s.grad), ods, "dshs")
    assert_bitexact(np_(d.grad), odd, "ddirs")


@pytest.mark.parametrize("deg,C", [(0, 3), (1, 3), (3, 3), (3, 4), (6, 35), (7, 3), (10, 3), (3, 70)])
def test_sh_points_bitexact(deg, C):
    """Fused view direction + SH against the oracle's restatement, bit for bit
    (gc kernels for degree <= 6 and C <= 64, LDS / per-Gaussian kernels otherwise)."""
    D = (deg + 1) ** 2
    gen = torch.Generator().manual_seed(300 + deg + C)
    n = 555
    xyz = torch.randn(n, 3, generator=gen) * 2.0
    campos = torch.tensor([0.5, -3.5, 2.0])
    shs = torch.randn(n, C, D, generator=gen)
    vis = torch.rand(n, generator=gen) > 0.1
    s = shs.to(DEV).requires_grad_()
    x = xyz.to(DEV).requires_grad_()
    c = campos.to(DEV).requires_grad_()
    val = ms.compute_sh_points(s, x, c, vis.to(DEV))
    assert_bitexact(np_(val), osh.sh_points_forward(np_(shs), np_(xyz), np_(campos), np_(vis)), "value")
    gv = torch.randn(n, C, generator=gen)
    (val * gv.to(DEV)).sum().backward()
    ods, odx = osh.sh_points_backward(np_(shs), np_(xyz), np_(campos), np_(gv), np_(vis))
    assert_bitexact(np_(s.grad), ods, "dshs")
    assert_bitexact(np_(x.grad), odx, "dxyz")
    assert rel_err(np_(c.grad), -odx.astype(np.float64).sum(0)) < 1e-5


def test_sh_points_equals_unfused_composition():
    """compute_sh_points == compute_sh(shs, normalize(xyz - campos)) with torch's
    normalisation, values and gradients (xyz, shs, campos) to fp32 rounding."""
    gen = torch.Generator().manual_seed(9)
    n, C, D = 4000, 3, 16
    xyz = (torch.rand(n, 3, generator=gen) * 2 - 1) * 3.0
    campos = torch.tensor([1.0, -4.0, 0.5])
    shs = torch.randn(n, C, D, generator=gen) * 0.3
    gv = torch.randn(n, C, generator=gen).to(DEV)
    outs = []
    for fused in (True, False):
        s = shs.to(DEV).requires_grad_()
        x = xyz.to(DEV).requires_grad_()
        c = campos.to(DEV).requires_grad_()
        if fused:
            val = ms.compute_sh_points(s, x, c)
        else:
            d = x - c[None]
            val = ms.compute_sh(s, d / d.norm(dim=1, keepdim=True))
        (val * gv).sum().backward()
        outs.append((val.detach(), s.grad, x.grad, c.grad))
    for a, b, name in zip(outs[0], outs[1], ("value", "dshs", "dxyz", "dcampos")):
        assert rel_err(np_(a), np_(b)) < 1e-5, name


def test_sort_kat():
    import json
    import os
    kat = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sort_kat.json")))
    uv = torch.tensor(kat["uv"], dtype=torch.float32, device=DEV)
    depth = torch.tensor(kat["depth"], dtype=torch.float32, device=DEV)[:, None]
    radius = torch.tensor(kat["radius"], dtype=torch.int32, device=DEV)[:, None]
    tiles = torch.tensor(kat["tiles"], dtype=torch.int32, device=DEV)[:, None]
    idx, rng = ms.sort_gaussian(uv, depth, kat["W"], kat["H"], radius, tiles)
    assert np_(idx).tolist() == kat["idx_sorted"]
    assert np_(rng).tolist() == kat["tile_range"]


@pytest.mark.parametrize("seed,W,H,n", [(0, 128, 96, 3000), (1, 640, 480, 20000), (2, 4000, 50, 5000),
                                        (3, 33, 17, 200), (4, 640, 480, 60000)])
def test_sort_bitexact(seed, W, H, n):
    """The last case has millions of pairs: the class sort takes the multi-launch passes (more than 320
    tiles of 4096 keys)."""
    rng = np.random.default_rng(seed)
    uv = rng.uniform(-40, max(W, H) + 40, size=(n, 2)).astype(np.float32)
    depth = rng.choice(np.float32([-2.5, -1.0, 0.5, 1.0, 1.0, 2.0, 3.25]), n) * rng.choice([1, 1.0001], n)
    depth = depth.astype(np.float32)
    radius = rng.integers(-2, 60, n).astype(np.int32)
    radius[rng.random(n) < 0.02] = 900  # a few huge splats
    oidx, orng, okeys = osort.sort_gaussian(uv, depth, radius, W, H)
    idx, trange, keys = sort_gaussian_keys(torch.from_numpy(uv).to(DEV), torch.from_numpy(depth).to(DEV), W, H,
                                           torch.from_numpy(radius).to(DEV))
    assert_bitexact(np_(keys), okeys, "key_sorted")
    assert_bitexact(np_(idx), oidx, "idx_sorted")
    assert_bitexact(np_(trange), orng, "tile_range")
    assert (okeys < 0).any()
    if n == 60000:
        assert len(okeys) > 320 * 4096, len(okeys)
    # the step API: reference-order keys, sorted by torch, ranges from keys
    area = np.asarray(osort.gaussian_keys(uv, depth, radius, W, H)[2], np.int32)
    cums = torch.from_numpy(np.cumsum(area).astype(np.int32)).to(DEV)
    k2, i2 = compute_gaussian_key(torch.from_numpy(uv).to(DEV), torch.from_numpy(depth).to(DEV), W, H,
                                  torch.from_numpy(radius).to(DEV), cums)
    ks, perm = torch.sort(k2, stable=True)
    assert_bitexact(np_(ks), okeys, "step keys")
    assert_bitexact(np_(i2[perm]), oidx, "step idx")
    assert_bitexact(np_(compute_tile_gaussian_range(W, H, cums, ks)), orng, "step range")


def test_sort_tiles_mismatch_raises():
    uv = torch.tensor([[8.0, 8.0]], device=DEV)
    depth = torch.ones(1, 1, device=DEV)
    radius = torch.tensor([4], dtype=torch.int32, device=DEV)
    with pytest.raises(RuntimeError):
        ms.sort_gaussian(uv, depth, 32, 32, radius, torch.tensor([3], dtype=torch.int32, device=DEV))


def _blend_inputs(seed, n=1500, W=96, H=80, C=3):
    g, cam = small_scene(seed, n=n, W=W, H=H, C=C)
    ouv, od = og.project_forward(np_(g["xyz"]), np_(cam.intr), np_(cam.extr), W, H)
    vis = od != 0
    ocov = og.cov3d_forward(np_(g["scale"]), np_(g["rotate"]), vis)
    oconic, orad, _ = og.ewa_forward(np_(g["xyz"]), ocov, np_(cam.intr), np_(cam.extr), ouv, W, H, vis)
    oidx, orng, _ = osort.sort_gaussian(ouv, od, orad, W, H)
    return g, ouv, oconic, oidx, orng


@pytest.mark.parametrize("C", [1, 3, 4, 7, 35, 70])
def test_blend_forward_backward(C):
    W, H = 96, 80
    g, ouv, oconic, oidx, orng = _blend_inputs(10 + C, C=C, W=W, H=H)
    feat = np.random.default_rng(C).normal(size=(len(ouv), C)).astype(np.float32)
    op = np_(g["opacity"])
    bg = 0.25
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    uv_ = t(ouv).requires_grad_()
    co_ = t(oconic).requires_grad_()
    op_ = t(op).requires_grad_()
    fe_ = t(feat).requires_grad_()
    out = ms.alpha_blending(uv_, co_, op_, fe_, t(oidx), t(orng), bg, W, H)
    o64, fT64, nc64, _ = oblend.blend_forward(ouv, oconic, op, feat, oidx, orng, bg, W, H, dtype=np.float64)
    o32, fT32, nc32, _ = oblend.blend_forward(ouv, oconic, op, feat, oidx, orng, bg, W, H, dtype=np.float32)
    err = np.abs(np_(out) - o64).max()
    assert err < 1e-4, f"image max-abs error vs fp64 oracle {err}"
    gout = np.random.default_rng(1).normal(size=out.shape).astype(np.float32)
    (out * t(gout)).sum().backward()
    duv, dco, dop, dfe = oblend.blend_backward(ouv, oconic, op, feat, oidx, orng, bg, W, H, fT64, nc64, gout,
                                               dtype=np.float64)
    for a, b, w in ((uv_.grad, duv, "duv"), (co_.grad, dco, "dconic"), (op_.grad, dop.reshape(-1, 1), "dopacity"),
                    (fe_.grad, dfe, "dfeature")):
        e = rel_err(np_(a), b)
        assert e < 1e-3, f"{w} rel err {e}"


def test_blend_empty_and_single():
    W, H = 40, 24
    uv = torch.tensor([[5.0, 6.0]], device=DEV)
    conic = torch.tensor([[0.5, 0.0, 0.5]], device=DEV)
    opac = torch.tensor([[0.6]], device=DEV)
    feat = torch.tensor([[2.0, -1.0]], device=DEV)
    idx = torch.zeros(1, dtype=torch.int32, device=DEV)
    rng = torch.zeros((6, 2), dtype=torch.int32, device=DEV)
    rng[0] = torch.tensor([0, 1])
    out = ms.alpha_blending(uv, conic, opac, feat, idx, rng, 0.5, W, H)
    o = np_(out)
    # pixel (5,6) is the centre: alpha = 0.6
    assert abs(o[0, 6, 5] - (0.6 * 2.0 + 0.4 * 0.5)) < 1e-6
    assert abs(o[1, 6, 5] - (0.6 * -1.0 + 0.4 * 0.5)) < 1e-6
    # every pixel outside tile 0 sees an empty list: background
    assert np.all(o[:, :, 16:] == 0.5) and np.all(o[:, 16:, :] == 0.5)
