"""Pin the blending oracle (CPU only): closed forms, an independent
cumulative-product formulation of front-to-back compositing, and fp64 finite
differences of the adjoint."""

import numpy as np
import pytest

from oracle import blend as ob


def one_tile_inputs(uv, conic, opac, feat, W=16, H=16):
    n = len(uv)
    idx = np.arange(n, dtype=np.int32)
    T = ((W + 15) // 16) * ((H + 15) // 16)
    rng = np.zeros((T, 2), np.int32)
    rng[0] = (0, n)
    return (np.asarray(uv, np.float64), np.asarray(conic, np.float64), np.asarray(opac, np.float64),
            np.asarray(feat, np.float64), idx, rng)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_single_gaussian_closed_form(dtype):
    uv, co, op, fe, idx, rng = one_tile_inputs([[5.0, 7.0]], [[0.3, 0.05, 0.2]], [0.7], [[2.0, -1.0, 0.5]])
    out, fT, nc, _ = ob.blend_forward(uv, co, op, fe, idx, rng, 0.25, 16, 16, dtype=dtype)
    # at the centre power = 0: alpha = opacity
    np.testing.assert_allclose(out[:, 7, 5], 0.7 * fe[0] + 0.3 * 0.25, rtol=1e-6)
    assert nc[7, 5] == 1 and abs(fT[7, 5] - 0.3) < 1e-7
    # elsewhere: alpha = o exp(-d^T conic d / 2) when above 1/255, else background
    for (x, y) in [(9, 7), (5, 12), (0, 0)]:
        d = np.array([5.0 - x, 7.0 - y])
        a = 0.7 * np.exp(-0.5 * (0.3 * d[0] ** 2 + 0.2 * d[1] ** 2) - 0.05 * d[0] * d[1])
        a = a if a >= 1 / 255 else 0.0
        np.testing.assert_allclose(out[:, y, x], a * fe[0] + (1 - a) * 0.25, rtol=2e-6, atol=1e-7)


def test_empty_tiles_are_background():
    uv, co, op, fe, idx, rng = one_tile_inputs([[5.0, 7.0]], [[0.3, 0.0, 0.3]], [0.5], [[1.0]], W=40, H=20)
    out, fT, nc, _ = ob.blend_forward(uv, co, op, fe, idx, rng, 0.75, 40, 20)
    assert np.all(out[:, :, 16:] == np.float32(0.75)) and np.all(out[:, 16:, :] == np.float32(0.75))
    assert np.all(fT[:, 16:] == 1) and np.all(nc[:, 16:] == 0)


def composite_reference(uv, conic, opac, feat, bg, W, H):
    """Front-to-back compositing written as cumulative products per pixel (one tile)."""
    n, C = feat.shape
    out = np.zeros((C, H, W))
    for y in range(H):
        for x in range(W):
            d = uv - np.array([x, y], np.float64)
            power = -0.5 * (conic[:, 0] * d[:, 0] ** 2 + conic[:, 2] * d[:, 1] ** 2) - conic[:, 1] * d[:, 0] * d[:, 1]
            a = np.minimum(0.99, opac * np.exp(power))
            use = (power <= 0) & (a >= 1 / 255)
            a = np.where(use, a, 0.0)
            trans = np.concatenate([[1.0], np.cumprod(1 - a)])
            # stop at the first Gaussian that would bring transmittance below 1e-4 (it is excluded)
            stop = np.nonzero(trans[1:] < 1e-4)[0]
            k = stop[0] if len(stop) else n
            w = a[:k] * trans[:k]
            out[:, y, x] = (w[:, None] * feat[:k]).sum(0) + trans[k] * bg
    return out


@pytest.mark.parametrize("seed", range(4))
def test_matches_cumprod_formulation(seed):
    rng = np.random.default_rng(seed)
    n = 40
    uv = rng.uniform(-4, 20, (n, 2))
    A = rng.normal(size=(n, 2, 2)) * 0.6
    cov = A @ A.transpose(0, 2, 1) + 0.3 * np.eye(2)
    inv = np.linalg.inv(cov)
    conic = np.stack([inv[:, 0, 0], inv[:, 0, 1], inv[:, 1, 1]], 1)
    opac = rng.uniform(0.2, 1.0, n)
    feat = rng.normal(size=(n, 3))
    args = one_tile_inputs(uv, conic, opac, feat)
    out, _, _, _ = ob.blend_forward(*args, 0.3, 16, 16, dtype=np.float64)
    bg = float(np.float32(0.3))  # bg is a float parameter of the op, as in the reference kernel
    np.testing.assert_allclose(out, composite_reference(uv, conic, opac, feat, bg, 16, 16), rtol=1e-12, atol=1e-12)


def test_backward_fd():
    rng = np.random.default_rng(5)
    n, C, W, H = 5, 3, 8, 8
    uv = rng.uniform(2, 6, (n, 2))
    conic = np.tile([1 / 400.0, 0.0, 1 / 300.0], (n, 1)) + rng.uniform(0, 1e-4, (n, 3))  # wide, smooth
    opac = rng.uniform(0.3, 0.6, n)
    feat = rng.normal(size=(n, C))
    g = rng.normal(size=(C, H, W))
    bg = 0.4
    idx = np.arange(n, dtype=np.int32)
    trange = np.array([[0, n]], np.int32)

    def loss(u=uv, c=conic, o=opac, f=feat):
        out, _, _, _ = ob.blend_forward(u, c, o, f, idx, trange, bg, W, H, dtype=np.float64)
        return (out * g).sum()

    out, fT, nc, _ = ob.blend_forward(uv, conic, opac, feat, idx, trange, bg, W, H, dtype=np.float64)
    assert np.all(nc == n)  # every Gaussian contributes everywhere: no threshold is crossed
    duv, dco, dop, dfe = ob.blend_backward(uv, conic, opac, feat, idx, trange, bg, W, H, fT, nc, g, dtype=np.float64)
    h = 1e-6

    def fd(name, arr, grad, hh=h):
        for i in np.ndindex(arr.shape):
            p, m = arr.copy(), arr.copy()
            p[i] += hh
            m[i] -= hh
            num = (loss(**{name: p}) - loss(**{name: m})) / (2 * hh)
            np.testing.assert_allclose(grad[i], num, rtol=1e-6, atol=1e-8, err_msg=f"{name}{i}")

    fd("u", uv, duv)
    fd("c", conic, dco, 1e-8)
    fd("o", opac, dop)
    fd("f", feat, dfe)


def test_threaded_oracle_equals_one_thread():
    """The CPU baseline's OpenMP variants: forward bit-exact with the
    one-thread oracle; adjoint equal up to the order of its per-thread sums."""
    from tests.helpers import small_scene
    from oracle import pipeline

    g, cam = small_scene(5, n=3000, W=96, H=80)
    npg = {k: v.numpy() for k, v in g.items()}
    args = (npg["xyz"], npg["scale"], npg["rotate"], npg["opacity"], npg["feature"], cam.intr.numpy(),
            cam.extr.numpy(), 96, 80, 0.2)
    _, inter = pipeline.rasterize_forward(*args)
    blend_args = (inter["uv"], inter["conic"], npg["opacity"], npg["feature"], inter["idx_sorted"],
                  inter["tile_range"], 0.2, 96, 80)
    a = ob.blend_forward(*blend_args)
    b = ob.blend_forward(*blend_args, threads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    dout = np.random.default_rng(0).normal(size=a[0].shape).astype(np.float32)
    ga = ob.blend_backward(*blend_args, a[1], a[2], dout)
    gb = ob.blend_backward(*blend_args, a[1], a[2], dout, threads=4)
    for x, y in zip(ga, gb):
        assert np.abs(x - y).max() <= 1e-5 * max(np.abs(x).max(), 1e-12)
