"""Pin the SH oracle (CPU only): against scipy's spherical harmonics with the
reference test's real-basis mapping, against golden vectors of the reference's
own polynomials (values and partials, tests/golden/sh_reference_basis.npz),
and against finite differences."""

import os

import numpy as np
import pytest
from scipy.special import sph_harm_y

from oracle import sh as osh
from tests.conftest import GOLDEN


def scipy_real_basis(d, deg):
    """Real SH from scipy's complex Y_n^m, mapped as msplat's test_compute_sh.py:328-355 does."""
    polar = np.arccos(d[2])
    azim = np.arctan2(d[1], d[0])
    out = np.zeros((deg + 1) ** 2)
    for n in range(deg + 1):
        for m in range(-n, n + 1):
            y1 = sph_harm_y(n, m, polar, azim)
            y2 = sph_harm_y(n, -m, polar, azim)
            if m < 0:
                v = (-1) ** (-m) * (1j * np.sqrt(0.5) * (y1 - (-1) ** (-m) * y2)).real
            elif m > 0:
                v = (-1) ** m * (np.sqrt(0.5) * (y2 + (-1) ** m * y1)).real
            else:
                v = y1.real
            out[n * n + n + m] = v
    return out


@pytest.mark.parametrize("deg", range(11))
def test_basis_matches_scipy(deg, rng):
    d = rng.normal(size=(8, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Y, _, _ = osh.basis(d, deg, np.float64)
    for i in range(len(d)):
        np.testing.assert_allclose(Y[i], scipy_real_basis(d[i], deg), rtol=1e-9, atol=1e-12)


def test_basis_matches_reference_polynomials():
    gold = np.load(os.path.join(GOLDEN, "sh_reference_basis.npz"))
    Y, dY, order = osh.basis(gold["dirs"], 10, np.float64, grad=True)
    np.testing.assert_allclose(Y, gold["Y"], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(dY, gold["dY"], rtol=1e-10, atol=1e-10)
    assert sorted(order) == list(range(121))


@pytest.mark.parametrize("deg", [0, 1, 3, 6, 10])
def test_sh_backward_fd(deg, rng):
    D = (deg + 1) ** 2
    n, C = 4, 3
    shs = rng.normal(size=(n, C, D))
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    dirs[1] *= 1.2  # the partials are those of the polynomials, unit or not
    g = rng.normal(size=(n, C))
    loss = lambda s, d: (osh.sh_forward(s, d, dtype=np.float64) * g).sum()
    dshs, ddirs = osh.sh_backward(shs, dirs, g, dtype=np.float64)
    h = 1e-6
    for idx in np.ndindex(dirs.shape):
        dp, dm = dirs.copy(), dirs.copy()
        dp[idx] += h
        dm[idx] -= h
        np.testing.assert_allclose(ddirs[idx], (loss(shs, dp) - loss(shs, dm)) / (2 * h), rtol=1e-6, atol=1e-7)
    for idx in [(0, 0, 0), (2, 1, D - 1), (3, 2, D // 2)]:
        sp, sm = shs.copy(), shs.copy()
        sp[idx] += h
        sm[idx] -= h
        np.testing.assert_allclose(dshs[idx], (loss(sp, dirs) - loss(sm, dirs)) / (2 * h), rtol=1e-6, atol=1e-9)


def test_visibility_and_fp32_accuracy(rng):
    n, C, D = 64, 2, 121
    shs = rng.normal(size=(n, C, D)).astype(np.float32)
    dirs = rng.normal(size=(n, 3))
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    vis = np.arange(n) % 4 != 0
    v32 = osh.sh_forward(shs, dirs, vis)
    v64 = osh.sh_forward(shs.astype(np.float64), dirs.astype(np.float64), vis, dtype=np.float64)
    assert np.all(v32[~vis] == 0)
    # the recurrence in fp32 stays far inside the reference test's 5e-4 tolerance at degree 10
    assert np.abs(v32 - v64).max() < 2e-5


@pytest.mark.parametrize("deg", [0, 1, 3, 6, 10])
def test_sh_points_backward_fd(deg, rng):
    """The fused view-direction adjoint against central differences of the fp64 forward."""
    D = (deg + 1) ** 2
    n, C = 5, 3
    shs = rng.normal(size=(n, C, D))
    xyz = rng.normal(size=(n, 3)) * 2.0
    campos = np.array([0.3, -4.0, 1.5])
    g = rng.normal(size=(n, C))
    loss = lambda p: (osh.sh_points_forward(shs, p, campos, dtype=np.float64) * g).sum()
    _, dxyz = osh.sh_points_backward(shs, xyz, campos, g, dtype=np.float64)
    h = 1e-6
    for idx in np.ndindex(xyz.shape):
        xp, xm = xyz.copy(), xyz.copy()
        xp[idx] += h
        xm[idx] -= h
        np.testing.assert_allclose(dxyz[idx], (loss(xp) - loss(xm)) / (2 * h), rtol=1e-6, atol=1e-7)


def test_sh_points_normalisation_adjoint_matches_torch(rng):
    """(g - d (d.g)) / r equals torch's fp64 autograd of v / |v| (an independent adjoint)."""
    import torch

    xyz = rng.normal(size=(50, 3)) * 3.0
    campos = rng.normal(size=3) * 4.0
    g = rng.normal(size=(50, 3))
    d, r = osh.view_dirs_from_points(xyz, campos, np.float64)
    v = torch.tensor(xyz - campos[None], dtype=torch.float64, requires_grad=True)
    dt = v / v.norm(dim=1, keepdim=True)
    np.testing.assert_allclose(d, dt.detach().numpy(), rtol=1e-15, atol=1e-15)
    dt.backward(torch.tensor(g))
    dot = (d * g).sum(1, keepdims=True)
    np.testing.assert_allclose((g - d * dot) / r[:, None], v.grad.numpy(), rtol=1e-12, atol=1e-14)


def test_sh_points_forward_is_sh_of_normalised_offsets(rng):
    n, C, D = 40, 3, 16
    shs = rng.normal(size=(n, C, D))
    xyz = rng.normal(size=(n, 3))
    campos = np.array([1.0, 2.0, -3.0])
    v = xyz - campos
    ref = osh.sh_forward(shs, v / np.linalg.norm(v, axis=1, keepdims=True), dtype=np.float64)
    np.testing.assert_allclose(osh.sh_points_forward(shs, xyz, campos, dtype=np.float64), ref, rtol=1e-13, atol=1e-14)
    vis = np.arange(n) % 3 != 0
    _, dxyz = osh.sh_points_backward(shs.astype(np.float32), xyz.astype(np.float32), campos, rng.normal(size=(n, C)),
                                     vis)
    assert np.all(dxyz[~vis] == 0) and dxyz.dtype == np.float32
