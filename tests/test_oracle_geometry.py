"""Pin the oracle's projection / covariance / EWA against closed forms, brute
force and fp64 finite differences (CPU only)."""

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import geometry as og

W, H = 128, 96


def rand_camera(rng):
    R = Rotation.random(random_state=int(rng.integers(1 << 30))).as_matrix()
    t = rng.normal(size=3)
    return np.array([300.0, 280.0, 64.0, 48.0]), np.concatenate([R, t[:, None]], 1)


def points_in_view(rng, extr, n, zlo=1.0, zhi=5.0):
    """World points whose camera coordinates are known (built independently of the oracle)."""
    pc = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.2, 0.2, n), rng.uniform(zlo, zhi, n)], 1)
    R, t = extr[:, :3], extr[:, 3]
    return (pc - t) @ R, pc  # p = R^T (pc - t)


def test_project_closed_form(rng):
    intr, extr = rand_camera(rng)
    p, pc = points_in_view(rng, extr, 200)
    uv, d = og.project_forward(p, intr, extr, W, H, extent=0.0, dtype=np.float64)
    u = intr[0] * pc[:, 0] / (pc[:, 2] + 1e-7) + intr[2] - 0.5
    v = intr[1] * pc[:, 1] / (pc[:, 2] + 1e-7) + intr[3] - 0.5
    np.testing.assert_allclose(uv[:, 0], u, rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(uv[:, 1], v, rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(d, pc[:, 2], rtol=1e-12)
    # a 4x4 extrinsic is read through its first three rows
    e4 = np.eye(4)
    e4[:3] = extr
    uv4, d4 = og.project_forward(p, intr, e4, W, H, extent=0.0, dtype=np.float64)
    np.testing.assert_array_equal(uv4, uv)


def test_project_culling():
    intr = np.array([100.0, 100.0, 50.0, 40.0])
    extr = np.concatenate([np.eye(3), np.zeros((3, 1))], 1)
    # u = 100 x / z + 49.5; extent 1.3 keeps u in [-15, 115] for W = 100
    p = np.array([[0.0, 0.0, 2.0], [0.6, 0.0, 1.0], [0.7, 0.0, 1.0], [0.0, 0.0, 0.1], [0.0, 0.0, -2.0]])
    uv, d = og.project_forward(p, intr, extr, 100, 80, nearest=0.0, extent=1.3, dtype=np.float64)
    assert d[0] == 2.0 and d[1] == 1.0 and d[2] == 0.0 and np.all(uv[2] == 0)
    assert d[3] == 0.1 and d[4] == -2.0  # no near culling by default, negative depth survives
    _, dn = og.project_forward(p, intr, extr, 100, 80, nearest=0.2, extent=1.3, dtype=np.float64)
    assert dn[3] == 0.0 and dn[4] == 0.0 and dn[0] == 2.0


def _fd(f, x, h, idx):
    xp = x.copy()
    xm = x.copy()
    xp[idx] += h
    xm[idx] -= h
    return (f(xp) - f(xm)) / (2 * h)


def test_project_backward_fd(rng):
    intr, extr = rand_camera(rng)
    p, _ = points_in_view(rng, extr, 6)
    guv = rng.normal(size=(6, 2))
    gd = rng.normal(size=6)

    def loss(xyz=p, it=intr, ex=extr):
        uv, d = og.project_forward(xyz, it, ex, W, H, extent=0.0, dtype=np.float64)
        return (uv * guv).sum() + (d * gd).sum()

    _, d = og.project_forward(p, intr, extr, W, H, extent=0.0, dtype=np.float64)
    dxyz, dintr, dextr = og.project_backward(p, intr, extr, d, guv, gd, dtype=np.float64)
    for idx in np.ndindex(p.shape):
        np.testing.assert_allclose(dxyz[idx], _fd(lambda x: loss(xyz=x), p, 1e-6, idx), rtol=1e-5, atol=1e-6)
    for i in range(4):
        np.testing.assert_allclose(dintr[i], _fd(lambda x: loss(it=x), intr, 1e-5, i), rtol=1e-5, atol=1e-6)
    for idx in np.ndindex(extr.shape):
        np.testing.assert_allclose(dextr[idx[0] * 4 + idx[1]], _fd(lambda x: loss(ex=x), extr, 1e-6, idx),
                                   rtol=1e-5, atol=1e-5)


def test_cov3d_closed_form(rng):
    n = 50
    s = np.exp(rng.uniform(-3, 1, (n, 3)))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    R = Rotation.from_quat(q[:, [1, 2, 3, 0]]).as_matrix()  # scipy is scalar-last
    sig = R @ (s[:, :, None] ** 2 * np.eye(3)) @ R.transpose(0, 2, 1)
    cov = og.cov3d_forward(s, q, dtype=np.float64)
    iu = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    np.testing.assert_allclose(cov, np.stack([sig[:, a, b] for a, b in iu], 1), rtol=1e-10, atol=1e-14)
    ident = og.cov3d_forward(s, np.tile([1.0, 0, 0, 0], (n, 1)), dtype=np.float64)
    np.testing.assert_allclose(ident[:, [0, 3, 5]], s ** 2, rtol=1e-15)
    assert np.all(ident[:, [1, 2, 4]] == 0)
    ev = np.linalg.eigvalsh(np.stack([np.array([[c[0], c[1], c[2]], [c[1], c[3], c[4]], [c[2], c[4], c[5]]])
                                      for c in cov]))
    np.testing.assert_allclose(np.sort(ev, 1), np.sort(s ** 2, 1), rtol=1e-8)
    vis = np.arange(n) % 3 != 0
    assert np.all(og.cov3d_forward(s, q, vis, dtype=np.float64)[~vis] == 0)


def test_cov3d_backward_fd(rng):
    n = 5
    s = np.exp(rng.uniform(-2, 0, (n, 3)))
    q = rng.normal(size=(n, 4))  # the op does not normalise: check the raw formula's derivative
    g = rng.normal(size=(n, 6))
    loss = lambda ss, qq: (og.cov3d_forward(ss, qq, dtype=np.float64) * g).sum()
    ds, dq = og.cov3d_backward(s, q, g, dtype=np.float64)
    for idx in np.ndindex(s.shape):
        np.testing.assert_allclose(ds[idx], _fd(lambda x: loss(x, q), s, 1e-5, idx), rtol=1e-6, atol=1e-8)
    for idx in np.ndindex(q.shape):
        np.testing.assert_allclose(dq[idx], _fd(lambda x: loss(s, x), q, 1e-5, idx), rtol=1e-6, atol=1e-8)


def _ewa_case(rng, n=40):
    intr, extr = rand_camera(rng)
    p, pc = points_in_view(rng, extr, n, 2.0, 6.0)
    s = np.exp(rng.uniform(-4, -1.5, (n, 3)))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    cov = og.cov3d_forward(s, q, dtype=np.float64)
    uv, _ = og.project_forward(p, intr, extr, W, H, extent=0.0, dtype=np.float64)
    return intr, extr, p, pc, cov, uv


def test_ewa_closed_form(rng):
    intr, extr, p, pc, cov, uv = _ewa_case(rng)
    conic, radius, tiles = og.ewa_forward(p, cov, intr, extr, uv, W, H, dtype=np.float64)
    R = extr[:, :3]
    fx, fy = intr[:2]
    for i in range(len(p)):
        x, y, z = pc[i]
        J = np.array([[fx / z, 0, -fx * x / z ** 2], [0, fy / z, -fy * y / z ** 2]])
        c = cov[i]
        S = np.array([[c[0], c[1], c[2]], [c[1], c[3], c[4]], [c[2], c[4], c[5]]])
        c2 = J @ R @ S @ R.T @ J.T + 0.3 * np.eye(2)
        inv = np.linalg.inv(c2)
        if tiles[i] == 0:
            continue
        np.testing.assert_allclose(conic[i], [inv[0, 0], inv[0, 1], inv[1, 1]], rtol=1e-8)
        lam = np.linalg.eigvalsh(c2)
        mid = lam.mean()
        disc = max(0.1, ((lam[1] - lam[0]) / 2) ** 2)
        assert radius[i] == int(np.ceil(3 * np.sqrt(mid + np.sqrt(disc))))
    assert (tiles > 0).sum() > 10


def test_ewa_isotropic_on_axis():
    # identity camera, point on the optical axis, isotropic sigma: cov2d = (f s / z)^2 + 0.3
    intr = np.array([200.0, 200.0, 64.0, 48.0])
    extr = np.concatenate([np.eye(3), np.zeros((3, 1))], 1)
    p = np.array([[0.0, 0.0, 4.0]])
    cov = np.array([[0.04, 0, 0, 0.04, 0, 0.04]])
    uv = np.array([[63.5, 47.5]])
    conic, radius, tiles = og.ewa_forward(p, cov, intr, extr, uv, W, H, dtype=np.float64)
    var = (200 * 0.2 / 4) ** 2 + 0.3
    np.testing.assert_allclose(conic[0], [1 / var, 0, 1 / var], rtol=1e-12)
    assert radius[0] == int(np.ceil(3 * np.sqrt(var + np.sqrt(0.1))))


@pytest.mark.parametrize("seed", range(5))
def test_tile_rect_brute_force(seed):
    # integer centre and radius: the box is the set of tiles meeting [p - r, p + r)
    rng = np.random.default_rng(seed)
    gx, gy = 9, 7
    px = rng.integers(-40, 180, 300)
    py = rng.integers(-40, 150, 300)
    r = rng.integers(0, 60, 300)
    x0, y0, x1, y1 = og.tile_rect(px.astype(np.float32), py.astype(np.float32), r, gx, gy)
    for i in range(300):
        xs = [t for t in range(gx) if 16 * t < px[i] + r[i] and 16 * t + 16 > px[i] - r[i]]
        ys = [t for t in range(gy) if 16 * t < py[i] + r[i] and 16 * t + 16 > py[i] - r[i]]
        area = (x1[i] - x0[i]) * (y1[i] - y0[i])
        assert area == len(xs) * len(ys), (px[i], py[i], r[i])
        if xs and ys:
            assert (x0[i], x1[i], y0[i], y1[i]) == (xs[0], xs[-1] + 1, ys[0], ys[-1] + 1)


def test_f2i_sat():
    v = np.array([np.nan, 3e9, -3e9, 2.7, -2.7, np.inf], np.float32)
    assert og.f2i_sat(v).tolist() == [0, 2147483647, -2147483648, 2, -2, 2147483647]


def test_ewa_backward_fd(rng):
    intr, extr, p, pc, cov, uv = _ewa_case(rng, n=5)
    _, radius, _ = og.ewa_forward(p, cov, intr, extr, uv, W, H, dtype=np.float64)
    radius = np.maximum(radius, 1)
    g = rng.normal(size=(len(p), 3))

    def loss(xyz=p, cv=cov, it=intr, ex=extr):
        c = og.ewa_forward(xyz, cv, it, ex, uv, W, H, dtype=np.float64, raw=True)
        return (c * g).sum()

    dxyz, dcov, dintr, dextr = og.ewa_backward(p, cov, intr, extr, radius, g, dtype=np.float64)
    for idx in np.ndindex(p.shape):
        np.testing.assert_allclose(dxyz[idx], _fd(lambda x: loss(xyz=x), p, 1e-6, idx), rtol=1e-5, atol=1e-5)
    for idx in np.ndindex(cov.shape):
        np.testing.assert_allclose(dcov[idx], _fd(lambda x: loss(cv=x), cov, 1e-8, idx), rtol=1e-5, atol=1e-3)
    for i in range(4):
        np.testing.assert_allclose(dintr[i], _fd(lambda x: loss(it=x), intr, 1e-4, i), rtol=1e-5, atol=1e-8)
    for idx in np.ndindex(extr.shape):
        np.testing.assert_allclose(dextr[idx[0] * 4 + idx[1]], _fd(lambda x: loss(ex=x), extr, 1e-7, idx),
                                   rtol=1e-5, atol=1e-4)
