"""Projection, 3D covariance and EWA splatting on the CPU (test infrastructure only).

Vectorised numpy restatement of msplat/src/project_point.cu,
msplat/src/compute_cov3d.cu, msplat/src/ewa_project.cu and the get_rect
helper of msplat/include/utils.h.  With ``dtype=np.float32`` every operation
is the fp32 IEEE operation the GPU kernels execute, in the same order.
"""

from __future__ import annotations

import numpy as np

TILE = 16


def _c(dt, v):
    return np.dtype(dt).type(v)


def f2i_sat(v: np.ndarray) -> np.ndarray:
    """float -> int32 as the GPU converts: truncate, saturate, NaN -> 0."""
    v = np.asarray(v, dtype=np.float64)
    out = np.where(np.isnan(v), 0.0, np.trunc(np.nan_to_num(v, nan=0.0, posinf=3e9, neginf=-3e9)))
    return np.clip(out, -2147483648.0, 2147483647.0).astype(np.int64)


def tile_rect(px, py, r, gx, gy, dtype=np.float32):
    """utils.h:17-37 get_rect: tile box [x0,x1) x [y0,y1) of integer radius r."""
    dt = np.dtype(dtype).type
    px = np.asarray(px, dtype=dt)
    py = np.asarray(py, dtype=dt)
    rf = np.asarray(r).astype(dt)
    s, one = dt(16.0), dt(1.0)
    x0 = np.clip(f2i_sat((px - rf) / s), 0, gx)
    y0 = np.clip(f2i_sat((py - rf) / s), 0, gy)
    x1 = np.clip(f2i_sat((((px + rf) + s) - one) / s), 0, gx)
    y1 = np.clip(f2i_sat((((py + rf) + s) - one) / s), 0, gy)
    return x0, y0, x1, y1


def to_camera(xyz, extr, dtype=np.float32):
    """t = R p + t0 with extr read as the first 12 floats, row-major (project_point.cu:27-30)."""
    dt = np.dtype(dtype).type
    e = np.asarray(extr, dtype=dt).reshape(-1)[:12]
    x = np.asarray(xyz, dtype=dt)
    px, py, pz = x[:, 0], x[:, 1], x[:, 2]
    tx = ((e[0] * px + e[1] * py) + e[2] * pz) + e[3]
    ty = ((e[4] * px + e[5] * py) + e[6] * pz) + e[7]
    tz = ((e[8] * px + e[9] * py) + e[10] * pz) + e[11]
    return tx, ty, tz


# ---------------------------------------------------------------- projection

def project_forward(xyz, intr, extr, W, H, nearest=0.0, extent=1.3, dtype=np.float32):
    """project_point.cu:13-57.  Returns uv [N,2], depth [N]."""
    dt = np.dtype(dtype).type
    fx, fy, cx, cy = np.asarray(intr, dtype=dt)
    tx, ty, tz = to_camera(xyz, extr, dtype)
    with np.errstate(all="ignore"):
        if dt is np.float32:
            norm1 = (1.0 / (tz.astype(np.float64) + 1e-7)).astype(np.float32)
        else:
            norm1 = 1.0 / (tz + 1e-7)
        u = ((fx * tx) * norm1 + cx) - dt(0.5)
        v = ((fy * ty) * norm1 + cy) - dt(0.5)
    cull = np.zeros(tz.shape, bool)
    if nearest > 0:
        cull |= tz <= dt(nearest)
    if extent > 0:
        e = dt(extent)
        xmin = ((dt(1) - e) * dt(W)) * dt(0.5)
        xmax = ((dt(1) + e) * dt(W)) * dt(0.5)
        ymin = ((dt(1) - e) * dt(H)) * dt(0.5)
        ymax = ((dt(1) + e) * dt(H)) * dt(0.5)
        cull |= (u < xmin) | (u > xmax) | (v < ymin) | (v > ymax)
    zero = dt(0)
    uv = np.stack([np.where(cull, zero, u), np.where(cull, zero, v)], 1)
    depth = np.where(cull, zero, tz)
    return uv, depth


def project_backward(xyz, intr, extr, depth, duv, ddepth=None, dtype=np.float32):
    """project_point.cu:59-145.  Returns dxyz [N,3], dintr [4], dextr [12] (camera sums in fp64)."""
    dt = np.dtype(dtype).type
    fx, fy = np.asarray(intr, dtype=dt)[:2]
    e = np.asarray(extr, dtype=dt).reshape(-1)[:12]
    x = np.asarray(xyz, dtype=dt)
    px, py, pz = x[:, 0], x[:, 1], x[:, 2]
    tx, ty, tz = to_camera(xyz, extr, dtype)
    live = np.asarray(depth).reshape(-1) != 0
    gu = np.asarray(duv, dtype=dt)[:, 0]
    gv = np.asarray(duv, dtype=dt)[:, 1]
    gd = np.zeros_like(gu) if ddepth is None else np.asarray(ddepth, dtype=dt).reshape(-1)
    with np.errstate(all="ignore"):
        if dt is np.float32:
            norm1 = (1.0 / tz.astype(np.float64)).astype(np.float32)
            norm2 = (1.0 / (tz * tz).astype(np.float64)).astype(np.float32)
        else:
            norm1 = 1.0 / tz
            norm2 = 1.0 / (tz * tz)
        d0 = (((fx * (e[0] * tz - tx * e[8])) * norm2) * gu + ((fy * (e[4] * tz - ty * e[8])) * norm2) * gv) + e[8] * gd
        d1 = (((fx * (e[1] * tz - tx * e[9])) * norm2) * gu + ((fy * (e[5] * tz - ty * e[9])) * norm2) * gv) + e[9] * gd
        d2 = (((fx * (e[2] * tz - tx * e[10])) * norm2) * gu + ((fy * (e[6] * tz - ty * e[10])) * norm2) * gv) + e[10] * gd
        cam = np.zeros((16, len(gu)), dtype=dt)
        cam[0] = (tx * norm1) * gu
        cam[1] = (ty * norm1) * gv
        cam[2] = gu
        cam[3] = gv
        for k, p in enumerate((px, py, pz)):
            cam[4 + k] = ((fx * p) * norm1) * gu
            cam[8 + k] = ((fy * p) * norm1) * gv
            cam[12 + k] = ((((-fx) * p) * tx) * norm2) * gu + ((((-fy) * p) * ty) * norm2) * gv + p * gd
        cam[7] = (fx * norm1) * gu
        cam[11] = (fy * norm1) * gv
        cam[15] = ((((-fx) * tx) * norm2) * gu + (((-fy) * ty) * norm2) * gv) + gd
    zero = dt(0)
    dxyz = np.stack([np.where(live, d0, zero), np.where(live, d1, zero), np.where(live, d2, zero)], 1)
    cams = np.where(live[None], cam, zero).astype(np.float64).sum(1)
    return dxyz, cams[:4], cams[4:]


# ---------------------------------------------------------------- covariance

def quat_rot(q, dtype=np.float32):
    """Rotation entries rg[col][row] of compute_cov3d.cu:24-40 (quaternion w,x,y,z, unnormalised)."""
    dt = np.dtype(dtype).type
    q = np.asarray(q, dtype=dt)
    r, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    one, two = dt(1), dt(2)
    return [
        [one - two * (y * y + z * z), two * (x * y - r * z), two * (x * z + r * y)],
        [two * (x * y + r * z), one - two * (x * x + z * z), two * (y * z - r * x)],
        [two * (x * z - r * y), two * (y * z + r * x), one - two * (x * x + y * y)],
    ]


def cov3d_forward(scale, quat, visible=None, dtype=np.float32):
    """compute_cov3d.cu:42-58: Sigma = (S R)^T (S R) as [N,6] = (00,01,02,11,12,22)."""
    dt = np.dtype(dtype).type
    s = np.asarray(scale, dtype=dt)
    rg = quat_rot(quat, dtype)
    m = [[s[:, i] * rg[j][i] for i in range(3)] for j in range(3)]

    def dot(a, b):
        return (m[a][0] * m[b][0] + m[a][1] * m[b][1]) + m[a][2] * m[b][2]

    cov = np.stack([dot(0, 0), dot(1, 0), dot(2, 0), dot(1, 1), dot(2, 1), dot(2, 2)], 1)
    if visible is not None:
        cov = np.where(np.asarray(visible).reshape(-1, 1), cov, dt(0))
    return cov


def cov3d_backward(scale, quat, dcov, visible=None, dtype=np.float32):
    """compute_cov3d.cu:60-117.  Returns dscale [N,3], dquat [N,4]."""
    dt = np.dtype(dtype).type
    s = np.asarray(scale, dtype=dt)
    q = np.asarray(quat, dtype=dt)
    dc = np.asarray(dcov, dtype=dt)
    rg = quat_rot(q, dtype)
    two, half, four = dt(2), dt(0.5), dt(4)
    m2 = [[two * (s[:, i] * rg[j][i]) for i in range(3)] for j in range(3)]
    g = [[dc[:, 0], half * dc[:, 1], half * dc[:, 2]],
         [half * dc[:, 1], dc[:, 3], half * dc[:, 4]],
         [half * dc[:, 2], half * dc[:, 4], dc[:, 5]]]
    dm = [[(m2[0][i] * g[j][0] + m2[1][i] * g[j][1]) + m2[2][i] * g[j][2] for i in range(3)] for j in range(3)]
    ds = np.stack([(rg[0][c] * dm[0][c] + rg[1][c] * dm[1][c]) + rg[2][c] * dm[2][c] for c in range(3)], 1)
    P = [[dm[j][i] * s[:, i] for i in range(3)] for j in range(3)]
    r, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    dr = (two * z * (P[1][0] - P[0][1]) + two * y * (P[0][2] - P[2][0])) + two * x * (P[2][1] - P[1][2])
    dx = ((two * y * (P[0][1] + P[1][0]) + two * z * (P[0][2] + P[2][0])) + two * r * (P[2][1] - P[1][2])) - \
        four * x * (P[2][2] + P[1][1])
    dy = ((two * x * (P[0][1] + P[1][0]) + two * r * (P[0][2] - P[2][0])) + two * z * (P[2][1] + P[1][2])) - \
        four * y * (P[2][2] + P[0][0])
    dz = ((two * r * (P[1][0] - P[0][1]) + two * x * (P[0][2] + P[2][0])) + two * y * (P[2][1] + P[1][2])) - \
        four * z * (P[1][1] + P[0][0])
    dq = np.stack([dr, dx, dy, dz], 1)
    if visible is not None:
        vis = np.asarray(visible).reshape(-1, 1)
        ds = np.where(vis, ds, dt(0))
        dq = np.where(vis, dq, dt(0))
    return ds, dq


# ---------------------------------------------------------------- EWA

def _ewa_cov2d(xyz, cov3d, intr, extr, dtype):
    dt = np.dtype(dtype).type
    fx, fy = np.asarray(intr, dtype=dt)[:2]
    e = np.asarray(extr, dtype=dt).reshape(-1)[:12]
    tx, ty, tz = to_camera(xyz, extr, dtype)
    v = np.asarray(cov3d, dtype=dt)
    tz2 = tz * tz
    j00 = fx / tz
    j11 = fy / tz
    j20 = -(fx * tx) / tz2
    j21 = -(fy * ty) / tz2
    T = [[j00 * e[j] + j20 * e[8 + j], j11 * e[4 + j] + j21 * e[8 + j]] for j in range(3)]
    V = [[v[:, 0], v[:, 1], v[:, 2]], [v[:, 1], v[:, 3], v[:, 4]], [v[:, 2], v[:, 4], v[:, 5]]]
    tv = [[(T[0][i] * V[k][0] + T[1][i] * V[k][1]) + T[2][i] * V[k][2] for i in range(2)] for k in range(3)]
    c00 = (tv[0][0] * T[0][0] + tv[1][0] * T[1][0]) + tv[2][0] * T[2][0]
    c01 = (tv[0][1] * T[0][0] + tv[1][1] * T[1][0]) + tv[2][1] * T[2][0]
    c11 = (tv[0][1] * T[0][1] + tv[1][1] * T[1][1]) + tv[2][1] * T[2][1]
    cx = c00 + dt(0.3)
    cy = c01
    cz = c11 + dt(0.3)
    det = cx * cz - cy * cy
    return dict(t=(tx, ty, tz), J=(j00, j11, j20, j21), T=T, V=V, cov=(cx, cy, cz), det=det, v=v, e=e,
                fx=fx, fy=fy)


def ewa_forward(xyz, cov3d, intr, extr, uv, W, H, visible=None, dtype=np.float32, raw=False):
    """ewa_project.cu:16-83.  Returns conic [N,3], radius [N] int32, tiles [N] int32.

    ``raw=True`` returns the conic before the det/area/visibility masks.
    """
    dt = np.dtype(dtype).type
    n = len(xyz)
    with np.errstate(all="ignore"):
        w = _ewa_cov2d(xyz, cov3d, intr, extr, dtype)
        cx, cy, cz = w["cov"]
        det = w["det"]
        mid = dt(0.5) * (cx + cz)
        disc = np.fmax(dt(0.1), mid * mid - det)
        l1 = mid + np.sqrt(disc)
        l2 = mid - np.sqrt(disc)
        rf = np.ceil(dt(3) * np.sqrt(np.fmax(l1, l2)))
        r = f2i_sat(rf)
        gx, gy = (W + TILE - 1) // TILE, (H + TILE - 1) // TILE
        uvv = np.asarray(uv, dtype=dt)
        x0, y0, x1, y1 = tile_rect(uvv[:, 0], uvv[:, 1], r, gx, gy, dtype)
        area = (x1 - x0) * (y1 - y0)
        det_inv = dt(1) / det
        conic = np.stack([cz * det_inv, -cy * det_inv, cx * det_inv], 1)
    if raw:
        return conic
    ok = (det != 0) & (area != 0)
    if visible is not None:
        ok &= np.asarray(visible).reshape(-1).astype(bool)
    conic = np.where(ok[:, None], conic, dt(0))
    radius = np.where(ok, r, 0).astype(np.int32)
    tiles = np.where(ok, area, 0).astype(np.int32)
    assert conic.shape == (n, 3)
    return conic, radius, tiles


def ewa_backward(xyz, cov3d, intr, extr, radius, dconic, dtype=np.float32):
    """ewa_project.cu:85-252.  Returns dxyz [N,3], dcov3d [N,6], dintr [4], dextr [12]."""
    dt = np.dtype(dtype).type
    two = dt(2)
    x = np.asarray(xyz, dtype=dt)
    px, py, pz = x[:, 0], x[:, 1], x[:, 2]
    gc = np.asarray(dconic, dtype=dt)
    with np.errstate(all="ignore"):
        w = _ewa_cov2d(xyz, cov3d, intr, extr, dtype)
        cx, cy, cz = w["cov"]
        det = w["det"]
        T, v, e = w["T"], w["v"], w["e"]
        fx, fy = w["fx"], w["fy"]
        tx, ty, tz = w["t"]
        j00, j11, j20, j21 = w["J"]
        nom = dt(1) / (det * det)
        dcx = nom * (((-cz) * cz * gc[:, 0] + (cy * cz) * gc[:, 1]) + (det - cx * cz) * gc[:, 2])
        dcy = nom * (((two * cy) * cz * gc[:, 0] - (det + (two * cy) * cy) * gc[:, 1]) + ((two * cx) * cy) * gc[:, 2])
        dcz = nom * (((det - cx * cz) * gc[:, 0] + (cx * cy) * gc[:, 1]) - (cx * cx) * gc[:, 2])
        dcov = np.stack([
            ((T[0][0] * T[0][0]) * dcx + (T[0][0] * T[0][1]) * dcy) + (T[0][1] * T[0][1]) * dcz,
            (((two * T[0][0]) * T[1][0]) * dcx + (T[0][0] * T[1][1] + T[0][1] * T[1][0]) * dcy) +
            ((two * T[0][1]) * T[1][1]) * dcz,
            (((two * T[0][0]) * T[2][0]) * dcx + (T[0][0] * T[2][1] + T[0][1] * T[2][0]) * dcy) +
            ((two * T[0][1]) * T[2][1]) * dcz,
            ((T[1][0] * T[1][0]) * dcx + (T[1][0] * T[1][1]) * dcy) + (T[1][1] * T[1][1]) * dcz,
            (((two * T[1][0]) * T[2][0]) * dcx + (T[1][0] * T[2][1] + T[1][1] * T[2][0]) * dcy) +
            ((two * T[1][1]) * T[2][1]) * dcz,
            ((T[2][0] * T[2][0]) * dcx + (T[2][0] * T[2][1]) * dcy) + (T[2][1] * T[2][1]) * dcz,
        ], 1)
        vi = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
        dT = []
        for j in range(3):
            a0 = (T[0][0] * v[:, vi[j][0]] + T[1][0] * v[:, vi[j][1]]) + T[2][0] * v[:, vi[j][2]]
            a1 = (T[0][1] * v[:, vi[j][0]] + T[1][1] * v[:, vi[j][1]]) + T[2][1] * v[:, vi[j][2]]
            dT.append(((two * a0) * dcx + a1 * dcy, a0 * dcy + (two * a1) * dcz))
        dj00 = (e[0] * dT[0][0] + e[1] * dT[1][0]) + e[2] * dT[2][0]
        dj20 = (e[8] * dT[0][0] + e[9] * dT[1][0]) + e[10] * dT[2][0]
        dj11 = (e[4] * dT[0][1] + e[5] * dT[1][1]) + e[6] * dT[2][1]
        dj21 = (e[8] * dT[0][1] + e[9] * dT[1][1]) + e[10] * dT[2][1]
        itz = dt(1) / tz
        tz2 = itz * itz
        tz3 = tz2 * itz
        dtx = ((-fx) * tz2) * dj20
        dty = ((-fy) * tz2) * dj21
        dtz = ((((-fx) * tz2) * dj00 - (fy * tz2) * dj11) + (((two * fx) * tx) * tz3) * dj20) + \
            (((two * fy) * ty) * tz3) * dj21
        cam = np.zeros((16, len(px)), dtype=dt)
        cam[0] = itz * dj00 + ((-tx) * tz2) * dj20
        cam[1] = itz * dj11 + ((-ty) * tz2) * dj21
        for k, p in enumerate((px, py, pz)):
            cam[4 + k] = j00 * dT[k][0] + p * dtx
            cam[8 + k] = j11 * dT[k][1] + p * dty
            cam[12 + k] = (j20 * dT[k][0] + j21 * dT[k][1]) + p * dtz
        cam[7] = dtx
        cam[11] = dty
        cam[15] = dtz
        dxyz = np.stack([(e[0] * dtx + e[4] * dty) + e[8] * dtz,
                         (e[1] * dtx + e[5] * dty) + e[9] * dtz,
                         (e[2] * dtx + e[6] * dty) + e[10] * dtz], 1)
    ok = (np.asarray(radius).reshape(-1) > 0) & (det != 0)
    zero = dt(0)
    dxyz = np.where(ok[:, None], dxyz, zero)
    dcov = np.where(ok[:, None], dcov, zero)
    cams = np.where(ok[None], cam, zero).astype(np.float64).sum(1)
    return dxyz, dcov, cams[:4], cams[4:]
