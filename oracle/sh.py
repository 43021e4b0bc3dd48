"""Real spherical harmonics, degree 0..10, on the CPU (test infrastructure only).

Restates msplat/src/compute_sh.cu:116-1694.  Each reference basis function is
K_lm * {Re, Im}(x + i y)^|m| * Q_l|m|(z, rho2), with Q the associated-Legendre
polynomial (-1)^m d^m P_l/dz^m, rho2 = x^2+y^2+z^2 in bands 0..3 and rho2 = 1
in bands 4..10 (the polynomial forms the reference writes out; their
partials are the reference's dL/ddir).  Evaluated with the Legendre and
complex-power recurrences; fp32 mode uses the kernel's operation order.
"""

from __future__ import annotations

import math

import numpy as np

MAX_DEG = 10


def knorm(l: int, m: int) -> float:
    k = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
    return k * math.sqrt(2.0) if m else k


def qmm(m: int) -> float:
    v = 1.0
    for i in range(1, m + 1):
        v *= -(2.0 * i - 1.0)
    return v


def degree_of(D: int) -> int:
    for l in range(MAX_DEG + 1):
        if (l + 1) ** 2 == D:
            return l
    raise ValueError(f"D={D} is not (deg+1)^2 with deg <= {MAX_DEG}")


def basis(dirs, deg: int, dtype=np.float32, grad: bool = False):
    """Y [N, D] (and dY [N, D, 3] if grad) at the directions dirs [N,3].

    Returns also the emission order (the kernel accumulates dL/ddir in it).
    """
    dt = np.dtype(dtype).type
    d = np.asarray(dirs, dtype=dt)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    n = len(d)
    D = (deg + 1) ** 2
    Y = np.zeros((n, D), dtype=dt)
    dY = np.zeros((n, D, 3), dtype=dt) if grad else None
    order = []
    cr = [np.ones(n, dt)]
    ci = [np.zeros(n, dt)]
    for m in range(1, deg + 1):
        cr.append(x * cr[m - 1] - y * ci[m - 1])
        ci.append(x * ci[m - 1] + y * cr[m - 1])
    rho2 = (x * x + y * y) + z * z
    zero = np.zeros(n, dt)
    for m in range(deg + 1):
        q1 = q2 = d1 = d2 = zero
        for l in range(m, deg + 1):
            if l == m:
                q = np.full(n, dt(qmm(m)))
                dq = zero
            elif l == m + 1:
                q = dt(2 * m + 1) * z * q1
                dq = dt(2 * m + 1) * q1
            else:
                a = dt(2 * l - 1) / dt(l - m)
                b = dt(l + m - 1) / dt(l - m)
                q = a * z * q1 - b * q2
                dq = a * (q1 + z * d1) - b * d2
            qv, qx, qy, qz = q, zero, zero, dq
            if l == 2 and m == 0:
                qv = dt(1.5) * z * z - dt(0.5) * rho2
                qx, qy, qz = -x, -y, dt(2) * z
            elif l == 3 and m == 1:
                qv = dt(-7.5) * z * z + dt(1.5) * rho2
                qx, qy, qz = dt(3) * x, dt(3) * y, dt(-12) * z
            elif l == 3 and m == 0:
                qv = dt(2.5) * z * z * z - dt(1.5) * z * rho2
                qx, qy, qz = dt(-3) * x * z, dt(-3) * y * z, dt(4.5) * z * z - dt(1.5) * rho2
            K = dt(knorm(l, m))
            if m == 0:
                k = l * l + l
                Y[:, k] = K * qv
                if grad:
                    dY[:, k] = np.stack([K * qx, K * qy, K * qz], 1)
                order.append(k)
            else:
                fm = dt(m)
                k = l * l + l + m
                Y[:, k] = K * (cr[m] * qv)
                if grad:
                    dY[:, k] = np.stack([K * (fm * cr[m - 1] * qv + cr[m] * qx),
                                         K * ((-fm) * ci[m - 1] * qv + cr[m] * qy),
                                         K * (cr[m] * qz)], 1)
                order.append(k)
                k = l * l + l - m
                Y[:, k] = K * (ci[m] * qv)
                if grad:
                    dY[:, k] = np.stack([K * (fm * ci[m - 1] * qv + ci[m] * qx),
                                         K * (fm * cr[m - 1] * qv + ci[m] * qy),
                                         K * (ci[m] * qz)], 1)
                order.append(k)
            q2, q1, d2, d1 = q1, q, d1, dq
    return Y, dY, order


def sh_forward(shs, dirs, visible=None, dtype=np.float32):
    """compute_sh.cu:1600-1637: value [N, C]."""
    dt = np.dtype(dtype).type
    s = np.asarray(shs, dtype=dt)
    n, C, D = s.shape
    Y, _, _ = basis(dirs, degree_of(D), dtype)
    out = np.zeros((n, C), dtype=dt)
    for c in range(C):
        acc = np.zeros(n, dt)
        for k in range(D):
            acc = acc + Y[:, k] * s[:, c, k]
        out[:, c] = acc
    if visible is not None:
        out = np.where(np.asarray(visible).reshape(-1, 1), out, dt(0))
    return out


def sh_backward(shs, dirs, dvalue, visible=None, dtype=np.float32):
    """compute_sh.cu:1639-1694: dL/dshs [N,C,D] and dL/ddirs [N,3]."""
    dt = np.dtype(dtype).type
    s = np.asarray(shs, dtype=dt)
    g = np.asarray(dvalue, dtype=dt)
    n, C, D = s.shape
    Y, dY, order = basis(dirs, degree_of(D), dtype, grad=True)
    dshs = Y[:, None, :] * g[:, :, None]
    # dL/ddir = sum_c sum_k (g_c shs[c,k]) grad Y_k: per channel in basis
    # order, channels summed first to last (the kernels' rounding order)
    dd = np.zeros((n, 3), dt)
    for c in range(C):
        part = np.zeros((n, 3), dt)
        for k in order:
            part = part + (g[:, c:c + 1] * s[:, c, k:k + 1]) * dY[:, k]
        dd = dd + part
    if visible is not None:
        vis = np.asarray(visible).reshape(-1)
        dshs = np.where(vis[:, None, None], dshs, dt(0))
        dd = np.where(vis[:, None], dd, dt(0))
    return dshs, dd


def view_dirs_from_points(xyz, campos, dtype=np.float32):
    """Fused view directions: v = xyz - campos, r = sqrt((vx vx + vy vy) + vz vz),
    d = v / r, in the kernel's operation order.  No reference counterpart (the
    reference's compute_sh takes the directions: compute_sh.py:8-33)."""
    dt = np.dtype(dtype).type
    v = np.asarray(xyz, dtype=dt) - np.asarray(campos, dtype=dt).reshape(1, 3)
    r = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
    return v / r[:, None], r


def sh_points_forward(shs, xyz, campos, visible=None, dtype=np.float32):
    """compute_sh of the fused view directions: value [N, C]."""
    d, _ = view_dirs_from_points(xyz, campos, dtype)
    return sh_forward(shs, d, visible, dtype)


def sh_points_backward(shs, xyz, campos, dvalue, visible=None, dtype=np.float32):
    """dL/dshs [N,C,D] and dL/dxyz [N,3]: the SH adjoint's dL/ddir g taken through
    the normalisation, (g - d (d . g)) / r, dot product summed x, y, z."""
    dt = np.dtype(dtype).type
    d, r = view_dirs_from_points(xyz, campos, dtype)
    dshs, g = sh_backward(shs, d, dvalue, visible, dtype)
    dot = (d[:, 0] * g[:, 0] + d[:, 1] * g[:, 1]) + d[:, 2] * g[:, 2]
    dxyz = (g - d * dot[:, None]) / r[:, None]
    if visible is not None:
        dxyz = np.where(np.asarray(visible).reshape(-1)[:, None], dxyz, dt(0))
    return dshs, dxyz.astype(dt)
