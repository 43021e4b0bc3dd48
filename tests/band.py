"""Finite differences of the squared-error loss of a rendered image, one
Gaussian at a time, on the CPU oracle in fp64 (test infrastructure).

A Gaussian appears only in the lists of the tiles its box covers, so the
derivative of the whole image's loss with respect to its parameters is the
derivative of the loss over those tiles.  ``WindowFD`` re-blends just that
window with the Gaussian's parameters perturbed: projection, covariance,
EWA and its SH colour (and depth channel) recomputed in fp64 for that one
Gaussian, every other Gaussian and every tile list held fixed (the fp32
discrete decisions, as the GPU takes them).
"""

from __future__ import annotations

import numpy as np

from oracle import blend, geometry, sh

TILE = 16


class WindowFD:
    def __init__(self, geo64, opacity, feature, idx, rng, W, H, target, bg, intr, extr, campos=None, shs=None,
                 depth_channel=False):
        """geo64: stage_geometry(..., dtype=float64); feature [N, C] fp64 (SH
        colours first when shs is given, the depth last when depth_channel);
        idx / rng: tile lists (full or band) of the fp32 decisions."""
        self.uv, self.conic = geo64["uv"], geo64["conic"]
        self.uv32, self.radius = geo64["uv32"], geo64["radius"]
        self.opacity = np.asarray(opacity, np.float64).reshape(-1)
        self.feature = np.asarray(feature, np.float64)
        self.idx, self.rng = idx, rng
        self.W, self.H = W, H
        self.gx, self.gy = (W + TILE - 1) // TILE, (H + TILE - 1) // TILE
        self.target = np.asarray(target, np.float64)
        self.bg = bg
        self.intr, self.extr = intr, extr
        self.campos, self.shs = campos, shs
        self.depth_channel = depth_channel

    def rect(self, g):
        x0, y0, x1, y1 = geometry.tile_rect(self.uv32[g:g + 1, 0], self.uv32[g:g + 1, 1], self.radius[g:g + 1],
                                            self.gx, self.gy, np.float32)
        return int(x0[0]), int(y0[0]), int(x1[0]), int(y1[0])

    def window(self, g):
        """The compact window of Gaussian g: its tiles' lists over the union of their Gaussians."""
        x0, y0, x1, y1 = self.rect(g)
        lists = []
        for ty in range(y0, y1):
            for tx in range(x0, x1):
                s, e = self.rng[ty * self.gx + tx]
                lists.append(self.idx[s:e])
        members = np.unique(np.concatenate(lists)) if lists else np.zeros(0, np.int64)
        cidx = np.concatenate([np.searchsorted(members, l) for l in lists]).astype(np.int32)
        lens = np.array([len(l) for l in lists], np.int64)
        starts = np.cumsum(lens) - lens
        crng = np.stack([starts, starts + lens], 1).astype(np.int32)
        ox, oy = TILE * x0, TILE * y0
        Wn, Hn = min(TILE * x1, self.W) - ox, min(TILE * y1, self.H) - oy
        return dict(g=g, members=members, me=int(np.searchsorted(members, g)), cidx=cidx, crng=crng, ox=ox, oy=oy,
                    Wn=Wn, Hn=Hn, target=self.target[:, oy:oy + Hn, ox:ox + Wn])

    def params(self, g, xyz, scale, rotate, opacity):
        """Gaussian g's parameter vector: xyz, scale, rotate, opacity, then the
        DC SH coefficient of each SH channel (when SH is rendered)."""
        p = [xyz[g], scale[g], rotate[g], [float(np.asarray(opacity).reshape(-1)[g])]]
        if self.shs is not None:
            p.append(self.shs[g, :, 0])
        return np.concatenate([np.asarray(v, np.float64) for v in p])

    def loss(self, win, p):
        """The window's loss with Gaussian g at parameter vector p."""
        g, m = win["g"], win["members"]
        uv = self.uv[m] - np.array([win["ox"], win["oy"]], np.float64)
        conic = self.conic[m].copy()
        op = self.opacity[m].copy()
        feat = self.feature[m].copy()
        xyz, scale, rot, o = p[None, 0:3], p[None, 3:6], p[None, 6:10], p[10]
        u, d = geometry.project_forward(xyz, self.intr, self.extr, self.W, self.H, nearest=0.0, extent=0.0,
                                        dtype=np.float64)
        cov = geometry.cov3d_forward(scale, rot, None, dtype=np.float64)
        c = geometry.ewa_forward(xyz, cov, self.intr, self.extr, u, self.W, self.H, None, dtype=np.float64, raw=True)
        k = win["me"]
        uv[k] = u[0] - np.array([win["ox"], win["oy"]], np.float64)
        conic[k] = c[0]
        op[k] = o
        if self.shs is not None:
            s = self.shs[g:g + 1].astype(np.float64).copy()
            s[0, :, 0] = p[11:11 + s.shape[1]]
            feat[k, :s.shape[1]] = sh.sh_points_forward(s, xyz, self.campos, None, dtype=np.float64)[0]
        if self.depth_channel:
            feat[k, -1] = d[0]
        out, _, _, _ = blend.blend_forward(uv, conic, op, feat, win["cidx"], win["crng"], self.bg, win["Wn"],
                                           win["Hn"], dtype=np.float64)
        return float(((out - win["target"]) ** 2).sum())

    def directional(self, g, p, v):
        """Central differences of the window loss along v at steps v and v/4:
        (derivative, derivative at the smaller step)."""
        win = self.window(g)
        d1 = (self.loss(win, p + v) - self.loss(win, p - v)) / 2.0
        d4 = (self.loss(win, p + v / 4) - self.loss(win, p - v / 4)) * 2.0
        return d1, d4


def sh_scene_forward(xyz, scale, rotate, opacity, shs, intr, extr, campos, W, H, bg, extra=None,
                     depth_channel=False, rows=None, dtype=np.float64):
    """The oracle's rasterization_sh: features = [SH colours of shs seen from
    campos, extra, depth], rendered over the full sort's tile lists, or over
    tile rows [rows[0], rows[1]) only (the rest of the image then holds bg).
    Returns (image, state for sh_scene_backward)."""
    from oracle import pipeline, sort
    geo = pipeline.stage_geometry(xyz, scale, rotate, intr, extr, W, H, dtype=dtype)
    uv32 = geo.get("uv32", geo["uv"])
    d32 = geo.get("depth32", geo["depth"])
    vis = geo["radius"] > 0
    if rows is not None:  # only the Gaussians whose tile box meets the band enter a list
        gx, gy = (W + 15) // 16, (H + 15) // 16
        _, y0, _, y1 = geometry.tile_rect(uv32[:, 0], uv32[:, 1], geo["radius"], gx, gy, np.float32)
        vis = vis & (y0 < rows[1]) & (y1 > rows[0])
    sub = np.nonzero(vis)[0]
    colours = np.zeros((len(xyz), shs.shape[1]), dtype)
    colours[sub] = sh.sh_points_forward(shs[sub], xyz[sub], campos, None, dtype=dtype)
    parts = [colours]
    if extra is not None:
        parts.append(np.asarray(extra, dtype))
    if depth_channel:
        parts.append(np.asarray(geo["depth"], dtype).reshape(-1, 1))
    feat = np.concatenate(parts, 1)
    if rows is None:
        idx, rng, _ = sort.sort_gaussian(uv32, d32, geo["radius"], W, H)
    else:
        idx, rng = sort.band_lists(uv32, d32, geo["radius"], W, H, *rows)
    out, fT, nc, _ = blend.blend_forward(geo["uv"], geo["conic"], opacity, feat, idx, rng, bg, W, H, dtype=dtype)
    return out, dict(geo=geo, feat=feat, idx=idx, rng=rng, final_T=fT, ncontrib=nc, vis=vis, sub=sub)


def sh_scene_backward(xyz, scale, rotate, opacity, shs, intr, extr, campos, W, H, bg, state, dout, extra=None,
                      depth_channel=False, dtype=np.float64):
    """Adjoint of sh_scene_forward for dL/dimage = dout: dict of dxyz,
    dscale, drotate, dopacity, dshs (and dextra)."""
    from oracle import pipeline
    geo, feat = state["geo"], state["feat"]
    duv, dconic, dop, dfeat = blend.blend_backward(geo["uv"], geo["conic"], opacity, feat, state["idx"],
                                                   state["rng"], bg, W, H, state["final_T"], state["ncontrib"],
                                                   dout, dtype=dtype)
    Csh = shs.shape[1]
    E = 0 if extra is None else extra.shape[1]
    ddepth = dfeat[:, -1] if depth_channel else None
    dxyz, dscale, drot, _, _ = pipeline.stage_geometry_backward(xyz, scale, rotate, intr, extr, geo, duv, dconic,
                                                                ddepth, dtype=dtype)
    sub = state["sub"]
    dshs = np.zeros(shs.shape, dtype)
    dxyz_sh = np.zeros((len(xyz), 3), dtype)
    dshs[sub], dxyz_sh[sub] = sh.sh_points_backward(shs[sub], xyz[sub], campos, dfeat[sub, :Csh], None, dtype=dtype)
    out = dict(dxyz=dxyz + dxyz_sh, dscale=dscale, drotate=drot, dopacity=dop, dshs=dshs)
    if E:
        out["dextra"] = dfeat[:, Csh:Csh + E]
    return out
