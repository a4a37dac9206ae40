"""Brute-force restatements used to pin the oracle (pure Python / scipy, one
voxel at a time).  Shared by the oracle pins and the golden-fixture checks.

* raymarch_bruteforce: scalar march with scipy map_coordinates(order=1,
  mode='grid-constant') as the zero-padded trilinear interpolant.
* smooth_bruteforce: the lateral mean with real-number neighbour coordinates
  and exact half-voxel ties to the even index.
* pb_bruteforce: the double-Gaussian pencil beam with np.interp for the LUT
  and a cross product for the distance to the spot axis.
"""

import math

import numpy as np
from scipy import ndimage

from oracle import impt_oracle as O


def raymarch_bruteforce(rho, s, g, vox):
    """Scalar march with scipy's zero-padded trilinear; CUDA axes (x = axis 2)."""
    nz, ny, nx = rho.shape
    k, j, i = vox
    iso = g[0:3]
    src = g[3:6]
    v = np.array([i * s, j * s, k * s]) - iso
    u = (src - v) / np.linalg.norm(src - v)
    wet = -0.05
    for step in range(0, 2000):
        p = v + u * step
        c = (p + iso) / s
        if not (0 <= c[0] < nx and 0 <= c[1] < ny and 0 <= c[2] < nz):
            continue
        val = ndimage.map_coordinates(rho.astype(np.float64), [[c[2]], [c[1]], [c[0]]], order=1,
                                      mode="grid-constant", cval=0.0)[0]
        wet += max(val, 0.0) * 0.1
    return wet


def round_tie_even(c):
    """Nearest integer; half-integers (to 1e-6: exact ties of the real-number
    coordinate) go to the even neighbour, like Python's round() on exact ties."""
    f = math.floor(c)
    if abs(c - f - 0.5) < 1e-6:
        return round(f + 0.5)
    return round(c)


def smooth_bruteforce(raw, s, g):
    nz, ny, nx = raw.shape
    out = np.empty_like(raw)
    iso = g[0:3]
    sg, cg, sc, cc = g[6:10]
    angles = [(1.0, 0.0), (0.5, math.sqrt(3) / 2), (-0.5, math.sqrt(3) / 2), (-1.0, 0.0),
              (-0.5, -math.sqrt(3) / 2), (0.5, -math.sqrt(3) / 2)]
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                wc = raw[k, j, i]
                p = np.array([i * s, j * s, k * s]) - iso
                h = O.image_to_head(g, p)
                tot, cnt = wc, 1
                for ca, sa in angles:
                    for dr in range(1, 11):
                        if not dr < min(wc * 10.0, 10.0):
                            continue
                        hx, hy, hz = h[0] + dr * ca, h[1] + dr * sa, h[2]
                        # head -> image: undo the axis swap, gantry, couch
                        xz, yz, zz = -hx, hz, hy
                        xg, yg = xz * cg - yz * sg, xz * sg + yz * cg
                        x, y, z = xg * cc + zz * sc, yg, -xg * sc + zz * cc
                        ii, jj, kk = (round_tie_even(c) for c in ((x + iso[0]) / s, (y + iso[1]) / s,
                                                                   (z + iso[2]) / s))
                        if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                            tot += raw[kk, jj, ii]
                            cnt += 1
                out[k, j, i] = tot / cnt
    return out


def pb_bruteforce(wet_cm, s, g, spots_sorted, layers, r80, marr):
    nz, ny, nx = wet_cm.shape
    iso, src = g[0:3], g[3:6]
    vx, vy = g[10], g[11]
    out = np.zeros_like(wet_cm)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                p = np.array([i * s, j * s, k * s]) - iso
                h = O.image_to_head(g, p)
                dsrc = np.linalg.norm(src - p)
                w_mm = wet_cm[k, j, i] * 10.0
                tot = 0.0
                for (st, ns, eid), R in zip(layers, r80):
                    if w_mm > 1.1 * R:
                        continue
                    dep = marr.depths[eid].astype(np.float64)
                    idd = np.interp(w_mm, dep, marr.idds[eid].astype(np.float64))
                    sms = np.interp(w_mm, dep, marr.sigmas[eid].astype(np.float64))
                    A, B, C = (float(c) for c in marr.dvp[eid, 2:5])
                    d = dsrc - w_mm + 0.7 * R
                    sig_t = A * d * d + B * d + C + sms
                    w = min(max(w_mm, 0.1), R - 0.1)
                    sn = max(2.85 + 0.0014 * R * math.log(w + 3) + 0.06 * w - 7.4e-5 * w * w
                             - 0.22 * R / (w - R - 5) ** 2, 0.0)
                    un = 0.052 * math.log(1.13 + w / (11.2 - 0.023 * R)) \
                        + 0.35 * (0.0017 * R * R - R) / ((R + 3) ** 2 - w * w) - 1.61e-9 * w * (R + 3) ** 2
                    un = min(max(un, 0.0), 0.9)
                    sig_h = math.hypot(sig_t, sn)
                    for q in range(st, st + ns):
                        xs, ys, mu = spots_sorted[q]
                        t = np.array([xs / vx, ys / vy, -1.0])
                        r = np.array([h[0] - xs, h[1] - ys, h[2]])
                        d2 = np.dot(np.cross(r, t), np.cross(r, t)) / np.dot(t, t)
                        pd = (1 - un) * idd / (2 * math.pi * sig_t ** 2 + 1e-8) * math.exp(-d2 / (2 * sig_t ** 2))
                        hd = un * idd / (2 * math.pi * sig_h ** 2 + 1e-8) * math.exp(-d2 / (2 * sig_h ** 2))
                        tot += mu * (pd + hd)
                out[k, j, i] = tot
    return out
