"""Seeded synthetic phantoms: sums of random ellipses (2D) and ellipsoids (3D).

Random numbers come from splitmix64 (Steele, Lea and Flood, "Fast splittable pseudorandom
number generators", OOPSLA 2014): state_i = seed + i * 0x9E3779B97F4A7C15, output the
state through the published mix; doubles take the top 53 bits. The same stream is easy to
reproduce in C or C++.

2D recipe (n x n, pixel spacing 1), per ellipse in draw order: centre x, centre y ~
U(-0.8, 0.8) * n/2; semi-axes a, b ~ U(2, 20) * n/64; rotation ~ U(0, pi); density ~
U(0, 1). 3D recipe (n^3): centre x, y, z ~ U(-0.8, 0.8) * n/2; semi-axes ~ U(8, 80) * n/256;
Euler angles (z, y, z) ~ U(0, pi); density ~ U(0, 1). Densities add where shapes overlap;
each pixel / voxel centre is point-sampled.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int) -> np.ndarray:
    """n doubles in [0, 1)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def ellipse_params(n: int, seed: int = 0, count: int = 10) -> np.ndarray:
    """(count, 6): cx, cy, a, b, angle, density in world units (pixel spacing 1)."""
    u = uniform(seed, 6 * count).reshape(count, 6)
    h = n / 2.0
    s = n / 64.0
    return np.stack([
        (-0.8 + 1.6 * u[:, 0]) * h, (-0.8 + 1.6 * u[:, 1]) * h,
        (2.0 + 18.0 * u[:, 2]) * s, (2.0 + 18.0 * u[:, 3]) * s,
        np.pi * u[:, 4], u[:, 5],
    ], axis=1)


def ellipses_2d(n: int, seed: int = 0, count: int = 10, params=None) -> np.ndarray:
    """n x n float64 image; pixel (iy, ix) centre at (ix - n/2 + .5, iy - n/2 + .5)."""
    p = ellipse_params(n, seed, count) if params is None else params
    c = np.arange(n, dtype=np.float64) - n / 2.0 + 0.5
    X, Y = np.meshgrid(c, c)
    img = np.zeros((n, n))
    for cx, cy, a, b, th, rho in p:
        ct, st = np.cos(th), np.sin(th)
        dx, dy = X - cx, Y - cy
        u = (dx * ct + dy * st) / a
        v = (-dx * st + dy * ct) / b
        img[u * u + v * v <= 1.0] += rho
    return img


def ellipsoid_params(n: int, seed: int = 0, count: int = 20) -> np.ndarray:
    """(count, 10): cx, cy, cz, a, b, c, alpha, beta, gamma, density."""
    u = uniform(seed, 10 * count).reshape(count, 10)
    h = n / 2.0
    s = n / 256.0
    return np.stack([
        (-0.8 + 1.6 * u[:, 0]) * h, (-0.8 + 1.6 * u[:, 1]) * h, (-0.8 + 1.6 * u[:, 2]) * h,
        (8.0 + 72.0 * u[:, 3]) * s, (8.0 + 72.0 * u[:, 4]) * s, (8.0 + 72.0 * u[:, 5]) * s,
        np.pi * u[:, 6], np.pi * u[:, 7], np.pi * u[:, 8], u[:, 9],
    ], axis=1)


def _rot_zyz(al, be, ga):
    def rz(t):
        c, s = np.cos(t), np.sin(t)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])

    def ry(t):
        c, s = np.cos(t), np.sin(t)
        return np.array([[c, 0, s], [0, 1.0, 0], [-s, 0, c]])

    return rz(al) @ ry(be) @ rz(ga)


def ellipsoids_3d(n: int, seed: int = 0, count: int = 20, params=None,
                  dtype=np.float32) -> np.ndarray:
    """(n, n, n) volume in (z, y, x) order; voxel centres at index - n/2 + .5. Each shape is
    evaluated inside its bounding box only."""
    p = ellipsoid_params(n, seed, count) if params is None else params
    c = np.arange(n, dtype=np.float64) - n / 2.0 + 0.5
    vol = np.zeros((n, n, n), dtype=dtype)
    for cx, cy, cz, a, b, cc, al, be, ga, rho in p:
        R = _rot_zyz(al, be, ga)
        r = max(a, b, cc)
        sl = []
        for ctr in (cz, cy, cx):
            lo = int(np.clip(np.floor(ctr - r + n / 2.0 - 0.5), 0, n))
            hi = int(np.clip(np.ceil(ctr + r + n / 2.0 - 0.5) + 1, 0, n))
            sl.append(slice(lo, hi))
        if any(s.stop <= s.start for s in sl):
            continue
        Z, Y, X = np.meshgrid(c[sl[0]], c[sl[1]], c[sl[2]], indexing="ij")
        d = np.stack([X - cx, Y - cy, Z - cz], axis=-1)
        q = d @ R  # coordinates in the ellipsoid frame
        inside = (q[..., 0] / a) ** 2 + (q[..., 1] / b) ** 2 + (q[..., 2] / cc) ** 2 <= 1.0
        vol[sl[0], sl[1], sl[2]][inside] += rho
    return vol


def ellipse_line_integrals(params: np.ndarray, p0: np.ndarray, d: np.ndarray) -> np.ndarray:
    """Exact line integrals of the ellipse phantom along rays p0 + t d (world units, d of any
    length; the integral is over arc length). p0, d: (..., 2). Used as an analytic check of
    the forward projector's scale."""
    p0 = np.asarray(p0, dtype=np.float64)
    d = np.asarray(d, dtype=np.float64)
    L = np.linalg.norm(d, axis=-1)
    dn = d / L[..., None]
    out = np.zeros(p0.shape[:-1])
    for cx, cy, a, b, th, rho in params:
        ct, st = np.cos(th), np.sin(th)
        qx, qy = p0[..., 0] - cx, p0[..., 1] - cy
        ux = (qx * ct + qy * st) / a
        uy = (-qx * st + qy * ct) / b
        ex = (dn[..., 0] * ct + dn[..., 1] * st) / a
        ey = (-dn[..., 0] * st + dn[..., 1] * ct) / b
        A = ex * ex + ey * ey
        B = 2 * (ux * ex + uy * ey)
        Cc = ux * ux + uy * uy - 1
        disc = B * B - 4 * A * Cc
        chord = np.where(disc > 0, np.sqrt(np.maximum(disc, 0)) / A, 0.0)
        out += rho * chord
    return out


def ellipsoid_line_integrals(params: np.ndarray, p0: np.ndarray, d: np.ndarray) -> np.ndarray:
    """Exact line integrals (over arc length) of the ellipsoid phantom along rays p0 + t d,
    world coordinates (x, y, z). p0, d: (..., 3)."""
    p0 = np.asarray(p0, dtype=np.float64)
    d = np.asarray(d, dtype=np.float64)
    dn = d / np.linalg.norm(d, axis=-1, keepdims=True)
    out = np.zeros(p0.shape[:-1])
    for cx, cy, cz, a, b, c, al, be, ga, rho in params:
        R = _rot_zyz(al, be, ga)
        q = (p0 - np.array([cx, cy, cz])) @ R / np.array([a, b, c])
        e = dn @ R / np.array([a, b, c])
        A = (e * e).sum(-1)
        B = 2 * (q * e).sum(-1)
        Cc = (q * q).sum(-1) - 1
        disc = B * B - 4 * A * Cc
        out += rho * np.where(disc > 0, np.sqrt(np.maximum(disc, 0)) / A, 0.0)
    return out
