"""Synthetic inputs: seeded random-ellipse phantoms and the Poisson/log
sinogram pipeline.

All randomness comes from numpy's counter-based Philox generator keyed by
(seed, stream), so an input is a pure function of its key.  The noise-free
sinogram is an FP64 forward projection through the device projector; the
noise is drawn on the host.

Conventions (shared with the tracers): image row index x, column index y,
pixel (x, y) centred at ((x + 0.5) * pixel_mm, (y + 0.5) * pixel_mm) with the
image corner at the origin; ellipse centres are given relative to the image
centre.
"""
from __future__ import annotations

import math

import numpy as np


def rng(seed: int, stream: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=[int(seed) & (2**64 - 1), int(stream) & (2**64 - 1)]))


def random_ellipses(n_ellipses, fov_mm, seed, stream=0, axes_mm=(5.0, 100.0), mu_max=0.02):
    """n_ellipses rows of (cx, cy, a, b, theta, mu): centres uniform in the
    inner 80% of the field of view, semi-axes uniform in axes_mm, rotation
    uniform in [0, pi), attenuation uniform in [0, mu_max] per mm."""
    g = rng(seed, stream)
    half = 0.4 * fov_mm
    e = np.empty((n_ellipses, 6))
    e[:, 0] = g.uniform(-half, half, n_ellipses)
    e[:, 1] = g.uniform(-half, half, n_ellipses)
    e[:, 2] = g.uniform(axes_mm[0], axes_mm[1], n_ellipses)
    e[:, 3] = g.uniform(axes_mm[0], axes_mm[1], n_ellipses)
    e[:, 4] = g.uniform(0.0, math.pi, n_ellipses)
    e[:, 5] = g.uniform(0.0, mu_max, n_ellipses)
    return e


def rasterize(ellipses, n, pixel_mm):
    """Sum of ellipse indicators times attenuation, sampled at pixel centres."""
    c = (np.arange(n) + 0.5) * pixel_mm - n * pixel_mm / 2.0
    X, Y = np.meshgrid(c, c, indexing="ij")
    img = np.zeros((n, n), np.float64)
    for cx, cy, a, b, th, mu in ellipses:
        dx, dy = X - cx, Y - cy
        ct, st = math.cos(th), math.sin(th)
        xr = ct * dx + st * dy
        yr = -st * dx + ct * dy
        img += mu * ((xr / a) ** 2 + (yr / b) ** 2 <= 1.0)
    return img


def ellipse_line_integral(ellipse, p0, p1):
    """Closed-form length of the chord of the line p0 -> p1 through the
    ellipse times its attenuation (coordinates relative to the image centre)."""
    cx, cy, a, b, th, mu = ellipse
    ct, st = math.cos(th), math.sin(th)
    d = np.asarray(p1, float) - np.asarray(p0, float)
    L = float(np.hypot(*d))
    u = d / L
    q = np.asarray(p0, float) - (cx, cy)
    qx, qy = ct * q[0] + st * q[1], -st * q[0] + ct * q[1]
    ux, uy = ct * u[0] + st * u[1], -st * u[0] + ct * u[1]
    A = (ux / a) ** 2 + (uy / b) ** 2
    B = 2 * (qx * ux / a**2 + qy * uy / b**2)
    Cc = (qx / a) ** 2 + (qy / b) ** 2 - 1
    disc = B * B - 4 * A * Cc
    if disc <= 0:
        return 0.0
    return mu * math.sqrt(disc) / A


def noisy_sinogram(p, i0, seed, stream=0):
    """counts ~ Poisson(i0 * exp(-p)); zero counts are clamped to one; returns
    -ln(counts / i0)."""
    g = rng(seed, 1_000_003 + stream)
    lam = i0 * np.exp(-np.asarray(p, np.float64))
    counts = g.poisson(lam).astype(np.float64)
    return -np.log(np.maximum(counts, 1.0) / i0)


def analytic_fan_sinogram(ellipses, g):
    """Exact line integrals of the ellipses along every ray of a flat-detector
    fan geometry (the rays the device tracer walks: source (0, -sod) and
    detector point ((bin - nbins/2 + 0.5 + offset) * bin_mm, sdd - sod),
    rotated by span * view / n_views), [n_views, n_bins] float64."""
    V, B = g.n_views, g.n_bins
    th = g.span_deg * (math.pi / 180.0) * np.arange(V) / V
    c, s = np.cos(th)[:, None], np.sin(th)[:, None]
    x1 = ((np.arange(B) - B // 2) + 0.5 + g.det_offset)[None, :] * g.bin_mm
    y1 = g.sdd_mm - g.sod_mm
    sx, sy = -s * (-g.sod_mm) * np.ones((1, B)), c * (-g.sod_mm) * np.ones((1, B))
    dx, dy = c * x1 - s * y1, s * x1 + c * y1
    ux, uy = dx - sx, dy - sy
    L = np.hypot(ux, uy)
    ux, uy = ux / L, uy / L
    out = np.zeros((V, B))
    for cx, cy, a, b, t, mu in ellipses:
        ct, st = math.cos(t), math.sin(t)
        qx0, qy0 = sx - cx, sy - cy
        qx, qy = ct * qx0 + st * qy0, -st * qx0 + ct * qy0
        vx, vy = ct * ux + st * uy, -st * ux + ct * uy
        A = (vx / a) ** 2 + (vy / b) ** 2
        Bq = 2 * (qx * vx / a**2 + qy * vy / b**2)
        Cq = (qx / a) ** 2 + (qy / b) ** 2 - 1
        disc = Bq * Bq - 4 * A * Cq
        out += mu * np.sqrt(np.maximum(disc, 0.0)) / A
    return out


def random_ellipsoids(n, fov_mm, seed, stream=0, axes_mm=(5.0, 100.0), mu_max=0.05):
    """n ellipsoids (cx, cy, cz, a, b, c, theta, mu): centres uniform in the
    inner 80% of the field of view, semi-axes uniform in axes_mm, rotation
    theta about z uniform in [0, pi), attenuation uniform in [0, mu_max] / mm."""
    g = rng(seed, 300 + stream)
    out = []
    for _ in range(n):
        cx, cy, cz = (float(t) for t in g.uniform(-0.4 * fov_mm, 0.4 * fov_mm, 3))
        a, b, c = (float(t) for t in g.uniform(axes_mm[0], axes_mm[1], 3))
        out.append((cx, cy, cz, a, b, c, float(g.uniform(0.0, math.pi)), float(g.uniform(0.0, mu_max))))
    return out


def rasterize3d(ellipsoids, nz, nx, ny, voxel_mm):
    """Volume [nz][nx][ny] (float32) of the ellipsoids sampled at voxel
    centres ((i + 0.5 - n / 2) * voxel_mm along each axis); overlaps add."""
    vol = np.zeros((nz, nx, ny), np.float32)
    cz_ = ((np.arange(nz) + 0.5) - nz / 2) * voxel_mm
    cx_ = ((np.arange(nx) + 0.5) - nx / 2) * voxel_mm
    cy_ = ((np.arange(ny) + 0.5) - ny / 2) * voxel_mm
    for cx, cy, cz, a, b, c, t, mu in ellipsoids:
        r = max(a, b)
        xs = np.nonzero(np.abs(cx_ - cx) <= r)[0]
        ys = np.nonzero(np.abs(cy_ - cy) <= r)[0]
        zs = np.nonzero(np.abs(cz_ - cz) <= c)[0]
        if not (len(xs) and len(ys) and len(zs)):
            continue
        X = (cx_[xs] - cx)[:, None]
        Y = (cy_[ys] - cy)[None, :]
        xr = X * math.cos(t) + Y * math.sin(t)
        yr = -X * math.sin(t) + Y * math.cos(t)
        q2 = (xr / a) ** 2 + (yr / b) ** 2
        zq = ((cz_[zs] - cz) / c) ** 2
        inside = (q2[None, :, :] + zq[:, None, None]) <= 1.0
        vol[zs[0]:zs[-1] + 1, xs[0]:xs[-1] + 1, ys[0]:ys[-1] + 1] += np.float32(mu) * inside
    return vol
