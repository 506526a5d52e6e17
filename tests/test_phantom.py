"""CPU checks of the synthetic-input generators (no GPU)."""
import math

import numpy as np

from medimage_amd import configs, phantom


def _endpoints(g, v, b):
    th = g.span_deg * (math.pi / 180) * v / g.n_views
    c, s = math.cos(th), math.sin(th)
    x0, y0 = 0.0, -g.sod_mm
    x1 = ((b - g.n_bins // 2) + 0.5 + g.det_offset) * g.bin_mm
    y1 = g.sdd_mm - g.sod_mm
    return (c * x0 - s * y0, s * x0 + c * y0), (c * x1 - s * y1, s * x1 + c * y1)


def test_analytic_sinogram_matches_scalar_chord():
    g = configs.mixed_geometry(384, 360, 576)
    ell = phantom.random_ellipses(10, g.nx * g.pixel_mm, 11)
    p = phantom.analytic_fan_sinogram(ell, g)
    rng = np.random.default_rng(0)
    for v, b in zip(rng.integers(0, g.n_views, 40), rng.integers(0, g.n_bins, 40)):
        p0, p1 = _endpoints(g, int(v), int(b))
        ref = sum(phantom.ellipse_line_integral(e, p0, p1) for e in ell)
        assert abs(ref - p[v, b]) <= 1e-12 * max(1.0, abs(ref))


def test_analytic_sinogram_disc_chord():
    """A centred disc of radius R: every ray at distance d from the centre
    carries mu * 2 sqrt(R^2 - d^2)."""
    g = configs.mixed_geometry(256, 90, 384)
    R, mu = 60.0, 0.02
    p = phantom.analytic_fan_sinogram(np.array([[0.0, 0.0, R, R, 0.3, mu]]), g)
    for v in (0, 17, 45, 89):
        for b in (0, 100, 191, 192, 250, 383):
            (x0, y0), (x1, y1) = _endpoints(g, v, b)
            d = abs(x0 * (y1 - y0) - y0 * (x1 - x0)) / math.hypot(x1 - x0, y1 - y0)
            want = mu * 2 * math.sqrt(R * R - d * d) if d < R else 0.0
            assert abs(p[v, b] - want) < 1e-12


def test_c5_requests_distribution():
    reqs = configs.c5_requests(400, seed=5, rate=20.0)
    t = np.array([r["t"] for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert abs(t[-1] / len(reqs) - 1 / 20.0) < 0.015          # mean inter-arrival
    assert {r["n"] for r in reqs} == {256, 384, 512, 768, 1024}
    assert {r["views"] for r in reqs} == {360, 720, 1152, 2304}
    assert all(r["bins"] == 3 * r["n"] // 2 for r in reqs)
    assert all(1e4 <= r["i0"] <= 1e5 for r in reqs)
