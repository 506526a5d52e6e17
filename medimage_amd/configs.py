"""The reconstruction workloads, with every size the descriptions leave open
fixed here.

* C1 plain: 256 x 256, 1.0 mm pixels, 360 views over 360 degrees, 512 bins of
  1.0 mm, SOD 541 mm, SDD 949 mm; 10 ellipses; I0 = 1e5; 50 AT iterations;
  alpha 1e-3, beta 1e-4, eps 1e-2.
* C2 precision: 512 x 512, 0.7 mm pixels, 720 views, 736 bins of 1.0 mm,
  same source/detector distances; I0 = 2e4; 12 subsets, 100 AT iterations.
* C3 batched: 64 slices, sizes from {256, 384, 512}, views from {360, 720,
  1152}, bins from {512, 736}; field of view 358.4 mm (pixel = 358.4 / N),
  detector width 736 mm (bin = 736 / bins); I0 uniform in [1e4, 1e5];
  8 subsets, 50 AT iterations.
* C4 cone beam: 512^3 volume of 0.5 mm voxels, 720 views over 360 degrees
  onto a 1024 x 768 flat panel (NDX = 1024 rows along z, NDY = 768 columns)
  of 0.4 mm pitch, SOD 600 mm, SDD 1000 mm; 20 ellipsoids with attenuation
  in [0, 0.05] / mm; I0 = 5e4; 20 subsets, 30 AT iterations.  industry3d
  centres both panel axes on NDX / 2 (HALFDET), so the columns span
  -204.6 .. +102.2 mm on the panel; the panel is narrower than the volume's
  projection, and rays outside it are not measured.
* C5 serving: sizes {256, 384, 512, 768, 1024}, views {360, 720, 1152, 2304},
  bins = 1.5 N, same field of view and detector width; I0 log-uniform in
  [1e4, 1e5]; stop at relative data-fidelity change < 1e-4 or 200 iterations.

One AT iteration is one pass over all subsets (per subset: forward
projection, residual, gather back projection, image update with an exact
line search), followed by one edge (v) step.
"""
from __future__ import annotations

import math

import numpy as np

from . import arc_geometry, at_params, cone_geometry, fan_geometry, phantom

SOD, SDD = 541.0, 949.0
FOV_MM = 358.4
DET_MM = 736.0
ALPHA, BETA, EPS = 1e-3, 1e-4, 1e-2


def c1_geometry():
    return fan_geometry(256, 1.0, 360, 512, 1.0, SOD, SDD)


def c2_geometry():
    return fan_geometry(512, 0.7, 720, 736, 1.0, SOD, SDD)


def c1_params(**kw):
    return at_params(ALPHA, BETA, EPS, n_subsets=kw.pop("n_subsets", 1), max_iters=kw.pop("max_iters", 50), **kw)


def c2_params(**kw):
    return at_params(ALPHA, BETA, EPS, n_subsets=kw.pop("n_subsets", 12), max_iters=kw.pop("max_iters", 100), **kw)


def mixed_geometry(n, views, bins):
    """C3/C5 geometry: FOV 358.4 mm and a 736 mm wide detector."""
    return fan_geometry(n, FOV_MM / n, views, bins, DET_MM / bins, SOD, SDD)


def c3_jobs(n_jobs=64, seed=3):
    g = phantom.rng(seed, 77)
    out = []
    for j in range(n_jobs):
        n = int(g.choice([256, 384, 512]))
        v = int(g.choice([360, 720, 1152]))
        b = int(g.choice([512, 736]))
        i0 = float(g.uniform(1e4, 1e5))
        out.append(dict(job=j, n=n, views=v, bins=b, i0=i0, seed=seed * 1000 + j))
    return out


def c5_requests(n_jobs=2000, seed=5, rate=20.0):
    """Serving stream: Poisson arrivals at `rate` jobs/s; N from {256, 384,
    512, 768, 1024}, views from {360, 720, 1152, 2304}, bins = 1.5 N, I0
    log-uniform on [1e4, 1e5]."""
    g = phantom.rng(seed, 55)
    t = 0.0
    out = []
    for j in range(n_jobs):
        t += float(g.exponential(1.0 / rate))
        n = int(g.choice([256, 384, 512, 768, 1024]))
        v = int(g.choice([360, 720, 1152, 2304]))
        i0 = float(math.exp(g.uniform(math.log(1e4), math.log(1e5))))
        out.append(dict(job=j, t=t, n=n, views=v, bins=3 * n // 2, i0=i0, seed=seed * 100000 + j))
    return out


def analytic_case(geom, i0, seed, n_ellipses=10, mu_max=0.02):
    """Noisy sinogram (float64) from the exact line integrals of the seeded
    ellipse phantom (no device projection; used where an FP64 system matrix
    of the geometry would not be worth building)."""
    fov = geom.nx * geom.pixel_mm
    ell = phantom.random_ellipses(n_ellipses, fov, seed, mu_max=mu_max)
    p = phantom.analytic_fan_sinogram(ell, geom)
    return phantom.noisy_sinogram(p, i0, seed)


def make_case(ctx, geom, i0, seed, n_ellipses=10, mu_max=0.02):
    """Phantom and noisy sinogram (float64) for a geometry."""
    fov = geom.nx * geom.pixel_mm
    ell = phantom.random_ellipses(n_ellipses, fov, seed, mu_max=mu_max)
    img = phantom.rasterize(ell, geom.nx, geom.pixel_mm)
    p = ctx.forward(geom, img, dtype=np.float64)
    sino = phantom.noisy_sinogram(p, i0, seed)
    return img, sino


def rays(geom):
    return geom.n_views * geom.n_bins


def subsets_for(views, want):
    s = min(want, views)
    return max(1, s)


__all__ = ["c1_geometry", "c2_geometry", "c1_params", "c2_params", "mixed_geometry", "c3_jobs", "c5_requests",
           "make_case", "analytic_case", "rays", "math"]


def arc_clinical_geometry(views=1152, n=874, pixel_mm=0.5859):
    """The clinical arc-detector slice geometry of the reference's helical
    rebinning configuration (helical/rebin/json/UCLA_PHANTOM/SCAN1_1.5mm.json:
    874 x 874 at 0.5859 mm, 736 elements of 0.06786112 degrees, OFF_CENTER
    1.75, SOURCE_TO_ISO 595 mm, SOURCE_TO_DET 1085.6 mm, 1152 views per turn)."""
    return arc_geometry(n, pixel_mm, views, 736, 0.06786112, 595.0, 1085.6, 1.75)



def c4_geometry(n=512, views=720, rows=1024, cols=768, sample_mm=0.5, pixel_mm=0.4, sod=600.0, sdd=1000.0):
    """C4 cone-beam geometry (industry3d Parameter, BEAM "Cone")."""
    return cone_geometry(n, n, n, views, rows, cols, sample_mm, pixel_mm, sod, sdd)
