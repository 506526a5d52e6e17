"""Pins of the CPU oracle (no GPU).

Each oracle function is checked against something other than itself: the
reference's own known-answer file, the reference compiled from its sources
(oracle/_ref, when present), fixtures that reference produced, closed forms
and brute force.
"""
import ctypes as C
import math
import os
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import oracle_py as O

REF = O.REF


def _reflib(name):
    p = REF / name
    if not p.exists():
        pytest.skip(f"{p} not built (make -C oracle ref)")
    return C.CDLL(str(p))


# --------------------------------------------------------------------------
# Parallel-beam tracer: std_g known-answer test.

def test_kat_std_g(golden):
    """A(std_f) vs blockrecon/phantom/std_g.dat (6 significant digits).

    The reference wrote std_g with FMA contraction enabled (-march=native,
    blockrecon/Makefile), which moves some tracer decisions; the oracle
    follows the source's strict IEEE semantics.  Required: at least 98% of
    the 138240 rays print identically, at least 99.7% agree within one unit
    of the 6th digit, and views 0 and 90 (pure row/column sums) agree
    exactly everywhere."""
    z = golden("blockrecon_std.npz")
    f, gref = z["f"], z["g"]
    g = O.par_fp_f32(f, 180, 768, flavour=1).astype(np.float64)
    same = np.array([("%g" % a) == ("%g" % b) for a, b in zip(g.ravel(), gref.ravel())])
    assert same.mean() >= 0.98
    unit = 10.0 ** (np.floor(np.log10(np.where(gref > 0, gref, 1.0))) - 5)
    close = np.abs(g - gref) <= unit * 1.0000001
    assert close.mean() >= 0.997
    assert close[0].all() and close[90].all()


def test_view0_view90_exact_sums(golden):
    """View-0 rays run along one image row with unit weights, view-90 rays
    along one column; their projections are the float sums in walk order.
    (The walk returns before the last pixel of an axis-aligned ray,
    msbeam/CPU_OpenMP/src/proj.c:194, so such rays hold 511 pixels.)"""
    f = golden("blockrecon_std.npz")["f"].ravel()
    for fl in (0, 1):
        g = O.par_fp_f32(f.reshape(512, 512), 180, 768, flavour=fl)
        for view, nr in [(0, 400), (0, 200), (90, 400), (90, 600)]:
            line, wt, _ = O.wray(view, nr, 180, 768, 512, fl)
            assert len(line) == 511 and np.all(wt == 1.0)
            if view == 0:
                assert len({i // 512 for i in line}) == 1
            else:
                assert len({i % 512 for i in line}) == 1
            acc = np.float32(0)
            for i in line:
                acc = np.float32(acc + f[i])
            assert acc == g[view, nr]


@pytest.mark.parametrize("flavour,lib", [(1, "libref_blockrecon.so"), (0, "libref_msbeam.so")])
def test_wray_matches_compiled_reference(flavour, lib, golden):
    """Every ray's (index, weight, count) equals the reference's wray (strict build)."""
    ref = _reflib(lib)
    fn = ref._Z4wrayiiPiPfS_S0_ if flavour == 1 else ref.wray
    line_r = np.zeros(1024, np.int32)
    wt_r = np.zeros(1024, np.float32)
    numb = C.c_int()
    sn = C.c_float()
    for npj in range(0, 180, 7):
        for nr in range(0, 768, 5):
            fn(npj, nr, O.P(line_r), O.P(wt_r), C.byref(numb), C.byref(sn))
            line, wt, s = O.wray(npj, nr, 180, 768, 512, flavour)
            n = numb.value
            assert n == len(line)
            assert np.array_equal(line, line_r[:n])
            assert np.array_equal(wt.view(np.uint32), wt_r[:n].view(np.uint32))
            assert np.float32(s).view(np.uint32) == np.float32(sn.value).view(np.uint32)


def test_par_fp_matches_compiled_reference(golden):
    ref = _reflib("libref_blockrecon.so")
    f = golden("blockrecon_std.npz")["f"]
    g = np.zeros(180 * 768, np.float32)
    ref._Z1APA768_fPA512_f(O.P(g), O.P(f))
    ours = O.par_fp_f32(f, 180, 768, 1).ravel()
    assert np.array_equal(ours.view(np.uint32), g.view(np.uint32))


def test_msbeam_projectors_match_compiled_reference(golden):
    """msbeam's A and AStar (msbeam/CPU_OpenMP/src/proj.c:14-52: float
    products added in ray order) against the oracle's restatement, bitwise;
    this pins the oracle the libmedimage_msbeam.so test compares with."""
    ref = _reflib("libref_msbeam.so")
    rs = np.random.RandomState(11)
    f = (golden("blockrecon_std.npz")["f"] + rs.rand(512, 512).astype(np.float32) * np.float32(0.25)).astype(np.float32)
    g = np.zeros(180 * 768, np.float32)
    ref.A(O.P(g), O.P(f))
    og = O.par_fp_f32(f, 180, 768, 0)
    assert np.array_equal(og.ravel().view(np.uint32), g.view(np.uint32))
    b = np.zeros(512 * 512, np.float32)
    ref.AStar(O.P(b), O.P(g))
    ob = O.par_bp_f32(og, 512, 0)
    assert np.array_equal(ob.ravel().view(np.uint32), b.view(np.uint32))


def test_wray_invariants():
    """numb <= 2N, snorm = sum w^2, sum of weights = chord length of the
    square for rays through its interior."""
    n = 64
    for npj in range(0, 36, 5):
        for nr in range(0, 96, 7):
            line, wt, sn = O.wray(npj, nr, 36, 96, n, 0)
            assert len(line) <= 2 * n
            assert abs(sn - float(np.sum(wt.astype(np.float64) ** 2))) <= 1e-4 * max(sn, 1)
            assert np.all((line >= 0) & (line < n * n))


# --------------------------------------------------------------------------
# 2D parametric Siddon.

def test_siddon_axis_aligned():
    ind, w = O.siddon2d(-10.0, 2.5, 20.0, 2.5, 8, 8, 1.0, 1.0)
    assert len(ind) == 8
    assert np.all(w == 1.0)
    assert list(ind) == [x * 8 + 2 for x in range(8)]


def test_siddon_45_degrees():
    """A diagonal through grid corners: each diagonal pixel gets sqrt(2), and
    every corner crossing (x plane first on the tie) adds a zero-length entry."""
    ind, w = O.siddon2d(-1.0, -1.0, 9.0, 9.0, 8, 8, 1.0, 1.0)
    nz = [(i, x) for i, x in zip(ind, w) if x > 1e-12]
    assert [i for i, _ in nz] == [k * 8 + k for k in range(8)]
    assert all(abs(x - math.sqrt(2)) < 1e-12 for _, x in nz)
    zero = [i for i, x in zip(ind, w) if x <= 1e-12]
    assert zero == [(k + 1) * 8 + k for k in range(7)]


def _exact_siddon(sx, sy, dx, dy, nx, ny):
    """Exact rational crossing merge on a unit grid (reference order: x
    crossing first on ties), returning (index, length^2) as Fractions."""
    sx, sy, dx, dy = map(Fraction, (sx, sy, dx, dy))
    rx, ry = dx - sx, dy - sy
    ts = []
    lo, hi = Fraction(0), Fraction(1)
    if rx != 0:
        t0, t1 = (0 - sx) / rx, (nx - sx) / rx
        lo, hi = max(lo, min(t0, t1)), min(hi, max(t0, t1))
    if ry != 0:
        t0, t1 = (0 - sy) / ry, (ny - sy) / ry
        lo, hi = max(lo, min(t0, t1)), min(hi, max(t0, t1))
    if lo >= hi:
        return []
    pts = {lo, hi}
    for k in range(nx + 1):
        if rx != 0:
            t = (k - sx) / rx
            if lo < t < hi:
                pts.add(t)
    for k in range(ny + 1):
        if ry != 0:
            t = (k - sy) / ry
            if lo < t < hi:
                pts.add(t)
    pts = sorted(pts)
    out = []
    L2 = rx * rx + ry * ry
    for a, b in zip(pts[:-1], pts[1:]):
        m = (a + b) / 2
        x, y = sx + m * rx, sy + m * ry
        out.append((int(math.floor(x)) * ny + int(math.floor(y)), (b - a) ** 2 * L2))
    return out


def test_siddon_exact_rational_small_grids():
    """Brute force with exact arithmetic: the double tracer's pixel sequence
    equals the exact one (zero-length corner entries aside) and every length
    is within 1e-9 of the exact segment length, up to the 5e-2 tail."""
    rs = np.random.RandomState(5)
    checked = 0
    for _ in range(400):
        nx = ny = int(rs.choice([4, 7, 8, 16]))
        ang = rs.uniform(0, 2 * math.pi)
        off = rs.uniform(-0.45, 0.45) * nx
        cx, cy = nx / 2 + off * -math.sin(ang), ny / 2 + off * math.cos(ang)
        R = 3.0 * nx
        sx, sy = cx - R * math.cos(ang), cy - R * math.sin(ang)
        dx, dy = cx + R * math.cos(ang), cy + R * math.sin(ang)
        ind, w = O.siddon2d(sx, sy, dx, dy, nx, ny, 1.0, 1.0)
        ex = _exact_siddon(sx, sy, dx, dy, nx, ny)
        keep = [(i, x) for i, x in zip(ind, w) if x > 1e-12]
        exk = [(i, math.sqrt(float(l2))) for i, l2 in ex if l2 > 0]
        # the tracer drops the final segment when it is shorter than 5e-2
        if len(exk) == len(keep) + 1 and exk[-1][1] < 5e-2 + 1e-9:
            exk = exk[:-1]
        assert [i for i, _ in keep] == [i for i, _ in exk]
        for (_, a), (_, b) in zip(keep, exk):
            assert abs(a - b) < 1e-9
        checked += 1
    assert checked == 400


def test_siddon_chord_length_and_count():
    fg = O.fan_geom(64, 1.0, 30, 96, 1.0)
    for view in range(0, 30, 3):
        for b in range(0, 96, 5):
            sx, sy, dx, dy = O.fan_endpoints(fg, view, b)
            ind, w = O.fan_ray(fg, view, b)
            if len(ind) == 0:
                continue
            ex = _exact_siddon(sx, sy, dx, dy, 64, 64)
            total = sum(math.sqrt(float(l2)) for _, l2 in ex)
            assert total - 5e-2 - 1e-9 <= w.sum() <= total + 1e-9
            xs = {i // 64 for i in ind}
            ys = {i % 64 for i in ind}
            assert len(ind) >= len(xs) + len(ys) - 1


def test_fan_geometry_closed_form():
    """Source at distance sod from the image centre, detector cell at
    distance sdd - sod beyond it, and the central ray through the centre."""
    fg = O.fan_geom(128, 0.5, 36, 64, 1.0, sod=541.0, sdd=949.0)
    c = 64 * 0.5
    for view in range(36):
        sx, sy, dx, dy = O.fan_endpoints(fg, view, 32)
        assert abs(math.hypot(sx - c, sy - c) - 541.0) < 1e-9
        th = 2 * math.pi * view / 36
        assert abs((sx - c) - 541.0 * math.sin(th)) < 1e-9
        assert abs((sy - c) + 541.0 * math.cos(th)) < 1e-9
        # bin 32 sits half a cell off axis
        dist = abs((dx - c) * (sy - c) - (dy - c) * (sx - c)) / math.hypot(dx - sx, dy - sy)
        assert abs(dist - 0.5 * 541.0 / math.hypot(949.0, 0.5)) < 1e-9


def test_arc_geometry_closed_form():
    """Arc detector (helical/rebin/src/ms2d.cpp:172-183): the source on the
    circle of radius SOD around the image centre at angle a_v, every detector
    element SDD away from it, ray k leaving the source at (k - bins/2 +
    OFF_CENTER) elements of LENGTH_PER_DET degrees off the line to the centre,
    so its distance from the centre is SOD |sin| of that angle."""
    n, px, bins, dpb, off = 100, 0.7, 64, 0.25, 1.75
    fg = O.arc_geom(n, px, 30, bins, dpb, sod=595.0, sdd=1085.6, off_center=off)
    c = n * px / 2
    for view in range(0, 30, 7):
        a = math.radians(360.0 * view / 30)
        for b in (0, 17, 31, 32, 63):
            sx, sy, dx, dy = O.fan_endpoints(fg, view, b)
            assert abs(sx - c - 595.0 * math.cos(a)) < 1e-9 and abs(sy - c - 595.0 * math.sin(a)) < 1e-9
            assert abs(math.hypot(dx - sx, dy - sy) - 1085.6) < 1e-9
            phi = math.radians((b - bins // 2 + off) * dpb)
            dist = abs((dx - sx) * (sy - c) - (dy - sy) * (sx - c)) / 1085.6
            assert abs(dist - 595.0 * abs(math.sin(phi))) < 1e-8
            # the ray points away from the source towards the centre side
            assert (dx - sx) * (c - sx) + (dy - sy) * (c - sy) > 0


def test_arc_rays_are_equiangular():
    fg = O.arc_geom(64, 1.0, 8, 40, 0.3, off_center=0.5)
    for view in range(8):
        prev = None
        for b in range(40):
            sx, sy, dx, dy = O.fan_endpoints(fg, view, b)
            ang = math.atan2(dy - sy, dx - sx)
            if prev is not None:
                d = (ang - prev + math.pi) % (2 * math.pi) - math.pi
                assert abs(d - math.radians(0.3)) < 1e-12
            prev = ang


def test_arc_projector_adjoint_and_disc():
    import medimage_amd.phantom as ph
    n, px = 96, 1.0
    fg = O.arc_geom(n, px, 18, 128, 0.12, sod=595.0, sdd=1085.6, off_center=1.75)
    rs = np.random.RandomState(4)
    x = rs.rand(n, n)
    y = rs.rand(18, 128)
    lhs = np.sum(O.fan_fp(fg, x) * y)
    rhs = np.sum(x * O.fan_bp(fg, y))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    e = np.array([[4.0, -2.0, 25.0, 25.0, 0.0, 0.01]])
    sino = O.fan_fp(fg, ph.rasterize(e, n, px))
    c = n * px / 2
    errs = []
    for view in range(18):
        for b in range(10, 118, 3):
            sx, sy, dx, dy = O.fan_endpoints(fg, view, b)
            ana = ph.ellipse_line_integral(e[0], (sx - c, sy - c), (dx - c, dy - c))
            if ana > 0.2:
                errs.append(abs(sino[view, b] - ana) / ana)
    assert len(errs) > 100 and np.median(errs) < 0.02


def test_fan_projector_adjoint():
    fg = O.fan_geom(48, 1.0, 24, 64, 1.0)
    rs = np.random.RandomState(1)
    x = rs.rand(48, 48)
    y = rs.rand(24, 64)
    lhs = np.sum(O.fan_fp(fg, x) * y)
    rhs = np.sum(x * O.fan_bp(fg, y))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_fan_projector_ellipse_closed_form():
    """Line integrals of a rasterised disc approach the analytic chord."""
    import medimage_amd.phantom as ph
    n, px = 128, 1.0
    fg = O.fan_geom(n, px, 12, 160, 1.0, sod=541.0, sdd=949.0)
    e = np.array([[5.0, -3.0, 30.0, 30.0, 0.0, 0.01]])
    img = ph.rasterize(e, n, px)
    sino = O.fan_fp(fg, img)
    c = n * px / 2
    errs = []
    for view in range(12):
        for b in range(40, 120, 4):
            sx, sy, dx, dy = O.fan_endpoints(fg, view, b)
            ana = ph.ellipse_line_integral(e[0], (sx - c, sy - c), (dx - c, dy - c))
            if ana > 0.2:
                errs.append(abs(sino[view, b] - ana) / ana)
    assert np.median(errs) < 0.02


# --------------------------------------------------------------------------
# GoldenC projector and AT steps.

def test_goldenc_projector_matches_compiled_reference(golden):
    ref = _reflib("libref_goldnc.so")
    f = golden("blockrecon_std.npz")["f"].astype(np.float64)
    g = np.zeros((180, 767))
    ref.A(O.P(g), O.P(f))
    assert np.array_equal(O.gc_fp(f), g)
    rs = np.random.RandomState(3)
    s = rs.rand(180, 767)
    b = np.zeros((512, 512))
    ref.AStar(O.P(b), O.P(s))
    assert np.array_equal(O.gc_bp(s), b)


def test_goldenc_astar_is_not_exact_adjoint():
    """proj.c:32 vs :225: the view-90 column differs by one between A and AStar."""
    rs = np.random.RandomState(4)
    x = rs.rand(512, 512)
    y = np.zeros((180, 767))
    y[90] = rs.rand(767)
    assert abs(np.sum(O.gc_fp(x) * y) - np.sum(x * O.gc_bp(y))) > 1e-6
    y[90] = 0
    y[7] = rs.rand(767)
    lhs, rhs = np.sum(O.gc_fp(x) * y), np.sum(x * O.gc_bp(y))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_goldenc_steps_match_compiled_reference(golden):
    ref = _reflib("libref_goldnc.so")
    z = golden("goldenc_ms.npz")
    g = z["g"]
    f0 = np.zeros((512, 512))
    v0 = np.ones((512, 512))
    out = np.zeros((512, 512))
    ref.minimize_in_IMAGE_variable.argtypes = [C.c_void_p] * 4 + [C.c_double, C.c_double, C.c_int]
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)
    try:
        ref.minimize_in_IMAGE_variable(O.P(out), O.P(g), O.P(f0), O.P(v0), 1000.0, 1e-4, 3)
        ours, k, why = O.image_step(O.op_goldenc(), g, f0, v0, 1000.0, 1e-4, 3)
        assert np.array_equal(ours, out)
        ref.minimize_in_EDGE_variable.argtypes = [C.c_void_p] * 3 + [C.c_double] * 3 + [C.c_int]
        vout = np.zeros((512, 512))
        ref.minimize_in_EDGE_variable(O.P(vout), O.P(out), O.P(v0), 1000.0, 0.002, 1e-4, 4)
        vo, _ = O.edge_step(out, v0, 1000.0, 0.002, 1e-4, 4)
        assert np.array_equal(vo, vout)
    finally:
        os.dup2(saved, 1)


@pytest.mark.slow
def test_goldenc_mumford_shah_fixture(golden):
    """Full MumfordShah(1000, 0.002, 1e-4), 10 x (10 + 10), against the
    reference program's own output (tests/golden/goldenc_ms.npz)."""
    z = golden("goldenc_ms.npz")
    f, v = O.mumford_shah(O.op_goldenc(), z["g"], 1000.0, 0.002, 1e-4, 10, 10, 10)
    assert np.array_equal(f, z["f"]) and np.array_equal(v, z["v"])


def test_medfilt2_matches_scipy():
    from scipy.ndimage import median_filter
    rs = np.random.RandomState(2)
    for shape in [(7, 9), (32, 32), (1, 5)]:
        a = rs.randint(-3, 4, size=shape).astype(np.float64)
        assert np.array_equal(O.medfilt2(a), median_filter(a, size=3, mode="constant", cval=0.0))


def test_ugrad_is_gradient_of_forward_difference_energy():
    """R(f, v) of origin/GoldenC/src/main.c:112-125 is -1/2 dQ/df for
    Q(f) = sum_{i,j >= -1} (V_ij + 0.01 eps^2) (|F_{i+1,j} - F_ij|^2 + |F_{i,j+1} - F_ij|^2)
    with F = f zero-padded and V = v^2 inside, 1 on the padded top row and
    left column (forward differences)."""
    rs = np.random.RandomState(7)
    h, w, eps = 6, 5, 0.3
    f = rs.rand(h, w)
    v = rs.rand(h, w)
    V = np.ones((h + 1, w + 1))
    V[1:, 1:] = v ** 2

    def Q(f):
        F = np.zeros((h + 2, w + 2))
        F[1:h + 1, 1:w + 1] = f
        dx = F[1:h + 2, 0:w + 1] - F[0:h + 1, 0:w + 1]
        dy = F[0:h + 1, 1:w + 2] - F[0:h + 1, 0:w + 1]
        return np.sum((V + 0.01 * eps ** 2) * (dx ** 2 + dy ** 2))

    R = O.ugrad_reg(np.zeros((h, w)), f, v, 1.0, eps)
    G = np.zeros((h, w))
    for i in range(h):
        for j in range(w):
            e = np.zeros((h, w))
            e[i, j] = 1e-4
            G[i, j] = -(Q(f + e) - Q(f - e)) / 4e-4
    assert np.allclose(R, G, rtol=0, atol=1e-9)
    # a dropped term, a swapped neighbour weight or a wrong sign breaks it
    R2 = O.ugrad_reg(np.zeros((h, w)), f, v.T.copy().T * 0.5, 1.0, eps)
    assert not np.allclose(R2, G, atol=1e-3)


def test_nv_closed_form():
    rs = np.random.RandomState(8)
    p = rs.rand(5, 6)
    v = rs.rand(5, 6)
    eps = 0.2
    pp = np.zeros((6, 7))
    pp[1:, 1:] = p
    a = p - pp[:-1, 1:]
    b = p - pp[1:, :-1]
    ref = np.sum((v ** 2 + 0.01 * eps ** 2) * (a ** 2 + b ** 2))
    assert abs(O.nV(p, v, eps) - ref) < 1e-12


# --------------------------------------------------------------------------
# Ordered subsets.

def test_subset_order_is_permutation():
    for S in [1, 2, 3, 8, 12, 20]:
        seen = np.concatenate([O.subset_views(720, S, s) for s in range(S)])
        assert sorted(seen.tolist()) == list(range(720))
    assert list(np.array([O.subset_views(12, 12, s)[0] for s in range(12)])) == [0, 8, 4, 2, 10, 6, 1, 9, 5, 3, 11, 7]


def test_os_one_subset_first_step_is_goldenc_first_cg_step():
    """At one subset and no edge step, the first AT pass is exactly the first
    (steepest-descent) step of minimize_in_IMAGE_variable."""
    fg = O.fan_geom(32, 1.0, 24, 48, 1.0)
    op = O.op_fan(fg)
    rs = np.random.RandomState(9)
    g = O.fan_fp(fg, rs.rand(32, 32))
    u, v, n, _ = O.os_at(op, g, 1e-3, 1e-4, 1e-2, 1, 1, iter_edge=0)
    gc, k, why = O.image_step(op, g, np.zeros((32, 32)), np.ones((32, 32)), 1e-3, 1e-2, 1)
    assert np.array_equal(u, gc)


def test_os_subsets_partition_gradient():
    """sum_s A_s^T r_s = A^T r (subset back projections add up)."""
    fg = O.fan_geom(24, 1.0, 20, 32, 1.0)
    rs = np.random.RandomState(10)
    r = rs.rand(20, 32)
    full = O.fan_bp(fg, r)
    parts = sum(O.fan_bp(fg, r, O.subset_views(20, 4, s)) for s in range(4))
    assert np.allclose(full, parts, rtol=1e-13, atol=1e-13)


def _dense(fg):
    rp, ind, ln = O.fan_csr(fg)
    A = np.zeros((fg.nviews * fg.nbins, fg.nx * fg.ny))
    for r in range(len(rp) - 1):
        np.add.at(A[r], ind[rp[r]:rp[r + 1]], ln[rp[r]:rp[r + 1]])
    return A


def _subset_energy(A, g, rows, S, alpha, eps, v):
    """S |g_s - A_s u|^2 plus alpha times the regulariser whose gradient the
    reference's image step uses: every grid edge p -> q (q the lower or right
    neighbour; u zero and v one outside the image) weighted by
    v_p^2 + 0.01 eps^2."""
    h, w = v.shape
    W = np.ones((h + 2, w + 2))
    W[1:-1, 1:-1] = v
    W = W ** 2 + 0.01 * eps ** 2

    def E(u):
        U = np.zeros((h + 2, w + 2))
        U[1:-1, 1:-1] = u
        r = g.ravel()[rows] - A[rows] @ u.ravel()
        dn = U[1:, :-1] - U[:-1, :-1]
        rt = U[:-1, 1:] - U[:-1, :-1]
        return S * r @ r + alpha * np.sum(W[:-1, :-1] * (dn ** 2 + rt ** 2))
    return E


def _rows(fg, S, s):
    return np.concatenate([np.arange(k * fg.nbins, (k + 1) * fg.nbins) for k in O.subset_views(fg.nviews, S, s)])


def test_os_step_direction_is_the_steepest_descent_of_the_subset_energy():
    """d = -1/2 grad(S |g_s - A_s u|^2 + alpha R(u; v)) at S = 4 for a random
    u and v (central differences are exact for a quadratic): a dropped S, a
    wrong alpha or sign, or a wrong subset fails it."""
    fg = O.fan_geom(12, 1.0, 16, 20, 1.0)
    op = O.op_fan(fg)
    A = _dense(fg)
    rs = np.random.RandomState(21)
    g = O.fan_fp(fg, rs.rand(12, 12))
    u = rs.rand(12, 12)
    v = 0.3 + 0.7 * rs.rand(12, 12)
    S, alpha, eps = 4, 0.8, 0.3
    for s in range(S):
        E = _subset_energy(A, g, _rows(fg, S, s), S, alpha, eps, v)
        grad = np.zeros(u.size)
        for k in range(u.size):
            e = np.zeros(u.size)
            e[k] = 1.0
            grad[k] = (E(u + e.reshape(u.shape)) - E(u - e.reshape(u.shape))) / 2.0
        _, d, _ = O.os_step(op, g, u, v, alpha, eps, S, s)
        assert np.allclose(d.ravel(), -0.5 * grad, rtol=1e-9, atol=1e-9 * np.abs(grad).max())


def _vertex(E, u, d, t):
    e0, e1, e2 = E(u), E(u + t * d), E(u - t * d)
    return t * (e2 - e1) / (2.0 * (e1 + e2 - 2.0 * e0))


def test_os_step_is_the_exact_minimiser_along_d():
    """The S-scaled step c = |d|^2 / (S |A_s d|^2 + alpha nV(d)) against a
    brute-force minimisation of the subset energy along d (the vertex of the
    parabola through three points; the energy is quadratic in c).
    (a) alpha = 0, random u and v: exact for any state.
    (b) alpha > 0: the reference's nV (backward differences, no edge to the
    zero padding below the last row or right of the last column) equals the
    regulariser's curvature along d when v = 1 and d vanishes on the last row
    and column.  That holds for g = A u with u zero on its last two rows and
    columns: the residual is zero and d is the regulariser term alone.  Every
    subset is checked, and the subset energy with S = 1 in place of S = 4
    has a different minimiser."""
    fg = O.fan_geom(12, 1.0, 16, 20, 1.0)
    op = O.op_fan(fg)
    A = _dense(fg)
    rs = np.random.RandomState(22)
    S, eps = 4, 0.3
    # (a)
    g = O.fan_fp(fg, rs.rand(12, 12))
    u = rs.rand(12, 12)
    v = 0.3 + 0.7 * rs.rand(12, 12)
    for s in range(S):
        E = _subset_energy(A, g, _rows(fg, S, s), S, 0.0, eps, v)
        un, d, c = O.os_step(op, g, u, v, 0.0, eps, S, s)
        assert abs(c - _vertex(E, u, d, c)) < 1e-9 * c
        assert np.allclose(un, u + _vertex(E, u, d, c) * d, rtol=1e-12, atol=1e-12)
    # (b)
    u = np.zeros((12, 12))
    u[:10, :10] = rs.rand(10, 10)
    g = O.fan_fp(fg, u)
    v = np.ones((12, 12))
    for alpha in (0.5, 5.0):
        for s in range(S):
            E = _subset_energy(A, g, _rows(fg, S, s), S, alpha, eps, v)
            un, d, c = O.os_step(op, g, u, v, alpha, eps, S, s)
            assert not d[-1, :].any() and not d[:, -1].any() and np.abs(d).max() > 0
            cv = _vertex(E, u, d, c)
            assert abs(c - cv) < 1e-9 * c, (alpha, s, c, cv)
            assert np.allclose(un, u + cv * d, rtol=1e-12, atol=1e-12)
            E1 = _subset_energy(A, g, _rows(fg, S, s), 1, alpha, eps, v)
            assert abs(c - _vertex(E1, u, d, c)) > 1e-3 * c


def test_os_at_pass_is_its_subset_steps():
    """or_os_at runs or_os_step over the subsets in order (no median filter,
    no edge step): bitwise the same u."""
    fg = O.fan_geom(16, 1.0, 24, 24, 1.0)
    op = O.op_fan(fg)
    rs = np.random.RandomState(23)
    g = O.fan_fp(fg, rs.rand(16, 16))
    u, v, n, _ = O.os_at(op, g, 1e-2, 1e-4, 1e-2, 4, 1, iter_edge=0)
    w = np.zeros((16, 16))
    for s in range(4):
        w, _, _ = O.os_step(op, g, w, np.ones((16, 16)), 1e-2, 1e-2, 4, s)
    assert np.array_equal(u, w)


def test_os_decreases_data_fidelity():
    fg = O.fan_geom(32, 1.0, 32, 48, 1.0)
    op = O.op_fan(fg)
    import medimage_amd.phantom as ph
    img = ph.rasterize(ph.random_ellipses(4, 32.0, 1, axes_mm=(2, 10)), 32, 1.0)
    g = O.fan_fp(fg, img)
    u, v, n, hist = O.os_at(op, g, 1e-3, 1e-4, 1e-2, 4, 6)
    assert n == 6 and np.all(np.diff(hist) < 0)
    assert np.all((v >= 0) & (v <= 1))


# --------------------------------------------------------------------------
# msbeam per-ray driver.

def test_msbeam_matches_reference_fixture(golden):
    """msbeam/CPU_OpenMP at one thread, 10 sweeps (tests/golden/msbeam_ref.npz)."""
    z = golden("msbeam_ref.npz")
    f, v = O.msbeam_solve(z["g"], 512, 10.0, 0.35, 0.001, 10, 0)
    assert np.array_equal(f.view(np.uint32), z["f"].view(np.uint32))
    assert np.array_equal(v.view(np.uint32), z["v"].view(np.uint32))


def test_cpu_baseline_matches_oracle():
    """The multi-threaded FP32 CPU baseline (oracle/cpu_at.c, what the
    benchmark's CPU leg times) runs the same iteration as the FP64 oracle:
    within FP32 rounding of it, and independent of the thread count."""
    fg = O.fan_geom(64, 4.0, 60, 96, 4.0)
    rs = np.random.RandomState(8)
    img = np.zeros((64, 64))
    img[20:44, 16:40] = 0.02
    g = O.fan_fp(fg, img) + rs.normal(0, 1e-3, (60, 96))
    op = O.op_fan(fg)
    ou, ov, n, _ = O.os_at(op, g, 1e-3, 1e-4, 1e-2, 4, 3)
    u1, v1, n1 = O.cpu_os_at_f32(op, g, 1e-3, 1e-4, 1e-2, 4, 3, threads=1)
    u4, v4, n4 = O.cpu_os_at_f32(op, g, 1e-3, 1e-4, 1e-2, 4, 3, threads=4)
    assert n1 == n4 == n == 3
    assert np.linalg.norm(u1 - ou) / np.linalg.norm(ou) < 1e-5
    assert np.linalg.norm(v1 - ov) / np.linalg.norm(ov) < 1e-5
    assert np.linalg.norm(u4 - u1) / np.linalg.norm(u1) < 1e-5



def test_arc_view_table_closed_forms():
    """Arc beam with a per-view angle table (a rebinned helical slice): a
    table equal to the uniform angles gives the uniform geometry's rays
    exactly; every source lies on the SOD circle at its table angle."""
    import medimage_amd as mi
    g = mi.arc_geometry(48, 6.0, 12, 64, 0.9)
    uni = O.fan_from(g)
    tab = O.fan_from(g.with_views(np.array([360.0 * v / 12 for v in range(12)])))
    assert np.array_equal(O.fan_csr(uni)[1], O.fan_csr(tab)[1])
    assert np.array_equal(O.fan_csr(uni)[2].view(np.uint64), O.fan_csr(tab)[2].view(np.uint64))
    ang = np.array([3.0, 17.5, 40.0, 44.0, 91.25, 130.0, 181.0, 200.0, 260.0, 299.0, 333.3, 719.0])
    fg = O.fan_from(g.with_views(ang))
    c = g.nx * g.pixel_mm / 2.0
    for v in range(12):
        sx, sy, dx, dy = O.fan_endpoints(fg, v, 10)
        th = np.deg2rad(ang[v])
        assert abs(sx - c - 595.0 * np.cos(th)) < 1e-9 and abs(sy - c - 595.0 * np.sin(th)) < 1e-9
        assert abs(np.hypot(dx - sx, dy - sy) - 1085.6) < 1e-9
