"""Pins of the 3D cone-beam oracle (oracle/cone3d.c), no GPU.

The reference's tracer (industry3d/src/tracing.cpp) needs boost through its
utility.cpp and is not compiled here, so the float restatement is pinned by
closed forms and exact rational arithmetic:

* cone() end points: source on the circle of radius SOD / SAMPLESIZE around
  the volume centre at the view angle, the panel plane perpendicular to the
  central ray at SDD, pixel centres on the (HALFDET - 0.5)-shifted grid;
* forward_proj: axis-aligned rays give unit lengths and one voxel per plane;
  for general rays the voxel sequence equals an exact enumeration with
  Fractions, lengths agree to float resolution, the count is the number of
  plane crossings inside the volume plus one (less a dropped final piece
  shorter than 0.05), and the lengths add up to the chord;
* the exact-length matrix (or_cone_exact, the back projector's) equals a
  per-voxel slab clip of every voxel of a small grid;
* the 26-neighbourhood AT stencils equal finite differences of the energies
  written independently in numpy (pair sums over the padded grid), the edge
  step is the exact minimiser along its direction, and one ordered-subset
  pass equals its restatement from dense matrices.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_py as O


def geom(nx=12, ny=10, nz=8, nproj=24, ndx=10, ndy=14, sample=1.0, pixel=2.0, sod=40.0, sdd=75.0):
    return O.ConeGeom(nx, ny, nz, nproj, ndx, ndy, sample, pixel, sod, sdd)


def test_cone_ray_closed_form():
    g = geom(sample=0.5, pixel=1.5)
    c, s = O.cone_tables(g)
    ctr = np.array([g.nx / 2, g.ny / 2, g.nz / 2])
    for v in (0, 5, 13):
        th = 2 * math.pi * v / g.nproj
        assert abs(c[v] - math.cos(th)) < 1e-6 and abs(s[v] - math.sin(th)) < 1e-6
        for r, k in ((0, 0), (3, 7), (g.ndx - 1, g.ndy - 1)):
            S, D = O.cone_ray(g, v, r, k)
            S = S.astype(np.float64)
            D = D.astype(np.float64)
            # source: (0, -SOD) rotated by theta, in voxels, z at the centre
            want_s = ctr + np.array([math.sin(th) * g.sod, -math.cos(th) * g.sod, 0.0]) / g.sample
            assert np.allclose(S, want_s, atol=1e-4)
            # panel point: ((k - H + .5) p, SDD - SOD, (H - r - .5) p) rotated
            H = g.ndx // 2
            px, py, pz = (k - H + 0.5) * g.pixel, g.sdd - g.sod, (H - r - 0.5) * g.pixel
            want_d = ctr + np.array([math.cos(th) * px - math.sin(th) * py,
                                     math.sin(th) * px + math.cos(th) * py, pz]) / g.sample
            assert np.allclose(D, want_d, atol=1e-4)
            # the central ray is perpendicular to the panel at distance SDD
            u = (ctr - S) / np.linalg.norm(ctr - S)
            assert abs(np.dot(D - S, u) - g.sdd / g.sample) < 1e-3


def test_parallel_ray_closed_form():
    """parallel(): start on the plane at -SOD, end on the panel at SDD - SOD,
    both at the column's offset and the row's height, every axis centred on
    NX / 2."""
    g = geom(nx=12, ny=10, nz=8, sample=0.5, pixel=1.5)
    g.parallel = 1
    half = g.nx / 2
    for v in (0, 7, 19):
        th = 2 * math.pi * v / g.nproj
        for r, k in ((0, 0), (4, 9)):
            S, D = O.cone_ray(g, v, r, k)
            H = g.ndx // 2
            px, pz = (k + 0.5 - H) * g.pixel, (H - r - 0.5) * g.pixel
            for y0, P in ((-g.sod, S), (g.sdd - g.sod, D)):
                want = np.array([math.cos(th) * px - math.sin(th) * y0, math.sin(th) * px + math.cos(th) * y0,
                                 pz]) / g.sample + half
                assert np.allclose(P, want, atol=1e-4)
            assert S[2] == D[2]


def test_parallel_walk_matches_exact_enumeration():
    g = geom(nx=10, ny=10, nz=6, ndx=8, ndy=12, pixel=1.0, sod=30.0, sdd=50.0)
    g.parallel = 1
    rng = np.random.default_rng(9)
    checked = 0
    for _ in range(300):
        v, r, k = int(rng.integers(g.nproj)), int(rng.integers(g.ndx)), int(rng.integers(g.ndy))
        S, D = O.cone_ray(g, v, r, k)
        N = (g.nx, g.ny, g.nz)
        if any(S[a] == D[a] and not (0 <= S[a] < N[a]) for a in range(3)):
            assert len(O.cone_walk(g, S, D)[0]) == 0     # parallel to a face, outside the volume
            continue
        ind, w = O.cone_walk(g, S, D)
        seq, lens, chord, crossings = _exact_walk(g, S, D)
        if not seq or min(float(x) for x in lens) < 1e-3:
            continue
        n = len(ind)
        assert n in (len(seq), len(seq) - 1) and list(ind) == seq[:n]
        assert np.allclose(w, [float(x) for x in lens[:n]], atol=2e-4)
        checked += 1
    assert checked > 100


def test_walk_axis_aligned():
    g = geom()
    # along +y through voxel column (x=3, z=5): ny unit lengths
    ind, w = O.cone_walk(g, [3.5, -20.0, 5.5], [3.5, 30.0, 5.5])
    assert len(ind) == g.ny
    assert np.all(w == 1.0)
    assert list(ind) == [5 * g.nx * g.ny + 3 * g.ny + y for y in range(g.ny)]
    # along -z through (x=7, y=2)
    ind, w = O.cone_walk(g, [7.5, 2.5, 40.0], [7.5, 2.5, -10.0])
    assert len(ind) == g.nz and np.all(w == 1.0)
    assert list(ind) == [z * g.nx * g.ny + 7 * g.ny + 2 for z in reversed(range(g.nz))]
    # along +x
    ind, w = O.cone_walk(g, [-3.0, 4.5, 1.5], [20.0, 4.5, 1.5])
    assert len(ind) == g.nx and np.all(w == 1.0)
    assert list(ind) == [1 * g.nx * g.ny + x * g.ny + 4 for x in range(g.nx)]


def _exact_walk(g, S, D):
    """Voxel sequence and lengths of segment S -> D through the grid,
    exactly (Fractions): merged plane crossings inside the volume."""
    S = [Fraction(float(a)) for a in S]
    D = [Fraction(float(a)) for a in D]
    N = (g.nx, g.ny, g.nz)
    r = [D[a] - S[a] for a in range(3)]
    lo, hi = Fraction(0), Fraction(1)
    for a in range(3):
        if r[a] == 0:
            assert 0 <= S[a] < N[a]
            continue
        t0, t1 = (0 - S[a]) / r[a], (N[a] - S[a]) / r[a]
        lo, hi = max(lo, min(t0, t1)), min(hi, max(t0, t1))
    if lo >= hi:
        return [], [], Fraction(0), 0
    ts = {lo, hi}
    crossings = 0
    for a in range(3):
        if r[a] == 0:
            continue
        for k in range(N[a] + 1):
            t = (k - S[a]) / r[a]
            if lo < t < hi:
                ts.add(t)
                crossings += 1
    ts = sorted(ts)
    L2 = sum(x * x for x in r)
    seq, lens = [], []
    for t1, t2 in zip(ts, ts[1:]):
        tm = (t1 + t2) / 2
        v = [math.floor(S[a] + tm * r[a]) for a in range(3)]
        seq.append(v[2] * g.nx * g.ny + v[0] * g.ny + v[1])
        lens.append((t2 - t1) * math.sqrt(L2))
    return seq, lens, (hi - lo) * math.sqrt(L2), crossings


def test_walk_matches_exact_enumeration():
    """Random rays of the cone geometry: the float walk visits exactly the
    voxels of the exact enumeration (the last one may be dropped when its
    piece is shorter than 0.05), lengths to float resolution."""
    g = geom()
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(400):
        v, r, k = int(rng.integers(g.nproj)), int(rng.integers(g.ndx)), int(rng.integers(g.ndy))
        S, D = O.cone_ray(g, v, r, k)
        ind, w = O.cone_walk(g, S, D)
        ind2, w2 = O.cone_trace(g, v, r, k)
        assert np.array_equal(ind, ind2) and np.array_equal(w, w2)
        seq, lens, chord, crossings = _exact_walk(g, S, D)
        if not seq:
            assert len(ind) == 0
            continue
        # rays through an edge or corner (two crossings within float
        # resolution) are left to the exact-length tests below
        if min(float(x) for x in lens) < 1e-3:
            continue
        n = len(ind)
        assert n in (len(seq), len(seq) - 1)
        if n == len(seq) - 1:
            assert float(lens[-1]) < 0.05 + 1e-3
        assert list(ind) == seq[:n]
        assert n >= crossings  # count = crossings + 1, less a dropped tail
        assert np.allclose(w, [float(x) for x in lens[:n]], atol=2e-4)
        assert abs(float(w.astype(np.float64).sum()) - float(chord)) < 0.05 + 1e-3
        checked += 1
    assert checked > 250


def test_exact_lengths_equal_slab_clip():
    g = geom(nx=6, ny=5, nz=4, nproj=6, ndx=6, ndy=8, pixel=3.0, sod=25.0, sdd=45.0)
    for v in range(g.nproj):
        for r in range(g.ndx):
            for k in range(0, g.ndy, 3):
                S, D = O.cone_ray(g, v, r, k)
                ind, ln = O.cone_exact(g, S, D)
                dense = np.zeros(g.nx * g.ny * g.nz)
                np.add.at(dense, ind, ln)
                brute = np.zeros_like(dense)
                for z in range(g.nz):
                    for x in range(g.nx):
                        for y in range(g.ny):
                            brute[z * g.nx * g.ny + x * g.ny + y] = O.cone_clip(S, D, x, y, z)
                assert np.allclose(dense, brute, rtol=0, atol=1e-12)


def test_fp_bp_pair():
    """or_cone_fp sums the traced entries in double; or_cone_bp is the
    transpose of the exact lengths.  On a small geometry the two agree as
    adjoints to the walk's float resolution and dropped tails."""
    g = geom()
    rng = np.random.default_rng(3)
    x = rng.random((g.nz, g.nx, g.ny)).astype(np.float32)
    y = rng.random((g.nproj, g.ndx, g.ndy))
    Ax = O.cone_fp(g, x)
    # or_cone_fp is the walk's entries summed in order
    v, r, k = 5, 4, 6
    ind, w = O.cone_trace(g, v, r, k)
    acc = 0.0
    for i, l in zip(ind, w):
        acc += float(x.reshape(-1)[i]) * float(l)
    assert Ax[v, r, k] == acc
    Bty = O.cone_bp(g, y)
    lhs = float(np.sum(Ax * y))
    rhs = float(np.sum(x.astype(np.float64) * Bty))
    assert abs(lhs - rhs) / abs(rhs) < 2e-3


# --------------------------------------------------------------------------
# 26-neighbourhood AT stencils and the 3D ordered-subset step, checked
# against an independent numpy statement of the energies (pair sums over the
# padded volume) and finite differences.
import ctypes as C

OFFS = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)
        if (a, b, c) > (0, 0, 0)]


def _pairs(p, wfun):
    """sum over offsets k of the 13 forward ones and all x of the one-voxel
    padded grid where x or x + k is inside: w_k * wfun(x) * (p(x+k) - p(x))^2,
    p zero outside."""
    nz, nx, ny = p.shape
    P = np.zeros((nz + 4, nx + 4, ny + 4))
    P[2:-2, 2:-2, 2:-2] = p
    inside = np.zeros(P.shape, bool)
    inside[2:-2, 2:-2, 2:-2] = True
    W = wfun(P.shape)
    tot = 0.0
    for (a, b, c) in OFFS:
        w = 1.0 / (a * a + b * b + c * c)
        sl0 = (slice(1, -2), slice(1, -2), slice(1, -2))
        x0 = P[1:-1, 1:-1, 1:-1]
        xk = P[1 + a:P.shape[0] - 1 + a, 1 + b:P.shape[1] - 1 + b, 1 + c:P.shape[2] - 1 + c]
        m = inside[1:-1, 1:-1, 1:-1] | inside[1 + a:P.shape[0] - 1 + a, 1 + b:P.shape[1] - 1 + b,
                                              1 + c:P.shape[2] - 1 + c]
        tot += w * float(np.sum((W[1:-1, 1:-1, 1:-1] * (xk - x0) ** 2)[m]))
    return tot


def _R(f, v, eps):
    eta = 0.01 * eps * eps
    def wf(shape):
        V = np.ones(shape)
        V[2:-2, 2:-2, 2:-2] = v
        return V * V + eta
    return _pairs(f, wf)


def _Q(p):
    return _pairs(p, lambda shape: np.ones(shape))


def _n(f):
    out = np.zeros_like(f)
    nz, nx, ny = f.shape
    F = np.zeros((nz + 2, nx + 2, ny + 2))
    F[1:-1, 1:-1, 1:-1] = f
    for (a, b, c) in OFFS:
        w = 1.0 / (a * a + b * b + c * c)
        out += w * (F[1 + a:nz + 1 + a, 1 + b:nx + 1 + b, 1 + c:ny + 1 + c] - f) ** 2
    return out


def _E_edge(v, f, alpha, beta, eps):
    V1 = v - 1.0       # Q of v with one-padding equals Q of (v - 1) with zero-padding
    return alpha * float(np.sum(v * v * _n(f))) + beta * eps * _Q(V1) + beta / (4 * eps) * float(np.sum(V1 ** 2))


def _fd_grad(E, x, h=1e-3):
    g = np.zeros_like(x)
    it = np.nditer(x, flags=["multi_index"])
    for _ in it:
        i = it.multi_index
        xp = x.copy(); xp[i] += h
        xm = x.copy(); xm[i] -= h
        g[i] = (E(xp) - E(xm)) / (2 * h)
    return g


def test_cone_stencils_match_energies():
    L = O._cone_lib()
    rng = np.random.default_rng(5)
    nz, nx, ny = 3, 4, 5
    f = rng.random((nz, nx, ny))
    v = 0.2 + 0.6 * rng.random((nz, nx, ny))
    eps, alpha = 0.3, 0.7
    d = np.zeros_like(f)
    L.or_cone_ugrad.argtypes = [C.c_void_p] * 3 + [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
    L.or_cone_ugrad(O.P(d), O.P(f), O.P(v), alpha, eps, nz, nx, ny)
    want = -0.5 * alpha * _fd_grad(lambda x: _R(x, v, eps), f)
    assert np.allclose(d, want, rtol=1e-8, atol=1e-10)
    L.or_cone_nV.restype = C.c_double
    L.or_cone_nV.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int]
    p = rng.standard_normal((nz, nx, ny))
    assert abs(L.or_cone_nV(O.P(p), O.P(v), eps, nz, nx, ny) - _R(p, v, eps)) < 1e-10 * _R(p, v, eps)
    # invariants: constant interior f with v = 1 has only the boundary terms;
    # f = 0 gives a zero gradient
    d0 = np.zeros_like(f)
    L.or_cone_ugrad(O.P(d0), O.P(np.zeros_like(f)), O.P(v), alpha, eps, nz, nx, ny)
    assert not d0.any()


def test_cone_edge_step_exact_line_search():
    L = O._cone_lib()
    L.or_cone_edge_step.argtypes = [C.c_void_p, C.c_void_p] + [C.c_double] * 3 + [C.c_int] * 4
    rng = np.random.default_rng(6)
    nz, nx, ny = 3, 4, 4
    f = 0.3 * rng.random((nz, nx, ny))
    v = 0.4 + 0.2 * rng.random((nz, nx, ny))
    alpha, beta, eps = 0.5, 0.05, 0.2
    E = lambda x: _E_edge(x, f, alpha, beta, eps)
    dv = -0.5 * _fd_grad(E, v)
    # exact minimiser of the quadratic E(v + c dv)
    e0, e1, e2 = E(v), E(v + dv), E(v - dv)
    c = (e2 - e1) / (2 * (e1 + e2 - 2 * e0))
    vn = v.copy()
    L.or_cone_edge_step(O.P(vn), O.P(f), alpha, beta, eps, 1, nz, nx, ny)
    assert np.all((vn > 0) & (vn < 1))      # no clamping in this case
    assert np.allclose(vn, v + c * dv, rtol=1e-7, atol=1e-9)
    # v = 1 with constant f = 0 is a fixed point
    one = np.ones_like(v)
    L.or_cone_edge_step(O.P(one), O.P(np.zeros_like(f)), alpha, beta, eps, 2, nz, nx, ny)
    assert np.all(one == 1.0)


def test_cone_os_at_step_from_dense_matrices():
    """One pass of two subsets, no clamp and no edge step, rebuilt from the
    dense matrices of the traced entries (A) and of the exact lengths (B),
    numpy's regulariser gradient and the exact line search."""
    g = geom(nx=5, ny=4, nz=3, nproj=8, ndx=5, ndy=7, pixel=2.5, sod=25.0, sdd=45.0)
    nvox = g.nx * g.ny * g.nz
    nr = g.ndx * g.ndy
    A = np.zeros((g.nproj * nr, nvox))
    B = np.zeros((g.nproj * nr, nvox))
    for a in range(g.nproj):
        for r in range(g.ndx):
            for k in range(g.ndy):
                row = (a * g.ndx + r) * g.ndy + k
                ind, w = O.cone_trace(g, a, r, k)
                np.add.at(A[row], ind, w.astype(np.float64))
                S, D = O.cone_ray(g, a, r, k)
                ind, ln = O.cone_exact(g, S, D)
                np.add.at(B[row], ind, ln)
    assert np.count_nonzero(A.sum(1)) > 40
    rng = np.random.default_rng(8)
    sino = A @ rng.random(nvox) + 0.01 * rng.standard_normal(A.shape[0])
    u0 = 0.1 * rng.random(nvox)
    v0 = 0.5 + 0.5 * rng.random(nvox)
    alpha, beta, eps, S = 0.05, 0.01, 0.3, 2
    u = u0.copy()
    shape = (g.nz, g.nx, g.ny)
    order = np.zeros(S, np.int32)
    O.lib().or_subset_order(S, O.P(order))
    for s in range(S):
        views = [a for a in range(g.nproj) if a % S == order[s]]
        rows = np.concatenate([np.arange(a * nr, (a + 1) * nr) for a in views])
        q = sino[rows] - A[rows] @ u
        d = S * (B[rows].T @ q)
        d += -0.5 * alpha * _fd_grad(lambda x: _R(x, v0.reshape(shape), eps), u.reshape(shape)).reshape(-1)
        c = float(d @ d) / (S * float(np.sum((A[rows] @ d) ** 2)) + alpha * _R(d.reshape(shape), v0.reshape(shape), eps))
        u = u + c * d
    p = O.AtParams(alpha, beta, eps, S, 1, 0, 0, 0.0, -1e30, 1e30)
    uo = u0.copy()
    vo = v0.copy()
    L = O._cone_lib()
    L.or_cone_os_at.argtypes = [C.c_void_p] * 5 + [C.c_void_p]
    n = L.or_cone_os_at(C.byref(g), O.P(sino), O.P(uo), O.P(vo), C.byref(p), None)
    assert n == 1
    assert np.allclose(uo, u, rtol=1e-6, atol=1e-9)
    assert np.array_equal(vo, v0)
