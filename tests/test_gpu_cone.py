"""GPU parity of the 3D cone-beam projectors (industry3d geometry) through
the C ABI, against the oracle (oracle/cone3d.c).

Bars: the traced (voxel, length) sequences bit-exact; the forward
projection bit-exact against the oracle's double sum over the same entries
rounded to float; the back projection within 1e-5 relative (L2 and max) of
the oracle's double transpose of the exact lengths.  At the full C4 size
(512^3, 720 views of a 1024 x 768 panel) sampled rays are checked bit-exactly, the back
projection of single rays against their exact lengths, and the pair
through the adjoint identity (the float walk and the exact lengths differ
by the walk's rounding and its dropped final pieces: 3e-6 at C4).  On the
small geometries the gather is also compared voxel by voxel with the
transpose of the float walk itself (the reference's scatter), within the
bound the dropped pieces set.
"""
from pathlib import Path

import numpy as np
import pytest

import medimage_amd as mi
from medimage_amd import configs, phantom
import oracle_py as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


SMALL = [
    mi.cone_geometry(20, 24, 16, 30, 20, 28, 1.0, 2.0, 60.0, 110.0),
    mi.cone_geometry(32, 32, 24, 45, 24, 40, 0.5, 0.8, 40.0, 70.0),
    mi.cone_geometry(24, 20, 16, 36, 24, 32, 1.0, 0.9, 30.0, 60.0, beam=mi.MI_CONE_PARALLEL),
    # odd sizes: a partial slice chunk (13 = 8 + 5), partial 16 x 16 tiles,
    # an odd panel (HALFDET = 8 of 17 rows), two back-projection launches
    mi.cone_geometry(18, 22, 13, 50, 17, 30, 1.0, 1.7, 45.0, 80.0),
    # two y tiles (70 > 64), a partial x tile (22 = 5 x 4 + 2) and two runs
    # of slices (40 > 32) in the volume stencils
    mi.cone_geometry(22, 70, 40, 24, 40, 80, 1.0, 1.6, 60.0, 110.0),
]


def _print(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("gi", range(len(SMALL)))
def test_cone_trace_bitexact(ctx, gi):
    g = SMALL[gi]
    cg = O.cone_from(g)
    rng = np.random.default_rng(11 + gi)
    hits = 0
    for _ in range(300):
        v, r, k = int(rng.integers(g.n_views)), int(rng.integers(g.n_rows)), int(rng.integers(g.n_cols))
        di, dl = ctx.cone_trace_ray(g, v, r, k)
        oi, ol = O.cone_trace(cg, v, r, k)
        assert np.array_equal(di, oi), (v, r, k)
        assert np.array_equal(dl.view(np.uint32), ol.view(np.uint32)), (v, r, k)
        hits += len(di) > 0
    assert hits > 100


@pytest.mark.parametrize("gi", range(len(SMALL)))
def test_cone_forward_bitexact(ctx, gi):
    g = SMALL[gi]
    _print(f"cone forward {g}")
    rng = np.random.default_rng(5 + gi)
    vol = rng.random((g.nz, g.nx, g.ny)).astype(np.float32)
    dev = ctx.cone_forward(g, vol)
    ref = O.cone_fp(O.cone_from(g), vol).astype(np.float32)
    assert np.count_nonzero(ref) > ref.size // 4
    assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("gi", range(len(SMALL)))
def test_cone_backproject_parity(ctx, gi):
    g = SMALL[gi]
    _print(f"cone back projection {g}")
    rng = np.random.default_rng(9 + gi)
    sino = rng.random((g.n_views, g.n_rows, g.n_cols)).astype(np.float32)
    dev = ctx.cone_backproject(g, sino)
    ref = O.cone_bp(O.cone_from(g), sino.astype(np.float64))
    assert rel_l2(dev, ref) < 1e-5
    assert rel_max(dev, ref) < 1e-5
    # deterministic: a second run is bitwise the same
    assert np.array_equal(ctx.cone_backproject(g, sino).view(np.uint32), dev.view(np.uint32))


def test_cone_pair_adjoint_small(ctx):
    """<A x, y> = <x, B y> with A the float walk and B the exact-length gather:
    3.2e-5 on this geometry (the walk drops each ray's final piece, below
    0.05 voxel, which is a larger share of these short rays than at C4)."""
    g = SMALL[1]
    rng = np.random.default_rng(1)
    x = rng.random((g.nz, g.nx, g.ny)).astype(np.float32)
    y = rng.random((g.n_views, g.n_rows, g.n_cols)).astype(np.float32)
    lhs = float(np.sum(ctx.cone_forward(g, x).astype(np.float64) * y))
    rhs = float(np.sum(x.astype(np.float64) * ctx.cone_backproject(g, y)))
    _print(f"small adjoint rel {abs(lhs - rhs) / abs(rhs):.2e}")
    assert abs(lhs - rhs) / abs(rhs) < 1e-4


def _walk_matrix(g):
    """The float walk's entries of every ray as a sparse matrix (the
    reference's forward_proj, industry3d/src/tracing.cpp:123-307, as the
    oracle restates it): the matrix whose transpose the reference applies
    when it scatters its gradient along the traced rays
    (industry3d/src/main.cpp:243-330, 385-406)."""
    import scipy.sparse as sp
    cg = O.cone_from(g)
    rows, cols, vals = [], [], []
    r = 0
    for v in range(g.n_views):
        for a in range(g.n_rows):
            for b in range(g.n_cols):
                ind, w = O.cone_trace(cg, v, a, b)
                rows.append(np.full(len(ind), r))
                cols.append(ind)
                vals.append(w.astype(np.float64))
                r += 1
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(r, g.nx * g.ny * g.nz))


@pytest.mark.parametrize("gi", range(len(SMALL)))
def test_cone_backproject_against_float_walk_transpose(ctx, gi):
    """The gather back projection (exact lengths) against the transpose of
    the float walk itself, voxel by voxel.  They differ by the walk's float
    rounding and by each ray's dropped final piece (below 0.05 voxel,
    industry3d/src/tracing.cpp:268): measured with the oracle on these five
    geometries, rel L2 1.9e-4 .. 4.1e-4 and rel max 1.0e-3 .. 4.7e-3."""
    g = SMALL[gi]
    rng = np.random.default_rng(9 + gi)
    y = rng.random((g.n_views, g.n_rows, g.n_cols)).astype(np.float32)
    bw = _walk_matrix(g).T @ y.reshape(-1).astype(np.float64)
    dev = ctx.cone_backproject(g, y).reshape(-1)
    el2, emax = rel_l2(dev, bw), rel_max(dev, bw)
    _print(f"cone bp vs float-walk transpose {g}: rel L2 {el2:.2e}, rel max {emax:.2e}")
    assert el2 < 6e-4 and emax < 7e-3


@pytest.mark.parametrize("g", [
    mi.cone_geometry(5, 7, 1, 1, 1, 3, 1.0, 1.0, 20.0, 40.0),                         # one slice, one view
    mi.cone_geometry(6, 6, 3, 2, 2, 2, 1.0, 1.0, 15.0, 30.0, beam=mi.MI_CONE_PARALLEL),
    mi.cone_geometry(9, 4, 5, 7, 9, 1, 0.7, 1.3, 12.0, 25.0),                         # one panel column
])
def test_cone_tiny_geometries(ctx, g):
    """Degenerate sizes: forward bitwise and back projection within 1e-5 of
    the oracle (or both zero)."""
    rng = np.random.default_rng(3)
    vol = rng.random((g.nz, g.nx, g.ny)).astype(np.float32)
    fp = ctx.cone_forward(g, vol)
    ofp = O.cone_fp(O.cone_from(g), vol).astype(np.float32)
    assert np.array_equal(fp.view(np.uint32), ofp.view(np.uint32))
    sino = rng.random((g.n_views, g.n_rows, g.n_cols)).astype(np.float32)
    bp = ctx.cone_backproject(g, sino)
    obp = O.cone_bp(O.cone_from(g), sino.astype(np.float64))
    if np.abs(obp).max() == 0:
        assert np.abs(bp).max() == 0
    else:
        assert rel_max(bp, obp) < 1e-5


class _DevBuf:
    """Device buffer through the HIP runtime the library itself uses."""

    def __init__(self, arr, offset=0):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = arr.nbytes
        self.base = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.base), C.c_size_t(self.n + 16)) == 0
        self.ptr = C.c_void_p(self.base.value + offset)
        assert self.hip.hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p), C.c_size_t(self.n), 1) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(self.C.c_void_p), self.ptr, self.C.c_size_t(self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.base)


@pytest.mark.parametrize("g", [
    mi.cone_geometry(5, 7, 1, 3, 2, 9, 1.0, 1.0, 20.0, 40.0),       # one slice, less than a tile
    mi.cone_geometry(3, 70, 2, 4, 3, 40, 1.0, 1.0, 60.0, 110.0),    # three x rows, two y tiles
])
def test_cone_at_tiny_volumes(ctx, g):
    """The volume solver where the stencil tiles are mostly outside the
    volume: u and v within 1e-5 of the FP64 oracle."""
    rng = np.random.default_rng(8)
    shape = (g.nz, g.nx, g.ny)
    sino = O.cone_fp(O.cone_from(g), 0.05 * rng.random(shape)).astype(np.float32)
    p = mi.at_params(0.5, 0.02, 0.05, n_subsets=2, max_iters=2)
    u, v, rep = ctx.cone_at_reconstruct(g, sino, p)
    ou, ov, n, _ = O.cone_os_at(O.cone_from(g), sino, np.zeros(shape), np.ones(shape), 0.5, 0.02, 0.05, 2, 2)
    assert rep.iterations == n == 2
    assert np.abs(ou).max() > 0 and ov.min() < 1
    assert rel_l2(u, ou) < 1e-5 and rel_max(u, ou) < 1e-5
    assert rel_l2(v, ov) < 1e-5 and rel_max(v, ov) < 1e-5


def test_cone_device_pointer_entry_points(ctx):
    """The _dev entry points on device buffers give the host ones' results
    bitwise."""
    g = SMALL[1]
    vol, sino = _cone_case(g, 7)
    dv, ds = _DevBuf(vol), _DevBuf(np.zeros((g.n_views, g.n_rows, g.n_cols), np.float32))
    ctx.cone_forward_dev(g, dv.ptr.value, ds.ptr.value)
    assert np.array_equal(ds.get(sino).view(np.uint32), ctx.cone_forward(g, vol).view(np.uint32))
    dv.free()
    ds.free()
    p = mi.at_params(0.5, 0.02, 0.05, n_subsets=3, max_iters=2)
    u, v, rep = ctx.cone_at_reconstruct(g, sino, p)
    shape = (g.nz, g.nx, g.ny)
    bs, bu, bv = _DevBuf(sino), _DevBuf(np.zeros(shape, np.float32)), _DevBuf(np.ones(shape, np.float32))
    rep2 = ctx.cone_at_reconstruct_dev(g, bs.ptr.value, bu.ptr.value, bv.ptr.value, p)
    assert rep2.iterations == rep.iterations and rep2.data_fidelity == rep.data_fidelity
    assert np.array_equal(bu.get(u).view(np.uint32), u.view(np.uint32))
    assert np.array_equal(bv.get(v).view(np.uint32), v.view(np.uint32))
    for b in (bs, bu, bv):
        b.free()


def test_cone_at_odd_volume_and_unaligned_buffers(ctx):
    """A volume whose voxel count is not a multiple of 4 (the updates' scalar
    tail) against the FP64 oracle, and the same solve on device buffers 4
    bytes past a 16-byte boundary (the updates' scalar path) bitwise equal."""
    g = mi.cone_geometry(17, 21, 11, 30, 20, 28, 1.0, 2.0, 60.0, 110.0)
    assert (g.nx * g.ny * g.nz) % 4 == 3
    vol, sino = _cone_case(g, 5)
    p = mi.at_params(0.5, 0.02, 0.05, n_subsets=3, max_iters=2)
    u, v, rep = ctx.cone_at_reconstruct(g, sino, p)
    shape = (g.nz, g.nx, g.ny)
    ou, ov, n, _ = O.cone_os_at(O.cone_from(g), sino, np.zeros(shape), np.ones(shape), 0.5, 0.02, 0.05, 3, 2)
    assert rel_l2(u, ou) < 1e-5 and rel_max(u, ou) < 1e-5
    assert rel_l2(v, ov) < 1e-5 and rel_max(v, ov) < 1e-5
    bs = _DevBuf(sino)
    bu, bv = _DevBuf(np.zeros(shape, np.float32), offset=4), _DevBuf(np.ones(shape, np.float32), offset=4)
    rep2 = ctx.cone_at_reconstruct_dev(g, bs.ptr.value, bu.ptr.value, bv.ptr.value, p)
    assert rep2.iterations == rep.iterations and rep2.data_fidelity == rep.data_fidelity
    assert np.array_equal(bu.get(u).view(np.uint32), u.view(np.uint32))
    assert np.array_equal(bv.get(v).view(np.uint32), v.view(np.uint32))
    for b in (bs, bu, bv):
        b.free()


def test_cone_geometry_beyond_index_width_rejected(ctx):
    """A volume under 2^31 voxels whose z nx exceeds 2^24 (the forward walk's
    24-bit offset multiplies) is refused with MI_E_LIMIT.  The call is
    mi_cone_trace_ray, which works on buffers of its own, so no caller
    pointer reaches the device whatever the check decides; at z nx = 2^24
    exactly (the largest offset is 2^24 - 1) the geometry is accepted."""
    import ctypes as C
    n = C.c_int()
    g = mi.cone_geometry(4097, 64, 4096, 4, 4, 4, 0.1, 1.0, 600.0, 1000.0)
    rc = ctx._lib.mi_cone_trace_ray(ctx._h, C.byref(g), 0, 0, 0, None, None, 0, C.byref(n))
    assert rc == mi.MI_E_LIMIT
    assert b"2^24" in ctx._lib.mi_last_error()
    g = mi.cone_geometry(4096, 64, 4096, 4, 4, 4, 0.1, 1.0, 600.0, 1000.0)
    rc = ctx._lib.mi_cone_trace_ray(ctx._h, C.byref(g), 0, 0, 0, None, None, 0, C.byref(n))
    assert rc == mi.MI_OK, ctx._lib.mi_last_error()
    ctx.trim()


def test_cone_tables_in_context_memory(ctx):
    g = SMALL[0]
    ctx.trim()
    m0 = ctx.mem_bytes()
    ctx.cone_forward(g, np.zeros((g.nz, g.nx, g.ny), np.float32))
    assert ctx.mem_bytes() > m0
    ctx.trim()
    assert ctx.mem_bytes() == 0


def test_cone_rejects_bad_geometry(ctx):
    g = mi.cone_geometry(64, 64, 64, 10, 8, 8, 1.0, 1.0, 30.0, 60.0)   # source inside the volume
    with pytest.raises(mi.MIError):
        ctx.cone_forward(g, np.zeros((64, 64, 64), np.float32))
    g = mi.cone_geometry(8, 8, 8, 10, 8, 8, 1.0, 1.0, 60.0, 50.0)      # sdd < sod
    with pytest.raises(mi.MIError):
        ctx.cone_backproject(g, np.zeros((10, 8, 8), np.float32))


def _cone_case(g, seed, n_ell=6):
    fov = g.nx * g.sample_mm
    ell = phantom.random_ellipsoids(n_ell, fov, seed, axes_mm=(0.05 * fov, 0.35 * fov), mu_max=0.05)
    vol = phantom.rasterize3d(ell, g.nz, g.nx, g.ny, g.sample_mm)
    p = O.cone_fp(O.cone_from(g), vol)
    return vol, phantom.noisy_sinogram(p, 5e4, seed).astype(np.float32)


@pytest.mark.parametrize("prm", [(1e-3, 1e-4, 1e-2, 5, 3, 1), (0.5, 0.02, 0.05, 3, 2, 1),
                                 (0.5, 0.02, 0.05, 4, 2, 2), (0.5, 0.02, 0.05, 4, 2, 4)])
def test_cone_at_parity(ctx, prm):
    """Ordered-subset AT on the volume: FP32 device run against the FP64
    oracle (1e-5 relative, u and v), the data fidelity of the last pass, and
    a second run bitwise identical."""
    alpha, beta, eps, S, iters, gi = prm
    g = SMALL[gi]
    _print(f"cone AT {g} alpha {alpha} beta {beta} eps {eps} subsets {S} passes {iters}")
    vol, sino = _cone_case(g, 12)
    p = mi.at_params(alpha, beta, eps, n_subsets=S, max_iters=iters)
    u, v, rep = ctx.cone_at_reconstruct(g, sino, p)
    shape = (g.nz, g.nx, g.ny)
    ou, ov, n, hist = O.cone_os_at(O.cone_from(g), sino, np.zeros(shape), np.ones(shape), alpha, beta, eps, S,
                                   iters)
    assert rep.iterations == n == iters
    _print(f"  u rel L2 {rel_l2(u, ou):.2e} max {rel_max(u, ou):.2e}; v rel L2 {rel_l2(v, ov):.2e} "
           f"max {rel_max(v, ov):.2e}; 1 - min v {1 - ov.min():.3f}")
    assert rel_l2(u, ou) < 1e-5 and rel_max(u, ou) < 1e-5
    assert rel_l2(v, ov) < 1e-5 and rel_max(v, ov) < 1e-5
    assert abs(rep.data_fidelity - hist[-1]) < 1e-5 * hist[-1]
    assert rel_l2(u, vol) < rel_l2(np.zeros_like(u), vol)
    u2, v2, _ = ctx.cone_at_reconstruct(g, sino, p)
    assert np.array_equal(u2.view(np.uint32), u.view(np.uint32))
    assert np.array_equal(v2.view(np.uint32), v.view(np.uint32))


def test_cone_at_rejects(ctx):
    g = SMALL[0]
    sino = np.zeros((g.n_views, g.n_rows, g.n_cols), np.float32)
    with pytest.raises(mi.MIError):
        ctx.cone_at_reconstruct(g, sino, mi.at_params(n_subsets=5, max_iters=1, median_every=2))
    with pytest.raises(mi.MIError):
        ctx.cone_at_reconstruct(g, sino, mi.at_params(n_subsets=5, max_iters=1, precision=mi.MI_FP64))
    with pytest.raises(mi.MIError):
        ctx.cone_at_reconstruct(g, sino, mi.at_params(n_subsets=g.n_views + 1, max_iters=1))


def test_cone_c4_full_size(ctx):
    """512^3, 720 views, 1024 x 768 panel: sampled rays bit-exact, single-ray
    back projections against their exact lengths, the adjoint identity."""
    g = configs.c4_geometry()
    cg = O.cone_from(g)
    _print(f"C4 full size {g}")
    rng = np.random.default_rng(4)
    vol = rng.random((g.nz, g.nx, g.ny), dtype=np.float32)
    sino = ctx.cone_forward(g, vol)
    vol_h = vol.reshape(-1)
    rays = [(int(rng.integers(g.n_views)), int(rng.integers(g.n_rows)), int(rng.integers(g.n_cols)))
            for _ in range(48)] + [(0, g.n_rows // 2, g.n_cols // 2 - 100), (359, 400, 300)]
    for (v, r, k) in rays:
        ind, w = O.cone_trace(cg, v, r, k)
        acc = 0.0
        for i, l in zip(ind, w):
            acc += float(vol_h[i]) * float(l)
        assert np.float32(acc) == sino[v, r, k], (v, r, k, acc, sino[v, r, k])
    _print("C4 forward sampled rays bit-exact")

    # back projection of a few single rays
    one = np.zeros_like(sino)
    want = {}
    for (v, r, k), val in zip(rays[:6], (1.0, 0.5, 2.0, 1.5, 0.25, 3.0)):
        one[v, r, k] = val
        S, D = O.cone_ray(cg, v, r, k)
        ind, ln = O.cone_exact(cg, S, D)
        for i, l in zip(ind, ln):
            want[int(i)] = want.get(int(i), 0.0) + val * l
    out_h = ctx.cone_backproject(g, one).reshape(-1)
    del one
    nz = np.nonzero(out_h)[0]
    keys = np.array(sorted(k for k, x in want.items() if x > 1e-9))
    assert set(nz.tolist()) <= set(want.keys())
    wv = np.array([want[int(i)] for i in keys])
    assert rel_max(out_h[keys], wv) < 1e-5
    _print("C4 single-ray back projections match the exact lengths")

    # adjoint identity on random data
    y = rng.random(sino.shape, dtype=np.float32)
    bty = ctx.cone_backproject(g, y)
    lhs = float(np.dot(sino.reshape(-1).astype(np.float64), y.reshape(-1).astype(np.float64)))
    rhs = float(np.dot(vol_h.astype(np.float64), bty.reshape(-1).astype(np.float64)))
    _print(f"C4 adjoint: <Ax,y> {lhs:.9e} <x,Bty> {rhs:.9e} rel {abs(lhs - rhs) / abs(rhs):.2e}")
    assert abs(lhs - rhs) / abs(rhs) < 1e-5


def test_cone_config_and_dr_file_into_device_solve(ctx, tmp_path):
    """industry3d's own configuration (tests/golden/industry3d_cone_taiji.json,
    BEAM "Cone") scaled down by four on every axis with the field of view
    kept, parsed by mi_cone_config_parse; its projections written as a .dr
    file (1024-byte header, ushort [NPROJ][NDX][NDY], then the blank scan;
    industry3d/src/ct3d.cpp:199-214) and read back by mi_read_dr; then the
    device's ordered-subset AT solve from them (ALPHA, BETA, EPSILON and
    LOWER_X / UPPER_X from the file) against the FP64 oracle, within 1e-5."""
    import json
    cfg = json.loads((Path(__file__).parent / "golden" / "industry3d_cone_taiji.json").read_text())
    for k in ("NX", "NY", "NZ", "NDX", "NDY"):
        cfg[k] //= 4
    cfg["NPROJ"] = 24
    cfg["SAMPLESIZE"] *= 4
    cfg["PIXELSIZE"] *= 4
    g, p = mi.cone_config_parse(json.dumps(cfg))
    assert (g.nx, g.nz, g.n_views, g.n_rows, g.n_cols) == (32, 32, 24, 64, 64)
    assert (p.alpha, p.beta, p.epsilon, p.lower, p.upper) == (2.0, 0.001, 0.125, 0.0, 10.0)
    rng = np.random.default_rng(12)
    vol = np.zeros((g.nz, g.nx, g.ny))
    vol[8:24, 6:26, 10:22] = 0.02 * (1.0 + rng.random((16, 20, 12)))
    line = O.cone_fp(O.cone_from(g), vol)
    counts = np.round(line * 1000.0).astype(np.uint16)          # the file's integer projections
    flat = np.full((g.n_rows, g.n_cols), 1000.0, np.float32)
    dr = tmp_path / "taiji_small.dr"
    dr.write_bytes(bytes(1024) + counts.tobytes() + flat.tobytes())
    proj, fl = mi.read_dr(dr, g.n_views, g.n_rows, g.n_cols)
    assert np.array_equal(proj, counts) and np.array_equal(fl, flat)
    sino = (proj.astype(np.float64) / 1000.0).astype(np.float32)   # the caller's unit conversion
    p.n_subsets = 4
    p.max_iters = 2
    u, v, rep = ctx.cone_at_reconstruct(g, sino, p)
    shape = (g.nz, g.nx, g.ny)
    ou, ov, n, _ = O.cone_os_at(O.cone_from(g), sino, np.zeros(shape), np.ones(shape), p.alpha, p.beta, p.epsilon,
                                4, 2, lower=p.lower, upper=p.upper)
    _print(f"taiji scaled: {rep.iterations} passes, rel L2 u {rel_l2(u, ou):.2e} v {rel_l2(v, ov):.2e}")
    assert rep.iterations == n == 2
    assert rel_l2(u, ou) < 1e-5 and rel_max(u, ou) < 1e-5
    assert rel_l2(v, ov) < 1e-5 and rel_max(v, ov) < 1e-5
