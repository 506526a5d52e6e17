"""The allocation-free cone back projection (mi_cone_backproject_dev_noalloc)
through the C ABI against mi_cone_backproject_dev, bit for bit, on two calls
in a row, the second with a smaller input; both geometries have more than 36
views, so the carried sums go through the device's fixed scratch array.
Also: a volume beyond the scratch is refused (MI_E_LIMIT), as is a geometry
whose tables are not on the device yet (MI_E_STATE)."""
import ctypes as C

import numpy as np
import pytest

import medimage_amd as mi

pytestmark = pytest.mark.gpu

HIP = C.CDLL("libamdhip64.so.7")


class DevBuf:
    def __init__(self, arr):
        self.n = arr.nbytes
        self.p = C.c_void_p()
        assert HIP.hipMalloc(C.byref(self.p), C.c_size_t(self.n)) == 0
        assert HIP.hipMemcpy(self.p, arr.ctypes.data_as(C.c_void_p), C.c_size_t(self.n), 1) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert HIP.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(self.n), 2) == 0
        return out

    def free(self):
        HIP.hipFree(self.p)


def _case(g, seed):
    rng = np.random.default_rng(seed)
    return rng.random((g.n_views, g.n_rows, g.n_cols)).astype(np.float32)


def _both(ctx, g, sino):
    lib = ctx._lib
    zeros = np.zeros((g.nz, g.nx, g.ny), np.float32)
    ds, da, db = DevBuf(sino), DevBuf(zeros), DevBuf(zeros)
    try:
        assert lib.mi_cone_backproject_dev(ctx._h, C.byref(g), ds.p, da.p) == mi.MI_OK
        rc = lib.mi_cone_backproject_dev_noalloc(ctx._h, C.byref(g), ds.p, db.p)
        assert rc == mi.MI_OK, lib.mi_last_error()
        assert HIP.hipDeviceSynchronize() == 0
        return da.get(zeros), db.get(zeros)
    finally:
        for b in (ds, da, db):
            b.free()


def test_noalloc_backprojection_matches_allocating_one(ctx):
    big = mi.cone_geometry(48, 40, 32, 90, 40, 56, 1.0, 1.5, 80.0, 150.0)     # 3 launches of views
    small = mi.cone_geometry(20, 24, 16, 50, 20, 28, 1.0, 2.0, 60.0, 110.0)   # 2 launches, fewer voxels
    for g, seed in ((big, 1), (small, 2)):
        print(f"[noalloc] {g}", flush=True)
        a, b = _both(ctx, g, _case(g, seed))
        assert np.abs(a).max() > 0
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noalloc_refusals(ctx):
    """Both refusals get device buffers of the geometry's full size, so even
    a wrong decision could not make a kernel touch memory it does not own."""
    lib = ctx._lib
    # tables of this geometry were never built on this context
    g = mi.cone_geometry(12, 12, 6, 40, 8, 12, 1.0, 2.0, 60.0, 110.0)
    ds, dv = DevBuf(_case(g, 3)), DevBuf(np.zeros((g.nz, g.nx, g.ny), np.float32))
    ctx.trim()
    assert lib.mi_cone_backproject_dev_noalloc(ctx._h, C.byref(g), ds.p, dv.p) == mi.MI_E_STATE
    ds.free()
    dv.free()
    # a volume above the 8 Mi-voxel scratch with more than 36 views: its
    # tables are built first (40 views of a 4 x 4 panel), then the call is
    # refused
    g = mi.cone_geometry(256, 256, 129, 40, 4, 4, 0.5, 1.0, 600.0, 1000.0)
    rc = lib.mi_cone_trace_ray(ctx._h, C.byref(g), 0, 0, 0, None, None, 0, C.byref(C.c_int()))
    assert rc == mi.MI_OK, lib.mi_last_error()
    ds, dv = DevBuf(_case(g, 4)), DevBuf(np.zeros((g.nz, g.nx, g.ny), np.float32))
    rc = lib.mi_cone_backproject_dev_noalloc(ctx._h, C.byref(g), ds.p, dv.p)
    assert rc == mi.MI_E_LIMIT
    assert b"scratch" in lib.mi_last_error()
    ds.free()
    dv.free()
    ctx.trim()
