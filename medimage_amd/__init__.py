"""medimage_amd: MI355X-native Ambrosio-Tortorelli low-dose CT reconstruction.

Python mirror of the C ABI in include/medimage_amd.h (ctypes).  The compute
path is libmedimage_amd.so (HIP, gfx950); importing this package without the
built library raises, and there is no CPU fallback.

Names follow the reference drivers: ``forward`` / ``backproject`` replace
``A`` / ``AStar`` (origin/GoldenC/src/proj.h:6-7, msbeam/CPU_OpenMP/src/proj.h:7-8),
``mumford_shah`` replaces ``MumfordShah`` (origin/GoldenC/src/main.c:35-49),
``image_step`` / ``edge_step`` replace ``minimize_in_IMAGE_variable`` /
``minimize_in_EDGE_variable`` (main.c:81-360), ``at_reconstruct`` is the
ordered-subset driver that takes the place of ``MSbeam``
(msbeam/CPU_OpenMP/src/main.c:364-394) on the GPU.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("MEDIMAGE_AMD_LIB") or str(_PKG / "libmedimage_amd.so"))

MI_OK, MI_E_ARG, MI_E_DEVICE, MI_E_NOMEM, MI_E_STATE, MI_E_LIMIT = 0, -1, -2, -3, -4, -5
MI_BEAM_FAN_FLAT, MI_BEAM_PARALLEL, MI_BEAM_GOLDENC, MI_BEAM_FAN_ARC = 0, 1, 2, 3
MI_FP32, MI_FP64 = 0, 1
MI_ORDER_REFERENCE, MI_ORDER_PARALLEL = 0, 1
MI_K_FP, MI_K_BP, MI_K_FPSQ, MI_K_STEP, MI_K_EDGE, MI_K_CONE_FP, MI_K_CONE_BP = 0, 1, 2, 3, 4, 5, 6
MI_BATCH_GRAPH, MI_BATCH_MIN_K, MI_BATCH_RESERVE_MB, MI_BATCH_MAX_K, MI_BATCH_MAX_ENTRIES, MI_BATCH_BP_LDS, \
    MI_BATCH_BP_SQ = 0, 1, 2, 3, 4, 5, 6
MI_JOB_PENDING, MI_JOB_ACTIVE, MI_JOB_DONE, MI_JOB_FAILED = 0, 1, 2, 3


class Geometry(C.Structure):
    _fields_ = [
        ("beam", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("pixel_mm", C.c_double),
        ("n_views", C.c_int32), ("span_deg", C.c_double), ("n_bins", C.c_int32), ("bin_mm", C.c_double),
        ("sod_mm", C.c_double), ("sdd_mm", C.c_double), ("det_offset", C.c_double),
        ("flavour", C.c_int32), ("reserved", C.c_int32), ("view_deg", C.POINTER(C.c_double)),
    ]

    def key(self):
        views = getattr(self, "_views", None)
        return tuple(getattr(self, f) for f, _ in self._fields_[:-1]) + \
            ((views.tobytes(),) if views is not None else ())

    def with_views(self, angles_deg):
        """A copy with per-view source angles (degrees; MI_BEAM_FAN_ARC), as a
        rebinned helical slice's _pos table gives them."""
        a = np.array(angles_deg, np.float64)   # an owned copy
        if a.shape != (self.n_views,):
            raise ValueError(f"need {self.n_views} angles, got {a.shape}")
        g = Geometry(self.beam, self.nx, self.ny, self.pixel_mm, self.n_views, self.span_deg, self.n_bins,
                     self.bin_mm, self.sod_mm, self.sdd_mm, self.det_offset, self.flavour, 0,
                     a.ctypes.data_as(C.POINTER(C.c_double)))
        g._views = a
        return g

    def __repr__(self):
        extra = ", view_deg=table" if getattr(self, "_views", None) is not None else ""
        return "Geometry(" + ", ".join(f"{f}={getattr(self, f)}" for f, _ in self._fields_[:-2]) + extra + ")"


class ConeGeometry(C.Structure):
    """industry3d Parameter for BEAM "Cone" (industry3d/include/ct3d.h:17-63):
    volume [nz][nx][ny], n_views over 360 degrees, panel n_rows (NDX) x
    n_cols (NDY); sinogram [view][row][col]."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("n_views", C.c_int32),
        ("n_rows", C.c_int32), ("n_cols", C.c_int32), ("sample_mm", C.c_float), ("pixel_mm", C.c_float),
        ("sod_mm", C.c_float), ("sdd_mm", C.c_float), ("beam", C.c_int32), ("reserved", C.c_int32),
    ]

    def __repr__(self):
        return "ConeGeometry(" + ", ".join(f"{f}={getattr(self, f)}" for f, _ in self._fields_[:-1]) + ")"


MI_CONE_BEAM, MI_CONE_PARALLEL = 0, 1


def cone_geometry(nx, ny, nz, views, rows, cols, sample_mm, pixel_mm, sod, sdd, beam=MI_CONE_BEAM):
    return ConeGeometry(nx, ny, nz, views, rows, cols, sample_mm, pixel_mm, sod, sdd, beam, 0)


class AtParams(C.Structure):
    _fields_ = [
        ("alpha", C.c_double), ("beta", C.c_double), ("epsilon", C.c_double),
        ("iter_alt", C.c_int32), ("iter_image", C.c_int32), ("iter_edge", C.c_int32),
        ("n_subsets", C.c_int32), ("max_iters", C.c_int32), ("median_every", C.c_int32),
        ("rel_tol", C.c_double), ("lower", C.c_double), ("upper", C.c_double),
        ("precision", C.c_int32), ("order", C.c_int32),
    ]


class Report(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("converged", C.c_int32), ("data_fidelity", C.c_double),
        ("rel_change", C.c_double), ("elapsed_ms", C.c_double), ("nnz", C.c_int64),
    ]


class Tuning(C.Structure):
    _fields_ = [("fp_vec", C.c_int32), ("bp_lanes", C.c_int32), ("fp_grid", C.c_int32), ("bp_grid", C.c_int32),
                ("fp_us", C.c_double), ("bp_us", C.c_double)]


class EngineTuning(C.Structure):
    _fields_ = [("k_cap", C.c_int32), ("bp_lds", C.c_int32), ("us_per_job_pass", C.c_double)]


class KernelStats(C.Structure):
    _fields_ = [("launches", C.c_int64), ("total_ms", C.c_double), ("bytes", C.c_double)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("iterations", C.c_int32), ("converged", C.c_int32), ("unit_k", C.c_int32),
                ("data_fidelity", C.c_double), ("rel_change", C.c_double), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("nnz", C.c_int64)]


def fan_geometry(n, pixel_mm, views, bins, bin_mm, sod=541.0, sdd=949.0, span_deg=360.0, det_offset=0.0):
    """2D fan beam, flat equidistant detector (MI_BEAM_FAN_FLAT)."""
    return Geometry(MI_BEAM_FAN_FLAT, n, n, pixel_mm, views, span_deg, bins, bin_mm, sod, sdd, det_offset, 0, 0)


def arc_geometry(n, pixel_mm, views, bins, deg_per_bin, sod=595.0, sdd=1085.6, off_center=0.0, span_deg=360.0):
    """2D fan beam on an equiangular (arc) detector (MI_BEAM_FAN_ARC): the
    helical/rebin per-slice geometry (SOURCE_TO_ISO, SOURCE_TO_DET,
    LENGTH_PER_DET degrees per element, OFF_CENTER elements)."""
    return Geometry(MI_BEAM_FAN_ARC, n, n, pixel_mm, views, span_deg, bins, deg_per_bin, sod, sdd, off_center, 0, 0)


def parallel_geometry(n=512, views=180, bins=768, flavour=0):
    """msbeam/blockrecon parallel beam (MI_BEAM_PARALLEL); flavour 0 msbeam, 1 blockrecon."""
    return Geometry(MI_BEAM_PARALLEL, n, n, 1.0, views, 180.0, bins, 1.0, 0.0, 0.0, 0.0, flavour, 0)


def goldenc_geometry():
    """GoldenC tiled DDA: 512 x 512, 180 views, 767 bins (MI_BEAM_GOLDENC)."""
    return Geometry(MI_BEAM_GOLDENC, 512, 512, 1.0, 180, 180.0, 767, 1.0, 0.0, 0.0, 0.0, 0, 0)


def at_params(alpha=1e-3, beta=1e-4, epsilon=1e-2, n_subsets=1, max_iters=50, iter_edge=1, rel_tol=0.0,
              median_every=0, lower=-math.inf, upper=math.inf, iter_alt=10, iter_image=10, precision=MI_FP32,
              order=0):
    """order (GoldenC form): MI_ORDER_REFERENCE (0) sums in the reference's
    order, MI_ORDER_PARALLEL (1) with tree reductions."""
    return AtParams(alpha, beta, epsilon, iter_alt, iter_image, iter_edge, n_subsets, max_iters, median_every,
                    rel_tol, lower, upper, precision, order)


class MIError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"medimage_amd error {code}: {msg}")
        self.code = code


_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    P = C.c_void_p
    G = C.POINTER(Geometry)
    CG = C.POINTER(ConeGeometry)
    sigs = {
        "mi_context_create": [C.c_int, C.POINTER(C.c_void_p)],
        "mi_context_destroy": [P],
        "mi_context_mem": [P, C.POINTER(C.c_int64)],
        "mi_context_trim": [P],
        "mi_geometry_prepare": [P, G, C.c_int, C.POINTER(C.c_int64)],
        "mi_trace_ray": [P, G, C.c_int, C.c_int, P, P, C.c_int, C.POINTER(C.c_int)],
        "mi_autotune": [P, G, C.c_int, C.POINTER(Tuning)],
        "mi_set_tuning": [P, G, C.c_int, C.POINTER(Tuning)],
        "mi_get_tuning": [P, G, C.c_int, C.POINTER(Tuning)],
        "mi_matrix_export": [P, G, P, P, P, C.c_int64],
        "mi_engine_autotune": [P, G, C.c_int, C.c_int, C.POINTER(EngineTuning)],
        "mi_engine_set_tuning": [P, G, C.c_int, C.c_int, C.POINTER(EngineTuning)],
        "mi_engine_get_tuning": [P, G, C.c_int, C.c_int, C.POINTER(EngineTuning)],
        "mi_forward": [P, G, P, P],
        "mi_backproject": [P, G, P, P],
        "mi_forward_f64": [P, G, P, P],
        "mi_backproject_f64": [P, G, P, P],
        "mi_forward_ordered": [P, G, C.c_int, P, P],
        "mi_backproject_ordered": [P, G, C.c_int, P, P],
        "mi_image_step_f64": [P, G, P, P, P, P, C.c_double, C.c_double, C.c_int],
        "mi_edge_step_f64": [P, C.c_int, C.c_int, P, P, P, C.c_double, C.c_double, C.c_double, C.c_int],
        "mi_medfilt2_f64": [P, C.c_int, C.c_int, P],
        "mi_mumford_shah": [P, G, C.POINTER(AtParams), P, P, P, C.POINTER(Report)],
        "mi_at_reconstruct": [P, G, C.POINTER(AtParams), P, P, P, C.POINTER(Report)],
        "mi_at_reconstruct_f64": [P, G, C.POINTER(AtParams), P, P, P, C.POINTER(Report)],
        "mi_job_create": [P, G, C.POINTER(AtParams), P, P, P, C.POINTER(C.c_void_p)],
        "mi_job_run": [P, C.c_int, C.POINTER(C.c_int)],
        "mi_job_enqueue": [P, C.c_int],
        "mi_job_synchronize": [P],
        "mi_job_report": [P, C.POINTER(Report)],
        "mi_job_read": [P, P, P],
        "mi_job_destroy": [P],
        "mi_batch_create": [P, C.c_int, C.c_int, C.POINTER(C.c_void_p)],
        "mi_batch_destroy": [P],
        "mi_batch_set": [P, C.c_int, C.c_int64],
        "mi_batch_submit": [P, G, C.POINTER(AtParams), P, P, P, C.POINTER(C.c_int64)],
        "mi_batch_step": [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "mi_batch_enqueue": [P, C.c_int],
        "mi_batch_poll": [P, C.POINTER(C.c_int)],
        "mi_batch_status": [P, C.c_int64, C.POINTER(JobStatus)],
        "mi_batch_read": [P, C.c_int64, P, P],
        "mi_batch_release": [P, C.c_int64],
        "mi_batch_units": [P, P, C.c_int, C.POINTER(C.c_int)],
        "mi_batch_retired": [P, P, C.c_int, C.POINTER(C.c_int)],
        "mi_stats_enable": [P, C.c_int],
        "mi_stats_select": [P, C.c_uint32],
        "mi_stats_get": [P, C.c_int, C.POINTER(KernelStats)],
        "mi_stats_reset": [P],
        "mi_device_copy": [P, C.c_int64, C.c_int, C.POINTER(C.c_double)],
        "mi_config_parse": [C.c_char_p, G, C.POINTER(AtParams)],
        "mi_read_dr": [C.c_char_p, C.c_int64, C.c_int64, C.c_int64, P, P],
        "mi_read_pos": [C.c_char_p, P, P, C.c_int, C.POINTER(C.c_int)],
        "mi_read_raw": [C.c_char_p, C.c_int64, C.c_int64, P],
        "mi_helical_reconstruct": [P, C.c_char_p, C.c_int, C.c_int, G, C.POINTER(AtParams), P, P, P],
        "mi_cone_trace_ray": [P, CG, C.c_int, C.c_int, C.c_int, P, P, C.c_int, C.POINTER(C.c_int)],
        "mi_cone_forward": [P, CG, P, P],
        "mi_cone_forward_dev": [P, CG, P, P],
        "mi_cone_backproject": [P, CG, P, P],
        "mi_cone_backproject_dev": [P, CG, P, P],
        "mi_cone_backproject_dev_noalloc": [P, CG, P, P],
        "mi_cone_at_reconstruct": [P, CG, C.POINTER(AtParams), P, P, P, C.POINTER(Report)],
        "mi_cone_at_reconstruct_dev": [P, CG, C.POINTER(AtParams), P, P, P, C.POINTER(Report)],
        "mi_cone_config_parse": [C.c_char_p, CG, C.POINTER(AtParams)],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.mi_last_error.restype = C.c_char_p
    lib.mi_last_error.argtypes = []
    _lib = lib
    return lib


def lib():
    return _load()


def config_parse(text: str):
    """(Geometry, AtParams) from the text of a reference JSON configuration
    (mi_config_parse: the helical slice driver's or industry3d's)."""
    g, p = Geometry(), AtParams()
    _check(_load().mi_config_parse(text.encode(), C.byref(g), C.byref(p)))
    return g, p


def cone_config_parse(text: str):
    """(ConeGeometry, AtParams) from an industry3d configuration with BEAM "Cone"."""
    g, p = ConeGeometry(), AtParams()
    _check(_load().mi_cone_config_parse(text.encode(), C.byref(g), C.byref(p)))
    return g, p


def read_dr(path, nproj: int, ndx: int, ndy: int, flat: bool = True):
    """Projections [nproj, ndx, ndy] uint16 (and the [ndx, ndy] float32 blank
    scan when flat) of an industry3d .dr file (mi_read_dr)."""
    proj = np.empty((nproj, ndx, ndy), np.uint16)
    fl = np.empty((ndx, ndy), np.float32) if flat else None
    _check(_load().mi_read_dr(str(path).encode(), nproj, ndx, ndy, _ptr(proj), _ptr(fl) if flat else None))
    return (proj, fl) if flat else proj


def read_pos(path):
    """(pos, angle_deg) columns of a rebinned helical slice's _pos table (mi_read_pos)."""
    n = C.c_int()
    _check(_load().mi_read_pos(str(path).encode(), None, None, 0, C.byref(n)))
    pos = np.empty(n.value, np.float64)
    ang = np.empty(n.value, np.float64)
    _check(_load().mi_read_pos(str(path).encode(), _ptr(pos), _ptr(ang), n.value, C.byref(n)))
    return pos, ang


def read_raw(path, n_views: int, n_bins: int):
    """[n_views, n_bins] sinogram of a slice's _raw file (mi_read_raw)."""
    out = np.empty((n_views, n_bins), np.float64)
    _check(_load().mi_read_raw(str(path).encode(), n_views, n_bins, _ptr(out)))
    return out


def _check(rc):
    if rc != MI_OK:
        raise MIError(rc, _load().mi_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _arr(a, dtype, shape=None):
    a = np.ascontiguousarray(a, dtype=dtype)
    if shape is not None and a.shape != tuple(shape):
        a = a.reshape(shape)
    return a


class Context:
    """One device, one HIP stream, a cache of system matrices (mi_context)."""

    def __init__(self, device: int = 0):
        self._lib = _load()
        h = C.c_void_p()
        _check(self._lib.mi_context_create(device, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mi_context_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- geometry ---------------------------------------------------------
    def prepare(self, g: Geometry, n_subsets: int = 1) -> int:
        nnz = C.c_int64()
        _check(self._lib.mi_geometry_prepare(self._h, C.byref(g), n_subsets, C.byref(nnz)))
        return nnz.value

    def autotune(self, g: Geometry, n_subsets: int = 1) -> Tuning:
        t = Tuning()
        _check(self._lib.mi_autotune(self._h, C.byref(g), n_subsets, C.byref(t)))
        return t

    def set_tuning(self, g: Geometry, n_subsets: int, fp_vec: int, bp_lanes: int, fp_grid: int = 0, bp_grid: int = 0):
        t = Tuning(fp_vec, bp_lanes, fp_grid, bp_grid, 0.0, 0.0)
        _check(self._lib.mi_set_tuning(self._h, C.byref(g), n_subsets, C.byref(t)))

    def get_tuning(self, g: Geometry, n_subsets: int = 1) -> Tuning:
        t = Tuning()
        _check(self._lib.mi_get_tuning(self._h, C.byref(g), n_subsets, C.byref(t)))
        return t

    def trace_ray(self, g: Geometry, view: int, bin: int, cap: int = 8192):
        ind = np.zeros(cap, np.int32)
        ln = np.zeros(cap, np.float64)
        n = C.c_int()
        _check(self._lib.mi_trace_ray(self._h, C.byref(g), view, bin, _ptr(ind), _ptr(ln), cap, C.byref(n)))
        k = min(n.value, cap)
        return ind[:k].copy(), ln[:k].copy()

    def engine_autotune(self, g: Geometry, n_subsets: int, precision: int = MI_FP32) -> EngineTuning:
        t = EngineTuning()
        _check(self._lib.mi_engine_autotune(self._h, C.byref(g), n_subsets, precision, C.byref(t)))
        return t

    def engine_tuning(self, g: Geometry, n_subsets: int, precision: int = MI_FP32, set_to: EngineTuning = None):
        """The geometry's engine settings; with set_to, installs them first."""
        if set_to is not None:
            _check(self._lib.mi_engine_set_tuning(self._h, C.byref(g), n_subsets, precision, C.byref(set_to)))
        t = EngineTuning()
        _check(self._lib.mi_engine_get_tuning(self._h, C.byref(g), n_subsets, precision, C.byref(t)))
        return t

    def matrix(self, g: Geometry):
        """(rowptr, ind, len) of the one-subset matrix, rays in [view][bin] order."""
        nnz = self.prepare(g, 1)
        rp = np.zeros(g.n_views * g.n_bins + 1, np.int64)
        ind = np.zeros(max(nnz, 1), np.int32)
        ln = np.zeros(max(nnz, 1), np.float64)
        _check(self._lib.mi_matrix_export(self._h, C.byref(g), _ptr(rp), _ptr(ind), _ptr(ln), nnz))
        return rp, ind[:nnz], ln[:nnz]

    def mem_bytes(self) -> int:
        b = C.c_int64()
        _check(self._lib.mi_context_mem(self._h, C.byref(b)))
        return b.value

    def trim(self):
        _check(self._lib.mi_context_trim(self._h))

    # -- projectors -------------------------------------------------------
    def forward(self, g: Geometry, img, dtype=np.float32):
        img = _arr(img, dtype, (g.nx, g.ny))
        out = np.empty((g.n_views, g.n_bins), dtype)
        fn = self._lib.mi_forward if dtype == np.float32 else self._lib.mi_forward_f64
        _check(fn(self._h, C.byref(g), _ptr(img), _ptr(out)))
        return out

    def backproject(self, g: Geometry, sino, dtype=np.float32):
        sino = _arr(sino, dtype, (g.n_views, g.n_bins))
        out = np.empty((g.nx, g.ny), dtype)
        fn = self._lib.mi_backproject if dtype == np.float32 else self._lib.mi_backproject_f64
        _check(fn(self._h, C.byref(g), _ptr(sino), _ptr(out)))
        return out

    def forward_ordered(self, g: Geometry, img, dtype=np.float64):
        """A in the reference's summation order and type (bitwise GoldenC's A
        in FP64, msbeam's A in FP32 on its parallel geometry)."""
        img = _arr(img, dtype, (g.nx, g.ny))
        out = np.empty((g.n_views, g.n_bins), dtype)
        prec = MI_FP64 if dtype == np.float64 else MI_FP32
        _check(self._lib.mi_forward_ordered(self._h, C.byref(g), prec, _ptr(img), _ptr(out)))
        return out

    def backproject_ordered(self, g: Geometry, sino, dtype=np.float64):
        sino = _arr(sino, dtype, (g.n_views, g.n_bins))
        out = np.empty((g.nx, g.ny), dtype)
        prec = MI_FP64 if dtype == np.float64 else MI_FP32
        _check(self._lib.mi_backproject_ordered(self._h, C.byref(g), prec, _ptr(sino), _ptr(out)))
        return out

    # -- 3D cone beam (industry3d) -----------------------------------------
    def cone_trace_ray(self, g: ConeGeometry, view: int, row: int, col: int, cap: int = 0):
        """(voxel index, length) sequence of ray (view, row, col) as traced on
        the device (forward_proj, industry3d/src/tracing.cpp:123-307)."""
        cap = cap or 4 * (g.nx + g.ny + g.nz) + 8
        ind = np.zeros(cap, np.int64)
        ln = np.zeros(cap, np.float32)
        n = C.c_int()
        _check(self._lib.mi_cone_trace_ray(self._h, C.byref(g), view, row, col, _ptr(ind), _ptr(ln), cap,
                                           C.byref(n)))
        k = min(n.value, cap)
        return ind[:k].copy(), ln[:k].copy()

    def cone_forward(self, g: ConeGeometry, vol):
        vol = _arr(vol, np.float32, (g.nz, g.nx, g.ny))
        out = np.empty((g.n_views, g.n_rows, g.n_cols), np.float32)
        _check(self._lib.mi_cone_forward(self._h, C.byref(g), _ptr(vol), _ptr(out)))
        return out

    def cone_backproject(self, g: ConeGeometry, sino):
        sino = _arr(sino, np.float32, (g.n_views, g.n_rows, g.n_cols))
        out = np.empty((g.nz, g.nx, g.ny), np.float32)
        _check(self._lib.mi_cone_backproject(self._h, C.byref(g), _ptr(sino), _ptr(out)))
        return out

    def cone_forward_dev(self, g: ConeGeometry, vol_ptr: int, sino_ptr: int):
        """Device pointers (e.g. torch tensors' data_ptr()); no copies."""
        _check(self._lib.mi_cone_forward_dev(self._h, C.byref(g), C.c_void_p(vol_ptr), C.c_void_p(sino_ptr)))

    def cone_backproject_dev(self, g: ConeGeometry, sino_ptr: int, vol_ptr: int):
        _check(self._lib.mi_cone_backproject_dev(self._h, C.byref(g), C.c_void_p(sino_ptr), C.c_void_p(vol_ptr)))

    def cone_at_reconstruct(self, g: ConeGeometry, sino, p: AtParams, u0=None, v0=None):
        """Ordered-subset AT on the cone-beam volume; returns (u, v, report)."""
        sino = _arr(sino, np.float32, (g.n_views, g.n_rows, g.n_cols))
        shape = (g.nz, g.nx, g.ny)
        u = np.zeros(shape, np.float32) if u0 is None else np.array(u0, np.float32).reshape(shape).copy()
        v = np.ones(shape, np.float32) if v0 is None else np.array(v0, np.float32).reshape(shape).copy()
        rep = Report()
        _check(self._lib.mi_cone_at_reconstruct(self._h, C.byref(g), C.byref(p), _ptr(sino), _ptr(u), _ptr(v),
                                                C.byref(rep)))
        return u, v, rep

    def cone_at_reconstruct_dev(self, g: ConeGeometry, sino_ptr: int, u_ptr: int, v_ptr: int, p: AtParams):
        """Device pointers: u and v hold the initial values and are updated in place."""
        rep = Report()
        _check(self._lib.mi_cone_at_reconstruct_dev(self._h, C.byref(g), C.byref(p), C.c_void_p(sino_ptr),
                                                    C.c_void_p(u_ptr), C.c_void_p(v_ptr), C.byref(rep)))
        return rep

    # -- GoldenC building blocks -----------------------------------------
    def image_step(self, g: Geometry, sino, f, v, alpha, epsilon, iter_image):
        sino = _arr(sino, np.float64, (g.n_views, g.n_bins))
        f = _arr(f, np.float64, (g.nx, g.ny))
        v = _arr(v, np.float64, (g.nx, g.ny))
        out = np.empty_like(f)
        _check(self._lib.mi_image_step_f64(self._h, C.byref(g), _ptr(out), _ptr(sino), _ptr(f), _ptr(v),
                                           alpha, epsilon, iter_image))
        return out

    def edge_step(self, f, v, alpha, beta, epsilon, iter_edge):
        f = _arr(f, np.float64)
        v = _arr(v, np.float64, f.shape)
        out = np.empty_like(v)
        _check(self._lib.mi_edge_step_f64(self._h, f.shape[0], f.shape[1], _ptr(out), _ptr(f), _ptr(v),
                                          alpha, beta, epsilon, iter_edge))
        return out

    def medfilt2(self, f):
        f = _arr(f, np.float64).copy()
        _check(self._lib.mi_medfilt2_f64(self._h, f.shape[0], f.shape[1], _ptr(f)))
        return f

    # -- drivers ----------------------------------------------------------
    def mumford_shah(self, g: Geometry, sino, p: AtParams):
        sino = _arr(sino, np.float64, (g.n_views, g.n_bins))
        f = np.empty((g.nx, g.ny), np.float64)
        v = np.empty((g.nx, g.ny), np.float64)
        rep = Report()
        _check(self._lib.mi_mumford_shah(self._h, C.byref(g), C.byref(p), _ptr(sino), _ptr(f), _ptr(v),
                                         C.byref(rep)))
        return f, v, rep

    def helical_reconstruct(self, prefix, first: int, count: int, g: Geometry, p: AtParams):
        """Per-slice loop over <prefix>i_pos / <prefix>i_raw (mi_helical_reconstruct);
        returns (u [count, nx, ny], v, reports)."""
        dt = np.float64 if p.precision == MI_FP64 else np.float32
        u = np.empty((count, g.nx, g.ny), dt)
        v = np.empty((count, g.nx, g.ny), dt)
        reps = (Report * count)()
        _check(self._lib.mi_helical_reconstruct(self._h, str(prefix).encode(), first, count, C.byref(g), C.byref(p),
                                                _ptr(u), _ptr(v), reps))
        return u, v, list(reps)

    def at_reconstruct(self, g: Geometry, sino, p: AtParams, u0=None, v0=None):
        dt = np.float64 if p.precision == MI_FP64 else np.float32
        sino = _arr(sino, dt, (g.n_views, g.n_bins))
        u = np.zeros((g.nx, g.ny), dt) if u0 is None else _arr(u0, dt, (g.nx, g.ny)).copy()
        v = np.ones((g.nx, g.ny), dt) if v0 is None else _arr(v0, dt, (g.nx, g.ny)).copy()
        rep = Report()
        fn = self._lib.mi_at_reconstruct_f64 if dt == np.float64 else self._lib.mi_at_reconstruct
        _check(fn(self._h, C.byref(g), C.byref(p), _ptr(sino), _ptr(u), _ptr(v), C.byref(rep)))
        return u, v, rep

    def job(self, g: Geometry, sino, p: AtParams, u0=None, v0=None):
        return Job(self, g, sino, p, u0, v0)

    def batch(self, n_subsets: int, precision: int = MI_FP32, graph: bool = True, min_k: int = 1):
        return Batch(self, n_subsets, precision, graph, min_k)

    # -- instrumentation --------------------------------------------------
    def stats_enable(self, on=True):
        _check(self._lib.mi_stats_enable(self._h, int(bool(on))))

    def stats_select(self, kernels):
        """Time only the given kernel ids (an empty list turns timing off)."""
        mask = 0
        for k in kernels:
            mask |= 1 << k
        _check(self._lib.mi_stats_select(self._h, mask))

    def stats(self, kernel: int) -> KernelStats:
        s = KernelStats()
        _check(self._lib.mi_stats_get(self._h, kernel, C.byref(s)))
        return s

    def stats_reset(self):
        _check(self._lib.mi_stats_reset(self._h))

    def device_copy(self, nbytes: int, reps: int = 1) -> float:
        """Plain device copy of nbytes, reps times; returns ms per copy."""
        ms = C.c_double()
        _check(self._lib.mi_device_copy(self._h, nbytes, reps, C.byref(ms)))
        return ms.value


class Job:
    """Device-resident ordered-subset job (mi_job)."""

    def __init__(self, ctx: Context, g: Geometry, sino, p: AtParams, u0=None, v0=None):
        self._lib = ctx._lib
        self.ctx = ctx
        self.g = g
        self.dtype = np.float64 if p.precision == MI_FP64 else np.float32
        sino = _arr(sino, self.dtype, (g.n_views, g.n_bins))
        u0 = None if u0 is None else _arr(u0, self.dtype, (g.nx, g.ny))
        v0 = None if v0 is None else _arr(v0, self.dtype, (g.nx, g.ny))
        h = C.c_void_p()
        _check(self._lib.mi_job_create(ctx._h, C.byref(g), C.byref(p), _ptr(sino),
                                       None if u0 is None else _ptr(u0), None if v0 is None else _ptr(v0),
                                       C.byref(h)))
        self._h = h

    def run(self, max_passes):
        n = C.c_int()
        _check(self._lib.mi_job_run(self._h, max_passes, C.byref(n)))
        return n.value

    def enqueue(self, passes):
        _check(self._lib.mi_job_enqueue(self._h, passes))

    def synchronize(self):
        _check(self._lib.mi_job_synchronize(self._h))

    def report(self) -> Report:
        r = Report()
        _check(self._lib.mi_job_report(self._h, C.byref(r)))
        return r

    def read(self):
        u = np.empty((self.g.nx, self.g.ny), self.dtype)
        v = np.empty((self.g.nx, self.g.ny), self.dtype)
        _check(self._lib.mi_job_read(self._h, _ptr(u), _ptr(v)))
        return u, v

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mi_job_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many ordered-subset jobs advanced together (mi_batch): units of K jobs
    sharing a system matrix, one launch sequence per pass for all of them,
    jobs admitted and retired between passes."""

    def __init__(self, ctx: Context, n_subsets: int, precision: int = MI_FP32, graph: bool = True, min_k: int = 1):
        self._lib = ctx._lib
        self.ctx = ctx
        self.S = n_subsets
        self.dtype = np.float64 if precision == MI_FP64 else np.float32
        self.precision = precision
        h = C.c_void_p()
        _check(self._lib.mi_batch_create(ctx._h, n_subsets, precision, C.byref(h)))
        self._h = h
        self._geom = {}
        self.set(MI_BATCH_GRAPH, int(graph))
        if min_k != 1:
            self.set(MI_BATCH_MIN_K, min_k)

    def set(self, option: int, value: int):
        _check(self._lib.mi_batch_set(self._h, option, value))

    def submit(self, g: Geometry, sino, p: AtParams, u0=None, v0=None) -> int:
        sino = _arr(sino, self.dtype, (g.n_views, g.n_bins))
        u0 = None if u0 is None else _arr(u0, self.dtype, (g.nx, g.ny))
        v0 = None if v0 is None else _arr(v0, self.dtype, (g.nx, g.ny))
        jid = C.c_int64()
        _check(self._lib.mi_batch_submit(self._h, C.byref(g), C.byref(p), _ptr(sino),
                                         None if u0 is None else _ptr(u0), None if v0 is None else _ptr(v0),
                                         C.byref(jid)))
        self._geom[jid.value] = (g.nx, g.ny)
        return jid.value

    def step(self, passes: int = 1):
        a, q = C.c_int(), C.c_int()
        _check(self._lib.mi_batch_step(self._h, passes, C.byref(a), C.byref(q)))
        return a.value, q.value

    def enqueue(self, passes: int):
        _check(self._lib.mi_batch_enqueue(self._h, passes))

    def poll(self) -> int:
        a = C.c_int()
        _check(self._lib.mi_batch_poll(self._h, C.byref(a)))
        return a.value

    def status(self, jid: int) -> JobStatus:
        st = JobStatus()
        _check(self._lib.mi_batch_status(self._h, jid, C.byref(st)))
        return st

    def read(self, jid: int):
        nx, ny = self._geom[jid]
        u = np.empty((nx, ny), self.dtype)
        v = np.empty((nx, ny), self.dtype)
        _check(self._lib.mi_batch_read(self._h, jid, _ptr(u), _ptr(v)))
        return u, v

    def release(self, jid: int):
        _check(self._lib.mi_batch_release(self._h, jid))
        self._geom.pop(jid, None)

    def retired(self):
        """Ids of the jobs retired since the last call."""
        out = []
        buf = np.zeros(256, np.int64)
        n = C.c_int()
        while True:
            _check(self._lib.mi_batch_retired(self._h, _ptr(buf), len(buf), C.byref(n)))
            out += [int(x) for x in buf[:n.value]]
            if n.value < len(buf):
                return out

    def units(self):
        k = np.zeros(64, np.int32)
        n = C.c_int()
        _check(self._lib.mi_batch_units(self._h, _ptr(k), 64, C.byref(n)))
        return [int(x) for x in k[:n.value]]

    def run(self, max_passes: int = 1 << 30, chunk: int = 1):
        """Steps until every submitted job finished (each job stops at its own
        max_iters / rel_tol); returns the number of passes run."""
        n = 0
        while n < max_passes:
            a, q = self.step(chunk)
            n += chunk
            if a == 0 and q == 0:
                break
        return n

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
