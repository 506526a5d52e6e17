/*
 * medimage_amd: MI355X-native Ambrosio-Tortorelli (Mumford-Shah) low-dose
 * CT reconstruction.  C ABI of libmedimage_amd.so.
 *
 * The call surface keeps the one of the reference drivers:
 *   - a geometry struct whose fields map onto industry3d's Parameter
 *     (industry3d/include/ct3d.h:17-63) and GoldenC's config.h
 *     (origin/GoldenC/src/config.h:4-17);
 *   - sinogram / image / edge buffers laid out as the reference lays them
 *     out (row-major [view][bin] sinogram, row-major [x][y] image, as
 *     origin/GoldenC/src/main.c:8-10 and msbeam/CPU_OpenMP/src/main.c:12-15);
 *   - alpha / beta / epsilon and the iteration counts of GoldenC
 *     (origin/GoldenC/src/main.c:25-35) plus the ordered-subset controls.
 *
 * Conventions.  Every entry point returns 0 on success or a negative
 * MI_E_* code; mi_last_error() returns a thread-local message for the last
 * failure.  Host pointers are caller-owned and copied in and out; *_dev
 * variants take device pointers and do not copy.  Calls on one context are
 * serialised on that context's HIP stream and are blocking unless named
 * *_async.  A context is not shared between threads without external
 * locking.  Nothing here ever falls back to a CPU path: without a usable
 * gfx950 device, mi_context_create fails with MI_E_DEVICE.
 */
#ifndef MEDIMAGE_AMD_H
#define MEDIMAGE_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_ABI_VERSION 1

enum mi_status {
    MI_OK = 0,
    MI_E_ARG = -1,        /* invalid argument or geometry                   */
    MI_E_DEVICE = -2,     /* no usable device / HIP runtime error           */
    MI_E_NOMEM = -3,      /* device memory exhausted                        */
    MI_E_STATE = -4,      /* call not valid in the object's current state   */
    MI_E_LIMIT = -5       /* size beyond a documented maximum               */
};

/* Beam / projector kinds. */
enum mi_beam {
    /* 2D fan beam, flat equidistant detector, parametric Siddon in double:
     * ray construction of industry3d/src/tracing.cpp:51-81 (cone(), mid-plane)
     * and tracer of helical/rebin/src/tracing.cpp:22-147. */
    MI_BEAM_FAN_FLAT = 0,
    /* 2D parallel beam over pi, unit detector cells, incremental tracer of
     * msbeam/CPU_OpenMP/src/proj.c:54-211 (= blockrecon/src/wray.cpp:10-167). */
    MI_BEAM_PARALLEL = 1,
    /* GoldenC symmetric tiled DDA, origin/GoldenC/src/proj.c:14-397
     * (512 x 512 image, 180 views, 767 bins only). */
    MI_BEAM_GOLDENC = 2,
    /* 2D fan beam, equiangular (arc) detector of the clinical scanner: ray
     * construction of helical/rebin/src/ms2d.cpp:50-56,172-183 (source at
     * angle a_v = span_deg v / n_views degrees, detector element j at
     * (j - n_bins/2 + det_offset) * bin_mm degrees + pi seen from the
     * source), same Siddon tracer as MI_BEAM_FAN_FLAT.  Here bin_mm holds
     * LENGTH_PER_DET (degrees per detector element), det_offset OFF_CENTER,
     * sod_mm SOURCE_TO_ISO, sdd_mm SOURCE_TO_DET; square images, even n_bins. */
    MI_BEAM_FAN_ARC = 3
};

enum mi_precision { MI_FP32 = 0, MI_FP64 = 1 };

/* Summation order of the GoldenC-form drivers (mi_mumford_shah).
 * MI_ORDER_REFERENCE (default): every sum is added in the order the
 * reference program adds it (A and AStar per ray / per pixel in the order
 * origin/GoldenC/src/proj.c:14-397 emits the products, energies and inner
 * products in the row-major loops of origin/GoldenC/src/main.c:81-360), so
 * results equal the reference's bit for bit.  MI_ORDER_PARALLEL: the same
 * algorithm with tree reductions (faster; rounding differs at 1e-16 and the
 * CG amplifies it, about 20x per alternation). */
enum mi_order { MI_ORDER_REFERENCE = 0, MI_ORDER_PARALLEL = 1 };

/* Geometry.  Field <- reference name (industry3d Parameter / GoldenC config.h). */
typedef struct mi_geometry {
    int32_t beam;          /* BEAM ("Parallel"/"Cone") -> enum mi_beam         */
    int32_t nx, ny;        /* NX, NY: image rows (slow) and columns (fast)     */
    double  pixel_mm;      /* SAMPLESIZE (vx = vy)                             */
    int32_t n_views;       /* NPROJ / nAlpha                                   */
    double  span_deg;      /* angular span; 360 for fan, 180 for parallel      */
    int32_t n_bins;        /* NDX / projWidth / NRAY                           */
    double  bin_mm;        /* PIXELSIZE (detector pitch at the detector)       */
    double  sod_mm;        /* SOD                                              */
    double  sdd_mm;        /* SDD                                              */
    double  det_offset;    /* OFF_CENTER, in bins                              */
    int32_t flavour;       /* MI_BEAM_PARALLEL only: 0 msbeam constants
                              (double pi/tolerances), 1 blockrecon constants
                              (float, blockrecon/src/config.h:35-38)           */
    int32_t reserved;
    /* MI_BEAM_FAN_ARC only: the source angle of each view in degrees
     * (n_views values: the angle column of a rebinned helical slice's _pos
     * table, helical/rebin/src/ms2d.cpp:40-47), or NULL for span_deg v /
     * n_views.  Read during the call; the library keeps its own copy. */
    const double *view_deg;
} mi_geometry;

/* AT parameters.  alpha/beta/epsilon as in
 * origin/MATLAB/MumfordShah_reg.m:4-7 and origin/GoldenC/src/config.h:6-8. */
typedef struct mi_at_params {
    double  alpha, beta, epsilon;
    int32_t iter_alt;       /* _IterAlt   (GoldenC form)                       */
    int32_t iter_image;     /* _IterImage (GoldenC form)                       */
    int32_t iter_edge;      /* _IterEdge  (both forms; edge SD steps per pass) */
    int32_t n_subsets;      /* ordered subsets (OS form)                       */
    int32_t max_iters;      /* AT iterations cap (OS form)                     */
    int32_t median_every;   /* OS form: 3x3 median of u every k passes, 0 off  */
    double  rel_tol;        /* OS form: stop when |df_k - df_{k-1}| < rel_tol*df_{k-1} */
    double  lower, upper;   /* LOWER_X / UPPER_X clamp of u (OS form)          */
    int32_t precision;      /* enum mi_precision                               */
    int32_t order;          /* enum mi_order (GoldenC form), 0 = reference     */
} mi_at_params;

/* Result report of a reconstruction. */
typedef struct mi_report {
    int32_t iterations;     /* AT passes (OS form) / alternations (GoldenC)    */
    int32_t converged;      /* 1 if the rel_tol test stopped the run           */
    double  data_fidelity;  /* last pass' sum_s ||g_s - A_s u||^2               */
    double  rel_change;     /* last relative data-fidelity change              */
    double  elapsed_ms;     /* device time of the solve (HIP events)           */
    int64_t nnz;            /* system-matrix entries of the geometry           */
} mi_report;

typedef struct mi_context mi_context;

/* ---- context ----------------------------------------------------------- */
int mi_context_create(int device, mi_context **out);
int mi_context_destroy(mi_context *ctx);
const char *mi_last_error(void);
/* Bytes of device memory the context holds (cached matrices and the cone-beam
 * geometry tables). */
int mi_context_mem(mi_context *ctx, int64_t *bytes);
/* Drop cached system matrices (all geometries) and the cone-beam tables. */
int mi_context_trim(mi_context *ctx);

/* ---- geometry / system matrix ------------------------------------------ */
/* Traces the geometry (if not cached) for the AT engine with n_subsets view
 * subsets and returns its entry count.  Jobs and batches of the same
 * geometry and subset count reuse it (it is kept in the engine's sliced
 * layout only; the projector calls below build their own copy on first use). */
int mi_geometry_prepare(mi_context *ctx, const mi_geometry *g, int n_subsets, int64_t *nnz);
/* Copies the (index, length) sequence of ray (view, bin) as traced on the
 * device: up to cap entries; *numb receives the full count.  Lengths are the
 * tracer's own type widened to double (float for MI_BEAM_PARALLEL). */
int mi_trace_ray(mi_context *ctx, const mi_geometry *g, int view, int bin,
                 int32_t *ind, double *len, int cap, int *numb);

/* Copies the whole one-subset system matrix of the geometry in the caller's
 * ray order (ray = view * n_bins + bin): rowptr[n_views*n_bins + 1] (entry
 * offsets), ind[nnz] pixel indices, len[nnz] lengths widened to double.
 * nnz must equal the value mi_geometry_prepare reports. */
int mi_matrix_export(mi_context *ctx, const mi_geometry *g, int64_t *rowptr, int32_t *ind, double *len,
                     int64_t nnz);

/* Launch shapes of the streaming kernels for one (geometry, subsets)
 * matrix.  mi_autotune times every candidate on the device and keeps the
 * fastest; mi_set_tuning installs a stored choice.  The AT engine (jobs and
 * batches) reads bp_lanes only; fp_vec / fp_grid / bp_grid shape the single
 * projector calls (mi_forward, mi_backproject and the GoldenC steps). */
typedef struct mi_tuning {
    int32_t fp_vec;         /* forward projector: 1 = four consecutive entries
                               per lane (16-byte loads), 0 = lane-strided      */
    int32_t bp_lanes;       /* back projector lanes per pixel: 4..64 (the AT
                               engine uses at least 8)                         */
    int32_t fp_grid;        /* forward projector grid cap (blocks), 0 default  */
    int32_t bp_grid;        /* back projector grid cap (blocks), 0 default     */
    double  fp_us, bp_us;   /* measured per-subset launch times (autotune)     */
} mi_tuning;
int mi_autotune(mi_context *ctx, const mi_geometry *g, int n_subsets, mi_tuning *out);
int mi_set_tuning(mi_context *ctx, const mi_geometry *g, int n_subsets, const mi_tuning *t);
int mi_get_tuning(mi_context *ctx, const mi_geometry *g, int n_subsets, mi_tuning *out);

/* Launch settings of the AT engine (jobs, batches, serving) for one
 * (geometry, subsets): how many jobs share a unit (one stream of the
 * geometry's matrix serves K interleaved jobs) and whether the back
 * projection stages its residual values in LDS.  mi_engine_autotune times
 * every candidate on the device (synthetic jobs, hipGraph replay) and keeps
 * the fastest for the geometry; batches then pack its jobs accordingly.
 * Neither setting changes a job's arithmetic: results are bitwise the same
 * as with the defaults.  The shapes that fix a job's summation order (rays
 * per forward block, lanes per back-projection pixel) are not tuned; they
 * are part of the result's definition.  The reference has no such tuner. */
typedef struct mi_engine_tuning {
    int32_t k_cap;            /* jobs per unit: 1, 2, 4 or 8; 0 = the default
                                 rule (K images within one XCD's L2)          */
    int32_t bp_lds;           /* -1 default rule (LDS up to 24 KiB of slots),
                                 0 global gathers, 1 LDS up to 56 KiB         */
    double  us_per_job_pass;  /* measured: one AT iteration of one job        */
} mi_engine_tuning;
int mi_engine_autotune(mi_context *ctx, const mi_geometry *g, int n_subsets, int precision, mi_engine_tuning *out);
int mi_engine_set_tuning(mi_context *ctx, const mi_geometry *g, int n_subsets, int precision,
                         const mi_engine_tuning *t);
int mi_engine_get_tuning(mi_context *ctx, const mi_geometry *g, int n_subsets, int precision, mi_engine_tuning *out);

/* ---- projectors (replace A / AStar) --------------------------------------
 * forward:     sino[view][bin] = sum over the ray of img[x][y] * len
 *              (origin/GoldenC/src/proj.c:14-204, msbeam proj.c:14-32)
 * backproject: img[x][y] = sum over rays through the pixel of sino * len,
 *              gather form, deterministic (proj.c:206-397, msbeam proj.c:34-52)
 * For MI_BEAM_GOLDENC the _f64 forms add each value's products in the order
 * proj.c does (one lane per ray / pixel), so they equal the reference's A
 * and AStar bit for bit.
 */
int mi_forward(mi_context *ctx, const mi_geometry *g, const float *img, float *sino);
int mi_backproject(mi_context *ctx, const mi_geometry *g, const float *sino, float *img);
int mi_forward_f64(mi_context *ctx, const mi_geometry *g, const double *img, double *sino);
int mi_backproject_f64(mi_context *ctx, const mi_geometry *g, const double *sino, double *img);

/* Projections in the reference's summation order and type: one lane per
 * ray (pixel) adds the products in the order of the reference's own loop.
 * MI_BEAM_GOLDENC with MI_FP64: A / AStar of origin/GoldenC/src/proj.c:14-397
 * bit for bit.  MI_BEAM_PARALLEL (flavour 0) with MI_FP32: msbeam's A /
 * AStar (msbeam/CPU_OpenMP/src/proj.c:14-52: rays in [view][bin] order,
 * float products and sums) bit for bit.  Other beams: the tracer's order
 * along each ray and ray order at each pixel.  Element type of img / sino:
 * float for MI_FP32, double for MI_FP64.  Slower than mi_forward (one lane
 * per sum); meant for drop-in parity. */
int mi_forward_ordered(mi_context *ctx, const mi_geometry *g, int precision, const void *img, void *sino);
int mi_backproject_ordered(mi_context *ctx, const mi_geometry *g, int precision, const void *sino, void *img);

/* ---- AT building blocks (GoldenC, origin/GoldenC/src/main.c) -------------
 * Both steps sum in the reference's order (MI_ORDER_REFERENCE): equal to the
 * reference's functions bit for bit on the GoldenC geometry. */
/* minimize_in_IMAGE_variable (main.c:81-257): CG image step, FP64. */
int mi_image_step_f64(mi_context *ctx, const mi_geometry *g, double *f_out, const double *sino,
                      const double *f, const double *v, double alpha, double epsilon,
                      int iter_image);
/* minimize_in_EDGE_variable (main.c:259-360): SD edge step, FP64. */
int mi_edge_step_f64(mi_context *ctx, int nx, int ny, double *v_out, const double *f,
                     const double *v, double alpha, double beta, double epsilon, int iter_edge);
/* medfilt2 (main.c:366-390): 3x3 median, zero padding. */
int mi_medfilt2_f64(mi_context *ctx, int nx, int ny, double *f);

/* ---- drivers ------------------------------------------------------------- */
/* MumfordShah (main.c:35-49): f = 0, v = 1, iter_alt x (image step, edge
 * step), FP64 throughout, in the summation order p->order selects. */
int mi_mumford_shah(mi_context *ctx, const mi_geometry *g, const mi_at_params *p,
                    const double *sino, double *f, double *v, mi_report *rep);
/* Ordered-subset AT reconstruction of one job (the plain per-job GPU form).
 * u and v are read as the starting point and overwritten. */
int mi_at_reconstruct(mi_context *ctx, const mi_geometry *g, const mi_at_params *p,
                      const float *sino, float *u, float *v, mi_report *rep);
int mi_at_reconstruct_f64(mi_context *ctx, const mi_geometry *g, const mi_at_params *p,
                          const double *sino, double *u, double *v, mi_report *rep);

/* The per-slice loop of the helical slice driver
 * (helical/rebin/src/main.cpp:158-195, ms2d.cpp:251-274): for slice i =
 * first .. first + count - 1 it reads <prefix>i_pos and <prefix>i_raw (as
 * the reference builds them from rebined_folder), takes the slice's view
 * count and per-view source angles from the table (g->n_views, span_deg and
 * view_deg are not used), and reconstructs it with the ordered-subset AT
 * form of p (u = 0, v = 1 at the start, as ms2d.cpp:275-276).  The files
 * are read on up to 16 host threads; an unreadable slice fails the call with
 * MI_E_ARG, naming the first such slice in order.  The slices run together
 * in one batch, admitted as device memory allows.  g: the arc
 * geometry every slice shares (MI_BEAM_FAN_ARC: image size, pixel, bins,
 * LENGTH_PER_DET, SOURCE_TO_ISO / _DET, OFF_CENTER).  u_stack / v_stack:
 * count slices of nx * ny, slice-major, float for MI_FP32 and double for
 * MI_FP64; reps (may be NULL): one report per slice.  The reference runs
 * its per-ray msbeam-style sweep there (ms2d.cpp:59-184); this runs the
 * ordered-subset form of the same energy. */
int mi_helical_reconstruct(mi_context *ctx, const char *prefix, int first, int count, const mi_geometry *g,
                           const mi_at_params *p, void *u_stack, void *v_stack, mi_report *reps);

/* ---- device-resident jobs (ordered-subset form) ----------------------------
 * A job keeps its sinogram, image u, edge v and scalars in HBM between
 * calls.  Element type of sino/u0/v0/u/v: float for MI_FP32, double for
 * MI_FP64 (p->precision).  u0 / v0 may be NULL (u = 0, v = 1, as
 * origin/GoldenC/src/main.c:36-40 and msbeam/CPU_OpenMP/src/main.c:366-369).
 * Convergence (p->rel_tol) is tested on the device: once met, the job's
 * remaining queued passes do no work. */
typedef struct mi_job mi_job;
int mi_job_create(mi_context *ctx, const mi_geometry *g, const mi_at_params *p, const void *sino,
                  const void *u0, const void *v0, mi_job **out);
/* Runs up to max_passes AT iterations (fewer if it converges); blocking. */
int mi_job_run(mi_job *job, int max_passes, int *passes_run);
/* Queues passes on the context's stream and returns without waiting. */
int mi_job_enqueue(mi_job *job, int passes);
int mi_job_synchronize(mi_job *job);
int mi_job_report(mi_job *job, mi_report *rep);
int mi_job_read(mi_job *job, void *u, void *v);
int mi_job_destroy(mi_job *job);

/* ---- batches: many jobs per launch, continuous batching ---------------------
 * A batch runs the ordered-subset AT iteration of many jobs at once: jobs
 * that share a geometry (and so a system matrix) are packed into units of
 * K = 1, 2, 4 or 8 jobs whose state is interleaved, and each pass -- all
 * subset steps of all units, the median and edge steps, the convergence
 * test -- is one sequence of launches that every unit shares, replayed from
 * a hipGraph.  Jobs join and leave between passes only.  All jobs of a batch
 * use the batch's subset count and precision; each keeps its own geometry
 * and parameters (p->max_iters caps its passes, 0 = no cap; p->rel_tol
 * stops it early).  A job's result does not depend on which other jobs
 * share the batch: every reduction is cut into fixed chunks combined in a
 * fixed order.  This replaces the reference's one-job-at-a-time drivers
 * (origin/GoldenC/src/main.c:35-49, msbeam/CPU_OpenMP/src/main.c:364-394);
 * the reference has no batched form. */
typedef struct mi_batch mi_batch;
int mi_batch_create(mi_context *ctx, int n_subsets, int precision, mi_batch **out);
int mi_batch_destroy(mi_batch *b);
enum mi_batch_option {
    MI_BATCH_GRAPH = 0,       /* 1 (default): replay passes from a hipGraph  */
    MI_BATCH_MIN_K = 1,       /* smallest K of a new unit (1, 2, 4, 8)       */
    MI_BATCH_RESERVE_MB = 2,  /* device memory left free at admission (1024) */
    MI_BATCH_MAX_K = 3,       /* largest K of a unit (1, 2, 4, 8; default 8);
                                 units are also kept to K images <= 4 MiB  */
    MI_BATCH_MAX_ENTRIES = 4, /* admission stops while the active jobs'
                                 matrices hold this many entries (0: no
                                 limit, the default); one job always runs  */
    MI_BATCH_BP_LDS = 5,      /* 1 (default): back projection gathers the
                                 residual from LDS when it fits; 0: from
                                 global memory (same results, bit for bit) */
    MI_BATCH_BP_SQ = 6        /* staged back projection: where the u / v
                                 squares of the regulariser go; -1
                                 (default): chosen by occupancy, 0: the
                                 residual's LDS after the walk, 1: an LDS
                                 region of their own, loaded before the walk
                                 where it fits (same results, bit for bit) */
};
int mi_batch_set(mi_batch *b, int option, int64_t value);
/* Queues a job (inputs are copied; u0 / v0 may be NULL).  It is admitted to
 * the device at the next mi_batch_step while free device memory allows
 * (idle cached matrices are evicted first, least recently used). */
int mi_batch_submit(mi_batch *b, const mi_geometry *g, const mi_at_params *p, const void *sino,
                    const void *u0, const void *v0, int64_t *job);
/* Admits what fits, runs `passes` passes of every active job, waits, and
 * retires finished jobs (their u, v are copied to host memory held by the
 * batch and their device slot is freed).  Blocking. */
int mi_batch_step(mi_batch *b, int passes, int *n_active, int *n_pending);
/* Queues passes without admission or waiting; mi_batch_poll waits and retires. */
int mi_batch_enqueue(mi_batch *b, int passes);
int mi_batch_poll(mi_batch *b, int *n_active);
enum mi_job_state { MI_JOB_PENDING = 0, MI_JOB_ACTIVE = 1, MI_JOB_DONE = 2, MI_JOB_FAILED = 3 };
typedef struct mi_job_status {
    int32_t state;          /* enum mi_job_state                               */
    int32_t iterations;     /* AT passes completed                             */
    int32_t converged;      /* rel_tol met                                     */
    int32_t unit_k;         /* K of the job's unit while active                */
    double  data_fidelity, rel_change;
    double  t_submit, t_admit, t_done;   /* seconds on the batch's clock      */
    int64_t nnz;
} mi_job_status;
int mi_batch_status(mi_batch *b, int64_t job, mi_job_status *out);
/* u, v of a finished (or active) job, element type of the batch precision. */
int mi_batch_read(mi_batch *b, int64_t job, void *u, void *v);
/* Forgets a job: cancels it if pending or active, frees its results if done. */
int mi_batch_release(mi_batch *b, int64_t job);
/* Ids of jobs retired since the last call (up to cap; *n = how many). */
int mi_batch_retired(mi_batch *b, int64_t *ids, int cap, int *n);
/* K of each unit currently on the device (up to cap entries); *n = count. */
int mi_batch_units(mi_batch *b, int32_t *k, int cap, int *n);

/* ---- instrumentation ----------------------------------------------------- */
enum mi_kernel_id {
    MI_K_FP = 0,          /* forward projection of a subset (residual form)  */
    MI_K_BP = 1,          /* gather back projection + image gradient         */
    MI_K_FPSQ = 2,        /* forward projection of the direction, ||A d||^2  */
    MI_K_STEP = 3,        /* nV, line-search step + update                   */
    MI_K_EDGE = 4,        /* edge (v) step kernels                           */
    MI_K_CONE_FP = 5,     /* 3D cone-beam forward projection (mi_cone_*)     */
    MI_K_CONE_BP = 6,     /* 3D cone-beam gather back projection             */
    MI_K_COUNT = 8
};
typedef struct mi_kernel_stats {
    int64_t launches;
    double  total_ms;     /* sum of HIP-event durations of the launches      */
    double  bytes;        /* algorithmic bytes moved by those launches       */
} mi_kernel_stats;
/* Enable per-launch HIP-event timing of the kernels above (adds one event
 * pair per launch on the context's stream). */
int mi_stats_enable(mi_context *ctx, int on);
/* Times only the kernels whose bit is set in mask (bit k = kernel id k);
 * each timed launch adds one event pair to the stream.  0 turns timing off. */
int mi_stats_select(mi_context *ctx, uint32_t mask);
int mi_stats_get(mi_context *ctx, int kernel, mi_kernel_stats *out);
int mi_stats_reset(mi_context *ctx);
/* A plain streaming copy of `bytes` (multiple of 16) between two fresh
 * device buffers, `reps` times; *ms = mean time per copy.  It has no
 * counterpart in the reference: it is the known-size kernel against which
 * hardware memory counters of the other kernels are read. */
int mi_device_copy(mi_context *ctx, int64_t bytes, int reps, double *ms);

/* ---- reference configuration and data files (host only, no context) ---- */
/* Fills *g and *p from the text of one of the reference's JSON configuration
 * files: the helical slice driver's (keys NDETECTOR, SOURCE_TO_ISO, ...;
 * helical/rebin/src/main.cpp:62-110; gives MI_BEAM_FAN_ARC, MS_ITERATIONS as
 * iterations) or industry3d's Parameter (industry3d/src/ct3d.cpp:20-93;
 * BEAM "Cone" gives the mid-plane flat fan, "Parallel" the parallel beam,
 * ITERATIONS, LOWER_X / UPPER_X).  A missing required key is MI_E_ARG, as a
 * missing key stops the reference's parser. */
int mi_config_parse(const char *json_text, mi_geometry *g, mi_at_params *p);
/* Reads a rebinned helical slice's table, <slice>_pos
 * (helical/rebin/src/ms2d.cpp:33-47): lines "pos,angle" (table position as
 * written, source angle in degrees) up to the first line that does not
 * parse.  Up to cap rows go to pos_mm / angle_deg (either may be NULL);
 * *n = rows in the file, the slice's view count (NPROJ). */
int mi_read_pos(const char *path, double *pos_mm, double *angle_deg, int cap, int *n);
/* Reads a slice's sinogram, <slice>_raw (ms2d.cpp:270-274): n_views *
 * n_bins whitespace-separated numbers in [view][bin] order.  Fewer is
 * MI_E_ARG. */
int mi_read_raw(const char *path, int64_t n_views, int64_t n_bins, double *sino);
/* Reads an industry3d projection file (industry3d/src/ct3d.cpp:199-214,
 * industry3d/readme.txt): a 1024-byte header, nproj * ndx * ndy unsigned
 * shorts in [proj][x][y] order into proj, then (flat != NULL) the ndx * ndy
 * float blank scan into flat. */
int mi_read_dr(const char *path, int64_t nproj, int64_t ndx, int64_t ndy, uint16_t *proj, float *flat);

/* ---- 3D cone beam (industry3d) ------------------------------------------ */
/* Geometry of industry3d's Parameter (industry3d/include/ct3d.h:17-63) for
 * BEAM "Cone": volume [nz][nx][ny] (ny fastest, get_img_addr of
 * industry3d/src/tracing.cpp:119-121) of sample_mm voxels; n_views views
 * over 360 degrees; a flat panel of n_rows (NDX, along z) by n_cols (NDY)
 * pixels of pixel_mm.  Both panel axes are centred on HALFDET = NDX / 2, as
 * industry3d/src/tracing.cpp:51-81 and ct3d.cpp:88-93 place them.  The
 * sinogram is [view][row][col] (CTInput's sino order, ct3d.h:65-105). */
typedef struct mi_cone_geometry {
    int32_t nx, ny, nz;        /* NX, NY, NZ                                  */
    int32_t n_views;           /* NPROJ                                       */
    int32_t n_rows, n_cols;    /* NDX, NDY                                    */
    float   sample_mm;         /* SAMPLESIZE                                  */
    float   pixel_mm;          /* PIXELSIZE                                   */
    float   sod_mm, sdd_mm;    /* SOD, SDD                                    */
    int32_t beam;              /* BEAM: MI_CONE_BEAM (0, "Cone") or
                                  MI_CONE_PARALLEL (1, "Parallel": parallel()
                                  of industry3d/src/tracing.cpp:9-49, rays
                                  from the plane at -SOD to the panel, every
                                  axis centred on HALFSIZE = NX / 2)          */
    int32_t reserved;
} mi_cone_geometry;
enum { MI_CONE_BEAM = 0, MI_CONE_PARALLEL = 1 };

/* Replaces get_tracing + forward_proj (industry3d/src/main.cpp:58-73,
 * industry3d/src/tracing.cpp:51-81,123-307): the (voxel index, length)
 * sequence of ray (view, row, col) as the device traces it, in float with
 * the reference's operation order; up to cap entries, *numb = full count. */
int mi_cone_trace_ray(mi_context *ctx, const mi_cone_geometry *g, int view, int row, int col, int64_t *ind,
                      float *len, int cap, int *numb);
/* A: sino[view][row][col] = sum over the ray's traced entries of
 * vol[idx] * len (the accum_Af of industry3d/src/main.cpp:254-266), summed
 * in double in walk order and rounded once. */
int mi_cone_forward(mi_context *ctx, const mi_cone_geometry *g, const float *vol, float *sino);
int mi_cone_forward_dev(mi_context *ctx, const mi_cone_geometry *g, const float *vol, float *sino);
/* Back projection in gather form: vol[z][x][y] = sum over the rays that
 * cross voxel (x, y, z) of sino * (exact length of the ray inside the
 * voxel), each voxel summed by one lane in a fixed order (no atomics).  The
 * reference applies the transpose by scattering along the traced rays
 * (industry3d/src/main.cpp:385-406); the lengths here are the exact ones
 * of the same rays, which the float tracer approximates (it also drops a
 * final piece shorter than 0.05 voxel, tracing.cpp:268). */
int mi_cone_backproject(mi_context *ctx, const mi_cone_geometry *g, const float *sino, float *vol);
int mi_cone_backproject_dev(mi_context *ctx, const mi_cone_geometry *g, const float *sino, float *vol);
/* mi_cone_backproject_dev without any device allocation.  It has no
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The double sums the back projection carries between its
 * launches of 36 views live in a fixed array of the library (64 MiB on each
 * device, held until the process ends): with more than 36 views the volume
 * may hold at most 8388608 voxels (MI_E_LIMIT above that).  The geometry's
 * tables must be on the device already (any earlier mi_cone_* call with the
 * same geometry on this context builds them), else MI_E_STATE.  The call is
 * queued on the context's stream and returns without waiting; the result is
 * in vol once the stream has run it.  Calls that overlap on one device
 * (other contexts, other threads) share the array and race. */
int mi_cone_backproject_dev_noalloc(mi_context *ctx, const mi_cone_geometry *g, const float *sino, float *vol);
/* Ordered-subset AT reconstruction of the volume: the OS form of
 * mi_at_reconstruct on [nz][nx][ny] with the projectors above and the
 * 26-neighbourhood widening of GoldenC's stencils (13 forward offsets k,
 * weight 1/|k|^2, u zero- and v one-padded); takes the place of industry3d's
 * per-row gdIMAGE / gdEDGE sweep (industry3d/src/main.cpp:233-532).  p uses
 * alpha, beta, epsilon, n_subsets, max_iters, iter_edge, rel_tol, lower,
 * upper; precision must be MI_FP32 and median_every 0.  u and v hold the
 * initial values and receive the result. */
int mi_cone_at_reconstruct(mi_context *ctx, const mi_cone_geometry *g, const mi_at_params *p, const float *sino,
                           float *u, float *v, mi_report *rep);
/* The same on device buffers (sinogram, u, v), updated in place. */
int mi_cone_at_reconstruct_dev(mi_context *ctx, const mi_cone_geometry *g, const mi_at_params *p,
                               const float *sino, float *u, float *v, mi_report *rep);
/* Fills *g and *p from the text of an industry3d configuration file
 * (Parameter::parse_config, industry3d/src/ct3d.cpp:20-93: BEAM "Cone" or
 * "Parallel", NX, NY, NZ, NPROJ, NDX, NDY, SAMPLESIZE, PIXELSIZE, SOD, SDD,
 * ALPHA, BETA, EPSILON, ITERATIONS; LOWER_X / UPPER_X optional).  A missing
 * key, or another BEAM, is MI_E_ARG.  Host only (no context). */
int mi_cone_config_parse(const char *json_text, mi_cone_geometry *g, mi_at_params *p);

#ifdef __cplusplus
}
#endif
#endif
