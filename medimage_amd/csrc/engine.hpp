// Batched ordered-subset AT engine (engine.hip): one launch per phase covers
// every batch unit, where a unit is K <= 8 jobs that share one system matrix
// and whose state is stored job-interleaved (element e of job j at e*K + j),
// so one matrix stream and one vector gather serve all K jobs.
//
// Every reduction that sets a step size is cut into fixed chunks (a block:
// four slices of the sliced forward / back projector, or 256 tiled pixel
// positions of the stencils) whose partial sums are combined in index order
// by a one-block-per-job finalize.  The chunking depends only on the
// geometry -- not on the grid, nor K, nor the other units of the launch -- so
// a job's result is bit-for-bit the same alone and in any batch.
#pragma once

#include "runtime.hpp"

namespace mi {

constexpr int kMaxK = 8;             // jobs per unit
constexpr int kMaxUnits = 64;        // units per batch
constexpr int kPixBlock = kBlock;    // tiled pixel positions per stencil block
// Largest dynamic LDS of a back-projection block (the staged residual
// slots); units above it run in the global-gather launch.  Measured with
// the slots staged by direct global-to-LDS loads: the 64-job batch (no
// tuning) runs 2829 AT iterations/s with a 24 KiB limit, 2903 with 40 KiB,
// 2911 with 56 KiB.  (With the slots staged through registers the copy cost
// more than the gathers saved above ~24 KiB: 2632 it/s all global, 2256 with
// 64 KiB.)
#ifndef MI_BP_LDS_KB
#define MI_BP_LDS_KB 56
#endif
constexpr size_t kBpLdsMax = (size_t)MI_BP_LDS_KB << 10;
// The most a tuned geometry may stage (within the 64 KiB a launch gets
// without opting in, with room for the kernel's static LDS).
constexpr size_t kBpLdsHw = (size_t)56 << 10;
static_assert(kBpLdsMax <= kBpLdsHw, "default staging limit above the tuned one");
// Staged back-projection units needing more than this many bytes of slots
// run in a launch of their own (a launch gives each block its largest unit's
// LDS).
#ifndef MI_BP_LDS_SPLIT_KB
#define MI_BP_LDS_SPLIT_KB 20
#endif
constexpr size_t kBpLdsSplit = (size_t)MI_BP_LDS_SPLIT_KB << 10;
// Dynamic LDS of a back-projection launch, slots and u / v squares together
// (the 64 KiB less the kernel's static LDS and a margin).
constexpr size_t kBpLdsLaunch = (size_t)63 << 10;

// Partial-sum slots of a unit (each K * cap doubles, see UnitDesc).
enum EngSlot { SL_DF = 0, SL_DD, SL_AD, SL_NV, SL_E0, SL_E1, SL_E2, SL_E3, SL_COUNT };

// Per-job parameters on the device.
struct JobParams {
    double alpha, beta, eps, lower, upper, rel_tol;
    int32_t iter_edge, median_every, max_iters, active;
};

// One unit.  Pointers to the sliced matrix (shared) and to the unit's
// interleaved state; images are tiled (npix_t positions).  Lengths are T.
template <typename T>
struct UnitDesc {
    const int64_t *fs_slice0, *fs_base;   // forward: S + 1 slice offsets, per-slice entry base
    const int32_t *fs_T;                   // per-slice entries per lane
    const uint32_t *fs_idx;                // tiled pixel indices (null when compact)
    const uint32_t *fs_first;              // compact form: first index per lane
    const int16_t *fs_d16;                 // compact form: index differences
    const T *fs_len;
    const int64_t *bs_slice0, *bs_base;   // back: same, items are tiled pixels
    const int32_t *bs_T;
    const uint32_t *bs_idx;                // rows (null when compact)
    const uint32_t *bs_first;
    const int16_t *bs_d16;
    const T *bs_len;
    const uint16_t *bs_slot16;             // LDS slot per entry (null: global gathers only)
    const int32_t *bs_vlo, *bs_voff;       // per (subset, block, view): first bin, slot offset
    const int64_t *bs_vtab0;               // S + 1 table offsets
    // checkpointed forward walk (fan beams; null otherwise): per lane of a
    // forward block of kOtfRays rays x kOtfSeg segments (runtime.hpp)
    const double *ot_lam, *ot_ax, *ot_ay, *ot_stx, *ot_sty;
    const int32_t *ot_cx, *ot_cy, *ot_sg, *ot_cnt;
    const int64_t *sub_row0;               // S + 1 row offsets
    T *g, *r;                              // rows * K
    T *u, *v, *d, *w;                      // npix_t * K
    int64_t npix_t, nrows;
    int64_t part_off, cap;                 // partials: part + part_off + (slot * K + j) * cap
    int nx, ny, tny, K, fs_P, bs_P;
    int job0;                              // job slots job0 .. job0 + K - 1 (stable)
    int nbins;
};

// Block ranges of one launch: unit u owns blocks [start[u], start[u+1]).
struct BlockMap {
    int nunits;
    int start[kMaxUnits + 1];
};

// Arguments common to the engine's launches.
template <typename T>
struct EngArgs {
    const UnitDesc<T> *units;
    const JobParams *jp;
    Scalars *sc;                // one per job slot
    double *part;               // partial sums, laid out per unit
    int nunits;                 // units in the table (finalize grid: nunits * kMaxK)
    int S;
};

// Launchers (all on `st`).  `s` is the subset step.  kmax: the largest K of
// the launch's units (selects code compiled for K <= 1, 4 or 8).  all_walk: every
// unit's forward projector is the checkpointed walk (selects code without
// the sliced forward branch).
template <typename T>
void eng_fp_resid(hipStream_t st, const EngArgs<T> &a, const BlockMap &m, int s, int kmax, bool all_walk);
template <typename T>
// lds_bytes: 0, or the dynamic LDS of the staged-residual form (every unit
// has slot codes and max_slots * K * sizeof(T) fits).  sq_mode, staged form
// only: where the u / v squares of the regulariser go (-1: by occupancy, 0:
// the slots' LDS after the walk, 1: a region of their own where it fits).
void eng_bp_grad(hipStream_t st, const EngArgs<T> &a, const BlockMap &m, int s, int kmax, size_t lds_bytes,
                 int sq_mode = -1);
// ||A_s d||^2 (rows map; the walk of the residual launch without the
// residual) and nV(d) (pixel map), as two launches.
template <typename T>
void eng_fp_sq(hipStream_t st, const EngArgs<T> &a, const BlockMap &rows, int s, int kmax, bool all_walk);
template <typename T>
void eng_nv(hipStream_t st, const EngArgs<T> &a, const BlockMap &pix, int kmax);
template <typename T>
void eng_axpy(hipStream_t st, const EngArgs<T> &a, const BlockMap &pix, int kmax);
template <typename T>
void eng_median(hipStream_t st, const EngArgs<T> &a, const BlockMap &pix, int kmax);
template <typename T>
void eng_edge_iter(hipStream_t st, const EngArgs<T> &a, const BlockMap &pix, int it, int kmax);

enum EngFin { EF_BEGIN = 0, EF_STEP, EF_EDGE_C, EF_EDGE_DONE, EF_END };
template <typename T>
void eng_finalize(hipStream_t st, const EngArgs<T> &a, int op, int s);

// Job I/O on one slot of a unit (n elements of the interleaved array).
template <typename T>
void eng_put_rows(hipStream_t st, T *dst, int K, int slot, int64_t n, const int64_t *perm, const T *src);
template <typename T>
void eng_put_img(hipStream_t st, T *dst, int K, int slot, int nx, int ny, int tny, const T *src, T fillv);
template <typename T>
void eng_get_img(hipStream_t st, T *dst, const T *src, int K, int slot, int nx, int ny, int tny);
// Copies slot `from` of a unit array to slot `to` of another (repacking).
template <typename T>
void eng_move(hipStream_t st, T *dst, int Kd, int sd, const T *src, int Ks, int ss, int64_t n);

// Plain device copy (bytes a multiple of 16).
void eng_copy(hipStream_t st, void *dst, const void *src, int64_t bytes);

}  // namespace mi
