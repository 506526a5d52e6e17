// Launchers of the projection and AT kernels (kernels.hip).  All launches
// go to the given stream; reductions land in device Scalars slots, never on
// the host.
#pragma once

#include "common.hpp"
#include "../../include/medimage_amd.h"

namespace mi {

// Per-block partial sums of the grid reductions: kSlots regions of kMaxGrid
// doubles.  Producers store one value per block (plain stores, no fences);
// a one-block finalize kernel sums a region in index order after the kernel
// boundary, so every reduction is deterministic.
constexpr int kMaxGrid = 8192;
constexpr int kSlots = 8;
struct RedBuf {
    double *partial = nullptr;   // kSlots * kMaxGrid doubles
    double *slot(int k) const { return partial + (size_t)k * kMaxGrid; }
};

enum FpMode {
    FP_WRITE = 0,   // out = A x;               partial sum out^2
    FP_RESID = 1,   // out = g - A x;           partial sum out^2
    FP_SQ = 2,      // (no output);             partial sum (A x)^2
    FP_DOT = 3,     // (no output);             partial sum (A x) q
    FP_ERR = 4      // out = A x;               partial sum (out - g)^2
};

struct CsrView {
    const int64_t *rowptr;
    const uint32_t *idx;
    const void *len;        // float* or double*
};
struct CscView {
    const int64_t *colptr;  // already offset to the subset's first column
    const uint32_t *crow;
    const void *len;
};

enum FinOp {
    FIN_STORE = 0,      // *dst = sum(in0)
    FIN_ACCUM = 1,      // *dst += sum(in0)
    FIN_STEP = 2,       // sc: dd, ad2, nv <- in0..2; c = dd / (S ad2 + alpha nv)
    FIN_EDGE_C = 3,     // sc: edge step size from in0..2 (origin/GoldenC/src/main.c:329)
    FIN_EDGE_DONE = 4   // sc: edge_done when max(in0) < 1e-5 (main.c:351)
};
struct FinArgs {
    int op = 0, nin = 0;
    const double *in[4] = {nullptr, nullptr, nullptr, nullptr};
    int n[4] = {0, 0, 0, 0};
    int is_max[4] = {0, 0, 0, 0};
    double *dst = nullptr;
    Scalars *sc = nullptr;
    const int32_t *skip = nullptr;
    double S = 1, alpha = 0, beta = 0, eps = 0;
};

// Grid sizes the producers use (the number of partials they write).
// grid_cap 0 selects the default cap (2048 blocks).
unsigned grid_rows(int64_t nrows, int grid_cap = 0);
unsigned grid_pixels(int64_t n);
unsigned grid_bp(int64_t npix, int lanes, int grid_cap = 0);

// Every producer takes an optional device flag `skip`: when *skip != 0 the
// kernel returns at once (a converged job's queued passes then cost one
// launch each and no memory traffic).  `part` may be null (no partials).
// Each returns its grid size.

// Forward projection of rows [row0, row0 + nrows); g/q/out are indexed by row.
// vec: 1 = each lane takes four consecutive entries (16-byte loads);
// 0 = lane-strided 4-byte loads, four entries in flight per lane.
template <typename T>
unsigned launch_fp(hipStream_t st, int mode, const CsrView &A, int64_t row0, int64_t nrows, const T *x, const T *g,
                   const T *q, T *out, double *part, const int32_t *skip = nullptr, int vec = 0,
                   int grid_cap = 0);

// d = scale * (A_s^T r) + alpha * R(u, v) over an nx x ny image (R is the
// GoldenC image-gradient regulariser; skipped when u == nullptr); partial
// sum d^2.  lanes: lanes per pixel column (4, 8, 16, 32 or 64).
template <typename T>
unsigned launch_bp_grad(hipStream_t st, const CscView &At, int nx, int ny, const T *r, double scale, const T *u,
                        const T *v, double alpha, double eps, T *d, double *part, int lanes,
                        const int32_t *skip = nullptr, int grid_cap = 0);

// partial sum (v^2 + 0.01 eps^2)((p - p_up)^2 + (p - p_left)^2), p zero
// padded (origin/GoldenC/src/main.c:152-158).
template <typename T>
unsigned launch_nv(hipStream_t st, int nx, int ny, const T *p, const T *v, double eps, double *part,
                   const int32_t *skip = nullptr);

// Interior energy sum of origin/GoldenC/src/main.c:135-139.
template <typename T>
unsigned launch_energy_reg(hipStream_t st, int nx, int ny, const T *f, const T *v, double alpha, double eps,
                           double *part);

template <typename T>
unsigned launch_dot(hipStream_t st, int64_t n, const T *a, const T *b, double *part);

// y = clamp(y + c x, lo, hi) with c = *cptr (or cval when cptr is null);
// save (optional) receives the old y; partial max |c x|.
template <typename T>
unsigned launch_axpy(hipStream_t st, int64_t n, T *y, const T *x, const double *cptr, double cval, double lo,
                     double hi, T *save, double *part, const int32_t *skip = nullptr);

// y = x + c y (c on the host).
template <typename T>
void launch_xpby(hipStream_t st, int64_t n, T *y, const T *x, double c);

// 3x3 median, zero padding (origin/GoldenC/src/main.c:366-390); out != in.
template <typename T>
void launch_median(hipStream_t st, int nx, int ny, const T *in, T *out);

void launch_finalize(hipStream_t st, const FinArgs &a);

// One SD edge iteration of origin/GoldenC/src/main.c:288-344 on v given f
// (five launches, partial slots 0-3).  Skips itself once sc->edge_done is
// set and sets it when max |c d| < 1e-5.
template <typename T>
void launch_edge_iter(hipStream_t st, const RedBuf &rb, int nx, int ny, const T *f, T *v, T *dwork, double alpha,
                      double beta, double eps, Scalars *sc);

// The edge step's parts: the descent direction d and n = |central grad f|^2
// (no sums), and the update v = clamp(v + e_c d, 0, 1) with the early-stop
// test (partial slot 3).  Both skip once sc->edge_done is set.
template <typename T>
void launch_edge_grad_n(hipStream_t st, int nx, int ny, const T *f, const T *v, T *d, T *n, double alpha,
                        double beta, double eps, const Scalars *sc);
template <typename T>
void launch_edge_update(hipStream_t st, const RedBuf &rb, int nx, int ny, T *v, const T *d, Scalars *sc);

// Pass bookkeeping of the ordered-subset form.
void launch_pass_begin(hipStream_t st, Scalars *sc);
void launch_pass_end(hipStream_t st, Scalars *sc, double rel_tol);

// Gathers a [view][bin] sinogram into subset row order: out[r] = in[perm[r]];
// scatter does the inverse.
template <typename T>
void launch_gather_rows(hipStream_t st, int64_t n, const int64_t *perm, const T *in, T *out);
template <typename T>
void launch_scatter_rows(hipStream_t st, int64_t n, const int64_t *perm, const T *in, T *out);

}  // namespace mi
