// Launchers of the projection and AT kernels (kernels.hip).  All launches
// go to the given stream; reductions land in device Scalars slots, never on
// the host.
#pragma once

#include "common.hpp"
#include "../../include/medimage_amd.h"

namespace mi {

// Work buffers of the grid reductions (one set per stream in flight).
struct RedBuf {
    double *partial = nullptr;   // kMaxGrid * 2 doubles
    unsigned *ticket = nullptr;  // one counter
};
constexpr int kMaxGrid = 8192;

// Where a reduced value goes: *dst = v (or += v when accumulate).
struct Slot {
    double *dst = nullptr;
    int accumulate = 0;
};

enum FpMode {
    FP_WRITE = 0,   // out = A x;               reduce sum out^2
    FP_RESID = 1,   // out = g - A x;           reduce sum out^2
    FP_SQ = 2,      // (no output);             reduce sum (A x)^2
    FP_DOT = 3,     // (no output);             reduce sum (A x) q
    FP_ERR = 4      // out = A x;               reduce sum (out - g)^2
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

// Every launcher takes an optional device flag `skip`: when *skip != 0 the
// kernel returns at once (a converged job's remaining passes cost one launch
// each and no memory traffic).
//
// Forward projection of rows [row0, row0 + nrows).  x: image; g/q/out are
// indexed by row.
template <typename T>
void launch_fp(hipStream_t st, const RedBuf &rb, int mode, const CsrView &A, int64_t row0, int64_t nrows,
               const T *x, const T *g, const T *q, T *out, Slot slot, const int32_t *skip = nullptr);

// d = scale * (A_s^T r) + alpha * R(u, v) over an nx x ny image (R is the
// GoldenC image-gradient regulariser; skipped when u == nullptr).  Reduces
// sum d^2 into slot.  lanes: lanes per pixel column (8, 16, 32 or 64).
template <typename T>
void launch_bp_grad(hipStream_t st, const RedBuf &rb, const CscView &At, int nx, int ny, const T *r, double scale,
                    const T *u, const T *v, double alpha, double eps, T *d, Slot slot, int lanes,
                    const int32_t *skip = nullptr);

// sum (v^2 + 0.01 eps^2)((p - p_up)^2 + (p - p_left)^2), p zero padded, v
// one padded (origin/GoldenC/src/main.c:152-158).  When sc != nullptr the
// last block also sets sc->c = sc->dd / (S * sc->ad2 + alpha * nv).
template <typename T>
void launch_nv(hipStream_t st, const RedBuf &rb, int nx, int ny, const T *p, const T *v, double eps, Slot slot,
               Scalars *sc, double S, double alpha, const int32_t *skip = nullptr);

// Interior energy sum of origin/GoldenC/src/main.c:135-139.
template <typename T>
void launch_energy_reg(hipStream_t st, const RedBuf &rb, int nx, int ny, const T *f, const T *v, double alpha,
                       double eps, Slot slot);

// sum a*b over n.
template <typename T>
void launch_dot(hipStream_t st, const RedBuf &rb, int64_t n, const T *a, const T *b, Slot slot);

// y = clamp(y + c x, lo, hi) with c = *cptr (or cval when cptr is null);
// save (optional) receives the old y; reduces max |c x| into slot.
template <typename T>
void launch_axpy(hipStream_t st, const RedBuf &rb, int64_t n, T *y, const T *x, const double *cptr, double cval,
                 double lo, double hi, T *save, Slot slot, const int32_t *skip = nullptr);

// y = x + c y (c on the host).
template <typename T>
void launch_xpby(hipStream_t st, int64_t n, T *y, const T *x, double c);

// 3x3 median, zero padding (origin/GoldenC/src/main.c:366-390); out != in.
template <typename T>
void launch_median(hipStream_t st, int nx, int ny, const T *in, T *out);

// One SD edge iteration of origin/GoldenC/src/main.c:288-344 on v given f.
// tmp: nx*ny work array.  Skips itself when sc->edge_done is set and sets it
// when max |c d| < 1e-5.
template <typename T>
void launch_edge_iter(hipStream_t st, const RedBuf &rb, int nx, int ny, const T *f, T *v, T *dwork, double alpha,
                      double beta, double eps, Scalars *sc);

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
