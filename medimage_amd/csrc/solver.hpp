// Solver drivers (solver.hip).
#pragma once

#include "kernels.hpp"
#include "runtime.hpp"

namespace mi {

// Device state of one reconstruction job.  Rows are in the matrix's subset
// order; images are row-major [nx][ny].
template <typename T>
struct JobState {
    int nx = 0, ny = 0;
    int64_t npix = 0, nrows = 0;
    DevBuf<T> g, r, q, af;         // rows: sinogram, residual, A p, A update
    DevBuf<T> u, v, d, p, t, w;    // images: update, edge, gradient, CG direction, saved update, scratch
    DevBuf<Scalars> sc;
    DevBuf<double> red_partial;
    explicit JobState(const Matrix &m);
    RedBuf rb() const;
    Scalars read_scalars(hipStream_t st) const;
};

template <typename T>
void fill(hipStream_t st, T *p, int64_t n, T v);

CsrView csr_view(const Matrix &m, bool f64);
CscView csc_view(const Matrix &m, int s, bool f64);
int bp_lanes(const Matrix &m);
double fp_bytes(const Matrix &m, int64_t nrows, size_t tsz, int mode);
double bp_bytes(const Matrix &m, int64_t nrows_sub, size_t tsz, bool reg);

// GoldenC steps on a job's device state (u: image, v: edge).
template <typename T>
int gc_image_step(mi_context *ctx, const Matrix &m, JobState<T> &js, double alpha, double eps, int iter_image,
                  int *why);
template <typename T>
void gc_edge_step(mi_context *ctx, JobState<T> &js, double alpha, double beta, double eps, int iter_edge);

// Measures the launch shapes of the streaming kernels for m and stores the
// fastest in m.tune.
mi_tuning autotune(mi_context *ctx, Matrix &m);

// Host <-> device sinogram transfer in the caller's [view][bin] order.
template <typename T>
void load_sino(hipStream_t st, const Matrix &m, const T *host, T *rows);

}  // namespace mi
