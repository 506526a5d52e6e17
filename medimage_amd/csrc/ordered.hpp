// Reference-order FP64 operators and GoldenC drivers (ordered.hip).
//
// The GoldenC program (origin/GoldenC/src/main.c, proj.c) adds every sum
// sequentially: each sinogram value of A and each pixel of AStar in the
// order its traversal emits the products, and every energy, inner product
// and norm in row-major loop order.  Its CG with median filters and a revert
// test amplifies rounding differences by about 20x per alternation, so the
// drop-in keeps those orders exactly: one lane per ray (or pixel) sums its
// products in the reference's order, and each reduction is one wave's
// sequential chain.  The results are then the reference's bit for bit.
#pragma once

#include "solver.hpp"

namespace mi {

// Per-item entry lists, items (rays or pixels) in groups of 64, one lane
// each: entry t of lane l of group gi sits at base[gi] + 64 t + l; a group
// holds steps[gi] entries per lane (its longest list rounded up to a
// multiple of four); lanes past their own count read padding (item 0,
// length 0) and do not add it.
struct SeqLists {
    int64_t items = 0;
    DevBuf<int64_t> base;
    DevBuf<int32_t> steps, cnt;
    DevBuf<uint32_t> other;   // rays: pixel index; pixels: row index
    DevBuf<double> len;
    size_t bytes() const {
        return base.bytes() + steps.bytes() + cnt.bytes() + other.bytes() + len.bytes();
    }
};
struct SeqMatrix {
    SeqLists rows, pixels;
    size_t bytes() const { return rows.bytes() + pixels.bytes(); }
};

// The matrix's lists in the reference's accumulation order, built once per
// matrix (one subset only).  MI_BEAM_GOLDENC: A's and AStar's emission
// order (proj.c:14-397).  Other beams: the tracer's order along each ray,
// and ray order for each pixel (a scatter loop over the rays, as msbeam's
// AStar, msbeam/CPU_OpenMP/src/proj.c:34-52).
const SeqMatrix &get_seq(hipStream_t st, Matrix &m);

// out[r] = sum of x[pix] * len over ray r in list order, in T; with g,
// also resid[r] = g[r] - out[r].
template <typename T>
void seq_forward(hipStream_t st, const SeqMatrix &q, const T *x, T *out, const T *g = nullptr,
                 T *resid = nullptr);
// d[p] = sum of r[row] * len over pixel p in list order, then (u != null)
// d[p] += alpha * R(u, v) (origin/GoldenC/src/main.c:112-125).
template <typename T>
void seq_backproject(hipStream_t st, const SeqMatrix &q, int nx, int ny, const T *r, const T *u, const T *v,
                     double alpha, double eps, T *d);

// minimize_in_IMAGE_variable / minimize_in_EDGE_variable
// (origin/GoldenC/src/main.c:81-257, 259-360) in the reference's order.
int gc_image_step_ordered(mi_context *ctx, Matrix &m, JobState<double> &js, double alpha, double eps,
                          int iter_image, int *why);
void gc_edge_step_ordered(mi_context *ctx, JobState<double> &js, double alpha, double beta, double eps,
                          int iter_edge);

}  // namespace mi
