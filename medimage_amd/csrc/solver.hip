// Solver drivers: the GoldenC CG/SD alternation (origin/GoldenC/src/main.c)
// and the ordered-subset AT iteration, both on device-resident state with
// FP64 reductions feeding every step size.
#include "solver.hpp"

#include <cmath>
#include <cstring>
#include <limits>

namespace mi {

namespace {

__global__ void k_fill_f32(float *p, int64_t n, float v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double *p, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

template <typename T>
void fill(hipStream_t st, T *p, int64_t n, T v) {
    const unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
    if (grid == 0) return;
    if constexpr (sizeof(T) == 4) k_fill_f32<<<grid, kBlock, 0, st>>>((float *)p, n, (float)v);
    else k_fill_f64<<<grid, kBlock, 0, st>>>((double *)p, n, (double)v);
    MI_HIP(hipGetLastError());
}
template void fill<float>(hipStream_t, float *, int64_t, float);
template void fill<double>(hipStream_t, double *, int64_t, double);

CsrView csr_view(const Matrix &m, bool f64) {
    return CsrView{m.rowptr.p, m.idx.p, f64 ? (const void *)m.len64.p : (const void *)m.len32.p};
}
CscView csc_view(const Matrix &m, int s, bool f64) {
    return CscView{m.colptr.p + (int64_t)s * m.npix, m.crow.p, f64 ? (const void *)m.clen64.p : (const void *)m.clen32.p};
}

// Lanes per pixel column of the gather back projector: each lane takes four
// entries per trip; aim at about three trips over the subset's mean column
// length, within [4, 64].
int bp_lanes(const Matrix &m) {
    if (m.tune.bp_lanes) return m.tune.bp_lanes;
    const double mean = (double)m.nnz / ((double)m.S * (double)m.npix);
    int l = 4;
    while (l < 64 && l * 12 < mean) l *= 2;
    return l;
}

double fp_bytes(const Matrix &m, int64_t nrows, size_t tsz, int mode) {
    // Mean entries per row of the geometry times the rows of the launch.
    const double nnz = (double)m.nnz * (double)nrows / (double)m.nrows;
    double b = nnz * (4.0 + tsz) + (double)(nrows + 1) * 8.0 + (double)m.npix * tsz;
    if (mode == FP_RESID || mode == FP_ERR) b += 2.0 * nrows * tsz;
    else if (mode == FP_WRITE || mode == FP_DOT) b += (double)nrows * tsz;
    return b;
}

double bp_bytes(const Matrix &m, int64_t nrows_sub, size_t tsz, bool reg) {
    const double nnz = (double)m.nnz * (double)nrows_sub / (double)m.nrows;
    double b = nnz * (4.0 + tsz) + (double)(m.npix + 1) * 8.0 + (double)nrows_sub * tsz + (double)m.npix * tsz;
    if (reg) b += 2.0 * m.npix * tsz;
    return b;
}

// ---------------------------------------------------------------------------

template <typename T>
JobState<T>::JobState(const Matrix &m) {
    nx = m.g.nx;
    ny = m.g.ny;
    npix = (int64_t)nx * ny;
    nrows = m.nrows;
    g.alloc(nrows);
    r.alloc(nrows);
    q.alloc(nrows);
    af.alloc(nrows);
    u.alloc(npix);
    v.alloc(npix);
    d.alloc(npix);
    p.alloc(npix);
    t.alloc(npix);
    w.alloc(npix);
    sc.alloc(1);
    red_partial.alloc((size_t)kSlots * kMaxGrid);
    MI_HIP(hipMemset(sc.p, 0, sizeof(Scalars)));
}

template <typename T>
RedBuf JobState<T>::rb() const {
    return RedBuf{red_partial.p};
}

template <typename T>
Scalars JobState<T>::read_scalars(hipStream_t st) const {
    Scalars h;
    MI_HIP(hipMemcpyAsync(&h, sc.p, sizeof(Scalars), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    return h;
}


// ---------------------------------------------------------------------------
// GoldenC: minimize_in_IMAGE_variable (main.c:81-257), host-driven CG with
// device reductions.  gen[] slots: 0 data term, 1 interior energy, 2 ||q||^2,
// 3 nV(p), 4 <p,d>, 5 max |c1 p|, 6 <A d, q>.
template <typename T>
int gc_image_step(mi_context *ctx, const Matrix &m, JobState<T> &js, double alpha, double eps, int iter_image,
                  int *why) {
    const hipStream_t st = ctx->stream;
    const bool f64 = sizeof(T) == 8;
    const RedBuf rb = js.rb();
    Scalars *sc = js.sc.p;
    const CsrView A = csr_view(m, f64);
    const CscView At = csc_view(m, 0, f64);
    const int lanes = bp_lanes(m);
    const double inf = std::numeric_limits<double>::infinity();
    const int cadence = (int)(iter_image / 2 + 0.0001);
    const size_t ib = js.npix * sizeof(T);
    int reason = 0, ran = 0;

    // Producer into partial slot k, then a one-block sum into sc->gen[k].
    auto to_gen = [&](unsigned grid, int k, bool is_max = false) {
        FinArgs f{};
        f.op = FIN_STORE;
        f.nin = 1;
        f.in[0] = rb.slot(k);
        f.n[0] = (int)grid;
        f.is_max[0] = is_max;
        f.dst = &sc->gen[k];
        launch_finalize(st, f);
    };
    to_gen(launch_fp<T>(st, FP_RESID, A, 0, m.nrows, js.u.p, js.g.p, nullptr, js.r.p, rb.slot(0)), 0);
    launch_bp_grad<T>(st, At, js.nx, js.ny, js.r.p, 1.0, js.u.p, js.v.p, alpha, eps, js.d.p, nullptr, lanes);
    to_gen(launch_energy_reg<T>(st, js.nx, js.ny, js.u.p, js.v.p, alpha, eps, rb.slot(1)), 1);
    MI_HIP(hipMemcpyAsync(js.p.p, js.d.p, ib, hipMemcpyDeviceToDevice, st));
    Scalars h = js.read_scalars(st);
    double at_old = h.gen[0] + h.gen[1];
    for (int k = 1; k <= iter_image; ++k) {
        ran = k;
        to_gen(launch_fp<T>(st, FP_WRITE, A, 0, m.nrows, js.p.p, nullptr, nullptr, js.q.p, rb.slot(2)), 2);
        to_gen(launch_nv<T>(st, js.nx, js.ny, js.p.p, js.v.p, eps, rb.slot(3)), 3);
        to_gen(launch_dot<T>(st, js.npix, js.p.p, js.d.p, rb.slot(4)), 4);
        h = js.read_scalars(st);
        const double nap = h.gen[2];
        const double c1 = h.gen[4] / (nap + alpha * h.gen[3]);
        to_gen(launch_axpy<T>(st, js.npix, js.u.p, js.p.p, nullptr, c1, -inf, inf, js.t.p, rb.slot(5)), 5, true);
        // The reference takes k % 0 when iter_image < 2 (undefined); no
        // median is applied then.
        if (cadence > 0 && k % cadence == 0) {
            launch_median<T>(st, js.nx, js.ny, js.u.p, js.w.p);
            MI_HIP(hipMemcpyAsync(js.u.p, js.w.p, ib, hipMemcpyDeviceToDevice, st));
        }
        h = js.read_scalars(st);
        if (k == iter_image) { reason = 0; break; }
        if (h.gen[5] < 1e-5) { reason = 1; break; }
        to_gen(launch_fp<T>(st, FP_ERR, A, 0, m.nrows, js.u.p, js.g.p, nullptr, js.af.p, rb.slot(0)), 0);
        to_gen(launch_energy_reg<T>(st, js.nx, js.ny, js.u.p, js.v.p, alpha, eps, rb.slot(1)), 1);
        h = js.read_scalars(st);
        const double at_new = h.gen[0] + h.gen[1];
        if (1.1 * at_old < at_new && 1 < k) {
            MI_HIP(hipMemcpyAsync(js.u.p, js.t.p, ib, hipMemcpyDeviceToDevice, st));
            reason = 2;
            break;
        }
        at_old = at_new;
        launch_axpy<T>(st, m.nrows, js.r.p, js.q.p, nullptr, -c1, -inf, inf, nullptr, nullptr);
        launch_bp_grad<T>(st, At, js.nx, js.ny, js.r.p, 1.0, js.u.p, js.v.p, alpha, eps, js.d.p, nullptr, lanes);
        to_gen(launch_fp<T>(st, FP_DOT, A, 0, m.nrows, js.d.p, nullptr, js.q.p, nullptr, rb.slot(6)), 6);
        h = js.read_scalars(st);
        const double c2 = -h.gen[6] / nap;
        launch_xpby<T>(st, js.npix, js.p.p, js.d.p, c2);
    }
    MI_HIP(hipStreamSynchronize(st));
    if (why) *why = reason;
    return ran;
}

// minimize_in_EDGE_variable (main.c:259-360) on js.v given js.u.
template <typename T>
void gc_edge_step(mi_context *ctx, JobState<T> &js, double alpha, double beta, double eps, int iter_edge) {
    const hipStream_t st = ctx->stream;
    launch_pass_begin(st, js.sc.p);
    for (int k = 0; k < iter_edge; ++k)
        launch_edge_iter<T>(st, js.rb(), js.nx, js.ny, js.u.p, js.v.p, js.d.p, alpha, beta, eps, js.sc.p);
    MI_HIP(hipStreamSynchronize(st));
}

template <typename T>
void load_sino(hipStream_t st, const Matrix &m, const T *host, T *rows) {
    DevBuf<T> user(m.nrows);
    MI_HIP(hipMemcpyAsync(user.p, host, m.nrows * sizeof(T), hipMemcpyHostToDevice, st));
    launch_gather_rows<T>(st, m.nrows, m.perm.p, user.p, rows);
    MI_HIP(hipStreamSynchronize(st));
}

template int gc_image_step<float>(mi_context *, const Matrix &, JobState<float> &, double, double, int, int *);
template int gc_image_step<double>(mi_context *, const Matrix &, JobState<double> &, double, double, int, int *);
template void gc_edge_step<float>(mi_context *, JobState<float> &, double, double, double, int);
template void gc_edge_step<double>(mi_context *, JobState<double> &, double, double, double, int);
template void load_sino<float>(hipStream_t, const Matrix &, const float *, float *);
template void load_sino<double>(hipStream_t, const Matrix &, const double *, double *);

// Times every launch shape of the two streaming kernels on subset 0 of the
// matrix (FP32, random data, HIP events) and keeps the fastest.
mi_tuning autotune(mi_context *ctx, Matrix &m) {
    const hipStream_t st = ctx->stream;
    DevBuf<float> x(m.npix), r(m.nrows), d(m.npix);
    fill<float>(st, x.p, m.npix, 0.5f);
    fill<float>(st, r.p, m.nrows, 0.25f);
    const int64_t nr = m.sub_row0[1];
    hipEvent_t a, b;
    MI_HIP(hipEventCreate(&a));
    MI_HIP(hipEventCreate(&b));
    auto time_us = [&](auto &&launch) {
        launch();  // warm
        MI_HIP(hipEventRecord(a, st));
        const int reps = 5;
        for (int i = 0; i < reps; ++i) launch();
        MI_HIP(hipEventRecord(b, st));
        MI_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        MI_HIP(hipEventElapsedTime(&ms, a, b));
        return 1e3 * ms / reps;
    };
    mi_tuning best{0, 4, 0, 0, 1e30, 1e30};
    const CsrView A = csr_view(m, false);
    const int grids[3] = {2048, 4096, 8192};
    for (int vec = 0; vec <= 1; ++vec)
        for (int gc : grids) {
            const double us = time_us([&] {
                launch_fp<float>(st, FP_SQ, A, 0, nr, x.p, nullptr, nullptr, nullptr, nullptr, nullptr, vec, gc);
            });
            if (us < best.fp_us) { best.fp_us = us; best.fp_vec = vec; best.fp_grid = gc; }
        }
    for (int lanes = 4; lanes <= 64; lanes *= 2)
        for (int gc : grids) {
            const double us = time_us([&] {
                launch_bp_grad<float>(st, csc_view(m, 0, false), m.g.nx, m.g.ny, r.p, 1.0, nullptr, nullptr, 0.0,
                                      0.0, d.p, nullptr, lanes, nullptr, gc);
            });
            if (us < best.bp_us) { best.bp_us = us; best.bp_lanes = lanes; best.bp_grid = gc; }
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    m.tune = best;
    return best;
}

template struct JobState<float>;
template struct JobState<double>;

}  // namespace mi
