// Projection and AT kernels (gfx950).  Memory-bound throughout: the forward
// projector streams the ray-major CSR once per launch, the back projector the
// subset's pixel-major CSC; image/sinogram gathers hit L2.  All sums are
// either Kahan per lane (FP32 runs) or plain FP64, then combined in double in
// a fixed order, so every result is deterministic.
#include "kernels.hpp"

#include <cfloat>
#include <cmath>

namespace mi {

namespace {

inline unsigned cap_grid(int64_t blocks) {
    if (blocks < 1) blocks = 1;
    return (unsigned)(blocks < kMaxGrid ? blocks : kMaxGrid);
}

__device__ __forceinline__ void slot_put(Slot s, double v) {
    if (!s.dst) return;
    if (s.accumulate) *s.dst += v;
    else *s.dst = v;
}

// Block partial in thread 0 -> grid total -> slot (+ optional extra action).
template <typename Fin>
__device__ __forceinline__ void reduce_to(const RedBuf &rb, double mine, double *lds, Fin fin) {
    double b = block_sum(mine, lds);
    GridRed g{rb.partial, rb.ticket};
    grid_finish(g, b, false, lds, fin);
}

// Two sums at once (partials interleaved).
template <typename Fin>
__device__ __forceinline__ void reduce2_to(const RedBuf &rb, double m0, double m1, double *lds, Fin fin) {
    double b0 = block_sum(m0, lds);
    double b1 = block_sum(m1, lds);
    __shared__ bool last;
    if (threadIdx.x == 0) {
        rb.partial[2 * blockIdx.x] = b0;
        rb.partial[2 * blockIdx.x + 1] = b1;
        __threadfence();
        unsigned t = atomicAdd(rb.ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    const volatile double *part = rb.partial;
    double a0 = 0.0, a1 = 0.0;
    for (unsigned i = threadIdx.x; i < gridDim.x; i += kBlock) {
        a0 += part[2 * i];
        a1 += part[2 * i + 1];
    }
    double t0 = block_sum(a0, lds);
    double t1 = block_sum(a1, lds);
    if (threadIdx.x == 0) {
        fin(t0, t1);
        *rb.ticket = 0u;
    }
}

template <typename T>
__device__ __forceinline__ T ld(const T *a, int i, int j, int nx, int ny, T pad) {
    return (i < 0 || j < 0 || i >= nx || j >= ny) ? pad : a[(int64_t)i * ny + j];
}

// ---------------------------------------------------------------------------
// Forward projection: one wave per ray, lanes stride the ray's entries.
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_fp(RedBuf rb, const int64_t *__restrict__ rowptr,
                                               const uint32_t *__restrict__ idx, const T *__restrict__ len,
                                               int64_t row0, int64_t nrows, const T *__restrict__ x,
                                               const T *__restrict__ g, const T *__restrict__ q,
                                               T *__restrict__ out, Slot slot, const int32_t *skip) {
    __shared__ double lds[kBlock / kWave];
    if (skip && *skip) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wid = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const int64_t nw = (int64_t)gridDim.x * (kBlock / kWave);
    double mine = 0.0;
    for (int64_t r = wid; r < nrows; r += nw) {
        const int64_t rr = row0 + r;
        const int64_t b = rowptr[rr], e = rowptr[rr + 1];
        Acc<T> acc;
        int64_t k = b + lane;
        // Two independent entries per lane per trip keep more loads in flight.
        for (; k + kWave < e; k += 2 * kWave) {
            const uint32_t i0 = idx[k], i1 = idx[k + kWave];
            const T l0 = len[k], l1 = len[k + kWave];
            const T x0 = x[i0], x1 = x[i1];
            acc.add(x0 * l0);
            acc.add(x1 * l1);
        }
        if (k < e) acc.add(x[idx[k]] * len[k]);
        const double s = group_sum<kWave>(acc.value());
        if (lane == 0) {
            if (MODE == FP_WRITE) {
                const T o = (T)s;
                out[rr] = o;
                mine += (double)o * (double)o;
            } else if (MODE == FP_RESID) {
                const T o = (T)((double)g[rr] - s);
                out[rr] = o;
                mine += (double)o * (double)o;
            } else if (MODE == FP_SQ) {
                mine += s * s;
            } else if (MODE == FP_DOT) {
                mine += s * (double)q[rr];
            } else {
                const T o = (T)s;
                out[rr] = o;
                const double dv = (double)o - (double)g[rr];
                mine += dv * dv;
            }
        }
    }
    reduce_to(rb, mine, lds, [&](double t) { slot_put(slot, t); });
}

// GoldenC image-gradient regulariser at pixel (i, j): f zero padded, v one
// padded, origin/GoldenC/src/main.c:112-125.
template <typename T>
__device__ __forceinline__ double gc_reg(const T *u, const T *v, int i, int j, int nx, int ny, double eps) {
    const double z = (double)u[(int64_t)i * ny + j];
    const double zd = (double)ld(u, i + 1, j, nx, ny, T(0)), zr = (double)ld(u, i, j + 1, nx, ny, T(0));
    const double zu = (double)ld(u, i - 1, j, nx, ny, T(0)), zl = (double)ld(u, i, j - 1, nx, ny, T(0));
    const double p = (double)v[(int64_t)i * ny + j];
    const double pu = (double)ld(v, i - 1, j, nx, ny, T(1)), pl = (double)ld(v, i, j - 1, nx, ny, T(1));
    return +(p * p) * (zd - z) + (p * p) * (zr - z) - (pu * pu) * (z - zu) - (pl * pl) * (z - zl) +
           0.01 * (eps * eps) * (zd + zr - 4 * z + zu + zl);
}

// Gather back projection: LPC lanes per pixel column of the subset's CSC,
// then the fused gradient d = scale * b + alpha * R(u, v).
template <typename T, int LPC>
__global__ __launch_bounds__(kBlock) void k_bp_grad(RedBuf rb, const int64_t *__restrict__ colptr,
                                                    const uint32_t *__restrict__ crow, const T *__restrict__ clen,
                                                    int nx, int ny, const T *__restrict__ r, double scale,
                                                    const T *__restrict__ u, const T *__restrict__ v, double alpha,
                                                    double eps, T *__restrict__ d, Slot slot,
                                                    const int32_t *skip) {
    __shared__ double lds[kBlock / kWave];
    if (skip && *skip) return;
    constexpr int G = kBlock / LPC;
    const int gl = threadIdx.x % LPC, grp = threadIdx.x / LPC;
    const int64_t npix = (int64_t)nx * ny;
    double mine = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * G + grp; p < npix; p += (int64_t)gridDim.x * G) {
        const int64_t b = colptr[p], e = colptr[p + 1];
        Acc<T> acc;
        int64_t k = b + gl;
        for (; k + LPC < e; k += 2 * LPC) {
            const uint32_t r0 = crow[k], r1 = crow[k + LPC];
            const T l0 = clen[k], l1 = clen[k + LPC];
            const T a0 = r[r0], a1 = r[r1];
            acc.add(a0 * l0);
            acc.add(a1 * l1);
        }
        if (k < e) acc.add(r[crow[k]] * clen[k]);
        const double s = group_sum<LPC>(acc.value());
        if (gl == 0) {
            double dv = scale * s;
            if (u) dv += alpha * gc_reg(u, v, (int)(p / ny), (int)(p % ny), nx, ny, eps);
            const T dt = (T)dv;
            d[p] = dt;
            mine += (double)dt * (double)dt;
        }
    }
    reduce_to(rb, mine, lds, [&](double t) { slot_put(slot, t); });
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_nv(RedBuf rb, int nx, int ny, const T *__restrict__ p,
                                               const T *__restrict__ v, double eps, Slot slot, Scalars *sc, double S,
                                               double alpha, const int32_t *skip) {
    __shared__ double lds[kBlock / kWave];
    if (skip && *skip) return;
    const int64_t npix = (int64_t)nx * ny;
    double mine = 0.0;
    const double e2 = 0.01 * (eps * eps);
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npix; q += (int64_t)gridDim.x * kBlock) {
        const int i = (int)(q / ny), j = (int)(q % ny);
        const double z = (double)p[q];
        const double a = z - (double)ld(p, i - 1, j, nx, ny, T(0));
        const double b = z - (double)ld(p, i, j - 1, nx, ny, T(0));
        const double w = (double)v[q];
        mine += (w * w + e2) * (a * a + b * b);
    }
    reduce_to(rb, mine, lds, [&](double t) {
        slot_put(slot, t);
        if (sc) {
            const double den = S * sc->ad2 + alpha * t;
            sc->nv = t;
            sc->c = (den > 0) ? sc->dd / den : 0.0;
        }
    });
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_energy_reg(RedBuf rb, int nx, int ny, const T *__restrict__ f,
                                                       const T *__restrict__ v, double alpha, double eps, Slot slot) {
    __shared__ double lds[kBlock / kWave];
    const int64_t npix = (int64_t)nx * ny;
    double mine = 0.0;
    const double e2 = 0.01 * (eps * eps);
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npix; q += (int64_t)gridDim.x * kBlock) {
        const int i = (int)(q / ny), j = (int)(q % ny);
        if (i < 1 || j < 1 || i > nx - 2 || j > ny - 2) continue;
        const double z = (double)f[q];
        const double a = z - (double)f[q - ny], b = z - (double)f[q - 1];
        const double w = (double)v[q];
        mine += alpha * (w * w + e2) * (a * a + b * b);
    }
    reduce_to(rb, mine, lds, [&](double t) { slot_put(slot, t); });
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(RedBuf rb, int64_t n, const T *__restrict__ a,
                                                const T *__restrict__ b, Slot slot) {
    __shared__ double lds[kBlock / kWave];
    double mine = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        mine += (double)a[i] * (double)b[i];
    reduce_to(rb, mine, lds, [&](double t) { slot_put(slot, t); });
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy(RedBuf rb, int64_t n, T *__restrict__ y, const T *__restrict__ x,
                                                 const double *cptr, double cval, double lo, double hi,
                                                 T *__restrict__ save, Slot slot, const int32_t *skip) {
    __shared__ double lds[kBlock / kWave];
    if (skip && *skip) return;
    const double c = cptr ? *cptr : cval;
    double mx = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double yo = (double)y[i];
        if (save) save[i] = y[i];
        const double cx = c * (double)x[i];
        double yn = yo + cx;
        if (yn < lo) yn = lo;
        if (yn > hi) yn = hi;
        y[i] = (T)yn;
        mx = fmax(mx, fabs(cx));
    }
    if (!slot.dst) return;
    double bm = block_max(mx, lds);
    GridRed g{rb.partial, rb.ticket};
    grid_finish(g, bm, true, lds, [&](double t) { slot_put(slot, t); });
}

template <typename T>
__global__ void k_xpby(int64_t n, T *__restrict__ y, const T *__restrict__ x, double c) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        y[i] = (T)((double)x[i] + c * (double)y[i]);
}

// Exact median of 9 by a min/max exchange network (any exact selection
// equals the reference's sort result).
template <typename T>
__device__ __forceinline__ void cswap(T &a, T &b) {
    T lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo; b = hi;
}

template <typename T>
__global__ void k_median(int nx, int ny, const T *__restrict__ in, T *__restrict__ out) {
    const int64_t npix = (int64_t)nx * ny;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npix; q += (int64_t)gridDim.x * kBlock) {
        const int i = (int)(q / ny), j = (int)(q % ny);
        T t[9];
        int k = 0;
        for (int di = -1; di <= 1; ++di)
            for (int dj = -1; dj <= 1; ++dj) t[k++] = ld(in, i + di, j + dj, nx, ny, T(0));
        // Paeth's 19-exchange median-of-9 network.
        cswap(t[1], t[2]); cswap(t[4], t[5]); cswap(t[7], t[8]);
        cswap(t[0], t[1]); cswap(t[3], t[4]); cswap(t[6], t[7]);
        cswap(t[1], t[2]); cswap(t[4], t[5]); cswap(t[7], t[8]);
        cswap(t[0], t[3]); cswap(t[5], t[8]); cswap(t[4], t[7]);
        cswap(t[3], t[6]); cswap(t[1], t[4]); cswap(t[2], t[5]);
        cswap(t[4], t[7]); cswap(t[4], t[2]); cswap(t[6], t[4]);
        cswap(t[4], t[2]);
        out[q] = t[4];
    }
}

// Edge step, origin/GoldenC/src/main.c:294-305: gradient and the two sums.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_edge_grad(RedBuf rb, int nx, int ny, const T *__restrict__ f,
                                                      const T *__restrict__ v, T *__restrict__ d, double alpha,
                                                      double beta, double eps, Scalars *sc) {
    __shared__ double lds[kBlock / kWave];
    if (sc->edge_done || sc->converged) return;
    const int64_t npix = (int64_t)nx * ny;
    double m0 = 0.0, m1 = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npix; q += (int64_t)gridDim.x * kBlock) {
        const int i = (int)(q / ny), j = (int)(q % ny);
        const double dxf = 0.5 * ((double)ld(f, i + 1, j, nx, ny, T(0)) - (double)ld(f, i - 1, j, nx, ny, T(0)));
        const double dyf = 0.5 * ((double)ld(f, i, j + 1, nx, ny, T(0)) - (double)ld(f, i, j - 1, nx, ny, T(0)));
        const double n = dxf * dxf + dyf * dyf;
        const double t = (double)v[q];
        const double lap = 0.25 * ((double)ld(v, i + 1, j, nx, ny, T(1)) + (double)ld(v, i, j + 1, nx, ny, T(1)) -
                                   4 * t + (double)ld(v, i - 1, j, nx, ny, T(1)) +
                                   (double)ld(v, i, j - 1, nx, ny, T(1)));
        const double gr = alpha * t * n - beta * eps * lap + beta * (t - 1) / (4 * eps);
        const T dt = (T)(-gr);
        d[q] = dt;
        const double dd = (double)dt * (double)dt;
        m0 += dd;
        m1 += n * dd;
    }
    reduce2_to(rb, m0, m1, lds, [&](double t0, double t1) {
        sc->e_nd = t0;
        sc->e_nd2 = t1;
    });
}

// main.c:311-329: |central gradient of d|^2 and the step size.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_edge_nabla(RedBuf rb, int nx, int ny, const T *__restrict__ d,
                                                       double alpha, double beta, double eps, Scalars *sc) {
    __shared__ double lds[kBlock / kWave];
    if (sc->edge_done || sc->converged) return;
    const int64_t npix = (int64_t)nx * ny;
    double mine = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npix; q += (int64_t)gridDim.x * kBlock) {
        const int i = (int)(q / ny), j = (int)(q % ny);
        const double a = 0.5 * ((double)ld(d, i + 1, j, nx, ny, T(0)) - (double)ld(d, i - 1, j, nx, ny, T(0)));
        const double b = 0.5 * ((double)ld(d, i, j + 1, nx, ny, T(0)) - (double)ld(d, i, j - 1, nx, ny, T(0)));
        mine += a * a + b * b;
    }
    reduce_to(rb, mine, lds, [&](double t) {
        sc->e_nnd = t;
        sc->e_c = sc->e_nd / (alpha * sc->e_nd2 + beta * eps * t + (beta / (4 * eps)) * sc->e_nd);
    });
}

// main.c:332-354: update, clamp to [0, 1], early-stop flag.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_edge_update(RedBuf rb, int nx, int ny, T *__restrict__ v,
                                                        const T *__restrict__ d, Scalars *sc) {
    __shared__ double lds[kBlock / kWave];
    if (sc->edge_done || sc->converged) return;
    const int64_t npix = (int64_t)nx * ny;
    const double c = sc->e_c;
    double mx = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npix; q += (int64_t)gridDim.x * kBlock) {
        const double cd = c * (double)d[q];
        double x = (double)v[q] + cd;
        mx = fmax(mx, fabs(cd));
        if (x > 1) x = 1;
        if (x < 0) x = 0;
        v[q] = (T)x;
    }
    double bm = block_max(mx, lds);
    GridRed g{rb.partial, rb.ticket};
    grid_finish(g, bm, true, lds, [&](double t) {
        sc->e_maxcd = t;
        if (t < 1e-5) sc->edge_done = 1;
    });
}

__global__ void k_pass_begin(Scalars *sc) {
    if (sc->converged) return;
    sc->df_pass = 0.0;
    sc->edge_done = 0;
}

__global__ void k_pass_end(Scalars *sc, double rel_tol) {
    if (sc->converged) return;
    const double df = sc->df_pass;
    if (sc->iters > 0) {
        const double diff = fabs(df - sc->df_prev);
        sc->rel = sc->df_prev > 0 ? diff / sc->df_prev : 0.0;
        if (diff < rel_tol * sc->df_prev) sc->converged = 1;
    }
    sc->df_prev = df;
    sc->iters += 1;
}

template <typename T>
__global__ void k_gather_rows(int64_t n, const int64_t *__restrict__ perm, const T *__restrict__ in,
                              T *__restrict__ out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = in[perm[r]];
}

template <typename T>
__global__ void k_scatter_rows(int64_t n, const int64_t *__restrict__ perm, const T *__restrict__ in,
                               T *__restrict__ out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[perm[r]] = in[r];
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.

template <typename T>
void launch_fp(hipStream_t st, const RedBuf &rb, int mode, const CsrView &A, int64_t row0, int64_t nrows,
               const T *x, const T *g, const T *q, T *out, Slot slot, const int32_t *skip) {
    const unsigned grid = cap_grid((nrows + 3) / 4);
    const T *len = (const T *)A.len;
#define MI_FP_CASE(M) \
    case M: k_fp<T, M><<<grid, kBlock, 0, st>>>(rb, A.rowptr, A.idx, len, row0, nrows, x, g, q, out, slot, skip); break;
    switch (mode) {
        MI_FP_CASE(FP_WRITE)
        MI_FP_CASE(FP_RESID)
        MI_FP_CASE(FP_SQ)
        MI_FP_CASE(FP_DOT)
        MI_FP_CASE(FP_ERR)
        default: fail(MI_E_ARG, "bad projection mode");
    }
#undef MI_FP_CASE
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_bp_grad(hipStream_t st, const RedBuf &rb, const CscView &At, int nx, int ny, const T *r, double scale,
                    const T *u, const T *v, double alpha, double eps, T *d, Slot slot, int lanes,
                    const int32_t *skip) {
    const int64_t npix = (int64_t)nx * ny;
    const T *len = (const T *)At.len;
#define MI_BP_CASE(L)                                                                                          \
    case L: {                                                                                                  \
        const unsigned grid = cap_grid((npix + kBlock / L - 1) / (kBlock / L));                                \
        k_bp_grad<T, L><<<grid, kBlock, 0, st>>>(rb, At.colptr, At.crow, len, nx, ny, r, scale, u, v, alpha,   \
                                                 eps, d, slot, skip);                                          \
        break;                                                                                                 \
    }
    switch (lanes) {
        MI_BP_CASE(4)
        MI_BP_CASE(8)
        MI_BP_CASE(16)
        MI_BP_CASE(32)
        MI_BP_CASE(64)
        default: fail(MI_E_ARG, "bad lanes-per-pixel");
    }
#undef MI_BP_CASE
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_nv(hipStream_t st, const RedBuf &rb, int nx, int ny, const T *p, const T *v, double eps, Slot slot,
               Scalars *sc, double S, double alpha, const int32_t *skip) {
    const int64_t npix = (int64_t)nx * ny;
    k_nv<T><<<cap_grid((npix + kBlock - 1) / kBlock), kBlock, 0, st>>>(rb, nx, ny, p, v, eps, slot, sc, S, alpha,
                                                                       skip);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_energy_reg(hipStream_t st, const RedBuf &rb, int nx, int ny, const T *f, const T *v, double alpha,
                       double eps, Slot slot) {
    const int64_t npix = (int64_t)nx * ny;
    k_energy_reg<T><<<cap_grid((npix + kBlock - 1) / kBlock), kBlock, 0, st>>>(rb, nx, ny, f, v, alpha, eps, slot);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_dot(hipStream_t st, const RedBuf &rb, int64_t n, const T *a, const T *b, Slot slot) {
    k_dot<T><<<cap_grid((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(rb, n, a, b, slot);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_axpy(hipStream_t st, const RedBuf &rb, int64_t n, T *y, const T *x, const double *cptr, double cval,
                 double lo, double hi, T *save, Slot slot, const int32_t *skip) {
    k_axpy<T><<<cap_grid((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(rb, n, y, x, cptr, cval, lo, hi, save, slot,
                                                                      skip);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_xpby(hipStream_t st, int64_t n, T *y, const T *x, double c) {
    k_xpby<T><<<cap_grid((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(n, y, x, c);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_median(hipStream_t st, int nx, int ny, const T *in, T *out) {
    const int64_t npix = (int64_t)nx * ny;
    k_median<T><<<cap_grid((npix + kBlock - 1) / kBlock), kBlock, 0, st>>>(nx, ny, in, out);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_edge_iter(hipStream_t st, const RedBuf &rb, int nx, int ny, const T *f, T *v, T *dwork, double alpha,
                      double beta, double eps, Scalars *sc) {
    const int64_t npix = (int64_t)nx * ny;
    const unsigned grid = cap_grid((npix + kBlock - 1) / kBlock);
    k_edge_grad<T><<<grid, kBlock, 0, st>>>(rb, nx, ny, f, v, dwork, alpha, beta, eps, sc);
    MI_HIP(hipGetLastError());
    k_edge_nabla<T><<<grid, kBlock, 0, st>>>(rb, nx, ny, dwork, alpha, beta, eps, sc);
    MI_HIP(hipGetLastError());
    k_edge_update<T><<<grid, kBlock, 0, st>>>(rb, nx, ny, v, dwork, sc);
    MI_HIP(hipGetLastError());
}

void launch_pass_begin(hipStream_t st, Scalars *sc) {
    k_pass_begin<<<1, 1, 0, st>>>(sc);
    MI_HIP(hipGetLastError());
}

void launch_pass_end(hipStream_t st, Scalars *sc, double rel_tol) {
    k_pass_end<<<1, 1, 0, st>>>(sc, rel_tol);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_gather_rows(hipStream_t st, int64_t n, const int64_t *perm, const T *in, T *out) {
    k_gather_rows<T><<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(n, perm, in, out);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_scatter_rows(hipStream_t st, int64_t n, const int64_t *perm, const T *in, T *out) {
    k_scatter_rows<T><<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(n, perm, in, out);
    MI_HIP(hipGetLastError());
}

#define MI_INST(T)                                                                                              \
    template void launch_fp<T>(hipStream_t, const RedBuf &, int, const CsrView &, int64_t, int64_t, const T *,  \
                               const T *, const T *, T *, Slot, const int32_t *);                               \
    template void launch_bp_grad<T>(hipStream_t, const RedBuf &, const CscView &, int, int, const T *, double,  \
                                    const T *, const T *, double, double, T *, Slot, int, const int32_t *);     \
    template void launch_nv<T>(hipStream_t, const RedBuf &, int, int, const T *, const T *, double, Slot,       \
                               Scalars *, double, double, const int32_t *);                                     \
    template void launch_energy_reg<T>(hipStream_t, const RedBuf &, int, int, const T *, const T *, double,     \
                                       double, Slot);                                                           \
    template void launch_dot<T>(hipStream_t, const RedBuf &, int64_t, const T *, const T *, Slot);              \
    template void launch_axpy<T>(hipStream_t, const RedBuf &, int64_t, T *, const T *, const double *, double,  \
                                 double, double, T *, Slot, const int32_t *);                                   \
    template void launch_xpby<T>(hipStream_t, int64_t, T *, const T *, double);                                 \
    template void launch_median<T>(hipStream_t, int, int, const T *, T *);                                      \
    template void launch_edge_iter<T>(hipStream_t, const RedBuf &, int, int, const T *, T *, T *, double,       \
                                      double, double, Scalars *);                                               \
    template void launch_gather_rows<T>(hipStream_t, int64_t, const int64_t *, const T *, T *);                 \
    template void launch_scatter_rows<T>(hipStream_t, int64_t, const int64_t *, const T *, T *);

MI_INST(float)
MI_INST(double)

}  // namespace mi
