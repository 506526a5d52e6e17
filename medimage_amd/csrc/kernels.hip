// Projection and AT kernels (gfx950).  Memory-bound throughout: the forward
// projector streams the ray-major CSR once per launch, the back projector the
// subset's pixel-major CSC; image/sinogram gathers hit L2.  All sums are
// either Kahan per lane (FP32 runs) or plain FP64, then combined in double in
// a fixed order, so every result is deterministic.
#include "kernels.hpp"
#include "gc_stencil.hpp"

#include <cfloat>
#include <cmath>

namespace mi {

namespace {

// Grids of the streaming kernels: at most kStreamGrid blocks of 4 waves (one
// full residency of the chip at 8 waves per SIMD), rounded to a multiple of
// the 8 XCDs so that xcd_block() is a permutation.
constexpr int kStreamGrid = 2048;
inline unsigned cap_grid(int64_t blocks, int cap = kStreamGrid) {
    if (blocks < 1) blocks = 1;
    if (blocks > cap) blocks = cap;
    if (blocks > 8) blocks = (blocks + 7) / 8 * 8;
    return (unsigned)blocks;
}

// Block partial: thread 0 of the block stores the block's sum (or max) at
// part[blockIdx.x].  No cross-block traffic inside the kernel; a one-block
// finalize kernel sums the partials in index order after the kernel boundary.
__device__ __forceinline__ void put_partial(double *part, double mine, double *lds) {
    const double b = block_sum(mine, lds);
    if (part && threadIdx.x == 0) part[blockIdx.x] = b;
}

__device__ __forceinline__ void put_partial_max(double *part, double mine, double *lds) {
    const double b = block_max(mine, lds);
    if (part && threadIdx.x == 0) part[blockIdx.x] = b;
}


// ---------------------------------------------------------------------------
// Forward projection: one wave per ray.  Each lane takes four consecutive
// entries per trip with 16-byte loads (the matrix arrays are padded and the
// walk starts at the 4-aligned entry below the row start; entries outside
// the row are loaded but weighted by zero), so a trip covers 256 entries and
// issues four independent image gathers per lane.
template <typename T, int MODE, int VEC>
__global__ __launch_bounds__(kBlock) void k_fp(const int64_t *__restrict__ rowptr,
                                               const uint32_t *__restrict__ idx, const T *__restrict__ len,
                                               int64_t row0, int64_t nrows, const T *__restrict__ x,
                                               const T *__restrict__ g, const T *__restrict__ q,
                                               T *__restrict__ out, double *part, const int32_t *skip) {
    __shared__ double lds[kBlock / kWave];
    if (skip && *skip) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wid = (int64_t)xcd_block() * (kBlock / kWave) + threadIdx.x / kWave;
    const int64_t nw = (int64_t)gridDim.x * (kBlock / kWave);
    double mine = 0.0;
    for (int64_t r = wid; r < nrows; r += nw) {
        const int64_t rr = row0 + r;
        const int64_t b = rowptr[rr], e = rowptr[rr + 1];
        Acc<T> acc;
        if (VEC) {
            for (int64_t k = (b & ~(int64_t)3) + 4 * lane; k < e; k += 4 * kWave) {
                uint32_t i4[4];
                Vec4<T> l4;
                load_idx4(idx + k, i4);
                l4.load(len + k);
                T xv[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) xv[j] = x[i4[j]];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (k + j >= b && k + j < e) acc.add(xv[j] * l4.v[j]);
            }
        } else {
            // Lane-strided: every load instruction is one contiguous 256-byte
            // (or 512-byte) segment and every gather instruction covers 64
            // consecutive pixels of the ray; four entries per lane in flight.
            // Entries past the row end read the padded tail or the next row
            // (valid indices) and are weighted by zero.
            for (int64_t k = b + lane; k < e; k += 4 * kWave) {
                uint32_t i4[4];
                T l4[4], xv[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int64_t kk = (k + j * kWave < e) ? k + j * kWave : k;
                    i4[j] = idx[kk];
                    l4[j] = len[kk];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) xv[j] = x[i4[j]];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (k + j * kWave < e) acc.add(xv[j] * l4[j]);
            }
        }
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
    put_partial(part, mine, lds);
}

// Gather back projection: LPC lanes per pixel column of the subset's CSC,
// four consecutive entries per lane per trip (16-byte loads, padded arrays),
// then the fused gradient d = scale * b + alpha * R(u, v).
template <typename T, int LPC>
__global__ __launch_bounds__(kBlock) void k_bp_grad(const int64_t *__restrict__ colptr,
                                                    const uint32_t *__restrict__ crow, const T *__restrict__ clen,
                                                    int nx, int ny, const T *__restrict__ r, double scale,
                                                    const T *__restrict__ u, const T *__restrict__ v, double alpha,
                                                    double eps, T *__restrict__ d, double *part,
                                                    const int32_t *skip) {
    __shared__ double lds[kBlock / kWave];
    if (skip && *skip) return;
    constexpr int G = kBlock / LPC;
    const int gl = threadIdx.x % LPC, grp = threadIdx.x / LPC;
    const int64_t npix = (int64_t)nx * ny;
    double mine = 0.0;
    for (int64_t p = (int64_t)xcd_block() * G + grp; p < npix; p += (int64_t)gridDim.x * G) {
        const int64_t b = colptr[p], e = colptr[p + 1];
        Acc<T> acc;
        for (int64_t k = (b & ~(int64_t)3) + 4 * gl; k < e; k += 4 * LPC) {
            uint32_t r4[4];
            Vec4<T> l4;
            load_idx4(crow + k, r4);
            l4.load(clen + k);
            T rv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) rv[j] = r[r4[j]];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (k + j >= b && k + j < e) acc.add(rv[j] * l4.v[j]);
        }
        const double s = group_sum<LPC>(acc.value());
        if (gl == 0) {
            double dv = scale * s;
            if (u) dv += alpha * gc_reg(u, v, (int)(p / ny), (int)(p % ny), nx, ny, eps);
            const T dt = (T)dv;
            d[p] = dt;
            mine += (double)dt * (double)dt;
        }
    }
    put_partial(part, mine, lds);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_nv(int nx, int ny, const T *__restrict__ p,
                                               const T *__restrict__ v, double eps, double *part,
                                               const int32_t *skip) {
    __shared__ double lds[kBlock / kWave];
    if (skip && *skip) return;
    const int64_t npix = (int64_t)nx * ny;
    double mine = 0.0;
    const double e2 = 0.01 * (eps * eps);
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npix; q += (int64_t)gridDim.x * kBlock) {
        const int i = (int)(q / ny), j = (int)(q % ny);
        const double z = (double)p[q];
        const double a = z - (double)gc_ld(p, i - 1, j, nx, ny, T(0));
        const double b = z - (double)gc_ld(p, i, j - 1, nx, ny, T(0));
        const double w = (double)v[q];
        mine += (w * w + e2) * (a * a + b * b);
    }
    put_partial(part, mine, lds);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_energy_reg(int nx, int ny, const T *__restrict__ f,
                                                       const T *__restrict__ v, double alpha, double eps,
                                                       double *part) {
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
    put_partial(part, mine, lds);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, const T *__restrict__ a, const T *__restrict__ b,
                                                double *part) {
    __shared__ double lds[kBlock / kWave];
    double mine = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        mine += (double)a[i] * (double)b[i];
    put_partial(part, mine, lds);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpy(int64_t n, T *__restrict__ y, const T *__restrict__ x,
                                                 const double *cptr, double cval, double lo, double hi,
                                                 T *__restrict__ save, double *part, const int32_t *skip) {
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
    if (part) put_partial_max(part, mx, lds);
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
            for (int dj = -1; dj <= 1; ++dj) t[k++] = gc_ld(in, i + di, j + dj, nx, ny, T(0));
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
__global__ __launch_bounds__(kBlock) void k_edge_grad(int nx, int ny, const T *__restrict__ f,
                                                      const T *__restrict__ v, T *__restrict__ d,
                                                      T *__restrict__ nout, double alpha, double beta, double eps,
                                                      const Scalars *sc, double *part0, double *part1) {
    __shared__ double lds[kBlock / kWave];
    if (sc->edge_done || sc->converged) return;
    const int64_t npix = (int64_t)nx * ny;
    double m0 = 0.0, m1 = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npix; q += (int64_t)gridDim.x * kBlock) {
        const int i = (int)(q / ny), j = (int)(q % ny);
        const double dxf = 0.5 * ((double)gc_ld(f, i + 1, j, nx, ny, T(0)) - (double)gc_ld(f, i - 1, j, nx, ny, T(0)));
        const double dyf = 0.5 * ((double)gc_ld(f, i, j + 1, nx, ny, T(0)) - (double)gc_ld(f, i, j - 1, nx, ny, T(0)));
        const double n = dxf * dxf + dyf * dyf;
        const double t = (double)v[q];
        const double lap = 0.25 * ((double)gc_ld(v, i + 1, j, nx, ny, T(1)) + (double)gc_ld(v, i, j + 1, nx, ny, T(1)) -
                                   4 * t + (double)gc_ld(v, i - 1, j, nx, ny, T(1)) +
                                   (double)gc_ld(v, i, j - 1, nx, ny, T(1)));
        const double gr = alpha * t * n - beta * eps * lap + beta * (t - 1) / (4 * eps);
        const T dt = (T)(-gr);
        d[q] = dt;
        if (nout) nout[q] = (T)n;
        const double dd = (double)dt * (double)dt;
        m0 += dd;
        m1 += n * dd;
    }
    put_partial(part0, m0, lds);
    put_partial(part1, m1, lds);
}

// main.c:311-329: |central gradient of d|^2 and the step size.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_edge_nabla(int nx, int ny, const T *__restrict__ d, const Scalars *sc,
                                                       double *part) {
    __shared__ double lds[kBlock / kWave];
    if (sc->edge_done || sc->converged) return;
    const int64_t npix = (int64_t)nx * ny;
    double mine = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npix; q += (int64_t)gridDim.x * kBlock) {
        const int i = (int)(q / ny), j = (int)(q % ny);
        const double a = 0.5 * ((double)gc_ld(d, i + 1, j, nx, ny, T(0)) - (double)gc_ld(d, i - 1, j, nx, ny, T(0)));
        const double b = 0.5 * ((double)gc_ld(d, i, j + 1, nx, ny, T(0)) - (double)gc_ld(d, i, j - 1, nx, ny, T(0)));
        mine += a * a + b * b;
    }
    put_partial(part, mine, lds);
}

// main.c:332-354: update, clamp to [0, 1], early-stop flag.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_edge_update(int nx, int ny, T *__restrict__ v, const T *__restrict__ d,
                                                        const Scalars *sc, double *part) {
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
    put_partial_max(part, mx, lds);
}

// One block: reduces partial arrays in index order and applies `op`.
constexpr int kFinBlock = 1024;
__global__ __launch_bounds__(kFinBlock) void k_finalize(FinArgs a) {
    __shared__ double lds[kFinBlock / kWave];
    Scalars *sc = a.sc;
    if (a.skip && *a.skip) return;
    if ((a.op == FIN_EDGE_C || a.op == FIN_EDGE_DONE) && (sc->edge_done || sc->converged)) return;
    double tot[4] = {0, 0, 0, 0};
    for (int k = 0; k < a.nin; ++k) {
        const bool mx = a.is_max[k];
        const int n = a.n[k];
        // Eight independent loads per thread before any add.
        double vals[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = threadIdx.x + u * kFinBlock;
            vals[u] = (i < n) ? a.in[k][i] : 0.0;
        }
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = mx ? fmax(acc, vals[u]) : acc + vals[u];
        // Fixed-order tree: wave butterfly, then the waves in order.
        acc = mx ? wave_max(acc) : group_sum<kWave>(acc);
        const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
        if (lane == 0) lds[w] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int i = 0; i < kFinBlock / kWave; ++i) t = mx ? fmax(t, lds[i]) : t + lds[i];
            tot[k] = t;
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    switch (a.op) {
        case FIN_STORE:
            *a.dst = tot[0];
            break;
        case FIN_ACCUM:
            *a.dst += tot[0];
            break;
        case FIN_STEP: {
            // in: <d,d>, ||A_s d||^2, nV(d), and the subset's ||r||^2
            sc->df_pass += tot[3];
            sc->dd = tot[0];
            sc->ad2 = tot[1];
            sc->nv = tot[2];
            const double den = a.S * tot[1] + a.alpha * tot[2];
            sc->c = (den > 0) ? tot[0] / den : 0.0;
            break;
        }
        case FIN_EDGE_C:
            // in: ||d||^2, sum n d^2, |grad d|^2 (origin/GoldenC/src/main.c:329)
            sc->e_nd = tot[0];
            sc->e_nd2 = tot[1];
            sc->e_nnd = tot[2];
            sc->e_c = tot[0] / (a.alpha * tot[1] + a.beta * a.eps * tot[2] + (a.beta / (4 * a.eps)) * tot[0]);
            break;
        case FIN_EDGE_DONE:
            sc->e_maxcd = tot[0];
            if (tot[0] < 1e-5) sc->edge_done = 1;
            break;
    }
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

unsigned grid_rows(int64_t nrows, int cap) { return cap_grid((nrows + 3) / 4, cap ? cap : kStreamGrid); }
unsigned grid_pixels(int64_t n) { return cap_grid((n + kBlock - 1) / kBlock); }
unsigned grid_bp(int64_t npix, int lanes, int cap) {
    return cap_grid((npix + kBlock / lanes - 1) / (kBlock / lanes), cap ? cap : kStreamGrid);
}

template <typename T>
unsigned launch_fp(hipStream_t st, int mode, const CsrView &A, int64_t row0, int64_t nrows, const T *x, const T *g,
                   const T *q, T *out, double *part, const int32_t *skip, int vec, int grid_cap) {
    const unsigned grid = grid_rows(nrows, grid_cap);
    const T *len = (const T *)A.len;
#define MI_FP_CASE(M)                                                                                          \
    case M:                                                                                                    \
        if (vec) k_fp<T, M, 1><<<grid, kBlock, 0, st>>>(A.rowptr, A.idx, len, row0, nrows, x, g, q, out, part, skip); \
        else k_fp<T, M, 0><<<grid, kBlock, 0, st>>>(A.rowptr, A.idx, len, row0, nrows, x, g, q, out, part, skip); \
        break;
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
    return grid;
}

template <typename T>
unsigned launch_bp_grad(hipStream_t st, const CscView &At, int nx, int ny, const T *r, double scale, const T *u,
                        const T *v, double alpha, double eps, T *d, double *part, int lanes, const int32_t *skip,
                        int grid_cap) {
    const int64_t npix = (int64_t)nx * ny;
    const T *len = (const T *)At.len;
    const unsigned grid = grid_bp(npix, lanes, grid_cap);
#define MI_BP_CASE(L)                                                                                          \
    case L:                                                                                                    \
        k_bp_grad<T, L><<<grid, kBlock, 0, st>>>(At.colptr, At.crow, len, nx, ny, r, scale, u, v, alpha, eps, d, \
                                                 part, skip);                                                  \
        break;
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
    return grid;
}

template <typename T>
unsigned launch_nv(hipStream_t st, int nx, int ny, const T *p, const T *v, double eps, double *part,
                   const int32_t *skip) {
    const unsigned grid = grid_pixels((int64_t)nx * ny);
    k_nv<T><<<grid, kBlock, 0, st>>>(nx, ny, p, v, eps, part, skip);
    MI_HIP(hipGetLastError());
    return grid;
}

template <typename T>
unsigned launch_energy_reg(hipStream_t st, int nx, int ny, const T *f, const T *v, double alpha, double eps,
                           double *part) {
    const unsigned grid = grid_pixels((int64_t)nx * ny);
    k_energy_reg<T><<<grid, kBlock, 0, st>>>(nx, ny, f, v, alpha, eps, part);
    MI_HIP(hipGetLastError());
    return grid;
}

template <typename T>
unsigned launch_dot(hipStream_t st, int64_t n, const T *a, const T *b, double *part) {
    const unsigned grid = grid_pixels(n);
    k_dot<T><<<grid, kBlock, 0, st>>>(n, a, b, part);
    MI_HIP(hipGetLastError());
    return grid;
}

template <typename T>
unsigned launch_axpy(hipStream_t st, int64_t n, T *y, const T *x, const double *cptr, double cval, double lo,
                     double hi, T *save, double *part, const int32_t *skip) {
    const unsigned grid = grid_pixels(n);
    k_axpy<T><<<grid, kBlock, 0, st>>>(n, y, x, cptr, cval, lo, hi, save, part, skip);
    MI_HIP(hipGetLastError());
    return grid;
}

template <typename T>
void launch_xpby(hipStream_t st, int64_t n, T *y, const T *x, double c) {
    k_xpby<T><<<grid_pixels(n), kBlock, 0, st>>>(n, y, x, c);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_median(hipStream_t st, int nx, int ny, const T *in, T *out) {
    k_median<T><<<grid_pixels((int64_t)nx * ny), kBlock, 0, st>>>(nx, ny, in, out);
    MI_HIP(hipGetLastError());
}

void launch_finalize(hipStream_t st, const FinArgs &a) {
    for (int k = 0; k < a.nin; ++k) MI_REQUIRE(a.n[k] <= 8 * kFinBlock, MI_E_LIMIT, "too many partials");
    k_finalize<<<1, kFinBlock, 0, st>>>(a);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_edge_iter(hipStream_t st, const RedBuf &rb, int nx, int ny, const T *f, T *v, T *dwork, double alpha,
                      double beta, double eps, Scalars *sc) {
    const unsigned grid = grid_pixels((int64_t)nx * ny);
    k_edge_grad<T><<<grid, kBlock, 0, st>>>(nx, ny, f, v, dwork, nullptr, alpha, beta, eps, sc, rb.slot(0),
                                            rb.slot(1));
    MI_HIP(hipGetLastError());
    k_edge_nabla<T><<<grid, kBlock, 0, st>>>(nx, ny, dwork, sc, rb.slot(2));
    MI_HIP(hipGetLastError());
    FinArgs c{};
    c.op = FIN_EDGE_C;
    c.nin = 3;
    for (int k = 0; k < 3; ++k) {
        c.in[k] = rb.slot(k);
        c.n[k] = (int)grid;
    }
    c.sc = sc;
    c.alpha = alpha;
    c.beta = beta;
    c.eps = eps;
    launch_finalize(st, c);
    launch_edge_update<T>(st, rb, nx, ny, v, dwork, sc);
}

template <typename T>
void launch_edge_grad_n(hipStream_t st, int nx, int ny, const T *f, const T *v, T *d, T *n, double alpha,
                        double beta, double eps, const Scalars *sc) {
    k_edge_grad<T><<<grid_pixels((int64_t)nx * ny), kBlock, 0, st>>>(nx, ny, f, v, d, n, alpha, beta, eps, sc,
                                                                     nullptr, nullptr);
    MI_HIP(hipGetLastError());
}

template <typename T>
void launch_edge_update(hipStream_t st, const RedBuf &rb, int nx, int ny, T *v, const T *d, Scalars *sc) {
    const unsigned grid = grid_pixels((int64_t)nx * ny);
    k_edge_update<T><<<grid, kBlock, 0, st>>>(nx, ny, v, d, sc, rb.slot(3));
    MI_HIP(hipGetLastError());
    FinArgs f{};
    f.op = FIN_EDGE_DONE;
    f.nin = 1;
    f.in[0] = rb.slot(3);
    f.n[0] = (int)grid;
    f.is_max[0] = 1;
    f.sc = sc;
    launch_finalize(st, f);
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
    template unsigned launch_fp<T>(hipStream_t, int, const CsrView &, int64_t, int64_t, const T *, const T *,   \
                                   const T *, T *, double *, const int32_t *, int, int);                        \
    template unsigned launch_bp_grad<T>(hipStream_t, const CscView &, int, int, const T *, double, const T *,   \
                                        const T *, double, double, T *, double *, int, const int32_t *, int);   \
    template unsigned launch_nv<T>(hipStream_t, int, int, const T *, const T *, double, double *,               \
                                   const int32_t *);                                                            \
    template unsigned launch_energy_reg<T>(hipStream_t, int, int, const T *, const T *, double, double,         \
                                           double *);                                                           \
    template unsigned launch_dot<T>(hipStream_t, int64_t, const T *, const T *, double *);                      \
    template unsigned launch_axpy<T>(hipStream_t, int64_t, T *, const T *, const double *, double, double,      \
                                     double, T *, double *, const int32_t *);                                   \
    template void launch_xpby<T>(hipStream_t, int64_t, T *, const T *, double);                                 \
    template void launch_median<T>(hipStream_t, int, int, const T *, T *);                                      \
    template void launch_edge_iter<T>(hipStream_t, const RedBuf &, int, int, const T *, T *, T *, double,       \
                                      double, double, Scalars *);                                               \
    template void launch_edge_grad_n<T>(hipStream_t, int, int, const T *, const T *, T *, T *, double, double,  \
                                        double, const Scalars *);                                               \
    template void launch_edge_update<T>(hipStream_t, const RedBuf &, int, int, T *, const T *, Scalars *);      \
    template void launch_gather_rows<T>(hipStream_t, int64_t, const int64_t *, const T *, T *);                 \
    template void launch_scatter_rows<T>(hipStream_t, int64_t, const int64_t *, const T *, T *);

MI_INST(float)
MI_INST(double)

}  // namespace mi
