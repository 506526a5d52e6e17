// Kernels of the batched ordered-subset AT engine (gfx950).
//
// One launch per phase serves every unit of a batch.  The global block index
// is first renumbered so that each of the 8 XCDs owns a contiguous range
// (xcd_block), then mapped to (unit, local block) through the BlockMap passed
// by value.  K, the number of jobs of a unit, selects a code path compiled
// for that K; a launch whose units all have K = 1 uses a kernel compiled for
// K = 1 alone (plain form), so it keeps the register budget of that case.
//
// Per-entry work per lane: entries are taken in groups of four consecutive
// entries (16-byte loads of the padded idx/len streams) from the 4-aligned
// entry at or below the row start; group g of a row goes to lane g % 64 (or
// g % lanes-per-pixel), and a lane visits its groups in increasing order.
// How many groups a lane loads per trip (G) only changes how many loads are
// in flight, never the order of the Kahan sums.
#include "engine.hpp"

#include <cmath>

namespace mi {

namespace {

// Global address space.  The unit table is read from device memory, so the
// compiler cannot infer that its pointers are global and would emit flat
// loads (64-bit addresses, counted against lgkmcnt as well as vmcnt, which
// serialises them with every scalar and LDS wait).  Every access of the
// engine goes through these pointers instead.
#define GL __attribute__((address_space(1)))

template <typename T>
struct UG {
    const GL int64_t *rowptr;
    const GL uint32_t *idx;
    const GL T *len;
    const GL int64_t *colptr;
    const GL uint32_t *crow;
    const GL T *clen;
    const GL int64_t *sub_row0;
    GL T *g, *r, *u, *v, *d, *w;
    int64_t npix, nrows, part_off, cap;
    int nx, ny, K, lanes, job0;
    __device__ __forceinline__ explicit UG(const UnitDesc<T> &D)
        : rowptr((const GL int64_t *)D.rowptr), idx((const GL uint32_t *)D.idx), len((const GL T *)D.len),
          colptr((const GL int64_t *)D.colptr), crow((const GL uint32_t *)D.crow), clen((const GL T *)D.clen),
          sub_row0((const GL int64_t *)D.sub_row0), g((GL T *)D.g), r((GL T *)D.r), u((GL T *)D.u),
          v((GL T *)D.v), d((GL T *)D.d), w((GL T *)D.w), npix(D.npix), nrows(D.nrows), part_off(D.part_off),
          cap(D.cap), nx(D.nx), ny(D.ny), K(D.K), lanes(D.lanes), job0(D.job0) {}
};

template <typename T>
__device__ __forceinline__ const GL JobParams &jpar(const EngArgs<T> &a, int js) {
    return ((const GL JobParams *)a.jp)[js];
}
template <typename T>
__device__ __forceinline__ GL Scalars &jsc(const EngArgs<T> &a, int js) {
    return ((GL Scalars *)a.sc)[js];
}

#ifndef MI_FP_G1
#define MI_FP_G1 2      // vec4 groups per lane per trip of the K = 1 forward projector
#endif
#ifndef MI_NT_FP
#define MI_NT_FP 1      // non-temporal matrix streams in the forward projector
#endif
#ifndef MI_NT_BP
#define MI_NT_BP 0      // ... and in the back projector (measured slower with it)
#endif

// Matrix streams are read once per launch: non-temporal 16-byte loads, so
// that they do not evict the image / residual lines the gathers reuse from L2.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ void stream_idx4(const GL uint32_t *p, uint32_t (&o)[4]) {
    const u32x4 a = NT ? __builtin_nontemporal_load((const GL u32x4 *)p) : *(const GL u32x4 *)p;
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}
template <bool NT>
__device__ __forceinline__ void stream4(const GL float *p, float (&o)[4]) {
    const f32x4 a = NT ? __builtin_nontemporal_load((const GL f32x4 *)p) : *(const GL f32x4 *)p;
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}
template <bool NT>
__device__ __forceinline__ void stream4(const GL double *p, double (&o)[4]) {
    const f64x2 a = NT ? __builtin_nontemporal_load((const GL f64x2 *)p) : *(const GL f64x2 *)p;
    const f64x2 b = NT ? __builtin_nontemporal_load((const GL f64x2 *)p + 1) : *((const GL f64x2 *)p + 1);
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}

// K consecutive values at p (p is aligned to K * sizeof(T)).
template <typename T, int K>
__device__ __forceinline__ void loadk(const GL T *__restrict__ p, T (&o)[K]) {
    constexpr int B = K * (int)sizeof(T);
    if constexpr (B >= 16) {
#pragma unroll
        for (int c = 0; c < B / 16; ++c) {
            const u32x4 a = ((const GL u32x4 *)p)[c];
            __builtin_memcpy(&o[c * (16 / (int)sizeof(T))], &a, 16);
        }
    } else if constexpr (B == 8) {
        const uint64_t a = *(const GL uint64_t *)p;
        __builtin_memcpy(&o[0], &a, 8);
    } else {
        o[0] = p[0];
    }
}

template <typename T, int K>
__device__ __forceinline__ void storek(GL T *__restrict__ p, const T (&o)[K]) {
    constexpr int B = K * (int)sizeof(T);
    if constexpr (B >= 16) {
#pragma unroll
        for (int c = 0; c < B / 16; ++c) {
            u32x4 a;
            __builtin_memcpy(&a, &o[c * (16 / (int)sizeof(T))], 16);
            ((GL u32x4 *)p)[c] = a;
        }
    } else if constexpr (B == 8) {
        uint64_t a;
        __builtin_memcpy(&a, &o[0], 8);
        *(GL uint64_t *)p = a;
    } else {
        p[0] = o[0];
    }
}

template <typename T, int K>
__device__ __forceinline__ void zerok(T (&o)[K]) {
#pragma unroll
    for (int j = 0; j < K; ++j) o[j] = T(0);
}

// Sums v[j] over each aligned group of L lanes (L >= K, powers of two) and
// leaves job (lane % K)'s total in v[0].  The levels pair lane with lane ^ o
// for o = L/2 .. 1, as a plain butterfly does, so every total is bit-for-bit
// the butterfly's; below o = K each lane keeps the half of its K values that
// its lane bit selects instead of all of them, which needs no dynamic
// register indexing.
template <int K>
__device__ __forceinline__ double reduce_scatter(double (&v)[K], int L) {
    for (int o = L >> 1; o >= K; o >>= 1)
#pragma unroll
        for (int j = 0; j < K; ++j) v[j] += __shfl_xor(v[j], o, kWave);
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int h = K / 2; h >= 1; h >>= 1) {
        const bool up = (lane & h) != 0;
#pragma unroll
        for (int j = 0; j < h; ++j) {
            const double send = up ? v[j] : v[j + h];
            const double keep = up ? v[j + h] : v[j];
            v[j] = keep + __shfl_xor(send, h, kWave);
        }
    }
    return v[0];
}

__device__ __forceinline__ int find_unit(const BlockMap &m, int b) {
    int lo = 0, hi = m.nunits;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (m.start[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <typename T>
__device__ __forceinline__ bool job_live(const EngArgs<T> &a, int js) {
    return jpar(a, js).active && !jsc(a, js).done;
}

template <typename T>
__device__ __forceinline__ bool unit_idle(const EngArgs<T> &a, const UG<T> &U) {
    for (int j = 0; j < U.K; ++j)
        if (job_live(a, U.job0 + j)) return false;
    return true;
}

template <typename T>
__device__ __forceinline__ GL double *part_at(const EngArgs<T> &a, const UG<T> &U, int slot, int j) {
    return (GL double *)a.part + U.part_off + ((int64_t)slot * U.K + j) * U.cap;
}

// Sum (or max) of one value per thread for each of K jobs; thread j < K
// stores job j's block total at part_at(slot, j)[lb].  Fixed order.
template <typename T, int K, bool MAX>
__device__ __forceinline__ void block_store(const EngArgs<T> &a, const UG<T> &U, int slot, int lb,
                                            const double (&m)[K], double *lds) {
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const double t = MAX ? wave_max(m[j]) : group_sum<kWave>(m[j]);
        if (lane == 0) lds[wv * K + j] = t;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t = MAX ? fmax(t, lds[w * K + threadIdx.x]) : t + lds[w * K + threadIdx.x];
        part_at(a, U, slot, threadIdx.x)[lb] = t;
    }
}

// ---------------------------------------------------------------------------
// Forward projection of subset s: one wave per ray, kRaysPerWave consecutive
// rays per wave.  RESID: r = g - A u (stored), partial sum r^2; else partial
// sum (A d)^2.
template <typename T, int K, bool RESID>
__device__ __forceinline__ void fp_unit(const EngArgs<T> &a, const UG<T> &U, int s, int lb, double *lds) {
    constexpr int G = (K == 1) ? MI_FP_G1 : (K == 2 ? 2 : 1);
    constexpr int64_t kTrip = 4 * kWave;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int64_t row0 = U.sub_row0[s], nr = U.sub_row0[s + 1] - row0;
    const GL T *__restrict__ x = RESID ? U.u : U.d;
    const GL uint32_t *__restrict__ idx = U.idx;
    const GL T *__restrict__ len = U.len;
    double wpart = 0.0;
    const int64_t r0 = (int64_t)lb * kRaysPerBlock + wv * kRaysPerWave;
    for (int i = 0; i < kRaysPerWave; ++i) {
        const int64_t r = r0 + i;
        if (r >= nr) break;
        const int64_t rr = row0 + r;
        const int64_t b = U.rowptr[rr], e = U.rowptr[rr + 1];
        Acc<T> acc[K];
        for (int64_t k0 = (b & ~(int64_t)3) + 4 * lane; k0 < e; k0 += kTrip * G) {
            uint32_t id[G][4];
            T ln[G][4];
            T xv[G][4][K];
#pragma unroll
            for (int q = 0; q < G; ++q) {
                if (k0 + kTrip * q < e) {
                    stream_idx4<MI_NT_FP>(idx + k0 + kTrip * q, id[q]);
                    stream4<MI_NT_FP>(len + k0 + kTrip * q, ln[q]);
                } else {
#pragma unroll
                    for (int m = 0; m < 4; ++m) { id[q][m] = 0; ln[q][m] = T(0); }
                }
            }
#pragma unroll
            for (int q = 0; q < G; ++q)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    if (k0 + kTrip * q < e) loadk<T, K>(x + (uint32_t)(id[q][m] * K), xv[q][m]);
                    else zerok<T, K>(xv[q][m]);
                }
#pragma unroll
            for (int q = 0; q < G; ++q)
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int64_t kk = k0 + kTrip * q + m;
                    if (kk >= b && kk < e)
#pragma unroll
                        for (int j = 0; j < K; ++j) acc[j].add(xv[q][m][j] * ln[q][m]);
                }
        }
        double tot[K];
#pragma unroll
        for (int j = 0; j < K; ++j) tot[j] = acc[j].value();
        const double sv = reduce_scatter<K>(tot, kWave);
        if (lane < K) {
            if (RESID) {
                const T o = (T)((double)U.g[rr * K + lane] - sv);
                U.r[rr * K + lane] = o;
                wpart += (double)o * (double)o;
            } else {
                wpart += sv * sv;
            }
        }
    }
    if (lane < K) lds[wv * K + lane] = wpart;
    __syncthreads();
    if (threadIdx.x < K) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += lds[w * K + threadIdx.x];
        part_at(a, U, RESID ? SL_DF : SL_AD, threadIdx.x)[lb] = t;
    }
}

// GoldenC image-gradient regulariser of job o of an interleaved pair (u, v)
// at pixel (i, j): u zero padded, v one padded (origin/GoldenC/src/main.c:112-125).
template <typename T>
__device__ __forceinline__ double gc_reg_k(const GL T *u, const GL T *v, int i, int j, int nx, int ny, double eps, int K,
                                           int o) {
    auto at = [&](const GL T *arr, int ii, int jj, double pad) -> double {
        return (ii < 0 || jj < 0 || ii >= nx || jj >= ny) ? pad : (double)arr[((int64_t)ii * ny + jj) * K + o];
    };
    const double z = at(u, i, j, 0.0);
    const double zd = at(u, i + 1, j, 0.0), zr = at(u, i, j + 1, 0.0);
    const double zu = at(u, i - 1, j, 0.0), zl = at(u, i, j - 1, 0.0);
    const double p = at(v, i, j, 1.0);
    const double pu = at(v, i - 1, j, 1.0), pl = at(v, i, j - 1, 1.0);
    return +(p * p) * (zd - z) + (p * p) * (zr - z) - (pu * pu) * (z - zu) - (pl * pl) * (z - zl) +
           0.01 * (eps * eps) * (zd + zr - 4 * z + zu + zl);
}

// Gather back projection of subset s over kBpPix pixels, U.lanes lanes per
// pixel column, fused with d = S * A_s^T r + alpha * R(u, v); partial sum d^2.
template <typename T, int K>
__device__ __forceinline__ void bp_unit(const EngArgs<T> &a, const UG<T> &U, int s, int lb, double *lds) {
    constexpr int G = (K <= 2) ? 2 : 1;
    const int L = U.lanes;
    const int gl = threadIdx.x & (L - 1), grp = threadIdx.x / L;
    const int per = kBlock / L;
    const int64_t trip = 4 * (int64_t)L;
    const GL int64_t *__restrict__ colptr = U.colptr + (int64_t)s * U.npix;
    const GL uint32_t *__restrict__ crow = U.crow;
    const GL T *__restrict__ clen = U.clen;
    const GL T *__restrict__ r = U.r;
    const double Ssc = (double)a.S;
    for (int p0 = 0; p0 < kBpPix; p0 += per) {
        const int pl = p0 + grp;
        const int64_t p = (int64_t)lb * kBpPix + pl;
        if (p < U.npix) {
            const int64_t b = colptr[p], e = colptr[p + 1];
            Acc<T> acc[K];
            for (int64_t k0 = (b & ~(int64_t)3) + 4 * gl; k0 < e; k0 += trip * G) {
                uint32_t id[G][4];
                T ln[G][4];
                T rv[G][4][K];
#pragma unroll
                for (int q = 0; q < G; ++q) {
                    if (k0 + trip * q < e) {
                        stream_idx4<MI_NT_BP>(crow + k0 + trip * q, id[q]);
                        stream4<MI_NT_BP>(clen + k0 + trip * q, ln[q]);
                    } else {
#pragma unroll
                        for (int m = 0; m < 4; ++m) { id[q][m] = 0; ln[q][m] = T(0); }
                    }
                }
#pragma unroll
                for (int q = 0; q < G; ++q)
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        if (k0 + trip * q < e) loadk<T, K>(r + (uint32_t)(id[q][m] * K), rv[q][m]);
                        else zerok<T, K>(rv[q][m]);
                    }
#pragma unroll
                for (int q = 0; q < G; ++q)
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const int64_t kk = k0 + trip * q + m;
                        if (kk >= b && kk < e)
#pragma unroll
                            for (int j = 0; j < K; ++j) acc[j].add(rv[q][m][j] * ln[q][m]);
                    }
            }
            double tot[K];
#pragma unroll
            for (int j = 0; j < K; ++j) tot[j] = acc[j].value();
            const double sv = reduce_scatter<K>(tot, L);
            const int pi = (int)(p / U.ny), pj = (int)(p % U.ny);
            if (gl < K) {
                const int j = gl;
                const GL JobParams &P = jpar(a, U.job0 + j);
                double dv = Ssc * sv;
                dv += P.alpha * gc_reg_k<T>(U.u, U.v, pi, pj, U.nx, U.ny, P.eps, K, j);
                const T dt = (T)dv;
                U.d[p * K + j] = dt;
                lds[pl * K + j] = (double)dt * (double)dt;
            }
        } else if (gl < K) {
            lds[pl * K + gl] = 0.0;
        }
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const double t = group_sum<kWave>(lds[threadIdx.x * K + j]);
            if (threadIdx.x == j) part_at(a, U, SL_DD, j)[lb] = t;
        }
    }
}

// nV(d) = sum (v^2 + 0.01 eps^2)((d - d_up)^2 + (d - d_left)^2), d zero
// padded (origin/GoldenC/src/main.c:152-158).
template <typename T, int K>
__device__ __forceinline__ void nv_unit(const EngArgs<T> &a, const UG<T> &U, int lb, double *lds) {
    const int64_t q = (int64_t)lb * kPixBlock + threadIdx.x;
    double m[K];
#pragma unroll
    for (int j = 0; j < K; ++j) m[j] = 0.0;
    if (q < U.npix) {
        const int i = (int)(q / U.ny), jj = (int)(q % U.ny);
        T dq[K], du[K], dl[K], vv[K];
        loadk<T, K>(U.d + q * K, dq);
        loadk<T, K>(U.v + q * K, vv);
        if (i > 0) loadk<T, K>(U.d + (q - U.ny) * K, du);
        else zerok<T, K>(du);
        if (jj > 0) loadk<T, K>(U.d + (q - 1) * K, dl);
        else zerok<T, K>(dl);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const double eps = jpar(a, U.job0 + j).eps;
            const double e2 = 0.01 * (eps * eps);
            const double z = (double)dq[j];
            const double A = z - (double)du[j], B = z - (double)dl[j];
            const double w = (double)vv[j];
            m[j] = (w * w + e2) * (A * A + B * B);
        }
    }
    block_store<T, K, false>(a, U, SL_NV, lb, m, lds);
}

// u = clamp(u + c d, lower, upper) for live jobs.
template <typename T, int K>
__device__ __forceinline__ void axpy_unit(const EngArgs<T> &a, const UG<T> &U, int lb) {
    const int64_t q = (int64_t)lb * kPixBlock + threadIdx.x;
    if (q >= U.npix) return;
    T uu[K], dd[K];
    loadk<T, K>(U.u + q * K, uu);
    loadk<T, K>(U.d + q * K, dd);
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int js = U.job0 + j;
        if (!job_live(a, js)) continue;
        const GL JobParams &P = jpar(a, js);
        double yn = (double)uu[j] + jsc(a, js).c * (double)dd[j];
        if (yn < P.lower) yn = P.lower;
        if (yn > P.upper) yn = P.upper;
        uu[j] = (T)yn;
    }
    storek<T, K>(U.u + q * K, uu);
}

template <typename T>
__device__ __forceinline__ void cswap(T &x, T &y) {
    const T lo = x < y ? x : y, hi = x < y ? y : x;
    x = lo;
    y = hi;
}

template <typename T>
__device__ __forceinline__ bool median_due(const EngArgs<T> &a, int js) {
    const GL JobParams &P = jpar(a, js);
    return job_live(a, js) && P.median_every > 0 && (jsc(a, js).iters + 1) % P.median_every == 0;
}

// w = 3x3 median of u (zero padding, origin/GoldenC/src/main.c:366-390) for
// jobs whose median is due this pass, else w = u.
template <typename T, int K>
__device__ __forceinline__ void median_unit(const EngArgs<T> &a, const UG<T> &U, int lb) {
    const int64_t q = (int64_t)lb * kPixBlock + threadIdx.x;
    if (q >= U.npix) return;
    const int i = (int)(q / U.ny), jj = (int)(q % U.ny);
    T nb[9][K];
    int n = 0;
#pragma unroll
    for (int di = -1; di <= 1; ++di)
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj, ++n) {
            const int ii = i + di, jc = jj + dj;
            if (ii < 0 || jc < 0 || ii >= U.nx || jc >= U.ny) zerok<T, K>(nb[n]);
            else loadk<T, K>(U.u + ((int64_t)ii * U.ny + jc) * K, nb[n]);
        }
    T out[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (!median_due(a, U.job0 + j)) {
            out[j] = nb[4][j];
            continue;
        }
        T t[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) t[k] = nb[k][j];
        // Paeth's 19-exchange median-of-9 network (exact selection).
        cswap(t[1], t[2]); cswap(t[4], t[5]); cswap(t[7], t[8]);
        cswap(t[0], t[1]); cswap(t[3], t[4]); cswap(t[6], t[7]);
        cswap(t[1], t[2]); cswap(t[4], t[5]); cswap(t[7], t[8]);
        cswap(t[0], t[3]); cswap(t[5], t[8]); cswap(t[4], t[7]);
        cswap(t[3], t[6]); cswap(t[1], t[4]); cswap(t[2], t[5]);
        cswap(t[4], t[7]); cswap(t[4], t[2]); cswap(t[6], t[4]);
        cswap(t[4], t[2]);
        out[j] = t[4];
    }
    storek<T, K>(U.w + q * K, out);
}

template <typename T, int K>
__device__ __forceinline__ void copy_wu_unit(const UG<T> &U, int lb) {
    const int64_t q = (int64_t)lb * kPixBlock + threadIdx.x;
    if (q >= U.npix) return;
    T t[K];
    loadk<T, K>(U.w + q * K, t);
    storek<T, K>(U.u + q * K, t);
}

template <typename T>
__device__ __forceinline__ bool edge_live(const EngArgs<T> &a, int js, int it) {
    return job_live(a, js) && !jsc(a, js).edge_done && it < jpar(a, js).iter_edge;
}

// Edge step, origin/GoldenC/src/main.c:294-305: w = -grad_v, partials
// ||w||^2 and sum n w^2 (f = u zero padded, v one padded).
template <typename T, int K>
__device__ __forceinline__ void edge_grad_unit(const EngArgs<T> &a, const UG<T> &U, int lb, int it,
                                               double *lds) {
    const int64_t q = (int64_t)lb * kPixBlock + threadIdx.x;
    double m0[K], m1[K];
#pragma unroll
    for (int j = 0; j < K; ++j) m0[j] = m1[j] = 0.0;
    if (q < U.npix) {
        const int i = (int)(q / U.ny), jj = (int)(q % U.ny);
        T fd[K], fu[K], fr[K], fl[K], vc[K], vd[K], vu[K], vr[K], vl[K];
        auto ldz = [&](const GL T *arr, int ii, int jc, T (&o)[K], T pad) {
            if (ii < 0 || jc < 0 || ii >= U.nx || jc >= U.ny) {
#pragma unroll
                for (int j = 0; j < K; ++j) o[j] = pad;
            } else {
                loadk<T, K>(arr + ((int64_t)ii * U.ny + jc) * K, o);
            }
        };
        ldz(U.u, i + 1, jj, fd, T(0));
        ldz(U.u, i - 1, jj, fu, T(0));
        ldz(U.u, i, jj + 1, fr, T(0));
        ldz(U.u, i, jj - 1, fl, T(0));
        loadk<T, K>(U.v + q * K, vc);
        ldz(U.v, i + 1, jj, vd, T(1));
        ldz(U.v, i - 1, jj, vu, T(1));
        ldz(U.v, i, jj + 1, vr, T(1));
        ldz(U.v, i, jj - 1, vl, T(1));
        T out[K];
        zerok<T, K>(out);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int js = U.job0 + j;
            if (!edge_live(a, js, it)) continue;
            const GL JobParams &P = jpar(a, js);
            const double dxf = 0.5 * ((double)fd[j] - (double)fu[j]);
            const double dyf = 0.5 * ((double)fr[j] - (double)fl[j]);
            const double n = dxf * dxf + dyf * dyf;
            const double t = (double)vc[j];
            const double lap = 0.25 * ((double)vd[j] + (double)vr[j] - 4 * t + (double)vu[j] + (double)vl[j]);
            const double gr = P.alpha * t * n - P.beta * P.eps * lap + P.beta * (t - 1) / (4 * P.eps);
            const T dt = (T)(-gr);
            out[j] = dt;
            const double dd = (double)dt * (double)dt;
            m0[j] = dd;
            m1[j] = n * dd;
        }
        storek<T, K>(U.w + q * K, out);
    }
    block_store<T, K, false>(a, U, SL_E0, lb, m0, lds);
    __syncthreads();
    block_store<T, K, false>(a, U, SL_E1, lb, m1, lds);
}

// main.c:311-329: |central gradient of w|^2.
template <typename T, int K>
__device__ __forceinline__ void edge_nabla_unit(const EngArgs<T> &a, const UG<T> &U, int lb, int it,
                                                double *lds) {
    const int64_t q = (int64_t)lb * kPixBlock + threadIdx.x;
    double m[K];
#pragma unroll
    for (int j = 0; j < K; ++j) m[j] = 0.0;
    if (q < U.npix) {
        const int i = (int)(q / U.ny), jj = (int)(q % U.ny);
        T d1[K], d2[K], d3[K], d4[K];
        auto ldz = [&](int ii, int jc, T (&o)[K]) {
            if (ii < 0 || jc < 0 || ii >= U.nx || jc >= U.ny) zerok<T, K>(o);
            else loadk<T, K>(U.w + ((int64_t)ii * U.ny + jc) * K, o);
        };
        ldz(i + 1, jj, d1);
        ldz(i - 1, jj, d2);
        ldz(i, jj + 1, d3);
        ldz(i, jj - 1, d4);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (!edge_live(a, U.job0 + j, it)) continue;
            const double A = 0.5 * ((double)d1[j] - (double)d2[j]);
            const double B = 0.5 * ((double)d3[j] - (double)d4[j]);
            m[j] = A * A + B * B;
        }
    }
    block_store<T, K, false>(a, U, SL_E2, lb, m, lds);
}

// main.c:332-354: v = clamp(v + c w, 0, 1); partial max |c w|.
template <typename T, int K>
__device__ __forceinline__ void edge_update_unit(const EngArgs<T> &a, const UG<T> &U, int lb, int it,
                                                 double *lds) {
    const int64_t q = (int64_t)lb * kPixBlock + threadIdx.x;
    double m[K];
#pragma unroll
    for (int j = 0; j < K; ++j) m[j] = 0.0;
    if (q < U.npix) {
        T vv[K], ww[K];
        loadk<T, K>(U.v + q * K, vv);
        loadk<T, K>(U.w + q * K, ww);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int js = U.job0 + j;
            if (!edge_live(a, js, it)) continue;
            const double cd = jsc(a, js).e_c * (double)ww[j];
            double x = (double)vv[j] + cd;
            m[j] = fabs(cd);
            if (x > 1) x = 1;
            if (x < 0) x = 0;
            vv[j] = (T)x;
        }
        storek<T, K>(U.v + q * K, vv);
    }
    block_store<T, K, true>(a, U, SL_E3, lb, m, lds);
}

// ---------------------------------------------------------------------------
// Kernels: block -> (unit, local block) and dispatch on K.

#define MI_KSWITCH(KMAX, CALL)                  \
    if constexpr (KMAX == 1) {                  \
        constexpr int K = 1;                    \
        CALL;                                   \
    } else {                                    \
        switch (U.K) {                          \
            case 1: { constexpr int K = 1; CALL; break; } \
            case 2: { constexpr int K = 2; CALL; break; } \
            case 4: { constexpr int K = 4; CALL; break; } \
            default: { constexpr int K = 8; CALL; break; } \
        }                                       \
    }

template <typename T, int KMAX, bool RESID>
__global__ __launch_bounds__(kBlock) void k_eng_fp(EngArgs<T> a, BlockMap m, int s) {
    __shared__ double lds[kBlock / kWave * kMaxK];
    const int b = (int)xcd_block();
    if (b >= m.start[m.nunits]) return;
    const int ui = find_unit(m, b);
    const UG<T> U(a.units[ui]);
    if (unit_idle(a, U)) return;
    const int lb = b - m.start[ui];
    MI_KSWITCH(KMAX, (fp_unit<T, K, RESID>(a, U, s, lb, lds)))
}

template <typename T, int KMAX>
__global__ __launch_bounds__(kBlock) void k_eng_bp(EngArgs<T> a, BlockMap m, int s) {
    __shared__ double lds[kBpPix * kMaxK];
    const int b = (int)xcd_block();
    if (b >= m.start[m.nunits]) return;
    const int ui = find_unit(m, b);
    const UG<T> U(a.units[ui]);
    if (unit_idle(a, U)) return;
    const int lb = b - m.start[ui];
    MI_KSWITCH(KMAX, (bp_unit<T, K>(a, U, s, lb, lds)))
}

// Two block maps in one launch: each XCD runs a contiguous 1/8 of the
// rows part and a contiguous 1/8 of the pixel part (both parts padded to a
// multiple of 8 blocks), so neither part is left to a few XCDs.
template <typename T, int KMAX>
__global__ __launch_bounds__(kBlock) void k_eng_fpsq_nv(EngArgs<T> a, BlockMap rows, BlockMap pix, int s) {
    __shared__ double lds[kBlock / kWave * kMaxK];
    const int nrb = rows.start[rows.nunits], npb = pix.start[pix.nunits];
    const int r8 = (nrb + 7) / 8, p8 = (npb + 7) / 8;
    const int x = blockIdx.x & 7, t = blockIdx.x >> 3;
    if (t < r8) {
        const int b = x * r8 + t;
        if (b >= nrb) return;
        const int ui = find_unit(rows, b);
        const UG<T> U(a.units[ui]);
        if (unit_idle(a, U)) return;
        const int lb = b - rows.start[ui];
        MI_KSWITCH(KMAX, (fp_unit<T, K, false>(a, U, s, lb, lds)))
    } else {
        const int b = x * p8 + (t - r8);
        if (b >= npb) return;
        const int ui = find_unit(pix, b);
        const UG<T> U(a.units[ui]);
        if (unit_idle(a, U)) return;
        const int lb = b - pix.start[ui];
        MI_KSWITCH(KMAX, (nv_unit<T, K>(a, U, lb, lds)))
    }
}

enum PixOp { PX_AXPY = 0, PX_MEDIAN, PX_COPY_WU, PX_EDGE_GRAD, PX_EDGE_NABLA, PX_EDGE_UPDATE };

template <typename T, int KMAX, int OP>
__global__ __launch_bounds__(kBlock) void k_eng_pix(EngArgs<T> a, BlockMap m, int it) {
    __shared__ double lds[kBlock / kWave * kMaxK];
    const int b = (int)xcd_block();
    if (b >= m.start[m.nunits]) return;
    const int ui = find_unit(m, b);
    const UG<T> U(a.units[ui]);
    if (unit_idle(a, U)) return;
    const int lb = b - m.start[ui];
    if constexpr (OP == PX_AXPY) { MI_KSWITCH(KMAX, (axpy_unit<T, K>(a, U, lb))) }
    else if constexpr (OP == PX_MEDIAN) { MI_KSWITCH(KMAX, (median_unit<T, K>(a, U, lb))) }
    else if constexpr (OP == PX_COPY_WU) { MI_KSWITCH(KMAX, (copy_wu_unit<T, K>(U, lb))) }
    else if constexpr (OP == PX_EDGE_GRAD) { MI_KSWITCH(KMAX, (edge_grad_unit<T, K>(a, U, lb, it, lds))) }
    else if constexpr (OP == PX_EDGE_NABLA) { MI_KSWITCH(KMAX, (edge_nabla_unit<T, K>(a, U, lb, it, lds))) }
    else { MI_KSWITCH(KMAX, (edge_update_unit<T, K>(a, U, lb, it, lds))) }
}

// Fixed-order block reduction of n partials (sum or max), 256 threads.
__device__ __forceinline__ double fin_reduce(const GL double *p, int64_t n, bool mx, double *lds) {
    double acc = 0.0;
    for (int64_t base = 0; base < n; base += (int64_t)kBlock * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int64_t i = base + u * kBlock + threadIdx.x;
            v[u] = (i < n) ? p[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = mx ? fmax(acc, v[u]) : acc + v[u];
    }
    acc = mx ? wave_max(acc) : group_sum<kWave>(acc);
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    if (lane == 0) lds[wv] = acc;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kBlock / kWave; ++w) t = mx ? fmax(t, lds[w]) : t + lds[w];
    __syncthreads();
    return t;
}

__host__ __device__ __forceinline__ int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// One block per job slot (units listed in m; m.nunits * kMaxK blocks).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_eng_fin(EngArgs<T> a, int op, int s) {
    __shared__ double lds[kBlock / kWave];
    const int ui = blockIdx.x / kMaxK, j = blockIdx.x % kMaxK;
    const UG<T> U(a.units[ui]);
    if (j >= U.K) return;
    const int js = U.job0 + j;
    if (!job_live(a, js)) return;
    GL Scalars *sc = &jsc(a, js);
    const GL JobParams &P = jpar(a, js);
    const int64_t npx = cdiv(U.npix, kPixBlock);
    switch (op) {
        case EF_BEGIN:
            if (threadIdx.x == 0) {
                sc->df_pass = 0.0;
                sc->edge_done = 0;
            }
            break;
        case EF_STEP: {
            const int64_t nr = cdiv(U.sub_row0[s + 1] - U.sub_row0[s], kRaysPerBlock);
            const double dd = fin_reduce(part_at(a, U, SL_DD, j), cdiv(U.npix, kBpPix), false, lds);
            const double ad2 = fin_reduce(part_at(a, U, SL_AD, j), nr, false, lds);
            const double nv = fin_reduce(part_at(a, U, SL_NV, j), npx, false, lds);
            const double df = fin_reduce(part_at(a, U, SL_DF, j), nr, false, lds);
            if (threadIdx.x == 0) {
                sc->df_pass += df;
                sc->dd = dd;
                sc->ad2 = ad2;
                sc->nv = nv;
                const double den = (double)a.S * ad2 + P.alpha * nv;
                sc->c = (den > 0) ? dd / den : 0.0;
            }
            break;
        }
        case EF_EDGE_C: {
            if (sc->edge_done || s >= P.iter_edge) return;
            const double nd = fin_reduce(part_at(a, U, SL_E0, j), npx, false, lds);
            const double nd2 = fin_reduce(part_at(a, U, SL_E1, j), npx, false, lds);
            const double nnd = fin_reduce(part_at(a, U, SL_E2, j), npx, false, lds);
            if (threadIdx.x == 0) {
                // origin/GoldenC/src/main.c:329
                sc->e_nd = nd;
                sc->e_nd2 = nd2;
                sc->e_nnd = nnd;
                sc->e_c = nd / (P.alpha * nd2 + P.beta * P.eps * nnd + (P.beta / (4 * P.eps)) * nd);
            }
            break;
        }
        case EF_EDGE_DONE: {
            if (sc->edge_done || s >= P.iter_edge) return;
            const double mx = fin_reduce(part_at(a, U, SL_E3, j), npx, true, lds);
            if (threadIdx.x == 0) {
                sc->e_maxcd = mx;
                if (mx < 1e-5) sc->edge_done = 1;   // main.c:351
            }
            break;
        }
        default: {  // EF_END
            if (threadIdx.x != 0) return;
            const double df = sc->df_pass;
            if (sc->iters > 0) {
                const double diff = fabs(df - sc->df_prev);
                sc->rel = sc->df_prev > 0 ? diff / sc->df_prev : 0.0;
                if (diff < P.rel_tol * sc->df_prev) sc->converged = 1;
            }
            sc->df_prev = df;
            sc->iters += 1;
            if (sc->converged || (P.max_iters > 0 && sc->iters >= P.max_iters)) sc->done = 1;
            break;
        }
    }
}

template <typename T>
__global__ void k_put_rows(T *dst, int K, int slot, int64_t n, const int64_t *perm, const T *src) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i * K + slot] = src[perm[i]];
}
template <typename T>
__global__ void k_put(T *dst, int K, int slot, int64_t n, const T *src, T fillv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i * K + slot] = src ? src[i] : fillv;
}
template <typename T>
__global__ void k_get(T *dst, const T *src, int K, int slot, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i * K + slot];
}
template <typename T>
__global__ void k_move(T *dst, int Kd, int sd, const T *src, int Ks, int ss, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i * Kd + sd] = src[i * Ks + ss];
}

// Plain streaming copy, 16 bytes per thread per step (the reference kernel
// against which memory counters are read).
__global__ __launch_bounds__(kBlock) void k_copy16(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                   int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        dst[i] = src[i];
}

unsigned grid8(int blocks) {
    if (blocks < 1) blocks = 1;
    return (unsigned)((blocks + 7) / 8 * 8);
}
unsigned grid_of(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

// ---------------------------------------------------------------------------

template <typename T>
void eng_fp_resid(hipStream_t st, const EngArgs<T> &a, const BlockMap &m, int s, int kmax) {
    const unsigned grid = grid8(m.start[m.nunits]);
    if (kmax == 1) k_eng_fp<T, 1, true><<<grid, kBlock, 0, st>>>(a, m, s);
    else k_eng_fp<T, kMaxK, true><<<grid, kBlock, 0, st>>>(a, m, s);
    MI_HIP(hipGetLastError());
}

template <typename T>
void eng_bp_grad(hipStream_t st, const EngArgs<T> &a, const BlockMap &m, int s, int kmax) {
    const unsigned grid = grid8(m.start[m.nunits]);
    if (kmax == 1) k_eng_bp<T, 1><<<grid, kBlock, 0, st>>>(a, m, s);
    else k_eng_bp<T, kMaxK><<<grid, kBlock, 0, st>>>(a, m, s);
    MI_HIP(hipGetLastError());
}

template <typename T>
void eng_fp_sq_nv(hipStream_t st, const EngArgs<T> &a, const BlockMap &rows, const BlockMap &pix, int s, int kmax) {
    const unsigned grid = 8 * (unsigned)((rows.start[rows.nunits] + 7) / 8 + (pix.start[pix.nunits] + 7) / 8);
    if (kmax == 1) k_eng_fpsq_nv<T, 1><<<grid, kBlock, 0, st>>>(a, rows, pix, s);
    else k_eng_fpsq_nv<T, kMaxK><<<grid, kBlock, 0, st>>>(a, rows, pix, s);
    MI_HIP(hipGetLastError());
}

template <typename T, int OP>
static void launch_pix(hipStream_t st, const EngArgs<T> &a, const BlockMap &m, int it, int kmax) {
    const unsigned grid = grid8(m.start[m.nunits]);
    if (kmax == 1) k_eng_pix<T, 1, OP><<<grid, kBlock, 0, st>>>(a, m, it);
    else k_eng_pix<T, kMaxK, OP><<<grid, kBlock, 0, st>>>(a, m, it);
    MI_HIP(hipGetLastError());
}

template <typename T>
void eng_axpy(hipStream_t st, const EngArgs<T> &a, const BlockMap &pix, int kmax) {
    launch_pix<T, PX_AXPY>(st, a, pix, 0, kmax);
}

template <typename T>
void eng_median(hipStream_t st, const EngArgs<T> &a, const BlockMap &pix, int kmax) {
    launch_pix<T, PX_MEDIAN>(st, a, pix, 0, kmax);
    launch_pix<T, PX_COPY_WU>(st, a, pix, 0, kmax);
}

template <typename T>
void eng_edge_iter(hipStream_t st, const EngArgs<T> &a, const BlockMap &pix, int it, int kmax) {
    launch_pix<T, PX_EDGE_GRAD>(st, a, pix, it, kmax);
    launch_pix<T, PX_EDGE_NABLA>(st, a, pix, it, kmax);
    eng_finalize<T>(st, a, EF_EDGE_C, it);
    launch_pix<T, PX_EDGE_UPDATE>(st, a, pix, it, kmax);
    eng_finalize<T>(st, a, EF_EDGE_DONE, it);
}

template <typename T>
void eng_finalize(hipStream_t st, const EngArgs<T> &a, int op, int s) {
    if (a.nunits <= 0) return;
    k_eng_fin<T><<<(unsigned)(a.nunits * kMaxK), kBlock, 0, st>>>(a, op, s);
    MI_HIP(hipGetLastError());
}

template <typename T>
void eng_put_rows(hipStream_t st, T *dst, int K, int slot, int64_t n, const int64_t *perm, const T *src) {
    if (n <= 0) return;
    k_put_rows<T><<<grid_of(n), kBlock, 0, st>>>(dst, K, slot, n, perm, src);
    MI_HIP(hipGetLastError());
}
template <typename T>
void eng_put(hipStream_t st, T *dst, int K, int slot, int64_t n, const T *src, T fillv) {
    if (n <= 0) return;
    k_put<T><<<grid_of(n), kBlock, 0, st>>>(dst, K, slot, n, src, fillv);
    MI_HIP(hipGetLastError());
}
template <typename T>
void eng_get(hipStream_t st, T *dst, const T *src, int K, int slot, int64_t n) {
    if (n <= 0) return;
    k_get<T><<<grid_of(n), kBlock, 0, st>>>(dst, src, K, slot, n);
    MI_HIP(hipGetLastError());
}
template <typename T>
void eng_move(hipStream_t st, T *dst, int Kd, int sd, const T *src, int Ks, int ss, int64_t n) {
    if (n <= 0) return;
    k_move<T><<<grid_of(n), kBlock, 0, st>>>(dst, Kd, sd, src, Ks, ss, n);
    MI_HIP(hipGetLastError());
}

void eng_copy(hipStream_t st, void *dst, const void *src, int64_t bytes) {
    MI_REQUIRE(bytes % 16 == 0, MI_E_ARG, "copy size must be a multiple of 16 bytes");
    const int64_t n = bytes / 16;
    if (n == 0) return;
    k_copy16<<<8192, kBlock, 0, st>>>((const u32x4 *)src, (u32x4 *)dst, n);
    MI_HIP(hipGetLastError());
}

#define MI_ENG_INST(T)                                                                                     \
    template void eng_fp_resid<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int, int);             \
    template void eng_bp_grad<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int, int);              \
    template void eng_fp_sq_nv<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, const BlockMap &, int,  \
                                  int);                                                                     \
    template void eng_axpy<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int);                      \
    template void eng_median<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int);                    \
    template void eng_edge_iter<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int, int);            \
    template void eng_finalize<T>(hipStream_t, const EngArgs<T> &, int, int);                               \
    template void eng_put_rows<T>(hipStream_t, T *, int, int, int64_t, const int64_t *, const T *);         \
    template void eng_put<T>(hipStream_t, T *, int, int, int64_t, const T *, T);                            \
    template void eng_get<T>(hipStream_t, T *, const T *, int, int, int64_t);                               \
    template void eng_move<T>(hipStream_t, T *, int, int, const T *, int, int, int64_t);

MI_ENG_INST(float)
MI_ENG_INST(double)

}  // namespace mi
