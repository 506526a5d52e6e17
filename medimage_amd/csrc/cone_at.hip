// Ordered-subset Ambrosio-Tortorelli reconstruction of a cone-beam volume.
//
// The 2D ordered-subset form of the engine, on the volume [nz][nx][ny]:
// per subset (views v with v % S == order[s], bit-reversed order)
//   q = g_s - A_s u                     (forward projection, residual form)
//   d = S A_s^T q + alpha r(u; v)       (gather back projection, stencil)
//   c = |d|^2 / (S |A_s d|^2 + alpha R(d; v)),  u = clamp(u + c d)
// and per pass iter_edge exact-line-search steps on v.  The regulariser
// widens GoldenC's 4-neighbour stencils (origin/GoldenC/src/main.c:112-125,
// 152-158, 296-329) to the 26-neighbourhood: the 13 forward offsets k with
// weight 1 / |k|^2, pairs (x, x + k) with at least one voxel inside, u
// zero-padded and v one-padded,
//   R(f; v) = sum w_k (v(x)^2 + eta) (f(x + k) - f(x))^2,  eta = 0.01 eps^2,
//   r = -1/2 dR/df,
// and the edge energy alpha sum v^2 n + beta eps Q(v) + beta / (4 eps)
// |1 - v|^2 with n(x) = sum_k w_k (u(x + k) - u(x))^2.  The reference's own
// 3D update is a per-row sweep (industry3d/src/main.cpp:233-532); this is
// its ordered-subset counterpart.
//
// Every sum that sets a step is a double partial per workgroup reduced in a
// fixed order (results do not depend on timing); the volumes are float.
// One thread per voxel in the stencil kernels: their traffic (a few volume
// reads) is small next to the projections.
#include "cone.hpp"

#include <cmath>
#include <vector>

namespace mi {

namespace {

struct Dims {
    int nz, nx, ny;
};

// The 13 forward offsets (dz, dx, dy) of the 26-neighbourhood.
__device__ __forceinline__ void koff(int k, int &a, int &b, int &c) {
    constexpr int T[13][3] = {{0, 0, 1},  {0, 1, -1}, {0, 1, 0},  {0, 1, 1},  {1, -1, -1},
                              {1, -1, 0}, {1, -1, 1}, {1, 0, -1}, {1, 0, 0},  {1, 0, 1},
                              {1, 1, -1}, {1, 1, 0},  {1, 1, 1}};
    a = T[k][0];
    b = T[k][1];
    c = T[k][2];
}

__device__ __forceinline__ float kwf(int a, int b, int c) {
    const int n = a * a + b * b + c * c;
    return n == 1 ? 1.0f : (n == 2 ? 0.5f : 1.0f / 3.0f);
}

__device__ __forceinline__ bool in3(const Dims &d, int z, int x, int y) {
    return (unsigned)z < (unsigned)d.nz && (unsigned)x < (unsigned)d.nx && (unsigned)y < (unsigned)d.ny;
}

__device__ __forceinline__ float at3(const float *f, const Dims &d, int z, int x, int y, float pad) {
    return in3(d, z, x, y) ? f[((size_t)z * d.nx + x) * d.ny + y] : pad;
}

__device__ __forceinline__ bool voxel(const Dims &d, size_t &i, int &z, int &x, int &y) {
    i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t n = (size_t)d.nz * d.nx * d.ny;
    y = (int)(i % d.ny);
    x = (int)((i / d.ny) % d.nx);
    z = (int)(i / ((size_t)d.ny * d.nx));
    return i < n;
}

__device__ __forceinline__ double block_sum(double x) {
    __shared__ double red[kBlock / kWave];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// d = S b + alpha r(u; v); part = sum d^2.
__global__ __launch_bounds__(kBlock) void k_cat_grad(Dims dm, const float *__restrict__ b, const float *__restrict__ u,
                                                     const float *__restrict__ v, float S, float alpha, float eta,
                                                     float *__restrict__ dout, double *__restrict__ part) {
    size_t i;
    int z, x, y;
    const bool ok = voxel(dm, i, z, x, y);
    double sq = 0.0;
    if (ok) {
        const float f0 = u[i], v0 = v[i], s0 = v0 * v0 + eta;
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            int a, bb, c;
            koff(k, a, bb, c);
            const float fp = at3(u, dm, z + a, x + bb, y + c, 0.0f);
            const float fm = at3(u, dm, z - a, x - bb, y - c, 0.0f);
            const float vm = at3(v, dm, z - a, x - bb, y - c, 1.0f);
            acc += kwf(a, bb, c) * (s0 * (fp - f0) - (vm * vm + eta) * (f0 - fm));
        }
        const float dd = S * b[i] + alpha * acc;
        dout[i] = dd;
        sq = (double)dd * (double)dd;
    }
    const double t = block_sum(sq);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// part = pair sum of p (zero outside) with weights (v^2 + eta) when WV, 1
// otherwise.
template <bool WV>
__global__ __launch_bounds__(kBlock) void k_cat_pairs(Dims dm, const float *__restrict__ p,
                                                      const float *__restrict__ v, float eta,
                                                      double *__restrict__ part) {
    size_t i;
    int z, x, y;
    const bool ok = voxel(dm, i, z, x, y);
    double acc = 0.0;
    if (ok) {
        const float p0 = p[i];
        const double s0 = WV ? (double)v[i] * v[i] + eta : 1.0, so = WV ? 1.0 + eta : 1.0;
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            int a, b, c;
            koff(k, a, b, c);
            const double w = kwf(a, b, c);
            const double q = (double)at3(p, dm, z + a, x + b, y + c, 0.0f) - (double)p0;
            acc += w * s0 * q * q;
            if (!in3(dm, z - a, x - b, y - c)) acc += w * so * (double)p0 * p0;
        }
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// u = clamp(u + c d), c = sc[1] / (S sc[2] + alpha sc[3]).
__global__ __launch_bounds__(kBlock) void k_cat_update(size_t n, float *__restrict__ u, const float *__restrict__ d,
                                                       const double *__restrict__ sc, double S, double alpha,
                                                       double lo, double hi) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double den = S * sc[2] + alpha * sc[3];
    const double c = den > 0 ? sc[1] / den : 0.0;
    double x = (double)u[i] + c * (double)d[i];
    x = x < lo ? lo : x;
    x = x > hi ? hi : x;
    u[i] = (float)x;
}

// n(x) = sum_k w_k (u(x + k) - u(x))^2.
__global__ __launch_bounds__(kBlock) void k_cat_edge_n(Dims dm, const float *__restrict__ u, float *__restrict__ nn) {
    size_t i;
    int z, x, y;
    if (!voxel(dm, i, z, x, y)) return;
    const float f0 = u[i];
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 13; ++k) {
        int a, b, c;
        koff(k, a, b, c);
        const float q = at3(u, dm, z + a, x + b, y + c, 0.0f) - f0;
        acc += kwf(a, b, c) * q * q;
    }
    nn[i] = acc;
}

// dv = -alpha v n + beta eps lap(v) - beta / (4 eps) (v - 1);
// part0 = sum dv^2, part1 = sum n dv^2.
__global__ __launch_bounds__(kBlock) void k_cat_edge_dir(Dims dm, const float *__restrict__ v,
                                                         const float *__restrict__ nn, float alpha, float beta,
                                                         float eps, float *__restrict__ dv, double *__restrict__ p0,
                                                         double *__restrict__ p1) {
    size_t i;
    int z, x, y;
    const bool ok = voxel(dm, i, z, x, y);
    double a0 = 0.0, a1 = 0.0;
    if (ok) {
        const float v0 = v[i];
        float lap = 0.0f;
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            int a, b, c;
            koff(k, a, b, c);
            lap += kwf(a, b, c) * ((at3(v, dm, z + a, x + b, y + c, 1.0f) - v0) -
                                   (v0 - at3(v, dm, z - a, x - b, y - c, 1.0f)));
        }
        const float d = -alpha * v0 * nn[i] + beta * eps * lap - beta / (4.0f * eps) * (v0 - 1.0f);
        dv[i] = d;
        a0 = (double)d * d;
        a1 = (double)nn[i] * d * d;
    }
    double t = block_sum(a0);
    if (threadIdx.x == 0) p0[blockIdx.x] = t;
    __syncthreads();
    t = block_sum(a1);
    if (threadIdx.x == 0) p1[blockIdx.x] = t;
}

// v = clamp(v + c dv, 0, 1), c = sc[4] / (alpha sc[5] + beta eps sc[6] + beta / (4 eps) sc[4]).
__global__ __launch_bounds__(kBlock) void k_cat_edge_update(size_t n, float *__restrict__ v,
                                                            const float *__restrict__ dv,
                                                            const double *__restrict__ sc, double alpha, double beta,
                                                            double eps) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double den = alpha * sc[5] + beta * eps * sc[6] + beta / (4.0 * eps) * sc[4];
    const double c = den > 0 ? sc[4] / den : 0.0;
    double x = (double)v[i] + c * (double)dv[i];
    x = x < 0.0 ? 0.0 : x;
    x = x > 1.0 ? 1.0 : x;
    v[i] = (float)x;
}

// slot = (acc ? slot : 0) + sum part[0 .. n), fixed order, one workgroup.
__global__ __launch_bounds__(kBlock) void k_cat_reduce(const double *__restrict__ part, size_t n,
                                                       double *__restrict__ slot, int acc) {
    // eight loads in flight per lane; the order of the additions is fixed
    double s = 0.0;
    for (size_t base = 0; base < n; base += (size_t)kBlock * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const size_t i = base + (size_t)u * kBlock + threadIdx.x;
            v[u] = i < n ? part[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    const double t = block_sum(s);
    if (threadIdx.x == 0) *slot = acc ? *slot + t : t;
}

__global__ void k_cat_zero(double *slot) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *slot = 0.0;
}

}  // namespace

void cone_at_solve(mi_context *ctx, const mi_cone_geometry &g, const mi_at_params &p, const float *sino_h,
                   float *u_h, float *v_h, mi_report *rep, bool dev) {
    MI_REQUIRE(sino_h && u_h && v_h, MI_E_ARG, "null argument");
    MI_REQUIRE(p.precision == MI_FP32, MI_E_ARG, "the cone-beam solver runs in FP32 (FP64 reductions)");
    MI_REQUIRE(p.n_subsets >= 1 && p.n_subsets <= g.n_views && p.max_iters >= 0 && p.iter_edge >= 0, MI_E_ARG,
               "bad AT parameters");
    MI_REQUIRE(p.median_every == 0, MI_E_ARG, "the cone-beam solver has no median step (median_every must be 0)");
    MI_REQUIRE(std::isfinite(p.alpha) && std::isfinite(p.beta) && p.epsilon > 0, MI_E_ARG,
               "alpha, beta finite and epsilon > 0 required");
    ConeTables &t = cone_tables(ctx, g);
    const hipStream_t st = ctx->stream;
    const Dims dm{g.nz, g.nx, g.ny};
    const size_t nvox = (size_t)g.nx * g.ny * g.nz, nsino = (size_t)g.n_views * g.n_rows * g.n_cols;
    const int S = p.n_subsets;

    // Subset view lists, concatenated.
    const std::vector<int> order = subset_order(S);
    std::vector<int> lists, off(S + 1, 0);
    for (int s = 0; s < S; ++s) {
        for (int a = 0; a < g.n_views; ++a)
            if (a % S == order[s]) lists.push_back(a);
        off[s + 1] = (int)lists.size();
    }
    int nv_max = 0;
    for (int s = 0; s < S; ++s) nv_max = std::max(nv_max, off[s + 1] - off[s]);

    const size_t vblocks = (nvox + kBlock - 1) / kBlock;
    const size_t fpb = (size_t)((g.n_cols + 15) / 16) * ((g.n_rows + 15) / 16) * nv_max;
    // dev: the caller's device buffers are used in place; otherwise host
    // copies in device buffers of the solve.
    DevBuf<float> sino(dev ? 0 : nsino), u(dev ? 0 : nvox), v(dev ? 0 : nvox);
    DevBuf<float> q(nsino), b(nvox), d(nvox), nn(nvox), dv(nvox);
    const float *SN = dev ? sino_h : sino.p;
    float *U = dev ? u_h : u.p, *V = dev ? v_h : v.p;
    DevBuf<double> part0(std::max(vblocks, fpb)), part1(vblocks), sc(8);
    DevBuf<int> dlists(lists.size());
    MI_HIP(hipMemcpyAsync(dlists.p, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice, st));
    if (!dev) {
        MI_HIP(hipMemcpyAsync(sino.p, sino_h, sino.bytes(), hipMemcpyHostToDevice, st));
        MI_HIP(hipMemcpyAsync(u.p, u_h, u.bytes(), hipMemcpyHostToDevice, st));
        MI_HIP(hipMemcpyAsync(v.p, v_h, v.bytes(), hipMemcpyHostToDevice, st));
    }

    const float eta = (float)(0.01 * p.epsilon * p.epsilon);
    const bool timed_fp = ctx->stats.on && ((ctx->stats.mask >> MI_K_CONE_FP) & 1u);
    const bool timed_bp = ctx->stats.on && ((ctx->stats.mask >> MI_K_CONE_BP) & 1u);
    if (timed_fp || timed_bp) cone_nnz(st, t);
    const dim3 vg((unsigned)vblocks), blk(kBlock);
    auto reduce = [&](const double *part, size_t n, int slot, int acc) {
        hipLaunchKernelGGL(k_cat_reduce, dim3(1), blk, 0, st, part, n, sc.p + slot, acc);
    };

    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    MI_HIP(hipEventRecord(e0, st));
    double df = 0.0, df_prev = 0.0, rel = 0.0;
    int it = 0, converged = 0;
    for (it = 1; it <= p.max_iters; ++it) {
        hipLaunchKernelGGL(k_cat_zero, dim3(1), dim3(64), 0, st, sc.p);
        for (int s = 0; s < S; ++s) {
            const int *views = dlists.p + off[s];
            const int nv = off[s + 1] - off[s];
            const double frac = (double)nv / g.n_views;
            size_t nb;
            {
                LaunchTimer lt(ctx, MI_K_CONE_FP, timed_fp ? frac * cone_fp_bytes(t) : 0.0, true);
                nb = cone_fp_views(st, t, views, nv, CFP_RESID, U, SN, q.p, part0.p);
            }
            reduce(part0.p, nb, 0, 1);
            {
                LaunchTimer lt(ctx, MI_K_CONE_BP, timed_bp ? frac * cone_bp_bytes(t) : 0.0, true);
                cone_bp_views(st, t, views, nv, q.p, b.p);
            }
            hipLaunchKernelGGL(k_cat_grad, vg, blk, 0, st, dm, b.p, U, V, (float)S, (float)p.alpha, eta, d.p,
                               part1.p);
            reduce(part1.p, vblocks, 1, 0);
            {
                LaunchTimer lt(ctx, MI_K_CONE_FP, timed_fp ? frac * cone_fp_bytes(t) : 0.0, true);
                nb = cone_fp_views(st, t, views, nv, CFP_SQ, d.p, nullptr, nullptr, part0.p);
            }
            reduce(part0.p, nb, 2, 0);
            hipLaunchKernelGGL(k_cat_pairs<true>, vg, blk, 0, st, dm, d.p, V, eta, part1.p);
            reduce(part1.p, vblocks, 3, 0);
            hipLaunchKernelGGL(k_cat_update, vg, blk, 0, st, nvox, U, d.p, sc.p, (double)S, p.alpha, p.lower,
                               p.upper);
        }
        if (p.iter_edge > 0) {
            hipLaunchKernelGGL(k_cat_edge_n, vg, blk, 0, st, dm, U, nn.p);
            for (int e = 0; e < p.iter_edge; ++e) {
                hipLaunchKernelGGL(k_cat_edge_dir, vg, blk, 0, st, dm, V, nn.p, (float)p.alpha, (float)p.beta,
                                   (float)p.epsilon, dv.p, part0.p, part1.p);
                reduce(part0.p, vblocks, 4, 0);
                reduce(part1.p, vblocks, 5, 0);
                hipLaunchKernelGGL(k_cat_pairs<false>, vg, blk, 0, st, dm, dv.p, nullptr, 0.0f, part0.p);
                reduce(part0.p, vblocks, 6, 0);
                hipLaunchKernelGGL(k_cat_edge_update, vg, blk, 0, st, nvox, V, dv.p, sc.p, p.alpha, p.beta,
                                   p.epsilon);
            }
        }
        MI_HIP(hipGetLastError());
        MI_HIP(hipMemcpyAsync(&df, sc.p, sizeof(double), hipMemcpyDeviceToHost, st));
        MI_HIP(hipStreamSynchronize(st));
        if (it > 1) {
            rel = std::fabs(df - df_prev) / (df_prev > 0 ? df_prev : 1.0);
            if (std::fabs(df - df_prev) < p.rel_tol * df_prev) {
                converged = 1;
                break;
            }
        }
        df_prev = df;
    }
    const int ran = it > p.max_iters ? p.max_iters : it;
    MI_HIP(hipEventRecord(e1, st));
    MI_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    MI_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!dev) {
        MI_HIP(hipMemcpyAsync(u_h, u.p, u.bytes(), hipMemcpyDeviceToHost, st));
        MI_HIP(hipMemcpyAsync(v_h, v.p, v.bytes(), hipMemcpyDeviceToHost, st));
    }
    MI_HIP(hipStreamSynchronize(st));
    stats_collect(ctx);
    if (rep) {
        *rep = mi_report{};
        rep->iterations = ran;
        rep->converged = converged;
        rep->data_fidelity = df;
        rep->rel_change = rel;
        rep->elapsed_ms = ms;
        rep->nnz = t.nnz;
    }
}

}  // namespace mi
