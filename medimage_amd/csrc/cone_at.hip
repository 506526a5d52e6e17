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

__device__ __forceinline__ double block_sum(double x) {
    __shared__ double red[kBlock / kWave];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// The stencil kernels march over z.  A workgroup owns a tile of kTx x kTy
// columns (one wave per x row, one lane per y) over up to kZrun slices; each
// plane of the tile and its one-voxel halo is staged once through LDS, and a
// lane keeps the 3 x 3 neighbourhood of the last planes in registers
// (rotating the arrays, no copies), so a volume is read about once per
// kernel.  A lane accumulates its column's double sums over the slices and
// the workgroup writes one partial.
constexpr int kTy = kWave, kTx = kBlock / kWave, kPy = kTy + 2, kPlane = (kTx + 2) * kPy, kZrun = 32, kRun = 64;
static_assert(kPlane > kBlock && kPlane <= 2 * kBlock, "a lane stages one or two entries of a plane");

// Workgroups of the stencil kernels for the volume (host side).
size_t tile_blocks(const Dims &d) {
    return (size_t)((d.ny + kTy - 1) / kTy) * ((d.nx + kTx - 1) / kTx) * ((d.nz + kZrun - 1) / kZrun);
}

struct Tile {
    int x0, y0, z0, z1;  // tile origin; slices [z0, z1)
    int x, y;            // this lane's column
    bool in;             // the column lies in the volume
    int e1;              // second LDS entry this lane stages (-1: none; the first is threadIdx.x)
    int off0, off1;      // in-plane offsets x * ny + y of the two entries (-1: outside the volume)
};

__device__ __forceinline__ int plane_off(const Dims &d, const Tile &t, int e) {
    const int r = e / kPy, c = e - r * kPy, x = t.x0 + r - 1, y = t.y0 + c - 1;
    return ((unsigned)x < (unsigned)d.nx && (unsigned)y < (unsigned)d.ny) ? x * d.ny + y : -1;
}

__device__ __forceinline__ Tile tile_of(const Dims &d) {
    const int nty = (d.ny + kTy - 1) / kTy, ntx = (d.nx + kTx - 1) / kTx;
    // tiles are numbered y fastest, then x, then slices; workgroups go to the
    // 8 XCDs round-robin, and are dealt to them in runs of kRun consecutive
    // tiles, so that neighbouring tiles, which share halo lines, mostly run in
    // one L2 while all XCDs work on the same slices
    int b = blockIdx.x;
    if (b < (int)(gridDim.x & ~(8u * kRun - 1))) {
        const int xcd = b & 7, j = b >> 3;
        b = ((j / kRun) * 8 + xcd) * kRun + j % kRun;
    }
    const int by = b % nty, bx = (b / nty) % ntx, bz = b / (nty * ntx);
    Tile t;
    t.y0 = by * kTy;
    t.x0 = bx * kTx;
    t.z0 = bz * kZrun;
    t.z1 = min(t.z0 + kZrun, d.nz);
    t.y = t.y0 + (int)(threadIdx.x % kTy);
    t.x = t.x0 + (int)(threadIdx.x / kTy);
    t.in = t.x < d.nx && t.y < d.ny;
    t.e1 = (int)threadIdx.x + kBlock < kPlane ? (int)threadIdx.x + kBlock : -1;
    t.off0 = plane_off(d, t, threadIdx.x);
    t.off1 = t.e1 >= 0 ? plane_off(d, t, t.e1) : -1;
    return t;
}

// The two entries of a plane of the tile that a lane stages, fetched one
// slice ahead of their use so that their latency overlaps a slice of work.
struct Ahead {
    float a, b;
};

// f(z, ...) at the lane's staged entries, pad outside the volume.
__device__ __forceinline__ Ahead fetch(const float *__restrict__ f, const Dims &d, const Tile &t, int z, float pad) {
    const bool zin = (unsigned)z < (unsigned)d.nz;
    const int zo = zin ? z * d.nx * d.ny : 0;
    Ahead r;
    r.a = (zin && t.off0 >= 0) ? f[zo + t.off0] : pad;
    r.b = (zin && t.off1 >= 0) ? f[zo + t.off1] : pad;
    return r;
}

// slot[r][c] = the fetched f(z, x0 + r - 1, y0 + c - 1); SQ stores f^2 + eta.
template <bool SQ>
__device__ __forceinline__ void put(float *slot, const Tile &t, const Ahead &r, float eta) {
    slot[threadIdx.x] = SQ ? r.a * r.a + eta : r.a;
    if (t.e1 >= 0) slot[t.e1] = SQ ? r.b * r.b + eta : r.b;
}

// n[b + 1][c + 1] = the staged value at (x + b, y + c) of this lane's column.
__device__ __forceinline__ void nbhd(float (&n)[3][3], const float *slot) {
    const float *p = slot + (int)(threadIdx.x / kTy + 1) * kPy + (int)(threadIdx.x % kTy) + 1;
#pragma unroll
    for (int b = -1; b <= 1; ++b)
#pragma unroll
        for (int c = -1; c <= 1; ++c) n[b + 1][c + 1] = p[b * kPy + c];
}

__device__ __forceinline__ int vidx(const Dims &d, const Tile &t, int z) { return (z * d.nx + t.x) * d.ny + t.y; }

using Nb = float[3][3];

// d = S b + alpha r(u; v); part = sum d^2.
// The planes a step stages are fetched three steps ahead (one register ring
// slot per unrolled step), so that their latency overlaps three slices of
// work.
__global__ __launch_bounds__(kBlock) void k_cat_grad(Dims dm, const float *__restrict__ b, const float *__restrict__ u,
                                                     const float *__restrict__ v, float S, float alpha, float eta,
                                                     float *__restrict__ dout, double *__restrict__ part) {
    __shared__ float lu[4][kPlane], ls[4][kPlane];  // plane p of u, v^2 + eta in slot (p + 1) & 3
    const Tile t = tile_of(dm);
    // ring slot of step z: u(z + 1), v(z), b(z)
    struct Ring {
        Ahead u, v;
        float b;
    };
    auto load = [&](Ring &r, int z) {
        r.u = fetch(u, dm, t, z + 1, 0.0f);
        r.v = fetch(v, dm, t, z, 1.0f);
        r.b = t.in ? b[vidx(dm, t, z)] : 0.0f;
    };
    put<false>(lu[t.z0 & 3], t, fetch(u, dm, t, t.z0 - 1, 0.0f), eta);
    put<false>(lu[(t.z0 + 1) & 3], t, fetch(u, dm, t, t.z0, 0.0f), eta);
    put<true>(ls[t.z0 & 3], t, fetch(v, dm, t, t.z0 - 1, 1.0f), eta);
    Ring R[3];
    load(R[0], t.z0);
    if (t.z0 + 1 < t.z1) load(R[1], t.z0 + 1);
    if (t.z0 + 2 < t.z1) load(R[2], t.z0 + 2);
    __syncthreads();
    Nb U0, U1, U2, S0, S1, S2;
    nbhd(U0, lu[t.z0 & 3]);
    nbhd(U1, lu[(t.z0 + 1) & 3]);
    nbhd(S0, ls[t.z0 & 3]);
    double sq = 0.0;
    // slice z from u on z - 1, z, z + 1 and v^2 + eta on z - 1, z; stages
    // u(z + 1) into up and s(z) into sc, refills the ring slot for z + 3
    auto step = [&](int z, const Nb &um, const Nb &uc, Nb &up, const Nb &sm, Nb &sc, Ring &r) {
        put<false>(lu[(z + 2) & 3], t, r.u, eta);
        put<true>(ls[(z + 1) & 3], t, r.v, eta);
        const float bz = r.b;
        if (z + 3 < t.z1) load(r, z + 3);
        __syncthreads();
        nbhd(up, lu[(z + 2) & 3]);
        nbhd(sc, ls[(z + 1) & 3]);
        if (t.in) {
            const float f0 = uc[1][1], s0 = sc[1][1];
            // per weight class |k|^2 = 1, 2, 3: sum of f(x + k) - f(x) and of
            // s(x - k) (f(x) - f(x - k))
            float A[3] = {0.0f, 0.0f, 0.0f}, B[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < 13; ++k) {
                int a, bb, c;
                koff(k, a, bb, c);
                const int n = a * a + bb * bb + c * c - 1;
                const float fp = a ? up[1 + bb][1 + c] : uc[1 + bb][1 + c];
                const float fm = a ? um[1 - bb][1 - c] : uc[1 - bb][1 - c];
                const float sk = a ? sm[1 - bb][1 - c] : sc[1 - bb][1 - c];
                A[n] += fp - f0;
                B[n] = fmaf(sk, f0 - fm, B[n]);
            }
            const float acc = (fmaf(s0, A[0], -B[0]) + 0.5f * fmaf(s0, A[1], -B[1])) +
                              (1.0f / 3.0f) * fmaf(s0, A[2], -B[2]);
            const float dd = S * bz + alpha * acc;
            dout[vidx(dm, t, z)] = dd;
            sq += (double)dd * (double)dd;
        }
    };
    for (int z = t.z0; z < t.z1; z += 3) {
        step(z, U0, U1, U2, S0, S1, R[0]);
        if (z + 1 < t.z1) step(z + 1, U1, U2, U0, S1, S2, R[1]);
        if (z + 2 < t.z1) step(z + 2, U2, U0, U1, S2, S0, R[2]);
    }
    const double tot = block_sum(sq);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// part = pair sum of p (zero outside) with weights (v^2 + eta) when WV, 1
// otherwise.  The squared differences of a weight class are summed in float
// and weighted in float (13 terms per voxel); the lane's sum over its
// slices and the workgroup's sum are double.  Planes are fetched two steps
// ahead.
template <bool WV>
__global__ __launch_bounds__(kBlock) void k_cat_pairs(Dims dm, const float *__restrict__ p,
                                                      const float *__restrict__ v, float eta,
                                                      double *__restrict__ part) {
    __shared__ float lp[2][kPlane];  // plane z in slot z & 1
    const Tile t = tile_of(dm);
    // ring slot of step z: p(z + 1) and v(z)
    struct Ring {
        Ahead p;
        float v;
    };
    auto load = [&](Ring &r, int z) {
        r.p = fetch(p, dm, t, z + 1, 0.0f);
        r.v = WV && t.in ? v[vidx(dm, t, z)] : 0.0f;
    };
    put<false>(lp[t.z0 & 1], t, fetch(p, dm, t, t.z0, 0.0f), 0.0f);
    Ring R[2];
    load(R[0], t.z0);
    if (t.z0 + 1 < t.z1) load(R[1], t.z0 + 1);
    // The pair weights of x whose other voxel x - k lies outside: per weight
    // class |k|^2 = 1, 2, 3 the count of such offsets (b = +-1 leaves through
    // the x faces, c = +-1 through the y faces, and on slice 0 a = 1 through
    // z = 0), summed with the weights; wout[1] on slice 0, wout[0] elsewhere.
    double wout[2];
    {
        const bool xlo = t.x == 0, xhi = t.x == dm.nx - 1, ylo = t.y == 0, yhi = t.y == dm.ny - 1;
        int c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            int a, b, c;
            koff(k, a, b, c);
            const int n = a * a + b * b + c * c - 1;
            const bool xy = (b == 1 && xlo) || (b == -1 && xhi) || (c == 1 && ylo) || (c == -1 && yhi);
            c0[n] += xy;
            c1[n] += xy || a == 1;
        }
        wout[0] = (c0[0] + 0.5 * c0[1]) + (1.0 / 3.0) * c0[2];
        wout[1] = (c1[0] + 0.5 * c1[1]) + (1.0 / 3.0) * c1[2];
    }
    __syncthreads();
    Nb P0, P1;
    nbhd(P0, lp[t.z0 & 1]);
    double acc = 0.0, acco = 0.0;
    // slice z from p on z, z + 1; stages p(z + 1) into pn, refills the ring
    // slot for z + 2
    auto step = [&](int z, const Nb &pc, Nb &pn, Ring &r) {
        put<false>(lp[(z + 1) & 1], t, r.p, 0.0f);
        const float vz = r.v;
        if (z + 2 < t.z1) load(r, z + 2);
        __syncthreads();
        nbhd(pn, lp[(z + 1) & 1]);
        if (t.in) {
            const float p0 = pc[1][1];
            float g[3] = {0.0f, 0.0f, 0.0f};   // per weight class: sum of (p(x + k) - p(x))^2
#pragma unroll
            for (int k = 0; k < 13; ++k) {
                int a, b, c;
                koff(k, a, b, c);
                const int n = a * a + b * b + c * c - 1;
                const float q = (a ? pn[1 + b][1 + c] : pc[1 + b][1 + c]) - p0;
                g[n] = fmaf(q, q, g[n]);
            }
            // the voxel's weighted pair sum in float (13 terms), added in double
            const float gs = fmaf(1.0f / 3.0f, g[2], fmaf(0.5f, g[1], g[0]));
            acc += WV ? (double)(fmaf(vz, vz, eta) * gs) : (double)gs;
            const double wz = wout[z == 0 ? 1 : 0];
            if (wz != 0.0) acco += wz * ((double)p0 * p0);   // the pairs with an outside voxel
        }
    };
    for (int z = t.z0; z < t.z1; z += 2) {
        step(z, P0, P1, R[0]);
        if (z + 1 < t.z1) step(z + 1, P1, P0, R[1]);
    }
    // outside voxels have v = 1: weight 1 + eta
    const double tot = block_sum(acc + (WV ? 1.0 + (double)eta : 1.0) * acco);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__device__ __forceinline__ float clamp_step(float u, float d, double c, double lo, double hi) {
    double x = (double)u + c * (double)d;
    x = x < lo ? lo : x;
    x = x > hi ? hi : x;
    return (float)x;
}

// u = clamp(u + c d) on four voxels per lane, with 16-byte accesses when both
// volumes are 16-byte aligned (vec).
__device__ __forceinline__ void step4(size_t n, float *__restrict__ u, const float *__restrict__ d, double c,
                                      double lo, double hi, bool vec) {
    const size_t i = ((size_t)blockIdx.x * kBlock + threadIdx.x) * 4;
    if (i >= n) return;
    if (vec && i + 4 <= n) {
        float4 a = *reinterpret_cast<const float4 *>(u + i);
        const float4 b = *reinterpret_cast<const float4 *>(d + i);
        a.x = clamp_step(a.x, b.x, c, lo, hi);
        a.y = clamp_step(a.y, b.y, c, lo, hi);
        a.z = clamp_step(a.z, b.z, c, lo, hi);
        a.w = clamp_step(a.w, b.w, c, lo, hi);
        *reinterpret_cast<float4 *>(u + i) = a;
    } else {
        for (size_t j = i; j < n && j < i + 4; ++j) u[j] = clamp_step(u[j], d[j], c, lo, hi);
    }
}

// u = clamp(u + c d), c = sc[1] / (S sc[2] + alpha sc[3]).
__global__ __launch_bounds__(kBlock) void k_cat_update(size_t n, float *__restrict__ u, const float *__restrict__ d,
                                                       const double *__restrict__ sc, double S, double alpha,
                                                       double lo, double hi, bool vec) {
    const double den = S * sc[2] + alpha * sc[3];
    step4(n, u, d, den > 0 ? sc[1] / den : 0.0, lo, hi, vec);
}

// n(x) = sum_k w_k (u(x + k) - u(x))^2.
__global__ __launch_bounds__(kBlock) void k_cat_edge_n(Dims dm, const float *__restrict__ u, float *__restrict__ nn) {
    __shared__ float lu[2][kPlane];  // plane z in slot z & 1
    const Tile t = tile_of(dm);
    put<false>(lu[t.z0 & 1], t, fetch(u, dm, t, t.z0, 0.0f), 0.0f);
    // ring slot of step z: u(z + 1), fetched two steps ahead
    Ahead R[2] = {fetch(u, dm, t, t.z0 + 1, 0.0f), fetch(u, dm, t, t.z0 + 2, 0.0f)};
    __syncthreads();
    Nb U0, U1;
    nbhd(U0, lu[t.z0 & 1]);
    // slice z from u on z, z + 1; stages u(z + 1) into un, refills the ring
    // slot for z + 2
    auto step = [&](int z, const Nb &uc, Nb &un, Ahead &r) {
        put<false>(lu[(z + 1) & 1], t, r, 0.0f);
        if (z + 2 < t.z1) r = fetch(u, dm, t, z + 3, 0.0f);
        __syncthreads();
        nbhd(un, lu[(z + 1) & 1]);
        if (t.in) {
            const float f0 = uc[1][1];
            float g[3] = {0.0f, 0.0f, 0.0f};  // per weight class |k|^2 = 1, 2, 3
#pragma unroll
            for (int k = 0; k < 13; ++k) {
                int a, b, c;
                koff(k, a, b, c);
                const float q = (a ? un[1 + b][1 + c] : uc[1 + b][1 + c]) - f0;
                g[a * a + b * b + c * c - 1] = fmaf(q, q, g[a * a + b * b + c * c - 1]);
            }
            nn[vidx(dm, t, z)] = (g[0] + 0.5f * g[1]) + (1.0f / 3.0f) * g[2];
        }
    };
    for (int z = t.z0; z < t.z1; z += 2) {
        step(z, U0, U1, R[0]);
        if (z + 1 < t.z1) step(z + 1, U1, U0, R[1]);
    }
}

// dv = -alpha v n + beta eps lap(v) - beta / (4 eps) (v - 1);
// part0 = sum dv^2, part1 = sum n dv^2.
__global__ __launch_bounds__(kBlock) void k_cat_edge_dir(Dims dm, const float *__restrict__ v,
                                                         const float *__restrict__ nn, float alpha, float beta,
                                                         float eps, float *__restrict__ dv, double *__restrict__ p0,
                                                         double *__restrict__ p1) {
    __shared__ float lv[4][kPlane];  // plane p in slot (p + 1) & 3
    const Tile t = tile_of(dm);
    // ring slot of step z: v(z + 1) and n(z), fetched three steps ahead
    struct Ring {
        Ahead v;
        float n;
    };
    auto load = [&](Ring &r, int z) {
        r.v = fetch(v, dm, t, z + 1, 1.0f);
        r.n = t.in ? nn[vidx(dm, t, z)] : 0.0f;
    };
    put<false>(lv[t.z0 & 3], t, fetch(v, dm, t, t.z0 - 1, 1.0f), 0.0f);
    put<false>(lv[(t.z0 + 1) & 3], t, fetch(v, dm, t, t.z0, 1.0f), 0.0f);
    Ring R[3];
    load(R[0], t.z0);
    if (t.z0 + 1 < t.z1) load(R[1], t.z0 + 1);
    if (t.z0 + 2 < t.z1) load(R[2], t.z0 + 2);
    __syncthreads();
    Nb V0, V1, V2;
    nbhd(V0, lv[t.z0 & 3]);
    nbhd(V1, lv[(t.z0 + 1) & 3]);
    double a0 = 0.0, a1 = 0.0;
    // slice z from v on z - 1, z, z + 1; stages v(z + 1) into vp, refills
    // the ring slot for z + 3
    auto step = [&](int z, const Nb &vm, const Nb &vc, Nb &vp, Ring &r) {
        put<false>(lv[(z + 2) & 3], t, r.v, 0.0f);
        const float n = r.n;
        if (z + 3 < t.z1) load(r, z + 3);
        __syncthreads();
        nbhd(vp, lv[(z + 2) & 3]);
        if (t.in) {
            const float v0 = vc[1][1];
            float lap = 0.0f;
#pragma unroll
            for (int k = 0; k < 13; ++k) {
                int a, b, c;
                koff(k, a, b, c);
                const float fp = a ? vp[1 + b][1 + c] : vc[1 + b][1 + c];
                const float fm = a ? vm[1 - b][1 - c] : vc[1 - b][1 - c];
                lap += kwf(a, b, c) * ((fp - v0) - (v0 - fm));
            }
            const float d = -alpha * v0 * n + beta * eps * lap - beta / (4.0f * eps) * (v0 - 1.0f);
            dv[vidx(dm, t, z)] = d;
            a0 += (double)d * d;
            a1 += (double)n * d * d;
        }
    };
    for (int z = t.z0; z < t.z1; z += 3) {
        step(z, V0, V1, V2, R[0]);
        if (z + 1 < t.z1) step(z + 1, V1, V2, V0, R[1]);
        if (z + 2 < t.z1) step(z + 2, V2, V0, V1, R[2]);
    }
    double s = block_sum(a0);
    if (threadIdx.x == 0) p0[blockIdx.x] = s;
    __syncthreads();
    s = block_sum(a1);
    if (threadIdx.x == 0) p1[blockIdx.x] = s;
}

// v = clamp(v + c dv, 0, 1), c = sc[4] / (alpha sc[5] + beta eps sc[6] + beta / (4 eps) sc[4]).
__global__ __launch_bounds__(kBlock) void k_cat_edge_update(size_t n, float *__restrict__ v,
                                                            const float *__restrict__ dv,
                                                            const double *__restrict__ sc, double alpha, double beta,
                                                            double eps, bool vec) {
    const double den = alpha * sc[5] + beta * eps * sc[6] + beta / (4.0 * eps) * sc[4];
    step4(n, v, dv, den > 0 ? sc[4] / den : 0.0, 0.0, 1.0, vec);
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

    const size_t tblocks = tile_blocks(dm);
    const size_t fpb = (size_t)((g.n_cols + 15) / 16) * ((g.n_rows + 15) / 16) * nv_max;
    // dev: the caller's device buffers are used in place; otherwise host
    // copies in device buffers of the solve.
    DevBuf<float> sino(dev ? 0 : nsino), u(dev ? 0 : nvox), v(dev ? 0 : nvox);
    DevBuf<float> q(nsino), b(nvox), d(nvox), nn(nvox), dv(nvox);
    const float *SN = dev ? sino_h : sino.p;
    float *U = dev ? u_h : u.p, *V = dev ? v_h : v.p;
    DevBuf<double> part0(std::max(tblocks, fpb)), part1(tblocks), sc(8);
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
    const dim3 v4g((unsigned)((nvox + 4 * kBlock - 1) / (4 * kBlock))), tg((unsigned)tblocks), blk(kBlock);
    const bool vec_u = (((uintptr_t)U | (uintptr_t)d.p) & 15) == 0, vec_v = (((uintptr_t)V | (uintptr_t)dv.p) & 15) == 0;
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
            hipLaunchKernelGGL(k_cat_grad, tg, blk, 0, st, dm, b.p, U, V, (float)S, (float)p.alpha, eta, d.p,
                               part1.p);
            reduce(part1.p, tblocks, 1, 0);
            {
                LaunchTimer lt(ctx, MI_K_CONE_FP, timed_fp ? frac * cone_fp_bytes(t) : 0.0, true);
                nb = cone_fp_views(st, t, views, nv, CFP_SQ, d.p, nullptr, nullptr, part0.p);
            }
            reduce(part0.p, nb, 2, 0);
            hipLaunchKernelGGL(k_cat_pairs<true>, tg, blk, 0, st, dm, d.p, V, eta, part1.p);
            reduce(part1.p, tblocks, 3, 0);
            hipLaunchKernelGGL(k_cat_update, v4g, blk, 0, st, nvox, U, d.p, sc.p, (double)S, p.alpha, p.lower,
                               p.upper, vec_u);
        }
        if (p.iter_edge > 0) {
            hipLaunchKernelGGL(k_cat_edge_n, tg, blk, 0, st, dm, U, nn.p);
            for (int e = 0; e < p.iter_edge; ++e) {
                hipLaunchKernelGGL(k_cat_edge_dir, tg, blk, 0, st, dm, V, nn.p, (float)p.alpha, (float)p.beta,
                                   (float)p.epsilon, dv.p, part0.p, part1.p);
                reduce(part0.p, tblocks, 4, 0);
                reduce(part1.p, tblocks, 5, 0);
                hipLaunchKernelGGL(k_cat_pairs<false>, tg, blk, 0, st, dm, dv.p, nullptr, 0.0f, part0.p);
                reduce(part0.p, tblocks, 6, 0);
                hipLaunchKernelGGL(k_cat_edge_update, v4g, blk, 0, st, nvox, V, dv.p, sc.p, p.alpha, p.beta,
                                   p.epsilon, vec_v);
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
