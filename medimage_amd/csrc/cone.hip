// 3D cone-beam projectors, industry3d geometry (flat panel, circular orbit).
//
// Rays.  cone() of industry3d/src/tracing.cpp:51-81 with rotate_2d of
// industry3d/src/utility.cpp:82-98, in float: the source (0, -SOD, 0) and
// the panel pixel ((col - HALFDET + 0.5) PIXELSIZE, SDD - SOD,
// (HALFDET - row - 0.5) PIXELSIZE) are rotated by the view angle, divided
// by SAMPLESIZE and shifted by N / 2 (in double, as the reference's
// `+= NX / 2.0` does).  The view's cos / sin come from host tables made
// with the C library's cosf / sinf, which the reference calls.  BEAM
// "Parallel" takes parallel() (tracing.cpp:9-49) instead: start and end share
// the column's x and the row's z, every axis shifted by HALFSIZE.
//
// Forward projection.  One lane per ray walks forward_proj
// (industry3d/src/tracing.cpp:123-307) in float, operation for operation,
// so the (voxel, length) sequence is the reference's bit for bit; the walk
// ends where the reference's does (lambda0 >= lambda_max - 0.05, tested in
// double) or when its voxel leaves the volume (where the reference would
// index outside its arrays; only rays through an edge or corner of the
// volume).  Products vol * len are summed in double in walk order and
// rounded once.  Each wave takes an 8 x 8 tile of one view's panel, so the
// 64 rays of a wave cross neighbouring voxels at each step; workgroups (16
// x 16 pixels) go by band of panel rows.
// Steps are taken in groups of 8: the group's voxel offsets and lengths are
// computed first, its 8 loads issued together, then accumulated in order.
//
// Back projection (gather form, no atomics).  One lane per voxel column
// (x, y) and a chunk of kZc slices along z.  For each view the four corners
// of the column are projected onto the panel: the columns whose rays can
// cross it, and per slice the rows.  For each candidate ray the exact length
// inside the voxel is the slab clip of the ray's parameter interval: the x / y
// slab in double once per (view, column) from a per-geometry table of
// 1 / (d - s), shared by the lane's kZc slices; the z slab per row from a
// per-row 1 / w table (copied to LDS), relative to the column's entry
// parameter so that the float differences keep full precision; times the ray
// length in float.  For the parallel beam the footprint is orthographic and
// a row's ray crosses a slice over its whole x / y chord or not at all.
// sino * length is summed in double, per (view, column, slice) over the rows
// in a register and then into the voxel's sum in LDS, in a fixed order.  A
// back projection over many views runs in chunks of kBpViewChunk views (their
// sinogram rows stay in the caches while every workgroup sweeps them), the
// double sums carried between chunks in a scratch volume.
#include "cone.hpp"

#include <mutex>

#include <cmath>
#include <cstring>
#include <vector>

namespace mi {

namespace {

constexpr int kConeTile = 16;   // forward: 16 x 16 panel pixels per workgroup
constexpr int kConeGroup = 8;   // forward: walk steps per load group
// Back projection: slices per lane.  A lane's per-view work on its column
// (footprint, panel columns, x / y clip) is shared by its kZc slices, whose
// double sums live in LDS (kZc x 2 KiB per workgroup, beside the rows' 1 / w):
// C4 (512^3, 720 views) back projection 644 ms with 6 (18 KiB, 8 workgroups
// per CU), 663 with 7, 674 with 5, 680 with 8 (22 KiB, 7 per CU), 689 with 4.
#ifndef MI_CONE_ZC
#define MI_CONE_ZC 6
#endif
constexpr int kZc = MI_CONE_ZC;
constexpr int kBpViewChunk = 36; // back projection: views per launch (their rows stay in the caches)
constexpr int kBpRowLds = 4096;  // back projection: panels up to this many rows keep 1 / w in LDS

__device__ __forceinline__ void cone_rot(float &x, float &y, float c, float s) {
    float d0 = 0.0f, d1 = 0.0f;
    d0 += c * x;
    d1 += s * x;
    d0 += -s * y;
    d1 += c * y;
    x = d0;
    y = d1;
}

// Source point of a view (voxel units, volume corner at the origin).
__device__ __forceinline__ void cone_source(const ConeDev &g, float c, float s, float &sx, float &sy, float &sz) {
    float X = 0.0f, Y = -g.sod, Z = 0.0f;
    cone_rot(X, Y, c, s);
    X /= g.sample;
    Z /= g.sample;
    Y /= g.sample;
    sx = (float)((double)X + g.nx / 2.0);
    sy = (float)((double)Y + g.ny / 2.0);
    sz = (float)((double)Z + g.nz / 2.0);
}

// In-plane detector point of column col (x, y).
__device__ __forceinline__ void cone_det_xy(const ConeDev &g, float c, float s, int col, float &dx, float &dy) {
    float X = (float)((double)(col - g.halfdet) + 0.5);
    X *= g.pixel;
    float Y = g.sdd - g.sod;
    cone_rot(X, Y, c, s);
    X /= g.sample;
    Y /= g.sample;
    dx = (float)((double)X + g.nx / 2.0);
    dy = (float)((double)Y + g.ny / 2.0);
}

// Height of row `row` of the panel.
__device__ __host__ __forceinline__ float cone_det_z(const ConeDev &g, int row) {
    float Z = (float)((double)(g.halfdet - row) - 0.5);
    Z *= g.pixel;
    Z /= g.sample;
    return (float)((double)Z + g.nz / 2.0);
}

// parallel() of industry3d/src/tracing.cpp:9-49: start and end share the
// panel column's x and the row's z; every axis is shifted by HALFSIZE (a
// float), in float.
__device__ __forceinline__ void par_endpoints(const ConeDev &g, float c, float s, int row, int col, float &sx,
                                              float &sy, float &sz, float &dx, float &dy, float &dz) {
    float X0 = (float)((double)col + 0.5 - g.halfdet);
    X0 *= g.pixel;
    float Y0 = -g.sod;
    float Z0 = (float)((double)(g.halfdet - row) - 0.5);
    Z0 *= g.pixel;
    float X1 = X0, Y1 = g.sdd - g.sod, Z1 = Z0;
    cone_rot(X0, Y0, c, s);
    cone_rot(X1, Y1, c, s);
    sx = X0 / g.sample + g.halfsize;
    sy = Y0 / g.sample + g.halfsize;
    sz = Z0 / g.sample + g.halfsize;
    dx = X1 / g.sample + g.halfsize;
    dy = Y1 / g.sample + g.halfsize;
    dz = Z1 / g.sample + g.halfsize;
}

// End points of ray (view with cos c / sin s, row, col) of either beam.
__device__ __forceinline__ void ray_points(const ConeDev &g, float c, float s, int row, int col, float &sx, float &sy,
                                           float &sz, float &dx, float &dy, float &dz) {
    if (g.parallel) {
        par_endpoints(g, c, s, row, col, sx, sy, sz, dx, dy, dz);
    } else {
        cone_source(g, c, s, sx, sy, sz);
        cone_det_xy(g, c, s, col, dx, dy);
        dz = cone_det_z(g, row);
    }
}

// LAMBDA_X of industry3d/src/tracing.cpp:83-85.
__device__ __forceinline__ float cone_lam(float i, float s, float d, float L) { return L * (i - s) / (d - s); }

// Walk state of forward_proj: lam0 the current parameter, l* the next
// crossings, st* their increments, (vx, vy, vz) the voxel.
struct ConeWalk {
    float lam0, lx, ly, lz, stx, sty, stz, stop;
    int vx, vy, vz, sgx, sgy, sgz;
};

// Entry part of forward_proj (tracing.cpp:142-249).  stop is the smallest
// float T with (double)T >= (double)lambda_max - 5.e-2, so that the float
// test lam0 < T equals the reference's double test.
__device__ __forceinline__ bool cone_walk_init(const ConeDev &g, float sx, float sy, float sz, float dx, float dy,
                                               float dz, ConeWalk &w) {
    const float rx = dx - sx, ry = dy - sy, rz = dz - sz;
    const float L = sqrtf(rx * rx + ry * ry + rz * rz);
    float lmin = 0.0f, lmax = L, l0, lN, t;
    const int sgx = (rx > 0) ? 1 : -1, sgy = (ry > 0) ? 1 : -1, sgz = (rz > 0) ? 1 : -1;
    const float ax = fabsf(rx), ay = fabsf(ry), az = fabsf(rz);
    const float stx = ((double)ax > 1.e-4) ? (L / ax) : 1.e6f;
    const float sty = ((double)ay > 1.e-4) ? (L / ay) : 1.e6f;
    const float stz = ((double)az > 1.e-4) ? (L / az) : 1.e6f;
    int index = 0;
    l0 = cone_lam(0.0f, sx, dx, L);
    lN = cone_lam((float)g.nx, sx, dx, L);
    t = (l0 < lN) ? l0 : lN;
    lmin = (lmin > t) ? lmin : t;
    if (lmin == t) index = 1;
    t = (l0 > lN) ? l0 : lN;
    lmax = (lmax < t) ? lmax : t;
    l0 = cone_lam(0.0f, sy, dy, L);
    lN = cone_lam((float)g.ny, sy, dy, L);
    t = (l0 < lN) ? l0 : lN;
    lmin = (lmin > t) ? lmin : t;
    if (lmin == t) index = 2;
    t = (l0 > lN) ? l0 : lN;
    lmax = (lmax < t) ? lmax : t;
    l0 = cone_lam(0.0f, sz, dz, L);
    lN = cone_lam((float)g.nz, sz, dz, L);
    t = (l0 < lN) ? l0 : lN;
    lmin = (lmin > t) ? lmin : t;
    if (lmin == t) index = 3;
    t = (l0 > lN) ? l0 : lN;
    lmax = (lmax < t) ? lmax : t;
    if (lmin >= lmax) return false;

    const float lam0 = lmin;
    float lx, ly, lz;
    int vx, vy, vz;
    if (index == 1) {
        vx = (sgx == 1) ? 0 : g.nx - 1;
        lx = lam0 + stx;
        vy = (int)(sy + lam0 * ry / L);
        ly = ((double)ay < 1.e-4) ? 1.e6f : cone_lam((float)vy + (float)(sgy > 0), sy, dy, L);
        vz = (int)(sz + lam0 * rz / L);
        lz = ((double)az < 1.e-4) ? 1.e6f : cone_lam((float)vz + (float)(sgz > 0), sz, dz, L);
    } else if (index == 2) {
        vy = (sgy == 1) ? 0 : g.ny - 1;
        ly = lam0 + sty;
        vx = (int)(sx + lam0 * rx / L);
        lx = ((double)ax < 1.e-4) ? 1.e6f : cone_lam((float)vx + (float)(sgx > 0), sx, dx, L);
        vz = (int)(sz + lam0 * rz / L);
        lz = ((double)az < 1.e-4) ? 1.e6f : cone_lam((float)vz + (float)(sgz > 0), sz, dz, L);
    } else {
        vz = (sgz == 1) ? 0 : g.nz - 1;
        lz = lam0 + stz;
        vx = (int)(sx + lam0 * rx / L);
        lx = ((double)ax < 1.e-4) ? 1.e6f : cone_lam((float)vx + (float)(sgx > 0), sx, dx, L);
        vy = (int)(sy + lam0 * ry / L);
        ly = ((double)ay < 1.e-4) ? 1.e6f : cone_lam((float)vy + (float)(sgy > 0), sy, dy, L);
    }
    const double hi = (double)lmax - 5.e-2;
    float stop = (float)hi;
    if ((double)stop < hi) stop = nextafterf(stop, INFINITY);
    w = ConeWalk{lam0, lx, ly, lz, stx, sty, stz, stop, vx, vy, vz, sgx, sgy, sgz};
    return true;
}

// One step when `live` (the walk has not ended): the current voxel's offset
// and length; state unchanged and length 0 otherwise.  The branches of the
// reference loop (tracing.cpp:268-305) become selects.
__device__ __forceinline__ void cone_step(const ConeDev &g, ConeWalk &w, bool &live, unsigned &idx, float &len) {
    // Bitwise, not short-circuit: no branch per step.  `live` never turns
    // back on, so after the end the state may keep moving: its offsets and
    // lengths are masked.
    live = live & (w.lam0 < w.stop) & ((unsigned)w.vx < (unsigned)g.nx) & ((unsigned)w.vy < (unsigned)g.ny) &
           ((unsigned)w.vz < (unsigned)g.nz);
    // The reference's branch choice (x if lx <= ly and lx <= lz, else y if
    // ly <= lz, else z) is "the first axis whose crossing equals the
    // minimum": the crossings are finite (parallel axes carry the 1e6
    // sentinel) and positive, so one three-way minimum and two equality
    // tests give the same choice and the same value.
    const float nxt = fminf(fminf(w.lx, w.ly), w.lz);
    const bool bx = w.lx == nxt;
    const bool by = !bx && w.ly == nxt;
    const bool bz = !bx && !by;
    // get_img_addr (tracing.cpp:119-121); 24-bit products (each factor is
    // below 2^24 and the volume below 2^31 voxels).
    const unsigned at = __umul24(__umul24((unsigned)w.vz, (unsigned)g.nx) + (unsigned)w.vx, (unsigned)g.ny) +
                        (unsigned)w.vy;
    idx = live ? at : 0u;
    len = live ? nxt - w.lam0 : 0.0f;
    w.lam0 = nxt;
    w.lx = bx ? w.lx + w.stx : w.lx;
    w.ly = by ? w.ly + w.sty : w.ly;
    w.lz = bz ? w.lz + w.stz : w.lz;
    w.vx += bx ? w.sgx : 0;
    w.vy += by ? w.sgy : 0;
    w.vz += bz ? w.sgz : 0;
}

__device__ __forceinline__ bool cone_ray_init(const ConeDev &g, const float *cs, int view, int row, int col,
                                              ConeWalk &w) {
    const float c = cs[2 * view], s = cs[2 * view + 1];
    float sx, sy, sz, dx, dy, dz;
    ray_points(g, c, s, row, col, sx, sy, sz, dx, dy, dz);
    return cone_walk_init(g, sx, sy, sz, dx, dy, dz, w);
}

__global__ void k_cone_trace(ConeDev g, const float *cs, int view, int row, int col, int64_t *ind, float *len,
                             int cap, int *numb) {
    if (threadIdx.x != 0) return;
    ConeWalk w;
    int n = 0;
    if (cone_ray_init(g, cs, view, row, col, w)) {
        bool live = true;
        for (;;) {
            unsigned idx;
            float l;
            cone_step(g, w, live, idx, l);
            if (!live) break;
            if (n < cap) {
                ind[n] = idx;
                len[n] = l;
            }
            ++n;
        }
    }
    *numb = n;
}

// Panel pixel of this lane: workgroup = 16 x 16 pixels of one view of the
// list (q-th entry), wave = 8 x 8 of them.  Workgroups are ordered by band of
// 16 panel rows first, then view, then column tile, so that the workgroups in
// flight trace rays through the same few slices of the volume.  Consecutive
// workgroups go to different XCDs (no XCD regrouping here): all eight XCDs
// then work on one band at a time and share its slices through the Infinity
// Cache (368 ms at C4 against 385 ms with each XCD on its own range of
// bands).  Returns false for the padding workgroups of the grid (all lanes
// alike); `ok` is false for lanes outside the panel.
__device__ __forceinline__ bool cone_fp_pixel(const ConeDev &g, int nv, unsigned &tile_id, int &q, int &row,
                                              int &col, bool &ok) {
    const int tc = (g.ncols + kConeTile - 1) / kConeTile, tr = (g.nrows + kConeTile - 1) / kConeTile;
    const unsigned b = blockIdx.x;
    const unsigned per_band = (unsigned)(nv * tc);
    tile_id = b;
    const int band = (int)(b / per_band);
    const unsigned rem = b % per_band;
    q = (int)(rem / tc);
    const int ct = (int)(rem % tc);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    row = band * kConeTile + (wv >> 1) * 8 + (lane >> 3);
    col = ct * kConeTile + (wv & 1) * 8 + (lane & 7);
    ok = row < g.nrows && col < g.ncols;
    return band < tr;
}

// Double sum over a workgroup of 256 lanes in a fixed pairing; lane 0 of
// wave 0 gets the total.
__device__ __forceinline__ double block_sum256(double x) {
    __shared__ double red[kBlock / kWave];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// Forward projection of the views views[0 .. nv) (sinogram [view][row][col]).
//   CFP_WRITE: out = A vol;
//   CFP_RESID: out = meas - A vol (float), part[tile] = sum of out^2;
//   CFP_SQ:    part[tile] = sum of (A vol)^2, nothing written.

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_cone_fp(ConeDev g, const float *cs, const int *views, int nv,
                                                    const float *__restrict__ vol, const float *__restrict__ meas,
                                                    float *__restrict__ out, double *__restrict__ part) {
    unsigned tile_id;
    int q, row, col;
    bool ok;
    if (!cone_fp_pixel(g, nv, tile_id, q, row, col, ok)) return;   // padding workgroup
    double acc = 0.0;
    int view = 0;
    ConeWalk w;
    if (ok && (view = views[q], cone_ray_init(g, cs, view, row, col, w))) {
        bool live = true;
        while (live) {
            unsigned id[kConeGroup];
            float ln[kConeGroup];
            bool lv[kConeGroup];
#pragma unroll
            for (int k = 0; k < kConeGroup; ++k) cone_step(g, w, live, id[k], ln[k]), lv[k] = live;
            float v[kConeGroup];
#pragma unroll
            for (int k = 0; k < kConeGroup; ++k) v[k] = vol[id[k]];
            // The product of two floats is exact in double, so the fused
            // multiply-add rounds once, as acc + (double)v * (double)len does.
#pragma unroll
            for (int k = 0; k < kConeGroup; ++k)
                acc = __builtin_fma((double)(lv[k] ? v[k] : 0.0f), (double)ln[k], acc);
        }
    }
    const size_t at = ((size_t)view * g.nrows + row) * g.ncols + col;
    if (MODE == CFP_WRITE) {
        if (ok) out[at] = (float)acc;
        return;
    }
    double sq = 0.0;
    if (MODE == CFP_RESID && ok) {
        const float r = (float)((double)meas[at] - acc);
        out[at] = r;
        sq = (double)r * (double)r;
    }
    if (MODE == CFP_SQ && ok) sq = acc * acc;
    const double tot = block_sum256(sq);
    if (threadIdx.x == 0) part[tile_id] = tot;
}

// Entry count of every ray (for the algorithmic-byte figure).
__global__ __launch_bounds__(kBlock) void k_cone_count(ConeDev g, const float *cs,
                                                       unsigned long long *total) {
    unsigned tile_id;
    int view, row, col;
    bool ok;
    unsigned long long n = 0;
    if (cone_fp_pixel(g, g.nviews, tile_id, view, row, col, ok) && ok) {
        ConeWalk w;
        if (cone_ray_init(g, cs, view, row, col, w)) {
            bool live = true;
            for (;;) {
                unsigned idx;
                float l;
                cone_step(g, w, live, idx, l);
                if (!live) break;
                ++n;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(total, n);
}

// Per (view, column) tables of the back projector: 1 / (dx - sx), 1 / (dy - sy)
// (double) and |(dx - sx, dy - sy)|^2 (float) of the column's rays; per view
// the source point.
__global__ void k_cone_vctab(ConeDev g, const float *cs, double2 *inv, float *l2, float4 *src, float2 *pos) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)g.nviews * g.ncols) return;
    const int v = (int)(i / g.ncols), col = (int)(i % g.ncols);
    const float c = cs[2 * v], s = cs[2 * v + 1];
    float sx, sy, sz, dx, dy, dz;
    ray_points(g, c, s, 0, col, sx, sy, sz, dx, dy, dz);   // x and y do not depend on the row
    const double rx = (double)dx - (double)sx, ry = (double)dy - (double)sy;
    inv[i] = make_double2(1.0 / rx, 1.0 / ry);
    l2[i] = (float)(rx * rx + ry * ry);
    pos[i] = make_float2(sx, sy);
    if (col == 0) src[v] = make_float4(sx, sy, sz, 0.0f);
}

// Slab interval of the unit slab [a, a + 1] along one axis for the ray
// s + t r given ir = 1 / r (infinite when r = 0: all or nothing).
__device__ __forceinline__ void cone_slab_inv(double s, double ir, double a, double &lo, double &hi) {
    if (isinf(ir)) {
        const bool in = (s >= a) && (s < a + 1.0);
        lo = in ? lo : 1.0;
        hi = in ? hi : 0.0;
        return;
    }
    const double t0 = (a - s) * ir, t1 = (a + 1.0 - s) * ir;
    lo = fmax(lo, fmin(t0, t1));
    hi = fmin(hi, fmax(t0, t1));
}

// PAR: the parallel beam (rays at constant z, orthographic footprint; the
// rows' z in rowtab[2 r], the columns' start points in vpos).
// RLDS: the rows' 1 / w (nrows doubles) are copied to LDS (dynamic shared
// memory) first.
template <bool PAR, bool RLDS>
__global__ __launch_bounds__(kBlock) void k_cone_bp(ConeDev g, const float *cs, const double *rowtab,
                                                    const double2 *__restrict__ vinv, const float *__restrict__ vl2,
                                                    const float4 *__restrict__ vsrc, const float2 *__restrict__ vpos,
                                                    const int *views, int nv, const float *__restrict__ sino,
                                                    double *carry, int first, int last, float *__restrict__ vol) {
    // Workgroup: 16 x 16 voxel columns, wave 8 x 8 of them; kZc slices.
    const int tx = (g.nx + 15) / 16, ty = (g.ny + 15) / 16;
    const unsigned b = xcd_block();
    const int per_z = tx * ty;
    const int zc = (int)(b / per_z), tile = (int)(b % per_z);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int x = (tile / ty) * 16 + (wv >> 1) * 8 + (lane >> 3);
    const int y = (tile % ty) * 16 + (wv & 1) * 8 + (lane & 7);
    const int z0 = zc * kZc;
    extern __shared__ double riw[];
    if (RLDS) {
        for (int r = threadIdx.x; r < g.nrows; r += kBlock) riw[r] = rowtab[2 * r + 1];
        __syncthreads();
    }
    if (x >= g.nx || y >= g.ny || z0 >= g.nz) return;

    // Per-lane sums of the kZc voxels in LDS (each lane owns its own slots,
    // so no barrier is needed); a register sums one (column, slice) first.
    __shared__ double acc[kZc][kBlock];
    const int nk = min(kZc, g.nz - z0);
    // carry: the double sums of the earlier view chunks of a back projection
    // split into several launches (nullptr for a single launch).
#pragma unroll
    for (int k = 0; k < kZc; ++k)
        acc[k][threadIdx.x] = (carry && !first && k < nk) ? carry[((size_t)(z0 + k) * g.nx + x) * g.ny + y] : 0.0;

    const float cxv = PAR ? g.halfsize : g.nx / 2.0f, cyv = PAR ? g.halfsize : g.ny / 2.0f;
    const float czv = PAR ? g.halfsize : g.nz / 2.0f;
    const float kpar = g.sample / g.pixel;       // parallel: panel pixels per voxel
    const float kcol = g.sdd / g.pixel;          // panel pixels per (voxel offset / voxel depth)
    const float sodv = g.sod / g.sample;         // source distance in voxels
    const float h5 = (float)g.halfdet - 0.5f;
    const float margin = 0.05f;                  // in pixels, covers the float rounding of the bounds

    for (int iv = 0; iv < nv; ++iv) {
        const int v = views[iv];
        const float c = cs[2 * v], s = cs[2 * v + 1];
        const float4 sp = vsrc[v];
        double sx = sp.x, sy = sp.y;
        const double sz = sp.z;
        // Corners of the column in the view's frame (u across the panel,
        // depth from the source): the panel column of a point is
        // u * kcol / depth + h5, the row of height zr is h5 - zr * kcol / depth.
        float cmin = 1e30f, cmax = -1e30f, imin = 1e30f, imax = -1e30f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float px = (float)(x + (q & 1)) - cxv, py = (float)(y + (q >> 1)) - cyv;
            const float u = c * px + s * py;
            const float id = PAR ? 1.0f : 1.0f / (-s * px + c * py + sodv);
            const float cf = PAR ? u * kpar : u * kcol * id;
            cmin = fminf(cmin, cf);
            cmax = fmaxf(cmax, cf);
            imin = fminf(imin, id);
            imax = fmaxf(imax, id);
        }
        const int c0 = max(0, (int)ceilf(cmin + h5 - margin));
        const int c1 = min(g.ncols - 1, (int)floorf(cmax + h5 + margin));
        for (int col = c0; col <= c1; ++col) {
            const size_t vc = (size_t)v * g.ncols + col;
            const double2 ir = vinv[vc];
            if (PAR) {
                const float2 p0 = vpos[vc];
                sx = p0.x;
                sy = p0.y;
            }
            double lo = 0.0, hi = 1.0;
            cone_slab_inv(sx, ir.x, (double)x, lo, hi);
            cone_slab_inv(sy, ir.y, (double)y, lo, hi);
            if (!(hi > lo)) continue;
            // z-slab parameters relative to lo, so that the float
            // differences below keep full relative precision.
            const float span = (float)(hi - lo);
            const float lxy2 = vl2[vc];
            const double nlo = -lo;
            const float *scol = sino + ((size_t)v * g.nrows) * g.ncols + col;
            if (PAR) {
                // Rays at constant height: row r crosses slice z when
                // z <= z_r < z + 1, over the whole x / y chord.
                const float len = span * sqrtf(lxy2);
#pragma unroll 1
                for (int k = 0; k < nk; ++k) {
                    const int z = z0 + k;
                    const float za = (float)z - czv;
                    const int r0 = max(0, (int)ceilf(h5 - (za + 1.0f) * kpar - margin));
                    const int r1 = min(g.nrows - 1, (int)floorf(h5 - za * kpar + margin));
                    double part = 0.0;
                    for (int r = r0; r <= r1; ++r) {
                        const double zr = rowtab[2 * r];
                        if (zr >= (double)z && zr < (double)z + 1.0)
                            part += (double)scol[(size_t)r * g.ncols] * (double)len;
                    }
                    acc[k][threadIdx.x] += part;
                }
                continue;
            }
#pragma unroll 1
            for (int k = 0; k < nk; ++k) {
                const int z = z0 + k;
                // Rows whose rays can meet slice z inside the column.
                const float za = (float)z - czv, zb = za + 1.0f;
                const float wlo = fminf(za * imin, za * imax) * kcol;
                const float whi = fmaxf(zb * imin, zb * imax) * kcol;
                const int r0 = max(0, (int)ceilf(h5 - whi - margin));
                const int r1 = min(g.nrows - 1, (int)floorf(h5 - wlo + margin));
                const double dz0 = (double)z - sz, dz1 = dz0 + 1.0;
                double part = 0.0;
                for (int r = r0; r <= r1; ++r) {
                    const double iw = RLDS ? riw[r] : rowtab[2 * r + 1];
                    const float ta = (float)fma(dz0, iw, nlo), tb = (float)fma(dz1, iw, nlo);
                    const float a = fmaxf(0.0f, fminf(ta, tb)), e = fminf(span, fmaxf(ta, tb));
                    if (!(e > a)) continue;
                    // w = 1 / iw and the ray length to about one float ulp
                    // (hardware reciprocal and square root).
                    const float wf = __builtin_amdgcn_rcpf((float)iw);
                    const float L = __builtin_amdgcn_sqrtf(lxy2 + wf * wf);
                    const float len = (e - a) * L;
                    // float x float is exact in double: the fused form rounds once
                    part = __builtin_fma((double)scol[(size_t)r * g.ncols], (double)len, part);
                }
                acc[k][threadIdx.x] += part;
            }
        }
    }
    for (int k = 0; k < nk; ++k) {
        const size_t i = ((size_t)(z0 + k) * g.nx + x) * g.ny + y;
        if (carry && !last) carry[i] = acc[k][threadIdx.x];
        else vol[i] = (float)acc[k][threadIdx.x];
    }
}

unsigned grid8(size_t blocks) { return (unsigned)((blocks + 7) / 8 * 8); }

}  // namespace

void validate_cone(const mi_cone_geometry &g) {
    MI_REQUIRE(g.nx > 0 && g.ny > 0 && g.nz > 0 && g.n_views > 0 && g.n_rows > 0 && g.n_cols > 0, MI_E_ARG,
               "cone geometry sizes must be positive");
    MI_REQUIRE((int64_t)g.nx * g.ny * g.nz < ((int64_t)1 << 31), MI_E_LIMIT, "volume beyond 2^31 voxels");
    // The forward walk forms a voxel's offset as ((z nx + x) ny) + y with
    // 24-bit multiplies: z nx + x and ny must stay below 2^24.
    MI_REQUIRE((int64_t)g.nz * g.nx <= ((int64_t)1 << 24) && g.ny < (1 << 24), MI_E_LIMIT,
               "cone volume needs nz * nx <= 2^24 and ny < 2^24");
    MI_REQUIRE((int64_t)g.n_views * g.n_rows * g.n_cols < ((int64_t)1 << 40), MI_E_LIMIT, "sinogram too large");
    MI_REQUIRE(std::isfinite(g.sample_mm) && std::isfinite(g.pixel_mm) && g.sample_mm > 0 && g.pixel_mm > 0,
               MI_E_ARG, "sample_mm and pixel_mm must be positive");
    MI_REQUIRE(std::isfinite(g.sdd_mm) && g.sod_mm > 0 && g.sdd_mm > g.sod_mm, MI_E_ARG, "need sdd > sod > 0");
    MI_REQUIRE(g.beam == MI_CONE_BEAM || g.beam == MI_CONE_PARALLEL, MI_E_ARG, "unknown cone-geometry beam");
    // The source circle must lie outside the volume (the back projector's
    // corner projection assumes every voxel is in front of the source).
    const double half = 0.5 * g.sample_mm * std::sqrt((double)g.nx * g.nx + (double)g.ny * g.ny);
    MI_REQUIRE(g.beam == MI_CONE_PARALLEL || g.sod_mm > half, MI_E_ARG, "the source orbit must lie outside the volume");
}

ConeTables &cone_tables(mi_context *ctx, const mi_cone_geometry &g) {
    validate_cone(g);
    if (ctx->cone && std::memcmp(&ctx->cone->g, &g, sizeof(g)) == 0) return *ctx->cone;
    auto t = std::make_shared<ConeTables>();
    t->g = g;
    t->d = ConeDev{g.nx, g.ny, g.nz, g.n_views, g.n_rows, g.n_cols, g.n_rows / 2,
                   g.sample_mm, g.pixel_mm, g.sod_mm, g.sdd_mm, g.beam == MI_CONE_PARALLEL ? 1 : 0,
                   (float)(g.nx / 2.0)};
    std::vector<float> cs(2 * (size_t)g.n_views);
    for (int a = 0; a < g.n_views; ++a) {
        // theta = (float)alpha / NPROJ * 2 * PI, industry3d/src/tracing.cpp:58.
        float th = (float)a / (float)g.n_views;
        th = th * 2.0f;
        th = (float)((double)th * 3.1415926535897932385);
        cs[2 * a] = cosf(th);
        cs[2 * a + 1] = sinf(th);
    }
    const float sz = (float)((double)(0.0f / g.sample_mm) + g.nz / 2.0);
    std::vector<double> rows(2 * (size_t)g.n_rows);
    for (int r = 0; r < g.n_rows; ++r) {
        if (t->d.parallel) {
            // the row's constant z: (HALFDET - r - 0.5) PIXELSIZE / SAMPLESIZE + HALFSIZE in float
            float Z = (float)((double)(t->d.halfdet - r) - 0.5);
            Z *= g.pixel_mm;
            rows[2 * r] = (double)(Z / g.sample_mm + t->d.halfsize);
            rows[2 * r + 1] = 0.0;
            continue;
        }
        const double w = (double)cone_det_z(t->d, r) - (double)sz;
        MI_REQUIRE(w != 0.0, MI_E_ARG, "a detector row lies in the plane of the source orbit");
        rows[2 * r] = w;
        rows[2 * r + 1] = 1.0 / w;
    }
    std::vector<int> all(g.n_views);
    for (int a = 0; a < g.n_views; ++a) all[a] = a;
    t->all.alloc(all.size());
    MI_HIP(hipMemcpyAsync(t->all.p, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    t->cs.alloc(cs.size());
    t->rows.alloc(rows.size());
    MI_HIP(hipMemcpyAsync(t->cs.p, cs.data(), cs.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    MI_HIP(hipMemcpyAsync(t->rows.p, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    const size_t nvc = (size_t)g.n_views * g.n_cols;
    t->vinv.alloc(nvc);
    t->vl2.alloc(nvc);
    t->vsrc.alloc(g.n_views);
    t->vpos.alloc(nvc);
    hipLaunchKernelGGL(k_cone_vctab, dim3((unsigned)((nvc + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, t->d,
                       t->cs.p, t->vinv.p, t->vl2.p, t->vsrc.p, t->vpos.p);
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(ctx->stream));
    ctx->cone = t;
    return *t;
}

void cone_trace(hipStream_t st, const ConeTables &t, int view, int row, int col, int64_t *ind, float *len, int cap,
                int *numb) {
    hipLaunchKernelGGL(k_cone_trace, dim3(1), dim3(64), 0, st, t.d, t.cs.p, view, row, col, ind, len, cap, numb);
    MI_HIP(hipGetLastError());
}

static size_t cone_fp_blocks(const ConeDev &d, int nv) {
    return (size_t)((d.ncols + kConeTile - 1) / kConeTile) * ((d.nrows + kConeTile - 1) / kConeTile) * nv;
}

size_t cone_fp_views(hipStream_t st, const ConeTables &t, const int *views, int nv, int mode, const float *vol,
                     const float *meas, float *out, double *part) {
    const size_t nb = cone_fp_blocks(t.d, nv);
    const dim3 grid(grid8(nb)), block(kBlock);
    switch (mode) {
        case CFP_WRITE:
            launch(k_cone_fp<CFP_WRITE>, grid, block, 0u, st, t.d, t.cs.p, views, nv, vol, meas, out, part);
            break;
        case CFP_RESID:
            launch(k_cone_fp<CFP_RESID>, grid, block, 0u, st, t.d, t.cs.p, views, nv, vol, meas, out, part);
            break;
        default:
            launch(k_cone_fp<CFP_SQ>, grid, block, 0u, st, t.d, t.cs.p, views, nv, vol, meas, out, part);
    }
    MI_HIP(hipGetLastError());
    return nb;
}

// The allocation-free back projection's carry (mi_cone_backproject_dev_noalloc):
// one copy per device, for the life of the process.
constexpr size_t kScratchCarry = (size_t)8 << 20;   // doubles (64 MiB): volumes up to 8 Mi voxels
static __device__ __align__(256) double scratch[kScratchCarry];

namespace {
constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
double *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];
}  // namespace

void cone_scratch_init(int device) {
    MI_REQUIRE(device >= 0 && device < kMaxDevices, MI_E_LIMIT, "device index beyond the scratch table");
    // The array's name is a host-side stand-in: its device address (one per
    // device) comes from the runtime, with that device current.
    std::call_once(g_scratch_once[device], [device] {
        g_scratch_err[device] = hipGetSymbolAddress((void **)&g_scratch_ptr[device], HIP_SYMBOL(scratch));
    });
    MI_HIP(g_scratch_err[device]);
}

int64_t cone_scratch_voxels() { return (int64_t)kScratchCarry; }

void cone_bp_views(hipStream_t st, const ConeTables &t, const int *views, int nv, const float *sino, float *vol,
                   int device_scratch) {
    const size_t blocks = (size_t)((t.d.nx + 15) / 16) * ((t.d.ny + 15) / 16) * ((t.d.nz + kZc - 1) / kZc);
    // Views in chunks whose sinogram rows stay cached while every workgroup
    // sweeps them; double sums carried between chunks (a buffer allocated
    // here, or the device's fixed scratch array).
    const int chunks = (nv + kBpViewChunk - 1) / kBpViewChunk;
    const size_t nvox = (size_t)t.d.nx * t.d.ny * t.d.nz;
    DevBuf<double> own(chunks > 1 && device_scratch < 0 ? nvox : 0);
    double *carry = own.p;
    if (chunks > 1 && device_scratch >= 0) {
        MI_REQUIRE(nvox <= kScratchCarry, MI_E_LIMIT,
                   "volume of " + std::to_string(nvox) + " voxels exceeds the allocation-free back projection's " +
                       std::to_string(kScratchCarry) + "-voxel scratch");
        carry = g_scratch_ptr[device_scratch];
        MI_REQUIRE(carry, MI_E_STATE, "device scratch not initialised");
    }
    for (int c = 0; c < chunks; ++c) {
        const int v0 = c * kBpViewChunk, n = std::min(kBpViewChunk, nv - v0);
        if (t.d.parallel)
            launch(k_cone_bp<true, false>, dim3(grid8(blocks)), dim3(kBlock), 0u, st, t.d, t.cs.p,
                               t.rows.p, t.vinv.p, t.vl2.p, t.vsrc.p, t.vpos.p, views + v0, n, sino, carry, c == 0,
                               c == chunks - 1, vol);
        else if (t.d.nrows <= kBpRowLds)
            launch(k_cone_bp<false, true>, dim3(grid8(blocks)), dim3(kBlock),
                               (uint32_t)(t.d.nrows * sizeof(double)), st, t.d, t.cs.p, t.rows.p, t.vinv.p, t.vl2.p,
                               t.vsrc.p, t.vpos.p, views + v0, n, sino, carry, c == 0, c == chunks - 1, vol);
        else
            launch(k_cone_bp<false, false>, dim3(grid8(blocks)), dim3(kBlock), 0u, st, t.d, t.cs.p,
                               t.rows.p, t.vinv.p, t.vl2.p, t.vsrc.p, t.vpos.p, views + v0, n, sino, carry, c == 0,
                               c == chunks - 1, vol);
        MI_HIP(hipGetLastError());
    }
    if (own.p) MI_HIP(hipStreamSynchronize(st));   // before the carry buffer is freed
}

void cone_forward(hipStream_t st, const ConeTables &t, const float *vol, float *sino) {
    cone_fp_views(st, t, t.all.p, t.d.nviews, CFP_WRITE, vol, nullptr, sino, nullptr);
}

void cone_backproject(hipStream_t st, const ConeTables &t, const float *sino, float *vol) {
    cone_bp_views(st, t, t.all.p, t.d.nviews, sino, vol);
}

double cone_fp_bytes(const ConeTables &t) {
    const double rays = (double)t.d.nviews * t.d.nrows * t.d.ncols;
    const double vox = (double)t.d.nx * t.d.ny * t.d.nz;
    return (double)t.nnz * 8.0 + rays * 4.0 + vox * 4.0;
}

double cone_bp_bytes(const ConeTables &t) { return cone_fp_bytes(t); }

int64_t cone_nnz(hipStream_t st, ConeTables &t) {
    if (t.nnz >= 0) return t.nnz;
    DevBuf<unsigned long long> n(1);
    MI_HIP(hipMemsetAsync(n.p, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_cone_count, dim3(grid8(cone_fp_blocks(t.d, t.d.nviews))), dim3(kBlock), 0, st, t.d, t.cs.p,
                       n.p);
    MI_HIP(hipGetLastError());
    unsigned long long h = 0;
    MI_HIP(hipMemcpyAsync(&h, n.p, sizeof(h), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    t.nnz = (int64_t)h;
    return t.nnz;
}

}  // namespace mi
