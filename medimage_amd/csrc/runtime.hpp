// Host runtime of medimage_amd: device buffers, system matrices, solver
// entry points.  Everything here is C++ behind the C ABI of
// include/medimage_amd.h.
#pragma once

#include "common.hpp"
#include "../../include/medimage_amd.h"

#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace mi {

// Owning device allocation.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; return *this; }
    ~DevBuf() { reset(); }
    void alloc(size_t count) {
        reset();
        if (count == 0) return;
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, count * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(e == hipErrorOutOfMemory ? MI_E_NOMEM : MI_E_DEVICE,
                 "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        p = (T *)q;
        n = count;
    }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Geometry + subset count identify one system matrix.
struct GeomKey {
    std::array<int32_t, 8> i;
    std::array<double, 6> d;
    std::vector<double> views;       // per-view source angles (arc beam), if given
    bool operator<(const GeomKey &o) const {
        if (i != o.i) return i < o.i;
        if (d != o.d) return d < o.d;
        return views < o.views;
    }
};
// The caller's per-view angle table as an owned copy (empty: none).
std::vector<double> view_table(const mi_geometry &g);
// layout: 0 = CSR / CSC (projector calls), 1 = sliced (AT engine).
GeomKey make_key(const mi_geometry &g, int n_subsets, int layout = 0);
void validate_geometry(const mi_geometry &g);

// Sliced storage of one projection direction for the AT engine.  Items
// (rays for the forward projector, pixels for the back projector) get P
// lanes each, lane p taking entries p, p + P, ... of its item; a wave's 64
// lanes are C = 64 / P consecutive items of one subset (a slice).  Steps go
// in groups of four: entry t of lane l of slice sl is at
// base[sl] + 256 (t / 4) + 4 l + t % 4 (sell_pos), so one 16-byte load per
// lane (one contiguous kilobyte per wave) brings four steps of lengths.
// Lanes past their last entry read padding (their last index, length 0); a
// slice's step count T is a multiple of four.
struct Sell {
    int P = 1, C = 64;
    std::vector<int64_t> slice0;     // S + 1: first slice of each subset
    DevBuf<int64_t> d_slice0;
    DevBuf<int64_t> base;            // per slice
    DevBuf<int32_t> T;               // per slice: entries per lane
    DevBuf<uint32_t> idx;            // forward: tiled pixel; back: row
    // Compact index stream (when every step fits 16 bits): the lane's first
    // index, then per entry the difference to the lane's previous entry
    // (0 at t = 0 and in the padding).  idx is freed then.
    DevBuf<uint32_t> first;          // per (slice, lane)
    DevBuf<int16_t> d16;
    DevBuf<float> l32;
    DevBuf<double> l64;
    int64_t entries = 0;             // stored, padding included
    // Back slices only: residual values staged in LDS.  For each (subset,
    // block of four slices, view of the subset) the block's first bin `vlo`
    // and slot offset `voff` (nv + 1 per block); every entry also carries the
    // LDS slot of its (view, bin) (slot16, same positions as l32).  Empty
    // when a block needs 65536 slots or more.
    DevBuf<uint16_t> slot16;
    DevBuf<int32_t> vlo, voff;
    std::vector<int64_t> vtab0;      // S + 1: first table entry of each subset
    DevBuf<int64_t> d_vtab0;
    int64_t max_slots = 0;           // slots of the largest block
    size_t bytes() const {
        return d_slice0.bytes() + base.bytes() + T.bytes() + idx.bytes() + first.bytes() + d16.bytes() +
               l32.bytes() + l64.bytes() + slot16.bytes() + vlo.bytes() + voff.bytes() + d_vtab0.bytes();
    }
};

// Back-projection items: a block of four slices covers a 16 x 16 pixel
// square (4 tile rows x 2 tile columns of the tiled layout); slice w of the
// block is tile row w of the square, lane l position l % 32 of tile column
// l / 32.  Returns the tiled position, or -1 outside the tile grid.
constexpr int kBpBlockTR = 4, kBpBlockTC = 2;
__host__ __device__ __forceinline__ int64_t bp_blocks_of(int tnx, int tny) {
    return (int64_t)((tnx + kBpBlockTR - 1) / kBpBlockTR) * ((tny + kBpBlockTC - 1) / kBpBlockTC);
}
// First voff entry of block blk of subset s (nv + 1 entries per block; vlo
// has nv per block from vtab0[s] + blk nv).
__host__ __device__ __forceinline__ int64_t voff_at(const int64_t *vtab0, int s, int64_t nblocks, int64_t blk, int nv) {
    return vtab0[s] + (int64_t)s * nblocks + blk * (nv + 1);
}
__host__ __device__ __forceinline__ int64_t bp_item(int64_t blk, int w, int lane, int tnx, int tny) {
    const int nbc = (tny + kBpBlockTC - 1) / kBpBlockTC;
    const int64_t tr = (blk / nbc) * kBpBlockTR + w;
    const int tc = (int)(blk % nbc) * kBpBlockTC + (lane >> 5);
    if (tr >= tnx || tc >= tny) return -1;
    return (tr * tny + tc) * 32 + (lane & 31);
}

// Checkpointed forward walk (fan beam).  A forward block (kBlock lanes)
// covers kOtfRays consecutive rays of a subset, each cut into kOtfSeg
// segments of equal entry count: segment p of ray a is lane kOtfRays h + a
// of the block's wave w, p = kOtfPerWave w + h, so a wave's lanes at one step
// sit on kOtfRays neighbouring rays (neighbouring pixels).  The lane walks the
// ray's entries [p Ls, min(L, (p + 1) Ls)), Ls = ceil(L / kOtfSeg), from the
// tracer's own state at entry p Ls, reproducing the stored CSR entries bit for
// bit while reading ~56 bytes per lane instead of 6 per entry.
#ifndef MI_OTF_RAYS
#define MI_OTF_RAYS 32
#endif
constexpr int kOtfRays = MI_OTF_RAYS, kOtfSeg = kBlock / kOtfRays, kOtfPerWave = kWave / kOtfRays,
              kOtfWaves = kBlock / kWave;
static_assert(kOtfRays >= 16 && kOtfRays <= kWave && kOtfSeg * kOtfRays == kBlock, "forward block shape");
struct Otf {
    std::vector<int64_t> slice0;     // S + 1
    DevBuf<int64_t> d_slice0;
    DevBuf<double> lam, ax, ay, stx, sty;
    DevBuf<int32_t> cx, cy, sg, cnt;  // sg: bit 0 x step > 0, bit 1 y step > 0
    size_t bytes() const {
        return d_slice0.bytes() + lam.bytes() + ax.bytes() + ay.bytes() + stx.bytes() + sty.bytes() + cx.bytes() +
               cy.bytes() + sg.bytes() + cnt.bytes();
    }
};

constexpr int kGroup = 4;
__host__ __device__ __forceinline__ int64_t sell_pos(int64_t base, int t, int lane) {
    return base + (int64_t)(t >> 2) * (kGroup * 64) + lane * kGroup + (t & 3);
}

// Images of the AT engine are stored in 4 x 8 pixel tiles (128 bytes of
// FP32): tile (ti, tj) at ti * tny + tj; inside a tile the two 4 x 4 halves
// follow each other (pixel (i & 3, j & 7) at (j >> 2) * 16 + (i & 3) * 4 +
// (j & 3)), so each 64-byte half line is a square.
constexpr int kTileH = 4, kTileW = 8;
constexpr int kStripTiles = 8;   // tiles of a stencil strip; each tile row holds a multiple of it
__host__ __device__ __forceinline__ int64_t tile_index(int i, int j, int tny) {
    return ((int64_t)(i >> 2) * tny + (j >> 3)) * (kTileH * kTileW) + (((j >> 2) & 1) << 4) + ((i & 3) << 2) +
           (j & 3);
}
__host__ __device__ __forceinline__ void tile_coords(int64_t q, int tny, int &i, int &j) {
    const int64_t t = q >> 5;
    const int w = (int)(q & 31);
    i = (int)(t / tny) * kTileH + ((w >> 2) & 3);
    j = (int)(t % tny) * kTileW + ((w >> 4) << 2) + (w & 3);
}

struct SeqMatrix;  // reference-order entry lists (ordered.hip)
size_t seq_bytes(const SeqMatrix *q);

// System matrix of one geometry, rows (rays) grouped by view subset.
//   Row order: subsets in order; inside a subset its views ascending; inside
//   a view the bins ascending.  Row r is ray (vlist[r / nbins], r % nbins).
//   CSR (ray-major) feeds the forward projector; one CSC per subset
//   (pixel-major, entries ordered by row) feeds the gather back projector.
struct Matrix {
    mi_geometry g{};                 // g.view_deg points into view_deg below
    std::vector<double> view_deg;
    int S = 1;
    int npix = 0;
    int64_t nrows = 0, nnz = 0;
    std::vector<int> vlist;          // subset-ordered view list
    std::vector<int64_t> sub_row0;   // S+1 row offsets
    std::vector<int64_t> sub_nnz;    // S entry counts (CSR rows of each subset)
    uint64_t last_use = 0;           // context tick of the last use (LRU eviction)
    bool has64 = false;
    mi_tuning tune{0, 0, 0, 0, 0.0, 0.0};  // bp_lanes 0: not set (heuristic)

    DevBuf<int> d_vlist;
    DevBuf<int64_t> d_sub_row0;      // S + 1
    DevBuf<int64_t> rowptr;          // nrows + 1
    DevBuf<uint32_t> idx;            // nnz, pixel index
    DevBuf<float> len32;             // nnz
    DevBuf<double> len64;            // nnz (FP64 runs only)
    DevBuf<int64_t> colptr;          // S * npix + 1
    DevBuf<uint32_t> crow;           // nnz, row index
    DevBuf<float> clen32;
    DevBuf<double> clen64;
    // Row permutation between the caller's [view][bin] sinogram and the
    // subset-ordered rows: perm[r] = view * nbins + bin of row r.
    DevBuf<int64_t> perm;
    // AT engine: sliced forward (rays, tiled pixel indices) and back
    // (pixels in tile order, per subset) projectors; tiled image size.
    Sell fs, bs;
    Otf otf;
    bool has_otf = false;
    int tny = 0;
    int64_t npix_t = 0;
    bool has_sell = false;
    bool has_csr = true;             // false once the engine layout replaced it
    // AT engine launch settings of this geometry (mi_engine_autotune):
    // jobs per unit (0: the L2 rule of Batch::k_cap) and the back
    // projection's residual staging (-1: LDS when the slots fit the default
    // limit, 0: global gathers, 1: LDS up to the hardware limit).  Neither
    // changes a job's arithmetic, so results stay bitwise the same.
    struct EngineTune {
        int k_cap = 0;
        int bp_lds = -1;
        double us_per_job_pass = 0.0;
    } etune;
    std::shared_ptr<SeqMatrix> seq;  // GoldenC-form drivers, built on first use

    size_t bytes() const;
};
// Builds m.fs / m.bs from the CSR / CSC (sell.hip); drop_csr frees each of
// the CSR and CSC as soon as its sliced copy exists.
void build_sell(Matrix &m, hipStream_t st, bool drop_csr = false);

// Builds the matrix for (g, S) on the stream; want64 also keeps double
// lengths.  Host-side GoldenC walk for MI_BEAM_GOLDENC, device tracers
// otherwise.
// otf: also record the checkpointed forward walk (fan beam only).
std::unique_ptr<Matrix> build_matrix(const mi_geometry &g, int S, bool want64, hipStream_t st, bool otf = false);
// Subset order: bit-reversed subset indices (values >= S dropped).
std::vector<int> subset_order(int S);

// Timing / byte counters of the instrumented kernels.
struct Stats {
    bool on = false;
    uint32_t mask = ~0u;             // instrumented kernel ids (bit k = kernel k)
    std::array<mi_kernel_stats, MI_K_COUNT> k{};
    struct Pending { int kernel; hipEvent_t a, b; double bytes; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
};

struct ConeTables;  // per-geometry tables of the 3D cone-beam projectors

}  // namespace mi

struct mi_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<mi::GeomKey, std::shared_ptr<mi::Matrix>> matrices;
    mi::Stats stats;
    uint64_t tick = 0;
    std::shared_ptr<mi::ConeTables> cone;  // tables of the last cone geometry used
};

namespace mi {
// Matrix for (g, S) from the context cache, building it if needed.  Jobs
// hold the shared pointer, so trimming the cache never frees a matrix in use.
std::shared_ptr<Matrix> get_matrix(mi_context *ctx, const mi_geometry &g, int S, bool want64);
// The same geometry in the AT engine's sliced layout only (CSR / CSC freed).
std::shared_ptr<Matrix> get_engine_matrix(mi_context *ctx, const mi_geometry &g, int S, bool want64);
// The cached engine matrix of (g, S), or null (nothing is built).
std::shared_ptr<Matrix> find_engine_matrix(mi_context *ctx, const mi_geometry &g, int S, bool want64);
// Frees cached matrices that no job holds, least recently used first, until
// at least `bytes` were freed (all idle ones when bytes < 0).  Returns bytes freed.
int64_t evict_idle(mi_context *ctx, int64_t bytes);
// Stats helpers: bracket a launch with events when enabled.  With
// `dispatch`, the scope's kernels launched through mi::launch carry the
// events themselves (kernel start to kernel end, as a profiler reads them);
// otherwise the events are recorded on the stream around the scope.
struct LaunchTimer {
    mi_context *ctx;
    int kernel;
    double bytes;
    hipEvent_t a = nullptr, b = nullptr;
    DispatchEvents de;
    DispatchEvents *prev = nullptr;
    bool dispatch = false;
    LaunchTimer(mi_context *c, int k, double by, bool dispatch = false);
    ~LaunchTimer();
};
void stats_collect(mi_context *ctx);
}  // namespace mi
