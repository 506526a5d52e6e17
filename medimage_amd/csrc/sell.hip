// Sliced (lane-per-segment) copies of a system matrix for the AT engine,
// built on the device from the ray-major CSR and the per-subset CSC.
//
// Forward direction: items are the rows of a subset in row order, so a
// slice holds C consecutive rays (neighbouring bins of one view); pixel
// indices are rewritten into the engine's tiled image layout.  Back
// direction: items are the pixels of the tiled layout (tile order), per
// subset; the entries keep their row order (the CSC order).  The entries of
// an item are dealt round-robin to its P lanes.
#include "runtime.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>

namespace mi {

namespace {

inline unsigned grid_of(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ int subset_of(const int64_t *slice0, int S, int64_t sl) {
    int lo = 0, hi = S;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (slice0[mid] <= sl) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Entry range [b, e) of forward item `lane` of slice sl (b == e: none).
struct FItems {
    const int64_t *rowptr, *sub_row0, *slice0;
    int S, P, C;
    __device__ __forceinline__ void range(int64_t sl, int item, int64_t &b, int64_t &e) const {
        const int s = subset_of(slice0, S, sl);
        const int64_t r = sub_row0[s] + (sl - slice0[s]) * C + item;
        if (r >= sub_row0[s + 1]) { b = e = 0; return; }
        b = rowptr[r];
        e = rowptr[r + 1];
    }
};

// Back item: lane `item` of slice sl (P = 1) -> its tiled pixel (bp_item:
// four slices per 16 x 16 block) -> its CSC column in subset s.
struct BItems {
    const int64_t *colptr, *slice0;
    int S, P, C, nx, ny, tnx, tny;
    int64_t npix, npix_t;
    __device__ __forceinline__ int64_t pixel(int64_t sl, int item, int &s, int &i, int &j) const {
        s = subset_of(slice0, S, sl);
        const int64_t li = sl - slice0[s];
        const int64_t q = bp_item(li / 4, (int)(li % 4), item, tnx, tny);
        if (q < 0) return -1;
        tile_coords(q, tny, i, j);
        return (i < nx && j < ny) ? q : -1;
    }
    __device__ __forceinline__ void range(int64_t sl, int item, int64_t &b, int64_t &e) const {
        int s, i, j;
        if (pixel(sl, item, s, i, j) < 0) { b = e = 0; return; }
        const int64_t c = (int64_t)s * npix + (int64_t)i * ny + j;
        b = colptr[c];
        e = colptr[c + 1];
    }
};

template <typename It>
__global__ void k_sell_T(It it, int64_t nslices, int32_t *T) {
    const int64_t sl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= nslices) return;
    int64_t t = 0;
    for (int item = 0; item < it.C; ++item) {
        int64_t b, e;
        it.range(sl, item, b, e);
        t = max(t, (e - b + it.P - 1) / it.P);
    }
    T[sl] = (int32_t)((t + kGroup - 1) / kGroup * kGroup);
}

__global__ void k_times64(const int32_t *T, int64_t n, int64_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int64_t)T[i] * kWave;
}

// One thread per (slice, lane).  FWD: rewrite pixel indices into tiles.
template <typename It, bool FWD>
__global__ void k_sell_fill(It it, int64_t nslices, const int64_t *base, const int32_t *T, const uint32_t *src_idx,
                            const float *src32, const double *src64, int ny, int tny, uint32_t *idx, float *l32,
                            double *l64) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t sl = g / kWave;
    const int lane = (int)(g % kWave);
    if (sl >= nslices) return;
    int64_t b, e;
    it.range(sl, lane / it.P, b, e);
    const int p = lane % it.P;
    // Lane p of an item takes its entries p, p + P, p + 2P, ... (step t holds
    // entries tP .. tP + P - 1 of the item across its P lanes).
    const int64_t cnt = (e - b > p) ? (e - b - p + it.P - 1) / it.P : 0;
    const int n = T[sl];
    uint32_t last = 0;
    for (int t = 0; t < n; ++t) {
        const int64_t pos = sell_pos(base[sl], t, lane);
        if (t < cnt) {
            const int64_t k = b + p + (int64_t)t * it.P;
            uint32_t v = src_idx[k];
            if (FWD) v = (uint32_t)tile_index((int)(v / (uint32_t)ny), (int)(v % (uint32_t)ny), tny);
            idx[pos] = v;
            l32[pos] = src32[k];
            if (l64) l64[pos] = src64[k];
            last = v;
        } else {
            idx[pos] = last;
            l32[pos] = 0.f;
            if (l64) l64[pos] = 0.0;
        }
    }
}

// Index deltas of one lane: first[] = its first index, d16 = difference to
// the previous entry; *overflow set when a difference does not fit 16 bits.
__global__ void k_sell_delta(int64_t nslices, const int64_t *base, const int32_t *T, const uint32_t *idx,
                             uint32_t *first, int16_t *d16, unsigned *overflow) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t sl = g / kWave;
    if (sl >= nslices) return;
    const int lane = (int)(g % kWave);
    const int n = T[sl];
    int64_t prev = n > 0 ? (int64_t)idx[sell_pos(base[sl], 0, lane)] : 0;
    first[g] = (uint32_t)prev;
    bool bad = false;
    for (int t = 0; t < n; ++t) {
        const int64_t pos = sell_pos(base[sl], t, lane);
        const int64_t cur = idx[pos];
        const int64_t d = cur - prev;
        if (d > 32767 || d < -32768) bad = true;
        d16[pos] = (int16_t)d;
        prev = cur;
    }
    if (bad) atomicOr(overflow, 1u);
}

// LDS staging tables of the back slices.  lo / hi: per (subset, block,
// view) the smallest and largest bin of the block's entries (integer
// atomics: the result does not depend on order).
__global__ void k_vtab_init(int32_t *lo, int32_t *hi, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        lo[i] = INT32_MAX;
        hi[i] = -1;
    }
}

__global__ void k_vtab_minmax(BItems it, int64_t nslices, const uint32_t *crow, const int64_t *sub_row0,
                              const int64_t *vtab0, int nbins, int32_t *lo, int32_t *hi) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nslices * kWave) return;
    const int64_t sl = g / kWave;
    int64_t b, e;
    it.range(sl, (int)(g % kWave), b, e);
    if (b == e) return;
    const int s = subset_of(it.slice0, it.S, sl);
    const int64_t blk = (sl - it.slice0[s]) / 4;
    const int nv = (int)((sub_row0[s + 1] - sub_row0[s]) / nbins);
    const int64_t tab = vtab0[s] + blk * nv;
    for (int64_t k = b; k < e; ++k) {
        const int64_t rl = (int64_t)crow[k] - sub_row0[s];
        const int v = (int)(rl / nbins), bin = (int)(rl % nbins);
        atomicMin(lo + tab + v, bin);
        atomicMax(hi + tab + v, bin);
    }
}

// One thread per (subset, block): slot offsets; *mx = largest block.
__global__ void k_vtab_off(const int64_t *vtab0, const int64_t *sub_row0, int S, int64_t nblocks, int nbins,
                           const int32_t *lo, const int32_t *hi, int32_t *vlo, int32_t *voff,
                           unsigned long long *mx) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)S * nblocks) return;
    const int s = (int)(g / nblocks);
    const int64_t blk = g % nblocks;
    const int nv = (int)((sub_row0[s + 1] - sub_row0[s]) / nbins);
    const int64_t t = vtab0[s] + blk * nv, to = voff_at(vtab0, s, nblocks, blk, nv);
    int64_t off = 0;
    for (int v = 0; v < nv; ++v) {
        const int c = hi[t + v] >= lo[t + v] ? hi[t + v] - lo[t + v] + 1 : 0;
        vlo[t + v] = c ? lo[t + v] : 0;
        voff[to + v] = (int32_t)off;
        off += c;
    }
    voff[to + nv] = (int32_t)off;
    atomicMax(mx, (unsigned long long)off);
}

// Slot of every stored back entry (same positions as the lengths); padding
// takes the lane's last slot.
__global__ void k_slot_fill(BItems it, int64_t nslices, const int64_t *base, const int32_t *T, const uint32_t *crow,
                            const int64_t *sub_row0, const int64_t *vtab0, int64_t nblocks, int nbins,
                            const int32_t *vlo, const int32_t *voff, uint16_t *slot) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nslices * kWave) return;
    const int64_t sl = g / kWave;
    const int lane = (int)(g % kWave);
    int64_t b, e;
    it.range(sl, lane, b, e);
    const int s = subset_of(it.slice0, it.S, sl);
    const int64_t blk = (sl - it.slice0[s]) / 4;
    const int nv = (int)((sub_row0[s + 1] - sub_row0[s]) / nbins);
    const int64_t t = vtab0[s] + blk * nv, to = voff_at(vtab0, s, nblocks, blk, nv);
    const int n = T[sl];
    uint16_t last = 0;
    for (int k = 0; k < n; ++k) {
        if (b + k < e) {
            const int64_t rl = (int64_t)crow[b + k] - sub_row0[s];
            const int v = (int)(rl / nbins), bin = (int)(rl % nbins);
            last = (uint16_t)(voff[to + v] + bin - vlo[t + v]);
        }
        slot[sell_pos(base[sl], k, lane)] = last;
    }
}

// Replaces sl.idx by the delta form when every difference fits.
void compact_sell(Sell &sl, hipStream_t st) {
    const int64_t ns = sl.slice0.back();
    if (ns == 0 || sl.entries == 0) return;
    DevBuf<unsigned> flag(1);
    MI_HIP(hipMemsetAsync(flag.p, 0, sizeof(unsigned), st));
    sl.first.alloc((size_t)ns * kWave);
    sl.d16.alloc((size_t)sl.entries);
    k_sell_delta<<<grid_of(ns * kWave), kBlock, 0, st>>>(ns, sl.base.p, sl.T.p, sl.idx.p, sl.first.p, sl.d16.p,
                                                         flag.p);
    MI_HIP(hipGetLastError());
    unsigned h = 0;
    MI_HIP(hipMemcpyAsync(&h, flag.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    if (h) {
        sl.first.reset();
        sl.d16.reset();
    } else {
        sl.idx.reset();
    }
}

// Lanes per forward item (stored forward slices: geometries without the
// checkpointed walk).  8, so a wave step covers 8 neighbouring rays x 8
// consecutive entries, an 8 x 8 patch of pixels (measured on the 512^2 /
// 720-view fan workload before the walk replaced it there: 60 us per subset
// against 76 us at 4 and 73 us for contiguous segments).  Fewer when rays
// are short.  Back slices always have one lane per pixel (a lane walks its
// pixel's column; 60 us against 82 us at 8 lanes in the same measurement).
#ifndef MI_SELL_PF
#define MI_SELL_PF 8
#endif
int choose_p(int P, double mean_len) {
    while (P > 1 && mean_len / P < 4.0) P /= 2;
    return P;
}

template <typename It>
void fill_sell(Sell &sl, It it, const uint32_t *src_idx, const float *s32, const double *s64, bool fwd, int ny,
               int tny, hipStream_t st) {
    const int64_t ns = sl.slice0.back();
    sl.d_slice0.alloc(sl.slice0.size());
    MI_HIP(hipMemcpyAsync(sl.d_slice0.p, sl.slice0.data(), sl.slice0.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                          st));
    it.slice0 = sl.d_slice0.p;
    sl.T.alloc(ns);
    sl.base.alloc(ns + 1);
    k_sell_T<<<grid_of(ns), kBlock, 0, st>>>(it, ns, sl.T.p);
    MI_HIP(hipGetLastError());
    DevBuf<int64_t> n64(ns + 1);
    MI_HIP(hipMemsetAsync(n64.p + ns, 0, sizeof(int64_t), st));
    k_times64<<<grid_of(ns), kBlock, 0, st>>>(sl.T.p, ns, n64.p);
    MI_HIP(hipGetLastError());
    size_t tmp = 0;
    MI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, n64.p, sl.base.p, (int)(ns + 1), st));
    {
        DevBuf<unsigned char> t(tmp ? tmp : 1);
        MI_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, n64.p, sl.base.p, (int)(ns + 1), st));
        MI_HIP(hipStreamSynchronize(st));
    }
    MI_HIP(hipMemcpy(&sl.entries, sl.base.p + ns, sizeof(int64_t), hipMemcpyDeviceToHost));
    sl.idx.alloc((size_t)std::max<int64_t>(sl.entries, 1));
    sl.l32.alloc((size_t)std::max<int64_t>(sl.entries, 1));
    if (s64) sl.l64.alloc((size_t)std::max<int64_t>(sl.entries, 1));
    if (fwd)
        k_sell_fill<It, true><<<grid_of(ns * kWave), kBlock, 0, st>>>(it, ns, sl.base.p, sl.T.p, src_idx, s32, s64, ny,
                                                                      tny, sl.idx.p, sl.l32.p,
                                                                      s64 ? sl.l64.p : nullptr);
    else
        k_sell_fill<It, false><<<grid_of(ns * kWave), kBlock, 0, st>>>(it, ns, sl.base.p, sl.T.p, src_idx, s32, s64,
                                                                       ny, tny, sl.idx.p, sl.l32.p,
                                                                       s64 ? sl.l64.p : nullptr);
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(st));
}

// LDS staging tables and slot codes of the back slices (before the CSC is
// dropped).  Left empty when a block would need 65536 slots or more.
void build_slots(Matrix &m, const BItems &bi, int64_t nblk, hipStream_t st) {
    Sell &q = m.bs;
    const int nbins = m.g.n_bins;
    q.vtab0.assign(m.S + 1, 0);
    for (int s = 0; s < m.S; ++s) q.vtab0[s + 1] = q.vtab0[s] + nblk * ((m.sub_row0[s + 1] - m.sub_row0[s]) / nbins);
    const int64_t ntab = q.vtab0.back();
    const int64_t ns = q.slice0.back();
    if (ntab == 0 || ns == 0) return;
    q.d_vtab0.alloc(q.vtab0.size());
    MI_HIP(hipMemcpyAsync(q.d_vtab0.p, q.vtab0.data(), q.vtab0.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    DevBuf<int32_t> lo(ntab), hi(ntab);
    k_vtab_init<<<grid_of(ntab), kBlock, 0, st>>>(lo.p, hi.p, ntab);
    MI_HIP(hipGetLastError());
    k_vtab_minmax<<<grid_of(ns * kWave), kBlock, 0, st>>>(bi, ns, m.crow.p, m.d_sub_row0.p, q.d_vtab0.p, nbins, lo.p,
                                                          hi.p);
    MI_HIP(hipGetLastError());
    q.vlo.alloc((size_t)ntab);
    q.voff.alloc((size_t)(ntab + (int64_t)m.S * nblk));
    DevBuf<unsigned long long> mx(1);
    MI_HIP(hipMemsetAsync(mx.p, 0, sizeof(unsigned long long), st));
    k_vtab_off<<<grid_of((int64_t)m.S * nblk), kBlock, 0, st>>>(q.d_vtab0.p, m.d_sub_row0.p, m.S, nblk, nbins, lo.p,
                                                               hi.p, q.vlo.p, q.voff.p, mx.p);
    MI_HIP(hipGetLastError());
    unsigned long long h = 0;
    MI_HIP(hipMemcpyAsync(&h, mx.p, sizeof(h), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    if (h >= 65536) {
        q.vlo.reset();
        q.voff.reset();
        q.d_vtab0.reset();
        q.max_slots = 0;
        return;
    }
    q.max_slots = (int64_t)h;
    q.slot16.alloc((size_t)std::max<int64_t>(q.entries, 1));
    k_slot_fill<<<grid_of(ns * kWave), kBlock, 0, st>>>(bi, ns, q.base.p, q.T.p, m.crow.p, m.d_sub_row0.p,
                                                        q.d_vtab0.p, nblk, nbins, q.vlo.p, q.voff.p, q.slot16.p);
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(st));
}

}  // namespace

void build_sell(Matrix &m, hipStream_t st, bool drop_csr) {
    if (m.has_sell) return;
    const mi_geometry &g = m.g;
    // tile rows padded to a multiple of 8 tiles: a stencil block (8 tiles) is
    // then a 4 x 64 pixel strip of one tile row
    m.tny = (g.ny + kStripTiles * kTileW - 1) / (kStripTiles * kTileW) * kStripTiles;
    m.npix_t = (int64_t)((g.nx + kTileH - 1) / kTileH) * kTileH * m.tny * kTileW;
    const bool f64 = m.has64;
    // forward
    if (!m.has_otf) {   // the checkpointed walk replaces the stored forward slices
        m.fs.P = choose_p(MI_SELL_PF, (double)m.nnz / std::max<int64_t>(m.nrows, 1));
        m.fs.C = kWave / m.fs.P;
        m.fs.slice0.assign(m.S + 1, 0);
        for (int s = 0; s < m.S; ++s)
            m.fs.slice0[s + 1] = m.fs.slice0[s] + (m.sub_row0[s + 1] - m.sub_row0[s] + m.fs.C - 1) / m.fs.C;
        FItems fi{m.rowptr.p, m.d_sub_row0.p, nullptr, m.S, m.fs.P, m.fs.C};
        fill_sell(m.fs, fi, m.idx.p, m.len32.p, f64 ? m.len64.p : nullptr, true, g.ny, m.tny, st);
        compact_sell(m.fs, st);
    }
    if (drop_csr) {
        m.rowptr.reset();
        m.idx.reset();
        m.len32.reset();
        m.len64.reset();
    }
    // back: one lane per pixel, four slices per 16 x 16 block (bp_item)
    m.bs.P = 1;
    m.bs.C = kWave;
    const int tnx = (g.nx + kTileH - 1) / kTileH;
    const int64_t nblk = bp_blocks_of(tnx, m.tny), per = nblk * 4;
    m.bs.slice0.assign(m.S + 1, 0);
    for (int s = 0; s < m.S; ++s) m.bs.slice0[s + 1] = m.bs.slice0[s] + per;
    BItems bi{m.colptr.p, nullptr, m.S, m.bs.P, m.bs.C, g.nx, g.ny, tnx, m.tny, m.npix, m.npix_t};
    fill_sell(m.bs, bi, m.crow.p, m.clen32.p, f64 ? m.clen64.p : nullptr, false, g.ny, m.tny, st);
    bi.slice0 = m.bs.d_slice0.p;
    build_slots(m, bi, nblk, st);
    compact_sell(m.bs, st);
    if (drop_csr) {
        m.colptr.reset();
        m.crow.reset();
        m.clen32.reset();
        m.clen64.reset();
        m.has_csr = false;
    }
    m.has_sell = true;
}

}  // namespace mi
