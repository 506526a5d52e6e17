// System-matrix construction: device ray tracing into a ray-major CSR, then
// a stable device radix sort into per-subset pixel-major CSC lists for the
// gather back projector.  Runs once per geometry; the matrices stay resident
// in HBM and every projection afterwards only streams them.
#include "runtime.hpp"
#include "tracers.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <numeric>

namespace mi {

namespace {

struct TraceArgs {
    int beam;
    int nviews, nbins, n;        // n: parallel-beam image size
    int flavour;
    FanGeom fan;
    const double *fcos, *fsin;   // fan: per-view double tables
    const double *dcos, *dsin;   // arc detector: per-bin double tables
    const float *pcos, *psin;    // parallel: per-view float tables
    const int *vlist;            // subset-ordered views
};

template <typename Emit>
__device__ __forceinline__ void trace_row(const TraceArgs &a, int64_t r, Emit &&emit) {
    const int view = a.vlist[r / a.nbins];
    const int bin = (int)(r % a.nbins);
    if (a.beam == MI_BEAM_FAN_FLAT || a.beam == MI_BEAM_FAN_ARC) {
        trace_fan(a.fan, a.fcos[view], a.fsin[view], a.dcos, a.dsin, bin, emit);
    } else if (a.flavour == 1) {
        trace_parallel<float>(a.psin[view], a.pcos[view], bin, a.nbins, a.n, emit);
    } else {
        trace_parallel<double>(a.psin[view], a.pcos[view], bin, a.nbins, a.n, emit);
    }
}

__global__ void k_trace_count(TraceArgs a, int64_t nrows, int64_t *counts) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int64_t n = 0;
    trace_row(a, r, [&](int, double) { ++n; });
    counts[r + 1] = n;
}

__global__ void k_trace_fill(TraceArgs a, int64_t nrows, const int64_t *rowptr, uint32_t *idx,
                             float *len32, double *len64) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int64_t k = rowptr[r];
    const int64_t end = rowptr[r + 1];
    trace_row(a, r, [&](int pix, double w) {
        if (k < end) {
            idx[k] = (uint32_t)pix;
            len32[k] = (float)w;
            if (len64) len64[k] = w;
        }
        ++k;
    });
}

// One wave per row of one subset: sort key (pixel), the entry's position in
// the subset's entry range and its row.
__global__ void k_csc_keys(const int64_t *rowptr, const uint32_t *idx, int64_t row0, int64_t nrows, int64_t e0,
                           uint32_t *keys, uint32_t *pos, uint32_t *rowid) {
    const int64_t r = row0 + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (r >= row0 + nrows) return;
    for (int64_t k = rowptr[r] + lane; k < rowptr[r + 1]; k += kWave) {
        keys[k - e0] = idx[k];
        pos[k - e0] = (uint32_t)(k - e0);
        rowid[k - e0] = (uint32_t)r;
    }
}

__global__ void k_csc_gather(const uint32_t *pos, int64_t n, int64_t e0, const uint32_t *rowid, const float *len32,
                             const double *len64, uint32_t *crow, float *clen32, double *clen64) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t k = pos[j];
    crow[e0 + j] = rowid[k];
    clen32[e0 + j] = len32[e0 + k];
    if (clen64) clen64[e0 + j] = len64[e0 + k];
}

// colptr of one subset from its sorted keys: entry j opens every column
// between the previous entry's key (exclusive) and its own (inclusive).
__global__ void k_csc_colptr(const uint32_t *keys, int64_t n, int64_t e0, int64_t ncols, int64_t *colptr) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t prev = (j == 0) ? -1 : (int64_t)keys[j - 1];
    for (int64_t c = prev + 1; c <= (int64_t)keys[j]; ++c) colptr[c] = e0 + j;
    if (j == n - 1)
        for (int64_t c = (int64_t)keys[j] + 1; c < ncols; ++c) colptr[c] = e0 + n;
}

// Counts entries whose pixel index is outside the image (must be zero before
// any kernel indexes an image with them).
__global__ void k_check_idx(const uint32_t *idx, int64_t nnz, uint32_t npix, unsigned *bad) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz && idx[k] >= npix) atomicAdd(bad, 1u);
}

__global__ void k_fill_i64(int64_t *p, int64_t n, int64_t v) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) p[j] = v;
}

__global__ void k_perm(const int *vlist, int nbins, int64_t nrows, int64_t *perm) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrows) perm[r] = (int64_t)vlist[r / nbins] * nbins + r % nbins;
}

inline unsigned grid_of(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// GoldenC symmetric tiled DDA (origin/GoldenC/src/proj.c:14-397), walked on
// the host: one traversal per (angle 1..45, tile, ray 0..383) feeds up to
// eight (view, bin) rays through the grid's rotations and mirror.  put(view,
// bin, pixel, length) is called in the order the reference accumulates the
// products (views 0 and 90 first, then the traversals).  The view-90 column
// differs between A (W/2 + k, proj.c:32) and AStar (W/2 - 1 + k, proj.c:225),
// so the adjoint list is walked separately.
template <typename Put>
void goldenc_emit(bool adjoint, Put &&put) {
    const int N = 512, NB = 767, MID = 383, TILE = 128, TX = 4, NT = 16;
    for (int ray = 0; ray < NB; ++ray) {
        int k = ray - NB / 2;
        if (k > N / 2 - 1 || k < 1 - N / 2) continue;
        for (int c = 0; c < N; ++c) put(0, ray, (N / 2 - 1 + k) * N + c, 1.0);
    }
    for (int ray = 0; ray < NB; ++ray) {
        int k = ray - NB / 2;
        if (k > N / 2 - 1 || k < 1 - N / 2) continue;
        int col = adjoint ? N / 2 - 1 + k : N / 2 + k;
        for (int r = 0; r < N; ++r) put(90, ray, r * N + col, 1.0);
    }
    auto eight = [&](int al, int ray, int ix, int iy, int dx, int dy, double w) {
        const int W = N;
        auto px = [&](int r, int c) { return r * N + c; };
        put(al, ray, px(iy + dy, ix + dx), w);
        put(90 + al, ray, px(W - 1 - ix - dx, iy + dy), w);
        if (ray != MID) {
            put(al, NB - 1 - ray, px(W - 1 - iy - dy, W - 1 - ix - dx), w);
            put(90 + al, NB - 1 - ray, px(ix + dx, W - 1 - iy - dy), w);
        }
        if (al != 45) {
            put(90 - al, ray, px(ix + dx, iy + dy), w);
            put(180 - al, ray, px(W - 1 - iy - dy, ix + dx), w);
            if (ray != MID) {
                put(90 - al, NB - 1 - ray, px(W - 1 - ix - dx, W - 1 - iy - dy), w);
                put(180 - al, NB - 1 - ray, px(iy + dy, W - 1 - ix - dx), w);
            }
        }
    };
    for (int al = 1; al < 46; ++al) {
        const double sf = std::sin(3.141593 / 180 * al), cf = std::cos(3.141593 / 180 * al);
        const double hx = 1.0 / sf, hy = 1.0 / cf, tn = sf / cf, rt = cf / sf;
        for (int tile = 0; tile < NT; ++tile)
            for (int ray = 0; ray < MID + 1; ++ray) {
                const int k = ray - NB / 2;
                const int dx = TILE * (TX - 1 - (tile % TX)), dy = TILE * (tile / TX);
                double c = cf * k + sf * k * tn;
                c = (N / 2 - dy) - tn * (N / 2 - TILE * (tile % TX)) - c;
                double fp, w, yc, xc;
                int ix, iy, step;
                const int incx = -1;
                if (c < 0) {
                    c = -c * rt;
                    ix = (int)c; iy = 0;
                    fp = ((double)(ix + 1)) - c;
                    ix = TILE - ix - 1;
                    w = hy * fp; yc = fp * tn; xc = -fp;
                    step = 2;
                } else {
                    ix = TILE - 1;
                    iy = (int)c;
                    fp = ((double)(iy + 1)) - c;
                    yc = -fp; w = hx * fp; xc = fp * rt;
                    step = 1;
                }
                while (ix >= 0 && ix < TILE && iy >= 0 && iy < TILE) {
                    if (step == 1) {
                        xc -= 1.0; yc += tn;
                        if (xc > 0) { eight(al, ray, ix, iy, dx, dy, hy); w -= hy; ix += incx; }
                        else { eight(al, ray, ix, iy, dx, dy, w); w = hy - w; iy++; step = 2; }
                    } else {
                        yc -= 1.0; xc += rt;
                        if (yc > 0) { w -= hx; eight(al, ray, ix, iy, dx, dy, hx); iy++; }
                        else { eight(al, ray, ix, iy, dx, dy, w); w = hx - w; ix += incx; step = 1; }
                    }
                }
            }
    }
}

struct GcRows {
    std::vector<std::vector<std::pair<uint32_t, double>>> rows;  // [view*767 + bin]
};

void goldenc_walk(GcRows &out, bool adjoint) {
    out.rows.assign((size_t)180 * 767, {});
    goldenc_emit(adjoint, [&](int view, int bin, int pix, double w) {
        out.rows[(size_t)view * 767 + bin].emplace_back((uint32_t)pix, w);
    });
}

// Host CSR in subset row order from per-(view, bin) lists.
void gc_to_csr(const GcRows &gr, const std::vector<int> &vlist, int nbins, std::vector<int64_t> &rowptr,
               std::vector<uint32_t> &idx, std::vector<double> &len) {
    const int64_t nrows = (int64_t)vlist.size() * nbins;
    rowptr.assign(nrows + 1, 0);
    for (int64_t r = 0; r < nrows; ++r)
        rowptr[r + 1] = rowptr[r] + (int64_t)gr.rows[(size_t)vlist[r / nbins] * nbins + r % nbins].size();
    idx.resize(rowptr[nrows]);
    len.resize(rowptr[nrows]);
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t k = rowptr[r];
        for (auto &e : gr.rows[(size_t)vlist[r / nbins] * nbins + r % nbins]) {
            idx[k] = e.first;
            len[k] = e.second;
            ++k;
        }
    }
}

// Matrix streams are read four entries at a time from the 4-aligned entry at
// or below a row's start, so every stream carries kPad zero entries at its end
// (index 0 and length 0: a valid pixel/row with no weight).
constexpr size_t kPad = 4;
template <typename T>
void alloc_padded(DevBuf<T> &b, size_t n, hipStream_t st) {
    b.alloc(n + kPad);
    MI_HIP(hipMemsetAsync(b.p + n, 0, kPad * sizeof(T), st));
}

template <typename T>
void upload_padded(DevBuf<T> &b, const std::vector<T> &h, hipStream_t st) {
    alloc_padded(b, h.size(), st);
    if (!h.empty()) MI_HIP(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

template <typename T>
void upload(DevBuf<T> &b, const std::vector<T> &h, hipStream_t st) {
    b.alloc(h.size());
    if (!h.empty()) MI_HIP(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

void inclusive_scan_i64(int64_t *d, int64_t n, hipStream_t st) {
    size_t tmp = 0;
    MI_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, d, d, (int)n, st));
    DevBuf<unsigned char> t(tmp ? tmp : 1);
    MI_HIP(hipcub::DeviceScan::InclusiveSum(t.p, tmp, d, d, (int)n, st));
    MI_HIP(hipStreamSynchronize(st));
}

void check_indices(const uint32_t *idx, int64_t nnz, int npix, hipStream_t st) {
    if (nnz == 0) return;
    DevBuf<unsigned> bad(1);
    MI_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned), st));
    k_check_idx<<<grid_of(nnz), kBlock, 0, st>>>(idx, nnz, (uint32_t)npix, bad.p);
    MI_HIP(hipGetLastError());
    unsigned h = 0;
    MI_HIP(hipMemcpyAsync(&h, bad.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    MI_REQUIRE(h == 0, MI_E_STATE, "tracer produced " + std::to_string(h) + " pixel indices outside the image");
}

// Per-subset CSC from a subset-ordered CSR (device arrays).  The rows of a
// subset hold a contiguous entry range; each range is sorted on its own by
// pixel (stable radix sort), so positions stay 32-bit however large the whole
// matrix is, and subset s's columns occupy the same entry range in the CSC.
void build_csc(Matrix &m, const int64_t *rowptr, const uint32_t *idx, const float *len32, const double *len64,
               hipStream_t st) {
    check_indices(idx, m.nnz, m.npix, st);
    const int64_t nnz = m.nnz, ncols = (int64_t)m.S * m.npix;
    m.colptr.alloc(ncols + 1);
    alloc_padded(m.crow, nnz, st);
    alloc_padded(m.clen32, nnz, st);
    if (len64) alloc_padded(m.clen64, nnz, st);
    k_fill_i64<<<grid_of(ncols + 1), kBlock, 0, st>>>(m.colptr.p, ncols + 1, nnz);
    MI_HIP(hipGetLastError());
    if (nnz == 0) return;
    std::vector<int64_t> e(m.S + 1);
    for (int s = 0; s <= m.S; ++s)
        MI_HIP(hipMemcpyAsync(&e[s], rowptr + m.sub_row0[s], sizeof(int64_t), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    int64_t most = 0;
    for (int s = 0; s < m.S; ++s) most = std::max(most, e[s + 1] - e[s]);
    MI_REQUIRE(most < (int64_t)INT32_MAX && m.nrows < (int64_t)UINT32_MAX, MI_E_LIMIT,
               "one subset of the system matrix exceeds 2^31 entries");
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < m.npix) ++bits;
    DevBuf<uint32_t> keys(most), keys2(most), pos(most), pos2(most), rowid(most);
    size_t tmp = 0;
    MI_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys.p, keys2.p, pos.p, pos2.p, (int)most, 0, bits, st));
    DevBuf<unsigned char> t(tmp ? tmp : 1);
    for (int s = 0; s < m.S; ++s) {
        const int64_t n = e[s + 1] - e[s], nr = m.sub_row0[s + 1] - m.sub_row0[s];
        if (n == 0) {
            k_fill_i64<<<grid_of(m.npix), kBlock, 0, st>>>(m.colptr.p + (int64_t)s * m.npix, m.npix, e[s]);
            MI_HIP(hipGetLastError());
            continue;
        }
        k_csc_keys<<<grid_of(nr * kWave), kBlock, 0, st>>>(rowptr, idx, m.sub_row0[s], nr, e[s], keys.p, pos.p,
                                                           rowid.p);
        MI_HIP(hipGetLastError());
        size_t tn = tmp;
        MI_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tn, keys.p, keys2.p, pos.p, pos2.p, (int)n, 0, bits, st));
        k_csc_gather<<<grid_of(n), kBlock, 0, st>>>(pos2.p, n, e[s], rowid.p, len32, len64, m.crow.p, m.clen32.p,
                                                    len64 ? m.clen64.p : nullptr);
        MI_HIP(hipGetLastError());
        k_csc_colptr<<<grid_of(n), kBlock, 0, st>>>(keys2.p, n, e[s], m.npix, m.colptr.p + (int64_t)s * m.npix);
        MI_HIP(hipGetLastError());
    }
    MI_HIP(hipStreamSynchronize(st));
}

// Checkpoints of the forward walk: one thread per row (subset order) walks
// its ray once and records the tracer state at the first entry of each of
// its kOtfSeg segments (segment p of the row's block: wave p / kOtfPerWave,
// lane kOtfRays (p % kOtfPerWave) + row % kOtfRays).
__global__ void k_otf_fill(TraceArgs a, int64_t nrows, const int64_t *rowptr, const int64_t *sub_row0,
                           const int64_t *slice0, int S, double *lam, double *ax, double *ay, double *stx,
                           double *sty, int32_t *cx, int32_t *cy, int32_t *sg, int32_t *cnt) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int s = 0;
    while (s + 1 < S && sub_row0[s + 1] <= r) ++s;
    const int64_t li = r - sub_row0[s];
    const int64_t lane0 = (slice0[s] + li / kOtfRays * kOtfWaves) * kWave + li % kOtfRays;
    const int64_t L = rowptr[r + 1] - rowptr[r];
    const int64_t Ls = (L + kOtfSeg - 1) / kOtfSeg;
    const int view = a.vlist[r / a.nbins];
    const int bin = (int)(r % a.nbins);
    double sx, sy, dx, dy;
    ray_endpoints(a.fan, a.fcos[view], a.fsin[view], a.dcos, a.dsin, bin, sx, sy, dx, dy);
    SiddonState st{};
    if (L > 0) siddon_init(sx, sy, dx, dy, a.fan.nx, a.fan.ny, a.fan.vx, a.fan.vy, st);
    int64_t k = 0;
    for (int p = 0; p < kOtfSeg; ++p) {
        const int64_t q = lane0 + (int64_t)(p / kOtfPerWave) * kWave + (p % kOtfPerWave) * kOtfRays;
        const int64_t b = min(L, p * Ls), e = min(L, (p + 1) * Ls);
        for (; k < b; ++k) {
            int x, y;
            double w;
            siddon_step(st, x, y, w);
        }
        lam[q] = st.lam;
        ax[q] = st.ax;
        ay[q] = st.ay;
        stx[q] = st.stx;
        sty[q] = st.sty;
        cx[q] = st.cx;
        cy[q] = st.cy;
        sg[q] = (st.sgx > 0 ? 1 : 0) | (st.sgy > 0 ? 2 : 0);
        cnt[q] = (int32_t)(e - b);
    }
}

void build_otf(Matrix &m, const TraceArgs &a, hipStream_t st) {
    Otf &o = m.otf;
    o.slice0.assign(m.S + 1, 0);
    for (int s = 0; s < m.S; ++s)
        o.slice0[s + 1] = o.slice0[s] + (m.sub_row0[s + 1] - m.sub_row0[s] + kOtfRays - 1) / kOtfRays * kOtfWaves;
    const int64_t lanes = o.slice0.back() * kWave;
    upload(o.d_slice0, o.slice0, st);
    for (auto *b : {&o.lam, &o.ax, &o.ay, &o.stx, &o.sty}) b->alloc((size_t)lanes);
    for (auto *b : {&o.cx, &o.cy, &o.sg, &o.cnt}) b->alloc((size_t)lanes);
    // Lanes of missing rows (the last slice of a subset) walk nothing.
    MI_HIP(hipMemsetAsync(o.cnt.p, 0, o.cnt.bytes(), st));
    MI_HIP(hipMemsetAsync(o.sg.p, 0, o.sg.bytes(), st));
    for (auto *b : {&o.lam, &o.ax, &o.ay, &o.stx, &o.sty}) MI_HIP(hipMemsetAsync(b->p, 0, b->bytes(), st));
    for (auto *b : {&o.cx, &o.cy}) MI_HIP(hipMemsetAsync(b->p, 0, b->bytes(), st));
    k_otf_fill<<<grid_of(m.nrows, 64), 64, 0, st>>>(a, m.nrows, m.rowptr.p, m.d_sub_row0.p, o.d_slice0.p, m.S,
                                                   o.lam.p, o.ax.p, o.ay.p, o.stx.p, o.sty.p, o.cx.p, o.cy.p,
                                                   o.sg.p, o.cnt.p);
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(st));
    m.has_otf = true;
}

// Entries of each subset's rows, from the device row pointer.
void subset_counts(Matrix &m, hipStream_t st) {
    std::vector<int64_t> at(m.S + 1);
    for (int s = 0; s <= m.S; ++s)
        MI_HIP(hipMemcpyAsync(&at[s], m.rowptr.p + m.sub_row0[s], sizeof(int64_t), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    m.sub_nnz.assign(m.S, 0);
    for (int s = 0; s < m.S; ++s) m.sub_nnz[s] = at[s + 1] - at[s];
}

}  // namespace

std::vector<int> subset_order(int S) {
    int bits = 0;
    while ((1 << bits) < S) ++bits;
    std::vector<int> o;
    for (int i = 0; i < (1 << bits); ++i) {
        int r = 0;
        for (int b = 0; b < bits; ++b)
            if (i & (1 << b)) r |= 1 << (bits - 1 - b);
        if (r < S) o.push_back(r);
    }
    return o;
}

size_t Matrix::bytes() const {
    return otf.bytes() + d_vlist.bytes() + d_sub_row0.bytes() + rowptr.bytes() + idx.bytes() + len32.bytes() + len64.bytes() + colptr.bytes() +
           crow.bytes() + clen32.bytes() + clen64.bytes() + perm.bytes() + fs.bytes() + bs.bytes() +
           seq_bytes(seq.get());
}

std::unique_ptr<Matrix> build_matrix(const mi_geometry &g, int S, bool want64, hipStream_t st, bool otf) {
    validate_geometry(g);
    MI_REQUIRE(S >= 1 && S <= g.n_views, MI_E_ARG, "n_subsets must be in [1, n_views]");
    auto m = std::make_unique<Matrix>();
    m->g = g;
    m->view_deg = view_table(g);
    m->g.view_deg = m->view_deg.empty() ? nullptr : m->view_deg.data();
    m->S = S;
    m->npix = g.nx * g.ny;
    m->nrows = (int64_t)g.n_views * g.n_bins;
    m->has64 = want64;
    const auto order = subset_order(S);
    m->sub_row0.assign(S + 1, 0);
    for (int s = 0; s < S; ++s) {
        int cnt = 0;
        for (int v = 0; v < g.n_views; ++v)
            if (v % S == order[s]) { m->vlist.push_back(v); ++cnt; }
        m->sub_row0[s + 1] = m->sub_row0[s] + (int64_t)cnt * g.n_bins;
    }
    upload(m->d_vlist, m->vlist, st);
    upload(m->d_sub_row0, m->sub_row0, st);
    m->perm.alloc(m->nrows);
    k_perm<<<grid_of(m->nrows), kBlock, 0, st>>>(m->d_vlist.p, g.n_bins, m->nrows, m->perm.p);
    MI_HIP(hipGetLastError());

    if (g.beam == MI_BEAM_GOLDENC) {
        GcRows fwd, adj;
        goldenc_walk(fwd, false);
        goldenc_walk(adj, true);
        std::vector<int64_t> rp, rpa;
        std::vector<uint32_t> ix, ixa;
        std::vector<double> ln, lna;
        gc_to_csr(fwd, m->vlist, g.n_bins, rp, ix, ln);
        gc_to_csr(adj, m->vlist, g.n_bins, rpa, ixa, lna);
        MI_REQUIRE(rp.back() == rpa.back(), MI_E_STATE, "GoldenC adjoint walk size mismatch");
        m->nnz = rp.back();
        std::vector<float> ln32(ln.begin(), ln.end()), lna32(lna.begin(), lna.end());
        upload(m->rowptr, rp, st);
        upload_padded(m->idx, ix, st);
        upload_padded(m->len32, ln32, st);
        if (want64) upload_padded(m->len64, ln, st);
        DevBuf<int64_t> arp;
        DevBuf<uint32_t> aix;
        DevBuf<float> aln32;
        DevBuf<double> aln64;
        upload(arp, rpa, st);
        upload(aix, ixa, st);
        upload(aln32, lna32, st);
        if (want64) upload(aln64, lna, st);
        build_csc(*m, arp.p, aix.p, aln32.p, want64 ? aln64.p : nullptr, st);
        subset_counts(*m, st);
        return m;
    }

    TraceArgs a{};
    a.beam = g.beam;
    a.nviews = g.n_views;
    a.nbins = g.n_bins;
    a.n = g.nx;
    a.flavour = g.flavour;
    a.vlist = m->d_vlist.p;
    DevBuf<double> dcos, dsin, bcos, bsin;
    DevBuf<float> pcos, psin;
    if (g.beam == MI_BEAM_FAN_FLAT) {
        std::vector<double> c(g.n_views), s(g.n_views);
        const double span = g.span_deg * (M_PI / 180.0);
        for (int v = 0; v < g.n_views; ++v) {
            double th = span * (double)v / (double)g.n_views;
            c[v] = std::cos(th);
            s[v] = std::sin(th);
        }
        upload(dcos, c, st);
        upload(dsin, s, st);
        a.fan = FanGeom{g.nx, g.ny, g.n_bins, g.pixel_mm, g.pixel_mm, g.bin_mm, g.sod_mm, g.sdd_mm, g.det_offset, 0};
        a.fcos = dcos.p;
        a.fsin = dsin.p;
    } else if (g.beam == MI_BEAM_FAN_ARC) {
        // load_table of helical/rebin/src/ms2d.cpp:30-62: source angle in
        // degrees -> cos(angle / 180 * pi); detector element j (from
        // -n_bins/2) -> cos((j + OFF_CENTER) * LENGTH_PER_DET * pi / 180 + pi)
        // (the angle of view v: the slice's _pos table, or uniform over span_deg)
        std::vector<double> c(g.n_views), s(g.n_views), dc(g.n_bins), ds(g.n_bins);
        for (int v = 0; v < g.n_views; ++v) {
            const double angle = g.view_deg ? g.view_deg[v] : g.span_deg * (double)v / (double)g.n_views;
            c[v] = std::cos(angle / 180.0 * M_PI);
            s[v] = std::sin(angle / 180.0 * M_PI);
        }
        const double lpd = g.bin_mm * M_PI / 180.;
        const int half = g.n_bins / 2;
        for (int j = -half; j < half; ++j) {
            dc[j + half] = std::cos((j + g.det_offset) * lpd + M_PI);
            ds[j + half] = std::sin((j + g.det_offset) * lpd + M_PI);
        }
        upload(dcos, c, st);
        upload(dsin, s, st);
        upload(bcos, dc, st);
        upload(bsin, ds, st);
        a.fan = FanGeom{g.nx, g.ny, g.n_bins, g.pixel_mm, g.pixel_mm, g.bin_mm, g.sod_mm, g.sdd_mm, g.det_offset, 1};
        a.fcos = dcos.p;
        a.fsin = dsin.p;
        a.dcos = bcos.p;
        a.dsin = bsin.p;
    } else {
        // posit() of msbeam/CPU_OpenMP/src/proj.c:55-56 (double pi, double
        // libm) or blockrecon/src/wray.cpp:11-12 (float pi, float overloads).
        std::vector<float> c(g.n_views), s(g.n_views);
        for (int v = 0; v < g.n_views; ++v) {
            if (g.flavour == 1) {
                const float pi = 3.14159265359f;
                s[v] = sinf(pi / g.n_views * v);
                c[v] = cosf(pi / g.n_views * v);
            } else {
                s[v] = (float)std::sin(3.14159265359 / g.n_views * v);
                c[v] = (float)std::cos(3.14159265359 / g.n_views * v);
            }
        }
        upload(pcos, c, st);
        upload(psin, s, st);
        a.pcos = pcos.p;
        a.psin = psin.p;
    }

    m->rowptr.alloc(m->nrows + 1);
    MI_HIP(hipMemsetAsync(m->rowptr.p, 0, sizeof(int64_t), st));
    k_trace_count<<<grid_of(m->nrows, 64), 64, 0, st>>>(a, m->nrows, m->rowptr.p);
    MI_HIP(hipGetLastError());
    inclusive_scan_i64(m->rowptr.p, m->nrows + 1, st);
    MI_HIP(hipMemcpy(&m->nnz, m->rowptr.p + m->nrows, sizeof(int64_t), hipMemcpyDeviceToHost));
    alloc_padded(m->idx, m->nnz, st);
    alloc_padded(m->len32, m->nnz, st);
    if (want64) alloc_padded(m->len64, m->nnz, st);
    k_trace_fill<<<grid_of(m->nrows, 64), 64, 0, st>>>(a, m->nrows, m->rowptr.p, m->idx.p, m->len32.p,
                                                       want64 ? m->len64.p : nullptr);
    MI_HIP(hipGetLastError());
    build_csc(*m, m->rowptr.p, m->idx.p, m->len32.p, want64 ? m->len64.p : nullptr, st);
    subset_counts(*m, st);
    if (otf && (g.beam == MI_BEAM_FAN_FLAT || g.beam == MI_BEAM_FAN_ARC)) build_otf(*m, a, st);
    return m;
}

void goldenc_lists(bool pixel_major, std::vector<int64_t> &ptr, std::vector<uint32_t> &other,
                   std::vector<double> &len) {
    const int64_t nitems = pixel_major ? (int64_t)512 * 512 : (int64_t)180 * 767;
    ptr.assign(nitems + 1, 0);
    // Count pass, then a fill pass that appends each item's entries in the
    // order the walk emits them.
    goldenc_emit(pixel_major, [&](int view, int bin, int pix, double) {
        ++ptr[1 + (pixel_major ? pix : (int64_t)view * 767 + bin)];
    });
    for (int64_t i = 0; i < nitems; ++i) ptr[i + 1] += ptr[i];
    other.resize(ptr[nitems]);
    len.resize(ptr[nitems]);
    std::vector<int64_t> at(ptr.begin(), ptr.end() - 1);
    goldenc_emit(pixel_major, [&](int view, int bin, int pix, double w) {
        const int64_t row = (int64_t)view * 767 + bin;
        const int64_t k = at[pixel_major ? pix : row]++;
        other[k] = (uint32_t)(pixel_major ? row : pix);
        len[k] = w;
    });
}

}  // namespace mi
