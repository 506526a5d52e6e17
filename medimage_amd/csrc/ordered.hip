// Reference-order FP64 operators and the GoldenC CG / SD drivers built on
// them (see ordered.hpp).  Every sum here is added in the order the
// reference program adds it, so results equal the reference's bit for bit:
//   * A: one lane per ray sums its products in the order proj.c emits them;
//   * AStar: one lane per pixel, likewise;
//   * energies, norms, inner products: one wave per sum, terms computed by
//     the 64 lanes in parallel and folded into one running double in
//     row-major order (v_readlane, no reassociation).
#include "ordered.hpp"
#include "gc_stencil.hpp"

#include <algorithm>
#include <cstring>
#include <limits>

namespace mi {

void goldenc_lists(bool pixel_major, std::vector<int64_t> &ptr, std::vector<uint32_t> &other,
                   std::vector<double> &len);

namespace {

struct SeqView {
    int64_t items;
    const int64_t *base;
    const int32_t *steps, *cnt;
    const uint32_t *other;
    const double *len;
};
SeqView view_of(const SeqLists &L) {
    return SeqView{L.items, L.base.p, L.steps.p, L.cnt.p, L.other.p, L.len.p};
}

// One lane per item: s = sum over the item's entries, in list order, of
// x[other] * len in T (product rounded, then added: no contraction; the
// lengths are stored widened to double, exactly, and narrowed back).
template <typename T>
__device__ __forceinline__ T seq_item_sum(const SeqView &L, int64_t i, const T *__restrict__ x) {
    const int64_t gi = i >> 6;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t b = L.base[gi];
    const int steps = L.steps[gi];
    const int c = i < L.items ? L.cnt[i] : 0;
    T s = T(0);
    for (int t = 0; t < steps; t += 4) {
        uint32_t k[4];
        T l[4], xv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t at = b + (int64_t)(t + j) * kWave + lane;
            k[j] = L.other[at];
            l[j] = (T)L.len[at];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) xv[j] = x[k[j]];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (t + j < c) s = s + xv[j] * l[j];
    }
    return s;
}

template <typename T, bool RESID>
__global__ __launch_bounds__(kBlock) void k_seq_fp(SeqView L, const T *__restrict__ x, T *__restrict__ out,
                                                   const T *__restrict__ g, T *__restrict__ resid) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (((i >> 6) << 6) >= L.items) return;
    const T s = seq_item_sum<T>(L, i, x);
    if (i < L.items) {
        out[i] = s;
        if (RESID) resid[i] = g[i] - s;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_seq_bp(SeqView L, int nx, int ny, const T *__restrict__ r,
                                                   const T *__restrict__ u, const T *__restrict__ v, double alpha,
                                                   double eps, T *__restrict__ d) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (((i >> 6) << 6) >= L.items) return;
    const T s = seq_item_sum<T>(L, i, r);
    if (i < L.items) {
        T dv = s;
        if (u) dv = (T)((double)dv + alpha * gc_reg(u, v, (int)(i / ny), (int)(i % ny), nx, ny, eps));
        d[i] = dv;
    }
}

// ---------------------------------------------------------------------------
// Sequential sums.  Terms (origin/GoldenC/src/main.c):
enum ChainTerm : int {
    CT_SQDIFF = 0,   // (a - b)^2, rows:  sqr(Af - g), :127-130
    CT_SQ = 1,       // a^2:              sqr(q), :162-164; sqr(d), :318-320
    CT_DOT = 2,      // a b:              p d, :159-161; Ad q, :241-243
    CT_REG_INT = 3,  // interior energy of f = a, v = b, :135-139
    CT_NV = 4,       // nV of p = a, v = b (p zero padded), :152-158
    CT_NABLA2 = 5,   // two terms per pixel of d = a (zero padded), :313-316
    CT_NDSQ = 6      // a b^2:            n d^2, :322-325
};
struct ChainSeg {
    int term;
    int64_t n;             // terms (pixels for CT_NABLA2)
    const double *a, *b;
};
struct ChainDesc {
    ChainSeg seg[2];       // the second segment continues the first's sum
    int nseg;
    double *out;
};
struct ChainArgs {
    ChainDesc c[4];
    int nchains;
    int nx, ny;
    double alpha, beta, eps;
    int edge_c;            // epilogue: chains are nNablaD, nD, sum n d^2 ->
    Scalars *sc;           //   sc->e_c (main.c:329); skipped once edge_done
};

__device__ __forceinline__ double lane_value(double x, int k) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), k);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), k);
    return __hiloint2double(hi, lo);
}

template <int TERM>
__device__ __forceinline__ void chain_term(const ChainSeg &s, const ChainArgs &a, int64_t e, double &t0, double &t1) {
    const double *A = s.a, *B = s.b;
    const int nx = a.nx, ny = a.ny;
    if constexpr (TERM == CT_SQDIFF) {
        const double x = A[e] - B[e];
        t0 = x * x;
    } else if constexpr (TERM == CT_SQ) {
        t0 = A[e] * A[e];
    } else if constexpr (TERM == CT_DOT) {
        t0 = A[e] * B[e];
    } else if constexpr (TERM == CT_REG_INT) {
        const int64_t q = (1 + e / (ny - 2)) * (int64_t)ny + 1 + e % (ny - 2);
        const double z = A[q], zu = A[q - ny], zl = A[q - 1], w = B[q];
        t0 = a.alpha * (w * w + 0.01 * (a.eps * a.eps)) * ((z - zu) * (z - zu) + (z - zl) * (z - zl));
    } else if constexpr (TERM == CT_NV) {
        const int i = (int)(e / ny), j = (int)(e % ny);
        const double z = A[e], zu = gc_ld(A, i - 1, j, nx, ny, 0.0), zl = gc_ld(A, i, j - 1, nx, ny, 0.0);
        const double w = B[e];
        t0 = (w * w + 0.01 * (a.eps * a.eps)) * ((z - zu) * (z - zu) + (z - zl) * (z - zl));
    } else if constexpr (TERM == CT_NABLA2) {
        const int i = (int)(e / ny), j = (int)(e % ny);
        const double x = 0.5 * (gc_ld(A, i + 1, j, nx, ny, 0.0) - gc_ld(A, i - 1, j, nx, ny, 0.0));
        const double y = 0.5 * (gc_ld(A, i, j + 1, nx, ny, 0.0) - gc_ld(A, i, j - 1, nx, ny, 0.0));
        t0 = x * x;
        t1 = y * y;
    } else {
        t0 = A[e] * (B[e] * B[e]);
    }
}

// Folds the segment's terms into acc in index order.  The next 64 terms
// are loaded before the current ones are folded.
template <int TERM>
__device__ double chain_segment(const ChainSeg &s, const ChainArgs &a, double acc) {
    constexpr bool two = TERM == CT_NABLA2;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t n = s.n;
    double c0 = 0.0, c1 = 0.0;
    if (lane < n) chain_term<TERM>(s, a, lane, c0, c1);
    for (int64_t b0 = 0; b0 < n; b0 += kWave) {
        double n0 = 0.0, n1 = 0.0;
        if (b0 + kWave + lane < n) chain_term<TERM>(s, a, b0 + kWave + lane, n0, n1);
        if (n - b0 >= kWave) {
#pragma unroll
            for (int k = 0; k < kWave; ++k) {
                acc = acc + lane_value(c0, k);
                if (two) acc = acc + lane_value(c1, k);
            }
        } else {
            const int m = (int)(n - b0);
            for (int k = 0; k < m; ++k) {
                acc = acc + lane_value(c0, k);
                if (two) acc = acc + lane_value(c1, k);
            }
        }
        c0 = n0;
        c1 = n1;
    }
    return acc;
}

__global__ void k_chains(ChainArgs a) {
    __shared__ double res[4];
    if (a.edge_c && (a.sc->edge_done || a.sc->converged)) return;
    const int w = threadIdx.x / kWave;
    if (w < a.nchains) {
        const ChainDesc &c = a.c[w];
        double acc = 0.0;
        for (int k = 0; k < c.nseg; ++k) {
            const ChainSeg &s = c.seg[k];
            switch (s.term) {
                case CT_SQDIFF: acc = chain_segment<CT_SQDIFF>(s, a, acc); break;
                case CT_SQ: acc = chain_segment<CT_SQ>(s, a, acc); break;
                case CT_DOT: acc = chain_segment<CT_DOT>(s, a, acc); break;
                case CT_REG_INT: acc = chain_segment<CT_REG_INT>(s, a, acc); break;
                case CT_NV: acc = chain_segment<CT_NV>(s, a, acc); break;
                case CT_NABLA2: acc = chain_segment<CT_NABLA2>(s, a, acc); break;
                default: acc = chain_segment<CT_NDSQ>(s, a, acc); break;
            }
        }
        if ((threadIdx.x & (kWave - 1)) == 0) {
            res[w] = acc;
            if (c.out) *c.out = acc;
        }
    }
    if (!a.edge_c) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        Scalars *sc = a.sc;
        const double nnd = res[0], nd = res[1], tmp = res[2];
        sc->e_nnd = nnd;
        sc->e_nd = nd;
        sc->e_nd2 = tmp;
        sc->e_c = nd / (a.alpha * tmp + a.beta * a.eps * nnd + (a.beta / (4 * a.eps)) * nd);
    }
}

void launch_chains(hipStream_t st, const ChainArgs &a) {
    MI_REQUIRE(a.nchains >= 1 && a.nchains <= 4, MI_E_ARG, "1 to 4 sums per launch");
    k_chains<<<1, kWave * a.nchains, 0, st>>>(a);
    MI_HIP(hipGetLastError());
}

ChainSeg seg(int term, int64_t n, const double *a, const double *b = nullptr) { return ChainSeg{term, n, a, b}; }
ChainDesc chain(double *out, ChainSeg s0) {
    ChainDesc c{};
    c.seg[0] = s0;
    c.nseg = 1;
    c.out = out;
    return c;
}
ChainDesc chain(double *out, ChainSeg s0, ChainSeg s1) {
    ChainDesc c = chain(out, s0);
    c.seg[1] = s1;
    c.nseg = 2;
    return c;
}

// ---------------------------------------------------------------------------
// Host construction of the lists.

template <typename T>
void upload_vec(DevBuf<T> &b, const std::vector<T> &h, hipStream_t st) {
    b.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty()) MI_HIP(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

void build_lists(SeqLists &L, const std::vector<int64_t> &ptr, const std::vector<uint32_t> &oth,
                 const std::vector<double> &len, hipStream_t st) {
    const int64_t items = (int64_t)ptr.size() - 1;
    const int64_t groups = (items + kWave - 1) / kWave;
    std::vector<int64_t> base(groups);
    std::vector<int32_t> steps(groups), cnt(items);
    int64_t total = 0;
    for (int64_t gi = 0; gi < groups; ++gi) {
        int64_t most = 0;
        for (int64_t i = gi * kWave; i < std::min(items, (gi + 1) * kWave); ++i) {
            cnt[i] = (int32_t)(ptr[i + 1] - ptr[i]);
            most = std::max<int64_t>(most, cnt[i]);
        }
        MI_REQUIRE(most < (int64_t)1 << 30, MI_E_LIMIT, "entry list too long");
        steps[gi] = (int32_t)((most + 3) / 4 * 4);
        base[gi] = total;
        total += (int64_t)steps[gi] * kWave;
    }
    std::vector<uint32_t> ho((size_t)total, 0u);
    std::vector<double> hl((size_t)total, 0.0);
    for (int64_t i = 0; i < items; ++i) {
        const int64_t b = base[i / kWave] + i % kWave;
        for (int64_t t = 0; t < cnt[i]; ++t) {
            ho[b + t * kWave] = oth[ptr[i] + t];
            hl[b + t * kWave] = len[ptr[i] + t];
        }
    }
    L.items = items;
    upload_vec(L.base, base, st);
    upload_vec(L.steps, steps, st);
    upload_vec(L.cnt, cnt, st);
    upload_vec(L.other, ho, st);
    upload_vec(L.len, hl, st);
    MI_HIP(hipStreamSynchronize(st));
}

template <typename T>
std::vector<T> download(const T *p, int64_t n, hipStream_t st) {
    std::vector<T> h((size_t)n);
    if (n > 0) MI_HIP(hipMemcpyAsync(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    return h;
}

}  // namespace

size_t seq_bytes(const SeqMatrix *q) { return q ? q->bytes() : 0; }

const SeqMatrix &get_seq(hipStream_t st, Matrix &m) {
    if (m.seq) return *m.seq;
    MI_REQUIRE(m.S == 1, MI_E_STATE, "reference-order lists need a one-subset matrix");
    auto q = std::make_shared<SeqMatrix>();
    std::vector<int64_t> ptr;
    std::vector<uint32_t> oth;
    std::vector<double> len;
    if (m.g.beam == MI_BEAM_GOLDENC) {
        goldenc_lists(false, ptr, oth, len);
        build_lists(q->rows, ptr, oth, len, st);
        goldenc_lists(true, ptr, oth, len);
        build_lists(q->pixels, ptr, oth, len, st);
    } else {
        // The tracer's lengths: double for the fan beams, float (stored
        // widened) for the parallel beam.
        MI_REQUIRE(m.has_csr && (m.has64 || m.g.beam == MI_BEAM_PARALLEL), MI_E_STATE,
                   "reference-order lists need the tracer's CSR / CSC");
        auto lengths = [&](const DevBuf<double> &l64, const DevBuf<float> &l32) {
            if (m.has64) return download(l64.p, m.nnz, st);
            const std::vector<float> f = download(l32.p, m.nnz, st);
            return std::vector<double>(f.begin(), f.end());
        };
        ptr = download(m.rowptr.p, m.nrows + 1, st);
        oth = download(m.idx.p, m.nnz, st);
        len = lengths(m.len64, m.len32);
        build_lists(q->rows, ptr, oth, len, st);
        ptr = download(m.colptr.p, (int64_t)m.npix + 1, st);
        oth = download(m.crow.p, m.nnz, st);
        len = lengths(m.clen64, m.clen32);
        build_lists(q->pixels, ptr, oth, len, st);
    }
    m.seq = q;
    return *m.seq;
}

template <typename T>
void seq_forward(hipStream_t st, const SeqMatrix &q, const T *x, T *out, const T *g, T *resid) {
    const SeqView L = view_of(q.rows);
    const unsigned grid = (unsigned)((L.items + kBlock - 1) / kBlock);
    if (g) k_seq_fp<T, true><<<grid, kBlock, 0, st>>>(L, x, out, g, resid);
    else k_seq_fp<T, false><<<grid, kBlock, 0, st>>>(L, x, out, nullptr, nullptr);
    MI_HIP(hipGetLastError());
}

template <typename T>
void seq_backproject(hipStream_t st, const SeqMatrix &q, int nx, int ny, const T *r, const T *u, const T *v,
                     double alpha, double eps, T *d) {
    const SeqView L = view_of(q.pixels);
    MI_REQUIRE(L.items == (int64_t)nx * ny, MI_E_STATE, "pixel lists do not match the image");
    const unsigned grid = (unsigned)((L.items + kBlock - 1) / kBlock);
    k_seq_bp<T><<<grid, kBlock, 0, st>>>(L, nx, ny, r, u, v, alpha, eps, d);
    MI_HIP(hipGetLastError());
}

template void seq_forward<float>(hipStream_t, const SeqMatrix &, const float *, float *, const float *, float *);
template void seq_forward<double>(hipStream_t, const SeqMatrix &, const double *, double *, const double *,
                                  double *);
template void seq_backproject<float>(hipStream_t, const SeqMatrix &, int, int, const float *, const float *,
                                     const float *, double, double, float *);
template void seq_backproject<double>(hipStream_t, const SeqMatrix &, int, int, const double *, const double *,
                                      const double *, double, double, double *);

// origin/GoldenC/src/main.c:81-257.  gen[] slots: 0 AT_f, 2 normAp, 3 nV,
// 4 pSkalard, 5 max |c1 p|, 6 AdSkalarAp.  Returns the CG iterations run;
// *why: 0 iteration cap, 1 IM_Tol, 2 unstable (reverted).
int gc_image_step_ordered(mi_context *ctx, Matrix &m, JobState<double> &js, double alpha, double eps,
                          int iter_image, int *why) {
    const hipStream_t st = ctx->stream;
    const SeqMatrix &q = get_seq(st, m);
    const int nx = js.nx, ny = js.ny;
    const int64_t npix = js.npix, nrows = js.nrows;
    const int64_t nint = (nx > 2 && ny > 2) ? (int64_t)(nx - 2) * (ny - 2) : 0;
    Scalars *sc = js.sc.p;
    const RedBuf rb = js.rb();
    const double inf = std::numeric_limits<double>::infinity();
    const int cadence = (int)(iter_image / 2 + 0.0001);
    const size_t ib = npix * sizeof(double);
    ChainArgs ca{};
    ca.nx = nx;
    ca.ny = ny;
    ca.alpha = alpha;
    ca.eps = eps;
    // AT_f = sum (Af - g)^2 over [view][bin], then + the interior energy.
    auto energy = [&] {
        ca.nchains = 1;
        ca.c[0] = chain(&sc->gen[0], seg(CT_SQDIFF, nrows, js.af.p, js.g.p), seg(CT_REG_INT, nint, js.u.p, js.v.p));
        launch_chains(st, ca);
    };
    int reason = 0, ran = 0;

    seq_forward(st, q, js.u.p, js.af.p, js.g.p, js.r.p);
    seq_backproject(st, q, nx, ny, js.r.p, js.u.p, js.v.p, alpha, eps, js.d.p);
    energy();
    MI_HIP(hipMemcpyAsync(js.p.p, js.d.p, ib, hipMemcpyDeviceToDevice, st));
    Scalars h = js.read_scalars(st);
    double at_old = h.gen[0];
    for (int k = 1; k <= iter_image; ++k) {
        ran = k;
        seq_forward(st, q, js.p.p, js.q.p);
        ca.nchains = 3;
        ca.c[0] = chain(&sc->gen[3], seg(CT_NV, npix, js.p.p, js.v.p));
        ca.c[1] = chain(&sc->gen[4], seg(CT_DOT, npix, js.p.p, js.d.p));
        ca.c[2] = chain(&sc->gen[2], seg(CT_SQ, nrows, js.q.p));
        launch_chains(st, ca);
        h = js.read_scalars(st);
        const double nap = h.gen[2];
        const double c1 = h.gen[4] / (nap + alpha * h.gen[3]);
        const unsigned grid = launch_axpy<double>(st, npix, js.u.p, js.p.p, nullptr, c1, -inf, inf, js.t.p,
                                                  rb.slot(5));
        FinArgs f{};
        f.op = FIN_STORE;
        f.nin = 1;
        f.in[0] = rb.slot(5);
        f.n[0] = (int)grid;
        f.is_max[0] = 1;
        f.dst = &sc->gen[5];
        launch_finalize(st, f);
        // The reference takes k % 0 when iter_image < 2 (undefined); no
        // median is applied then.
        if (cadence > 0 && k % cadence == 0) {
            launch_median<double>(st, nx, ny, js.u.p, js.w.p);
            MI_HIP(hipMemcpyAsync(js.u.p, js.w.p, ib, hipMemcpyDeviceToDevice, st));
        }
        h = js.read_scalars(st);
        if (k == iter_image) { reason = 0; break; }
        if (h.gen[5] < 1e-5) { reason = 1; break; }
        seq_forward(st, q, js.u.p, js.af.p);
        energy();
        h = js.read_scalars(st);
        const double at_new = h.gen[0];
        if (1.1 * at_old < at_new && 1 < k) {
            MI_HIP(hipMemcpyAsync(js.u.p, js.t.p, ib, hipMemcpyDeviceToDevice, st));
            reason = 2;
            break;
        }
        at_old = at_new;
        // r -= c1 q  (r + (-c1) q rounds the same way)
        launch_axpy<double>(st, nrows, js.r.p, js.q.p, nullptr, -c1, -inf, inf, nullptr, nullptr);
        seq_backproject(st, q, nx, ny, js.r.p, js.u.p, js.v.p, alpha, eps, js.d.p);
        seq_forward(st, q, js.d.p, js.af.p);
        ca.nchains = 1;
        ca.c[0] = chain(&sc->gen[6], seg(CT_DOT, nrows, js.af.p, js.q.p));
        launch_chains(st, ca);
        h = js.read_scalars(st);
        const double c2 = -h.gen[6] / nap;
        launch_xpby<double>(st, npix, js.p.p, js.d.p, c2);
    }
    MI_HIP(hipStreamSynchronize(st));
    if (why) *why = reason;
    return ran;
}

// origin/GoldenC/src/main.c:259-360 on js.v given js.u; n in js.w, the
// direction in js.d.  No host round trip: the step size and the early stop
// live in the job's Scalars.
void gc_edge_step_ordered(mi_context *ctx, JobState<double> &js, double alpha, double beta, double eps,
                          int iter_edge) {
    const hipStream_t st = ctx->stream;
    const int nx = js.nx, ny = js.ny;
    const int64_t npix = js.npix;
    ChainArgs ca{};
    ca.nx = nx;
    ca.ny = ny;
    ca.alpha = alpha;
    ca.beta = beta;
    ca.eps = eps;
    ca.edge_c = 1;
    ca.sc = js.sc.p;
    ca.nchains = 3;
    ca.c[0] = chain(nullptr, seg(CT_NABLA2, npix, js.d.p));
    ca.c[1] = chain(nullptr, seg(CT_SQ, npix, js.d.p));
    ca.c[2] = chain(nullptr, seg(CT_NDSQ, npix, js.w.p, js.d.p));
    launch_pass_begin(st, js.sc.p);
    for (int k = 0; k < iter_edge; ++k) {
        launch_edge_grad_n<double>(st, nx, ny, js.u.p, js.v.p, js.d.p, js.w.p, alpha, beta, eps, js.sc.p);
        launch_chains(st, ca);
        launch_edge_update<double>(st, js.rb(), nx, ny, js.v.p, js.d.p, js.sc.p);
    }
    MI_HIP(hipStreamSynchronize(st));
}

}  // namespace mi
