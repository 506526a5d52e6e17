// Kernels of the batched ordered-subset AT engine (gfx950).
//
// One launch per phase serves every unit of a batch.  The global block index
// is first renumbered so that each of the 8 XCDs owns a contiguous range
// (xcd_block), then mapped to (unit, local block) through the BlockMap passed
// by value.  K, the number of jobs of a unit, selects a code path compiled
// for that K; a launch uses the kernel compiled for K <= 1, <= 4 or <= 8
// that covers its units, so it keeps the register budget of that case.
//
// Forward projection of fan beams is the checkpointed walk (walk_otf): the
// tracer re-run from stored per-segment states, a block holding 32
// neighbouring rays x 8 segments.  Back projection (and the forward
// projection of other geometries) reads the sliced matrices (Sell,
// runtime.hpp): a lane owns one pixel column (back) or one segment of one
// ray (forward) and walks it in entry order with a Kahan sum, four entries
// per 16-byte load, a wave's loads contiguous; neighbouring lanes own
// neighbouring rays / pixels, so one gather instruction touches few cache
// lines.  The back projection stages each 16 x 16 pixel block's residual
// values in LDS first where they fit (bp_unit).  Images are stored in 4 x 8
// tiles (tile_index).  Partial sums of an item are combined in a fixed
// order.
#include "engine.hpp"
#include "tracers.hpp"

#include <cmath>
#include <type_traits>

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
    const GL int64_t *fs_slice0, *fs_base;
    const GL int32_t *fs_T;
    const GL uint32_t *fs_idx, *fs_first;
    const GL int16_t *fs_d16;
    const GL T *fs_len;
    const GL int64_t *bs_slice0, *bs_base;
    const GL int32_t *bs_T;
    const GL uint32_t *bs_idx, *bs_first;
    const GL int16_t *bs_d16;
    const GL T *bs_len;
    const GL uint16_t *bs_slot16;
    const GL int32_t *bs_vlo, *bs_voff;
    const GL int64_t *bs_vtab0;
    const GL double *ot_lam, *ot_ax, *ot_ay, *ot_stx, *ot_sty;
    const GL int32_t *ot_cx, *ot_cy, *ot_sg, *ot_cnt;
    const GL int64_t *sub_row0;
    GL T *g, *r, *u, *v, *d, *w;
    int64_t npix_t, nrows, part_off, cap;
    int nx, ny, tny, K, fs_P, bs_P, job0, nbins;
    __device__ __forceinline__ explicit UG(const UnitDesc<T> &D)
        : fs_slice0((const GL int64_t *)D.fs_slice0), fs_base((const GL int64_t *)D.fs_base),
          fs_T((const GL int32_t *)D.fs_T), fs_idx((const GL uint32_t *)D.fs_idx),
          fs_first((const GL uint32_t *)D.fs_first), fs_d16((const GL int16_t *)D.fs_d16),
          fs_len((const GL T *)D.fs_len), bs_slice0((const GL int64_t *)D.bs_slice0),
          bs_base((const GL int64_t *)D.bs_base), bs_T((const GL int32_t *)D.bs_T),
          bs_idx((const GL uint32_t *)D.bs_idx), bs_first((const GL uint32_t *)D.bs_first),
          bs_d16((const GL int16_t *)D.bs_d16), bs_len((const GL T *)D.bs_len),
          bs_slot16((const GL uint16_t *)D.bs_slot16), bs_vlo((const GL int32_t *)D.bs_vlo),
          bs_voff((const GL int32_t *)D.bs_voff), bs_vtab0((const GL int64_t *)D.bs_vtab0),
          ot_lam((const GL double *)D.ot_lam), ot_ax((const GL double *)D.ot_ax), ot_ay((const GL double *)D.ot_ay),
          ot_stx((const GL double *)D.ot_stx), ot_sty((const GL double *)D.ot_sty),
          ot_cx((const GL int32_t *)D.ot_cx), ot_cy((const GL int32_t *)D.ot_cy),
          ot_sg((const GL int32_t *)D.ot_sg), ot_cnt((const GL int32_t *)D.ot_cnt),
          sub_row0((const GL int64_t *)D.sub_row0), g((GL T *)D.g), r((GL T *)D.r), u((GL T *)D.u),
          v((GL T *)D.v), d((GL T *)D.d), w((GL T *)D.w), npix_t(D.npix_t), nrows(D.nrows),
          part_off(D.part_off), cap(D.cap), nx(D.nx), ny(D.ny), tny(D.tny), K(D.K), fs_P(D.fs_P), bs_P(D.bs_P),
          job0(D.job0), nbins(D.nbins) {}
    // Tiled index of pixel (i, j), or -1 outside the image.
    __device__ __forceinline__ int64_t at(int i, int j) const {
        return (i < 0 || j < 0 || i >= nx || j >= ny) ? -1 : tile_index(i, j, tny);
    }
};

template <typename T>
__device__ __forceinline__ const GL JobParams &jpar(const EngArgs<T> &a, int js) {
    return ((const GL JobParams *)a.jp)[js];
}
template <typename T>
__device__ __forceinline__ GL Scalars &jsc(const EngArgs<T> &a, int js) {
    return ((GL Scalars *)a.sc)[js];
}

#ifndef MI_NT_FP
#define MI_NT_FP 1      // non-temporal matrix streams in the forward projector
#endif
#ifndef MI_NT_BP
#define MI_NT_BP 1      // ... and in the back projector
#endif
// Staged back projection: the walk's first stream loads issued with the
// residual copies, before their wait (0: never, 1: K = 1 only, 2: every K).
// Measured on the 64-job batch: with every K the K <= 4 staged launch took
// 678 us against 617 us unprimed; K = 1 only: 622 us.
#ifndef MI_BP_PRIME
#define MI_BP_PRIME 1
#endif
#ifndef MI_STEPS
#define MI_STEPS 8      // steps of a projector lane in flight (K <= 2; 4 above)
#endif

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Matrix streams are read once per launch; non-temporal loads keep them from
// evicting the image / residual lines that the gathers reuse from L2.
template <bool NT, typename V>
__device__ __forceinline__ V stream_ld(const GL V *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
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

// Gathers through a raw buffer resource: a 32-bit byte offset per lane against
// a base held in scalar registers (the image is < 4 GiB).
#ifndef MI_BUF_GATHER
#define MI_BUF_GATHER 1
#endif
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

template <typename T, int K>
__device__ __forceinline__ void loadk_buf(__amdgpu_buffer_rsrc_t rs, uint32_t off, T (&o)[K]) {
    constexpr int B = K * (int)sizeof(T);
    if constexpr (B >= 16) {
#pragma unroll
        for (int c = 0; c < B / 16; ++c) {
            const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(off + 16 * c), 0, 0);
            __builtin_memcpy(&o[c * (16 / (int)sizeof(T))], &a, 16);
        }
    } else if constexpr (B == 8) {
        typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
        const u32x2 a = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off, 0, 0);
        __builtin_memcpy(&o[0], &a, 8);
    } else {
        const uint32_t a = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0);
        __builtin_memcpy(&o[0], &a, 4);
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

// Sums v[j] over aligned groups of P lanes (fixed butterfly).
template <int K>
__device__ __forceinline__ void group_sums(double (&v)[K], int P) {
    for (int o = P >> 1; o > 0; o >>= 1)
#pragma unroll
        for (int j = 0; j < K; ++j) v[j] += __shfl_xor(v[j], o, kWave);
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
// One projector lane: its entries t = 0 .. n-1 (sell_pos), Kahan-summed in
// entry order (zero-length padding skipped), gathering K interleaved values of
// `x` per entry.  Index stream: absolute (ip) or, when D16, 16-bit
// differences (dp) accumulated from the lane's first index.  ip / dp / lp
// point at the lane's first group; a group of 4 steps is one 16-byte length
// load and one 16-byte (absolute) or 8-byte (differences) index load.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

#define LDSP __attribute__((address_space(3)))
enum IdxMode { IM_ABS32 = 0, IM_DELTA16 = 1, IM_SLOT16 = 2 };

// Stream registers of one pipeline stage of walk: GR groups of kGroup
// entries (index word and length) per lane.
template <typename T, int K>
struct Stage {
    static constexpr int GR = (K <= 2) ? MI_STEPS / kGroup : 1;   // groups per pipeline stage
    uint32_t sv[GR][kGroup];
    T lv[GR][kGroup];
};

// Issues the stream loads of groups g0 .. g0 + GR - 1 of a lane into b.
template <typename T, int K, bool NT, int IM>
__device__ __forceinline__ void stage_fetch(Stage<T, K> &b, const GL uint32_t *__restrict__ ip,
                                            const GL int16_t *__restrict__ dp, const GL T *__restrict__ lp, int g0) {
    constexpr bool D16 = IM != IM_ABS32;
    constexpr int64_t kStride = kGroup * kWave;   // entries between a lane's groups
#pragma unroll
    for (int r = 0; r < Stage<T, K>::GR; ++r) {
        const int64_t o = (int64_t)(g0 + r) * kStride;
        if constexpr (IM == IM_SLOT16) {
            const u32x2 a = stream_ld<NT>((const GL u32x2 *)(dp + o));
            b.sv[r][0] = a.x & 0xffffu;
            b.sv[r][1] = a.x >> 16;
            b.sv[r][2] = a.y & 0xffffu;
            b.sv[r][3] = a.y >> 16;
        } else if constexpr (D16) {
            const u32x2 a = stream_ld<NT>((const GL u32x2 *)(dp + o));
            b.sv[r][0] = (uint32_t)(int32_t)(int16_t)(a.x & 0xffffu);
            b.sv[r][1] = (uint32_t)(int32_t)(int16_t)(a.x >> 16);
            b.sv[r][2] = (uint32_t)(int32_t)(int16_t)(a.y & 0xffffu);
            b.sv[r][3] = (uint32_t)(int32_t)(int16_t)(a.y >> 16);
        } else {
            const u32x4 a = stream_ld<NT>((const GL u32x4 *)(ip + o));
            b.sv[r][0] = a.x; b.sv[r][1] = a.y; b.sv[r][2] = a.z; b.sv[r][3] = a.w;
        }
        if constexpr (sizeof(T) == 4) {
            const u32x4 a = stream_ld<NT>((const GL u32x4 *)(lp + o));
            __builtin_memcpy(&b.lv[r][0], &a, 16);
        } else {
            const u32x4 a = stream_ld<NT>((const GL u32x4 *)(lp + o));
            const u32x4 c = stream_ld<NT>((const GL u32x4 *)(lp + o) + 1);
            __builtin_memcpy(&b.lv[r][0], &a, 16);
            __builtin_memcpy(&b.lv[r][2], &c, 16);
        }
    }
}

// The first stage of a walk over n entries (none when the walk has no full
// stage); a walk with PRIMED set starts from it.
template <typename T, int K, bool NT, int IM>
__device__ __forceinline__ void walk_prime(Stage<T, K> &b, const GL uint32_t *ip, const GL int16_t *dp,
                                           const GL T *lp, int n) {
    if (n / kGroup / Stage<T, K>::GR > 0) stage_fetch<T, K, NT, IM>(b, ip, dp, lp, 0);
}

// IM: absolute 32-bit indices (ip), 16-bit differences (dp, from `cur`), or
// 16-bit LDS slots (dp read unsigned).  LDSG: gather from `lx` (LDS) instead
// of the global array x.  b holds the stage in flight; with PRIMED the caller
// has issued the first one (walk_prime).
template <typename T, int K, bool NT, int IM, bool LDSG = false, bool PRIMED = false>
__device__ __forceinline__ void walk(const GL uint32_t *__restrict__ ip, const GL int16_t *__restrict__ dp,
                                     uint32_t cur, const GL T *__restrict__ lp, int n, const GL T *__restrict__ x,
                                     uint32_t xbytes, Acc<T> (&acc)[K], Stage<T, K> &b,
                                     const LDSP T *lx = nullptr) {
    constexpr bool D16 = IM != IM_ABS32;
    constexpr int GR = Stage<T, K>::GR;
    constexpr int64_t kStride = kGroup * kWave;             // entries between a lane's groups
#if MI_BUF_GATHER
    const __amdgpu_buffer_rsrc_t xr = make_rsrc((const void *)x, xbytes);
#else
    (void)xbytes;
#endif
    auto fetch = [&](int g0) { stage_fetch<T, K, NT, IM>(b, ip, dp, lp, g0); };
    auto gather = [&](uint32_t id, T (&o)[K]) {
        if constexpr (LDSG) {
#pragma unroll
            for (int j = 0; j < K; ++j) o[j] = lx[id * K + j];
        } else {
#if MI_BUF_GATHER
            loadk_buf<T, K>(xr, id * (uint32_t)(K * sizeof(T)), o);
#else
            loadk<T, K>(x + (uint32_t)(id * K), o);
#endif
        }
    };
    const int ng = n / kGroup, full = ng / GR;
    // Software pipeline: the next stage's stream loads are issued after the
    // current stage's gathers (the barriers keep that order) and before they
    // are consumed, so a lane waits for one memory round trip per stage.
    if (!PRIMED && full > 0) fetch(0);
    int g = 0;
    for (int c = 0; c < full; ++c, g += GR) {
        uint32_t id[GR][kGroup];
        T ln[GR][kGroup], xv[GR][kGroup][K];
#pragma unroll
        for (int r = 0; r < GR; ++r)
#pragma unroll
            for (int q = 0; q < kGroup; ++q) {
                if constexpr (IM == IM_DELTA16) {
                    cur += b.sv[r][q];
                    id[r][q] = cur;
                } else {
                    id[r][q] = b.sv[r][q];
                }
                ln[r][q] = b.lv[r][q];
            }
#pragma unroll
        for (int r = 0; r < GR; ++r)
#pragma unroll
            for (int q = 0; q < kGroup; ++q) gather(id[r][q], xv[r][q]);
        __builtin_amdgcn_sched_barrier(0);
        fetch(c + 1 < full ? g + GR : g);   // the last stage re-reads its own groups
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < GR; ++r)
#pragma unroll
            for (int q = 0; q < kGroup; ++q)
#pragma unroll
                for (int j = 0; j < K; ++j) acc[j].add_prod_if(xv[r][q][j], ln[r][q], ln[r][q] != T(0));
    }
    for (; g < ng; ++g) {
        const int64_t o = (int64_t)g * kStride;
        uint32_t id[kGroup];
        T ln[kGroup], xv[kGroup][K];
        if constexpr (IM == IM_SLOT16) {
            const u32x2 a = stream_ld<NT>((const GL u32x2 *)(dp + o));
            const uint32_t w[2] = {a.x, a.y};
#pragma unroll
            for (int q = 0; q < kGroup; ++q) id[q] = (w[q >> 1] >> (16 * (q & 1))) & 0xffffu;
        } else if constexpr (D16) {
            const u32x2 a = stream_ld<NT>((const GL u32x2 *)(dp + o));
            const uint32_t w[2] = {a.x, a.y};
#pragma unroll
            for (int q = 0; q < kGroup; ++q) {
                cur += (uint32_t)(int32_t)(int16_t)((w[q >> 1] >> (16 * (q & 1))) & 0xffffu);
                id[q] = cur;
            }
        } else {
            const u32x4 a = stream_ld<NT>((const GL u32x4 *)(ip + o));
            id[0] = a.x; id[1] = a.y; id[2] = a.z; id[3] = a.w;
        }
#pragma unroll
        for (int q = 0; q < kGroup; ++q) ln[q] = lp[o + q];
#pragma unroll
        for (int q = 0; q < kGroup; ++q) gather(id[q], xv[q]);
#pragma unroll
        for (int q = 0; q < kGroup; ++q)
#pragma unroll
            for (int j = 0; j < K; ++j) acc[j].add_prod_if(xv[q][j], ln[q], ln[q] != T(0));
    }
}

// Walks lane `lane` of slice sl of one sliced matrix in whichever index form
// it is stored.
template <typename T, int K, bool NT>
__device__ __forceinline__ void walk_slice(const GL int64_t *base, const GL int32_t *Tn, const GL uint32_t *idx,
                                           const GL uint32_t *first, const GL int16_t *d16, const GL T *len,
                                           int64_t sl, int lane, const GL T *x, int64_t xn, Acc<T> (&acc)[K]) {
    const int64_t p0 = sell_pos(base[sl], 0, lane);
    const int n = Tn[sl];
    const uint32_t xb = (uint32_t)(xn * K * (int64_t)sizeof(T));
    Stage<T, K> b;
    if (d16) walk<T, K, NT, IM_DELTA16>(nullptr, d16 + p0, first[sl * kWave + lane], len + p0, n, x, xb, acc, b);
    else walk<T, K, NT, IM_ABS32>(idx + p0, nullptr, 0u, len + p0, n, x, xb, acc, b);
}

// One lane of the checkpointed forward walk: n Siddon steps from the stored
// tracer state, each step's pixel gathered from the tiled image x and its
// length (rounded to T as the stored matrix rounds it) Kahan-summed in step
// order.  The step is siddon_step (tracers.hpp) written with selects instead
// of a branch: the same comparisons and the same additions on the taken side,
// so lam / ax / ay and the lengths are bit-identical.  The pixel position is
// carried as byte offsets into the tiled layout, X = 4 E cx and Y = E cy
// (E bytes per pixel), so the tiled offset is
//   E tile_index(cx, cy) = (cx >> 2) (32 tny - 16) E + X + (cy >> 2) 12 E + Y.
// Every step of a lane is summed, zero-length crossings included (as the
// oracle sums every matrix entry).  Steps go in groups of OS, software
// pipelined: the next group is traced and its gathers issued before this
// group's gathers are consumed.
template <typename T, int K>
__device__ __forceinline__ void walk_otf(const UG<T> &U, int64_t q, const GL T *__restrict__ x, Acc<T> (&acc)[K]) {
#ifndef MI_OTF_OS8
#define MI_OTF_OS8 2   // steps in flight at K = 8 (4: 116 VGPRs, slower at every image size)
#endif
    constexpr int OS = (K <= 2) ? 8 : (K == 4 ? 4 : MI_OTF_OS8);
    constexpr uint32_t E = K * sizeof(T), LE = __builtin_ctz(E);
    double lam = U.ot_lam[q], ax = U.ot_ax[q], ay = U.ot_ay[q];
    const double stx = U.ot_stx[q], sty = U.ot_sty[q];
    const int sgb = U.ot_sg[q];
    uint32_t X = (uint32_t)U.ot_cx[q] * (4 * E), Y = (uint32_t)U.ot_cy[q] * E;
    const uint32_t dX = (sgb & 1) ? 4 * E : 0u - 4 * E, dY = (sgb & 2) ? E : 0u - E;
    const uint32_t rowE = (uint32_t)(U.tny * (kTileH * kTileW) - 16) * E;
    const int n = U.ot_cnt[q];
#if MI_BUF_GATHER
    const __amdgpu_buffer_rsrc_t xr = make_rsrc((const void *)x, (uint32_t)(U.npix_t * (int64_t)E));
#endif
    auto step = [&](uint32_t &off, T &ln) {
        const bool bx = ax <= ay;
        const double nxt = bx ? ax : ay;
        ln = (T)(nxt - lam);
        lam = nxt;
        off = (uint32_t)__umul24(X >> (4 + LE), rowE) + X + (uint32_t)__umul24(Y >> (2 + LE), 12 * E) + Y;
        double nax = ax + stx, nay = ay + sty;
        // both sums exist on every lane: keeps the selects below from being
        // turned back into a divergent branch
        asm volatile("" : "+v"(nax), "+v"(nay));
        ax = bx ? nax : ax;
        ay = bx ? ay : nay;
        X += bx ? dX : 0u;
        Y += bx ? 0u : dY;
    };
    auto gather = [&](uint32_t off, T (&o)[K]) {
#if MI_BUF_GATHER
        loadk_buf<T, K>(xr, off, o);
#else
        loadk<T, K>((const GL T *)((const GL char *)x + off), o);
#endif
    };
    auto stage = [&](T (&ln)[OS], T (&xv)[OS][K]) {
        uint32_t off[OS];
#pragma unroll
        for (int k = 0; k < OS; ++k) step(off[k], ln[k]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < OS; ++k) gather(off[k], xv[k]);
        __builtin_amdgcn_sched_barrier(0);
    };
    auto consume = [&](const T (&ln)[OS], const T (&xv)[OS][K]) {
#pragma unroll
        for (int k = 0; k < OS; ++k)
#pragma unroll
            for (int j = 0; j < K; ++j) acc[j].add_prod(xv[k][j], ln[k]);
    };
    const int ng = n / OS;
    T la[OS], lb[OS], xa[OS][K], xb[OS][K];
    int g = 0;
    if (ng > 0) {
        stage(la, xa);
        for (g = 1; g + 1 < ng; g += 2) {
            stage(lb, xb);
            consume(la, xa);
            stage(la, xa);
            consume(lb, xb);
        }
        if (g < ng) {
            stage(lb, xb);
            consume(la, xa);
            consume(lb, xb);
        } else {
            consume(la, xa);
        }
    }
    // the last n % OS steps: traced and gathered together, one round trip
    const int rem = n - ng * OS;
    if (rem > 0) {
        uint32_t off[OS - 1];
        T ln[OS - 1], xv[OS - 1][K];
#pragma unroll
        for (int k = 0; k < OS - 1; ++k)
            if (k < rem) step(off[k], ln[k]);
#pragma unroll
        for (int k = 0; k < OS - 1; ++k)
            if (k < rem) gather(off[k], xv[k]);
#pragma unroll
        for (int k = 0; k < OS - 1; ++k)
            if (k < rem)
#pragma unroll
                for (int j = 0; j < K; ++j) acc[j].add_prod(xv[k][j], ln[k]);
    }
}

// Forward projection of subset s: a block is 4 slices (one per wave).
// RESID: r = g - A u (stored), partial sum r^2; else partial sum (A d)^2.
static_assert(kOtfWaves == 4, "the forward walk's cross-wave combination is written for 4 waves");

// FM: 1 when every unit of the launch has the checkpointed walk (the sliced
// branch is not compiled), 3 when either form may occur.
template <typename T, int K, bool RESID, int FM>
__device__ __forceinline__ void fp_unit(const EngArgs<T> &a, const UG<T> &U, int s, int lb, double *lds,
                                        double *ldr /* kBlock / kOtfPerWave * K */) {
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int P = U.fs_P, C = kWave / P;
    const int64_t sl0 = U.fs_slice0[s], li = (int64_t)lb * (kBlock / kWave) + wv;
    double m[K];
#pragma unroll
    for (int j = 0; j < K; ++j) m[j] = 0.0;
    if (li < U.fs_slice0[s + 1] - sl0) {
        const int64_t sl = sl0 + li;
        Acc<T> acc[K];
        double tot[K];
        int64_t row;
        bool owner;
        if (FM == 1 || U.ot_cnt) {
            // the block is kOtfRays rays x kOtfSeg segments (runtime.hpp);
            // a ray's segment sums combine in a fixed tree: inside the wave
            // (lanes a, a + kOtfRays, ...), then over the block's waves
            walk_otf<T, K>(U, sl * kWave + lane, RESID ? U.u : U.d, acc);
#pragma unroll
            for (int j = 0; j < K; ++j) tot[j] = acc[j].value();
            for (int o = kOtfRays; o < kWave; o <<= 1)
#pragma unroll
                for (int j = 0; j < K; ++j) tot[j] += __shfl_xor(tot[j], o, kWave);
            if (lane < kOtfRays)
#pragma unroll
                for (int j = 0; j < K; ++j) ldr[(wv * kOtfRays + lane) * K + j] = tot[j];
            __syncthreads();
            if (wv == 0 && lane < kOtfRays)
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const double *q = ldr + lane * K + j;
                    tot[j] = (q[0] + q[kOtfRays * K]) + (q[2 * kOtfRays * K] + q[3 * kOtfRays * K]);
                }
            row = U.sub_row0[s] + (int64_t)lb * kOtfRays + lane;
            owner = wv == 0 && lane < kOtfRays;
        } else if constexpr (FM != 1) {
            walk_slice<T, K, MI_NT_FP>(U.fs_base, U.fs_T, U.fs_idx, U.fs_first, U.fs_d16, U.fs_len, sl, lane,
                                       RESID ? U.u : U.d, U.npix_t, acc);
#pragma unroll
            for (int j = 0; j < K; ++j) tot[j] = acc[j].value();
            group_sums<K>(tot, P);
            row = U.sub_row0[s] + li * C + lane / P;
            owner = lane % P == 0;
        }
        if (owner && row < U.sub_row0[s + 1]) {
            if (RESID) {
                T gg[K], o[K];
                loadk<T, K>(U.g + row * K, gg);
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    o[j] = (T)((double)gg[j] - tot[j]);
                    m[j] = (double)o[j] * (double)o[j];
                }
                storek<T, K>(U.r + row * K, o);
            } else {
#pragma unroll
                for (int j = 0; j < K; ++j) m[j] = tot[j] * tot[j];
            }
        }
    }
    block_store<T, K, false>(a, U, RESID ? SL_DF : SL_AD, lb, m, lds);
}

// K interleaved values at pixel (i, j), or `pad` outside the image.
template <typename T, int K>
__device__ __forceinline__ void ldn(const UG<T> &U, const GL T *arr, int i, int j, T (&o)[K], T pad) {
    const int64_t q = U.at(i, j);
    if (q < 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) o[k] = pad;
    } else {
        loadk<T, K>(arr + q * K, o);
    }
}

// The back projection's 16 x 16 pixel square (rows i0 .. i0 + 15, columns
// j0 .. j0 + 15) and its one-pixel halo, 18 x 18 pixels of K values, for the
// regulariser of its epilogue: staged through LDS once instead of eight
// neighbour loads per pixel.
constexpr int kSqW = kBpBlockTC * kTileW + 2, kSqN = kSqW * kSqW;
static_assert(kBpBlockTR * kTileH == kBpBlockTC * kTileW, "the back-projection block is square");
__device__ __forceinline__ int sq_at(int r, int c) { return (r + 1) * kSqW + c + 1; }

// L[sq_at(r, c) * K + k] = arr at (i0 + r, j0 + c), `pad` outside the image.
// Thread (w, lane) stages pixel (tile row w, tile column lane / 32, tile
// position lane % 32) of the square; threads 0 .. 67 the halo.  Every thread
// calls it; ends with a barrier.
template <typename T, int K>
__device__ __forceinline__ void stage_square(const UG<T> &U, const GL T *arr, int i0, int j0, T pad, LDSP T *L) {
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave, w = lane & 31;
    const int r = wv * kTileH + ((w >> 2) & 3), c = (lane >> 5) * kTileW + ((w >> 4) << 2) + (w & 3);
    T x[K];
    ldn<T, K>(U, arr, i0 + r, j0 + c, x, pad);
#pragma unroll
    for (int k = 0; k < K; ++k) L[sq_at(r, c) * K + k] = x[k];
    constexpr int E = kSqW - 2, kHalo = 2 * kSqW + 2 * E;   // 68
    if (threadIdx.x < kHalo) {
        const int e = threadIdx.x;
        const int hr = e < kSqW ? -1 : (e < 2 * kSqW ? E : (e - 2 * kSqW) % E);
        const int hc = e < 2 * kSqW ? e % kSqW - 1 : (e < 2 * kSqW + E ? -1 : E);
        ldn<T, K>(U, arr, i0 + hr, j0 + hc, x, pad);
#pragma unroll
        for (int k = 0; k < K; ++k) L[sq_at(hr, hc) * K + k] = x[k];
    }
    __syncthreads();
}

// The same square as stage_square, in the same layout, moved by direct
// global-to-LDS loads (positions outside the image get `pad` by LDS stores).
// Nothing waits here: the caller waits for every load (s_waitcnt 0) and
// synchronises the block before reading L.
template <typename T, int K>
__device__ __forceinline__ void dma_square(const UG<T> &U, const GL T *arr, int i0, int j0, T pad, LDSP T *L) {
    constexpr int DW = (int)(sizeof(T) / 4), PD = K * DW, N = kSqN * PD;   // dwords per pixel, per square
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    uint32_t pw[DW];
    __builtin_memcpy(pw, &pad, sizeof(T));
    const GL uint32_t *src = (const GL uint32_t *)arr;
    LDSP uint32_t *dst = (LDSP uint32_t *)L;
    for (int b = wv * kWave; b < N; b += kBlock) {
        const int t = b + lane;
        if (t < N) {
            const int p = t / PD, e = t - p * PD;
            const int i = i0 + p / kSqW - 1, j = j0 + p % kSqW - 1;
            if (i >= 0 && j >= 0 && i < U.nx && j < U.ny)
                __builtin_amdgcn_global_load_lds((const GL void *)(src + tile_index(i, j, U.tny) * PD + e),
                                                 (LDSP void *)(dst + b), 4, 0, 0);
            else
                dst[t] = (DW == 1 || (e & 1) == 0) ? pw[0] : pw[DW - 1];
        }
    }
}

// GoldenC image-gradient regulariser of K interleaved jobs at square pixel
// (r, c): u zero padded, v one padded (origin/GoldenC/src/main.c:112-125).
template <typename T, int K>
__device__ __forceinline__ void gc_reg(const LDSP T *Lu, const LDSP T *Lv, int r, int c, const double (&eps)[K],
                                       double (&out)[K]) {
    const LDSP T *u = Lu + sq_at(r, c) * K, *v = Lv + sq_at(r, c) * K;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double Z = u[k], Zd = u[kSqW * K + k], Zr = u[K + k], Zu = u[-kSqW * K + k], Zl = u[-K + k];
        const double Pp = v[k], Pu = v[-kSqW * K + k], Pl = v[-K + k];
        out[k] = +(Pp * Pp) * (Zd - Z) + (Pp * Pp) * (Zr - Z) - (Pu * Pu) * (Z - Zu) - (Pl * Pl) * (Z - Zl) +
                 0.01 * (eps[k] * eps[k]) * (Zd + Zr - 4 * Z + Zu + Zl);
    }
}

// Gather back projection of subset s over 4 slices of tiled pixels, fused
// with d = S * A_s^T r + alpha * R(u, v); partial sum d^2.
// Back projection of subset s over one 16 x 16 block of pixels (bp_item;
// wave w holds tile row w of the block).  LDS: the block first copies the
// residual values of every (view, bin) its entries use into `ldr` (slots of
// the block's table, `vlo` / `voff`), then gathers from there.
template <typename T, int K, bool LDS, bool SQ>
__device__ __forceinline__ void bp_unit(const EngArgs<T> &a, const UG<T> &U, int s, int lb, double *lds,
                                        LDSP T *ldr, int sq_off) {
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int64_t sl0 = U.bs_slice0[s], li = (int64_t)lb * (kBlock / kWave) + wv, sl = sl0 + li;
    const bool has = li < U.bs_slice0[s + 1] - sl0;
    const int tnx = (U.nx + kTileH - 1) / kTileH;
    const int nbc = (U.tny + kBpBlockTC - 1) / kBpBlockTC;
    const int i0 = (lb / nbc) * (kBpBlockTR * kTileH), j0 = (lb % nbc) * (kBpBlockTC * kTileW);
    double m[K];
#pragma unroll
    for (int j = 0; j < K; ++j) m[j] = 0.0;
    // u and v of the square for the regulariser: with SQ in their own LDS
    // region at sq_off, loaded now, behind the walk; otherwise after the
    // walk, in the LDS the residual slots used.
    LDSP T *Lu = ldr + (SQ ? sq_off : 0), *Lv = Lu + kSqN * K;
    if constexpr (SQ) {
        dma_square<T, K>(U, U.u, i0, j0, T(0), Lu);
        dma_square<T, K>(U, U.v, i0, j0, T(1), Lv);
    }
    Stage<T, K> st;
    int64_t p0 = 0;
    int n = 0;
    if (LDS) {
        if (has) {
            p0 = sell_pos(U.bs_base[sl], 0, lane);
            n = U.bs_T[sl];
        }
        const int nv = (int)((U.sub_row0[s + 1] - U.sub_row0[s]) / U.nbins);
        const int64_t nblk = bp_blocks_of(tnx, U.tny);
        const GL int32_t *vlo = U.bs_vlo + U.bs_vtab0[s] + (int64_t)lb * nv;
        // the same offsets as voff_at (runtime.hpp)
        const GL int32_t *voff = U.bs_voff + (U.bs_vtab0[s] + (int64_t)s * nblk + (int64_t)lb * (nv + 1));
        // Wave w copies views w, w + 4, ...: a view's slots are c consecutive
        // residual rows (K values each), so the copy is one contiguous run of
        // dwords, moved by direct global-to-LDS loads (no register round
        // trip; every load of the block in flight until the one wait).
        // The views' slot offsets and first rows are read 64 views at a time,
        // one view per lane, and taken from the lanes in turn, so that no
        // load waits between the direct loads.
        constexpr int NW = kBlock / kWave, DW = (int)(sizeof(T) / 4) * K;
        for (int vb = wv; vb < nv; vb += NW * kWave) {
            const int vl = vb + NW * lane;
            int o_l = 0, c_l = 0, row_l = 0;   // row within the subset's residual rows
            if (vl < nv) {
                o_l = voff[vl];
                c_l = voff[vl + 1] - o_l;
                row_l = vl * U.nbins + vlo[vl];
            }
            const int nb = min(kWave, (nv - vb + NW - 1) / NW);
            const GL uint32_t *r0 = (const GL uint32_t *)(U.r + U.sub_row0[s] * K);
            for (int i = 0; i < nb; ++i) {
                const int o = __builtin_amdgcn_readlane(o_l, i), nd = __builtin_amdgcn_readlane(c_l, i) * DW;
                const GL uint32_t *src = r0 + (int64_t)__builtin_amdgcn_readlane(row_l, i) * DW;
                LDSP uint32_t *dst = (LDSP uint32_t *)(ldr + (int64_t)o * K);
                for (int e0 = 0; e0 < nd; e0 += kWave)
                    if (e0 + lane < nd)
                        __builtin_amdgcn_global_load_lds((const GL void *)(src + e0 + lane), (LDSP void *)(dst + e0), 4,
                                                         0, 0);
            }
        }
        // The walk's first stream loads go out with the copies, so that the
        // one wait covers both.
        if (has && (MI_BP_PRIME == 2 || (MI_BP_PRIME == 1 && K == 1)))
            walk_prime<T, K, MI_NT_BP, IM_SLOT16>(st, nullptr, (const GL int16_t *)(U.bs_slot16 + p0),
                                                  U.bs_len + p0, n);
        __builtin_amdgcn_s_waitcnt(0);   // the block's direct loads have landed in LDS
        __syncthreads();
    }
    double tot[K];
#pragma unroll
    for (int j = 0; j < K; ++j) tot[j] = 0.0;
    if (has) {
        Acc<T> acc[K];
        if (LDS) {
            if (MI_BP_PRIME == 2 || (MI_BP_PRIME == 1 && K == 1))
                walk<T, K, MI_NT_BP, IM_SLOT16, true, true>(nullptr, (const GL int16_t *)(U.bs_slot16 + p0), 0u,
                                                            U.bs_len + p0, n, U.r, 0u, acc, st, ldr);
            else
                walk<T, K, MI_NT_BP, IM_SLOT16, true, false>(nullptr, (const GL int16_t *)(U.bs_slot16 + p0), 0u,
                                                             U.bs_len + p0, n, U.r, 0u, acc, st, ldr);
        } else {
            walk_slice<T, K, MI_NT_BP>(U.bs_base, U.bs_T, U.bs_idx, U.bs_first, U.bs_d16, U.bs_len, sl, lane, U.r,
                                       U.nrows, acc);
        }
#pragma unroll
        for (int j = 0; j < K; ++j) tot[j] = acc[j].value();
    }
    if constexpr (!SQ) {
        __syncthreads();   // every wave has finished its walk
        stage_square<T, K>(U, U.u, i0, j0, T(0), Lu);
        stage_square<T, K>(U, U.v, i0, j0, T(1), Lv);
    } else if constexpr (!LDS) {
        __builtin_amdgcn_s_waitcnt(0);   // the squares' direct loads
        __syncthreads();
    }
    if (has) {
        const int64_t q = bp_item(lb, wv, lane, tnx, U.tny);
        if (q >= 0) {
            int i, j;
            tile_coords(q, U.tny, i, j);
            T dv[K];
            if (i < U.nx && j < U.ny) {
                double eps[K], reg[K];
#pragma unroll
                for (int k = 0; k < K; ++k) eps[k] = jpar(a, U.job0 + k).eps;
                gc_reg<T, K>(Lu, Lv, i - i0, j - j0, eps, reg);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const double d = (double)a.S * tot[k] + jpar(a, U.job0 + k).alpha * reg[k];
                    dv[k] = (T)d;
                    m[k] = (double)dv[k] * (double)dv[k];
                }
            } else {
                zerok<T, K>(dv);
            }
            storek<T, K>(U.d + q * K, dv);
        }
    }
    block_store<T, K, false>(a, U, SL_DD, lb, m, lds);
}

// Stencil kernels: one thread per tiled position q (256 per block); the
// padding positions of the tiles do no work.

// Stencil strips.  Tile rows hold a multiple of kStripTiles tiles, so the 256
// tiled positions of stencil block lb are the 4 x 64 pixels of one strip
// (rows i0 .. i0 + 3, columns j0 .. j0 + 63).  A stencil stages the strip of
// each array it reads around, with a one-pixel halo (6 x 66 pixels, K values
// each), through LDS once, and reads the neighbours from there.
constexpr int kStripW = kStripTiles * kTileW, kHaloW = kStripW + 2, kHaloN = (kTileH + 2) * kHaloW;
static_assert(kPixBlock == kStripTiles * kTileH * kTileW, "a stencil block is one strip");

struct Strip {
    int64_t q;        // this thread's tiled position
    int i0, j0;       // strip origin
    int li, lj;       // this thread's pixel, relative to the origin
    bool real;        // the pixel lies in the image
};
template <typename T>
__device__ __forceinline__ Strip strip_of(const UG<T> &U, int lb) {
    const int t0 = lb * kStripTiles;
    Strip s;
    s.q = (int64_t)lb * kPixBlock + threadIdx.x;
    s.i0 = (t0 / U.tny) * kTileH;
    s.j0 = (t0 % U.tny) * kTileW;
    int i, j;
    tile_coords(s.q, U.tny, i, j);
    s.li = i - s.i0;
    s.lj = j - s.j0;
    s.real = i < U.nx && j < U.ny;
    return s;
}
__device__ __forceinline__ int halo_at(int li, int lj) { return (li + 1) * kHaloW + lj + 1; }

// L[halo_at(r, c) * K + k] = arr at pixel (i0 + r, j0 + c) for the strip and
// its halo, `pad` outside the image.  Every thread calls it; ends with a
// barrier.
template <typename T, int K>
__device__ __forceinline__ void stage_strip(const UG<T> &U, const GL T *arr, const Strip &s, T pad, T *L) {
    T x[K];
    if (s.real) {
        loadk<T, K>(arr + s.q * K, x);
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) x[k] = pad;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) L[halo_at(s.li, s.lj) * K + k] = x[k];
    // the halo: rows -1 and 4 (66 pixels each, corners included), columns
    // -1 and 64 of rows 0 .. 3
    constexpr int kHalo = 2 * kHaloW + 2 * kTileH;
    static_assert(kHalo <= kBlock, "one halo pixel per thread");
    if (threadIdx.x < kHalo) {
        const int e = threadIdx.x, f = e - 2 * kHaloW;
        const int li = e < 2 * kHaloW ? (e < kHaloW ? -1 : kTileH) : f % kTileH;
        const int lj = e < 2 * kHaloW ? e % kHaloW - 1 : (f < kTileH ? -1 : kStripW);
        ldn<T, K>(U, arr, s.i0 + li, s.j0 + lj, x, pad);
#pragma unroll
        for (int k = 0; k < K; ++k) L[halo_at(li, lj) * K + k] = x[k];
    }
    __syncthreads();
}

// The staged values at offset (di, dj) from this thread's pixel.
template <typename T, int K>
__device__ __forceinline__ void ldl(const T *L, const Strip &s, int di, int dj, T (&o)[K]) {
    const T *p = L + halo_at(s.li + di, s.lj + dj) * K;
#pragma unroll
    for (int k = 0; k < K; ++k) o[k] = p[k];
}

// nV(d) = sum (v^2 + 0.01 eps^2)((d - d_up)^2 + (d - d_left)^2), d zero
// padded (origin/GoldenC/src/main.c:152-158).
template <typename T, int K>
__device__ __forceinline__ void nv_unit(const EngArgs<T> &a, const UG<T> &U, int lb, double *lds, T *L) {
    const Strip st = strip_of(U, lb);
    stage_strip<T, K>(U, U.d, st, T(0), L);
    double m[K];
#pragma unroll
    for (int j = 0; j < K; ++j) m[j] = 0.0;
    if (st.real) {
        T dq[K], du[K], dl[K], vv[K];
        ldl<T, K>(L, st, 0, 0, dq);
        ldl<T, K>(L, st, -1, 0, du);
        ldl<T, K>(L, st, 0, -1, dl);
        loadk<T, K>(U.v + st.q * K, vv);
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

__device__ __forceinline__ bool real_pixel(int64_t q, int64_t npix_t, int tny, int nx, int ny) {
    if (q >= npix_t) return false;
    int i, j;
    tile_coords(q, tny, i, j);
    return i < nx && j < ny;
}

// u = clamp(u + c d, lower, upper) for live jobs.
template <typename T, int K>
__device__ __forceinline__ void axpy_unit(const EngArgs<T> &a, const UG<T> &U, int lb) {
    const int64_t q = (int64_t)lb * kPixBlock + threadIdx.x;
    if (!real_pixel(q, U.npix_t, U.tny, U.nx, U.ny)) return;
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
__device__ __forceinline__ void median_unit(const EngArgs<T> &a, const UG<T> &U, int lb, T *L) {
    const Strip st = strip_of(U, lb);
    stage_strip<T, K>(U, U.u, st, T(0), L);
    if (!st.real) return;
    const int64_t q = st.q;
    T nb[9][K];
    int n = 0;
#pragma unroll
    for (int di = -1; di <= 1; ++di)
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj, ++n) ldl<T, K>(L, st, di, dj, nb[n]);
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
    if (!real_pixel(q, U.npix_t, U.tny, U.nx, U.ny)) return;
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
                                               double *lds, T *L) {
    const Strip st = strip_of(U, lb);
    T *Lu = L, *Lv = L + kHaloN * K;
    stage_strip<T, K>(U, U.u, st, T(0), Lu);
    stage_strip<T, K>(U, U.v, st, T(1), Lv);
    const int64_t q = st.q;
    double m0[K], m1[K];
#pragma unroll
    for (int j = 0; j < K; ++j) m0[j] = m1[j] = 0.0;
    if (st.real) {
        T fd[K], fu[K], fr[K], fl[K], vc[K], vd[K], vu[K], vr[K], vl[K];
        ldl<T, K>(Lu, st, 1, 0, fd);
        ldl<T, K>(Lu, st, -1, 0, fu);
        ldl<T, K>(Lu, st, 0, 1, fr);
        ldl<T, K>(Lu, st, 0, -1, fl);
        ldl<T, K>(Lv, st, 0, 0, vc);
        ldl<T, K>(Lv, st, 1, 0, vd);
        ldl<T, K>(Lv, st, -1, 0, vu);
        ldl<T, K>(Lv, st, 0, 1, vr);
        ldl<T, K>(Lv, st, 0, -1, vl);
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
                                                double *lds, T *L) {
    const Strip st = strip_of(U, lb);
    stage_strip<T, K>(U, U.w, st, T(0), L);
    double m[K];
#pragma unroll
    for (int j = 0; j < K; ++j) m[j] = 0.0;
    if (st.real) {
        T d1[K], d2[K], d3[K], d4[K];
        ldl<T, K>(L, st, 1, 0, d1);
        ldl<T, K>(L, st, -1, 0, d2);
        ldl<T, K>(L, st, 0, 1, d3);
        ldl<T, K>(L, st, 0, -1, d4);
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
    if (real_pixel(q, U.npix_t, U.tny, U.nx, U.ny)) {
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
    } else if constexpr (KMAX == 4) {           \
        switch (U.K) {                          \
            case 1: { constexpr int K = 1; CALL; break; } \
            case 2: { constexpr int K = 2; CALL; break; } \
            default: { constexpr int K = 4; CALL; break; } \
        }                                       \
    } else {                                    \
        switch (U.K) {                          \
            case 1: { constexpr int K = 1; CALL; break; } \
            case 2: { constexpr int K = 2; CALL; break; } \
            case 4: { constexpr int K = 4; CALL; break; } \
            default: { constexpr int K = 8; CALL; break; } \
        }                                       \
    }

template <typename T, int KMAX, bool RESID, int FM>
__global__ __launch_bounds__(kBlock) void k_eng_fp(EngArgs<T> a, BlockMap m, int s) {
    __shared__ double lds[kBlock / kWave * kMaxK];
    __shared__ double ldr[kBlock / kOtfPerWave * KMAX];
    const int b = (int)xcd_block();
    if (b >= m.start[m.nunits]) return;
    const int ui = find_unit(m, b);
    const UG<T> U(a.units[ui]);
    if (unit_idle(a, U)) return;
    const int lb = b - m.start[ui];
    MI_KSWITCH(KMAX, (fp_unit<T, K, RESID, FM>(a, U, s, lb, lds, ldr)))
}

template <typename T, int KMAX, bool LDS, bool SQ>
__global__ __launch_bounds__(kBlock) void k_eng_bp(EngArgs<T> a, BlockMap m, int s, int sq_off) {
    __shared__ double lds[kBlock / kWave * kMaxK];
    extern __shared__ __align__(16) unsigned char ldr_raw[];   // LDS: the block's residual slots
    const int b = (int)xcd_block();
    if (b >= m.start[m.nunits]) return;
    const int ui = find_unit(m, b);
    const UG<T> U(a.units[ui]);
    if (unit_idle(a, U)) return;
    const int lb = b - m.start[ui];
    MI_KSWITCH(KMAX, (bp_unit<T, K, LDS, SQ>(a, U, s, lb, lds, (LDSP T *)ldr_raw, sq_off)))
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

// Fixed-order block sums of up to four partial arrays at once (all their
// loads in flight together), 256 threads; out[i] = sum of p[i][0 .. n[i]).
template <int NA>
__device__ __forceinline__ void fin_reduce_n(const GL double *const (&p)[NA], const int64_t (&n)[NA],
                                             double (&out)[NA], double *lds) {
    double acc[NA];
#pragma unroll
    for (int q = 0; q < NA; ++q) acc[q] = 0.0;
    int64_t nmax = 0;
#pragma unroll
    for (int q = 0; q < NA; ++q) nmax = n[q] > nmax ? n[q] : nmax;
    for (int64_t base = 0; base < nmax; base += (int64_t)kBlock * 4) {
        double v[NA][4];
#pragma unroll
        for (int q = 0; q < NA; ++q)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t i = base + u * kBlock + threadIdx.x;
                v[q][u] = (i < n[q]) ? p[q][i] : 0.0;
            }
#pragma unroll
        for (int q = 0; q < NA; ++q)
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[q] += v[q][u];
    }
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < NA; ++q) {
        acc[q] = group_sum<kWave>(acc[q]);
        if (lane == 0) lds[q * (kBlock / kWave) + wv] = acc[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NA; ++q) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += lds[q * (kBlock / kWave) + w];
        out[q] = t;
    }
    __syncthreads();
}

__host__ __device__ __forceinline__ int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// The subset step's sums of job j of unit U (fixed order over the blocks of
// the forward walks, back projection and nV) and its line search
// c = <d,d> / (S |A_s d|^2 + alpha nV(d)).  All threads of the block.
template <typename T>
__device__ __forceinline__ void fin_step(const EngArgs<T> &a, const UG<T> &U, int j, int s, double *lds) {
    const int js = U.job0 + j;
    if (!job_live(a, js)) return;
    GL Scalars *sc = &jsc(a, js);
    const GL JobParams &P = jpar(a, js);
    const int64_t nr = cdiv(U.fs_slice0[s + 1] - U.fs_slice0[s], kBlock / kWave);
    const int64_t nb = cdiv(U.bs_slice0[s + 1] - U.bs_slice0[s], kBlock / kWave);
    const int64_t npx = cdiv(U.npix_t, kPixBlock);
    const GL double *const p[4] = {part_at(a, U, SL_DD, j), part_at(a, U, SL_AD, j), part_at(a, U, SL_NV, j),
                                   part_at(a, U, SL_DF, j)};
    const int64_t n[4] = {nb, nr, npx, nr};
    double r[4];
    fin_reduce_n<4>(p, n, r, lds);
    const double dd = r[0], ad2 = r[1], nv = r[2], df = r[3];
    if (threadIdx.x == 0) {
        sc->df_pass += df;
        sc->dd = dd;
        sc->ad2 = ad2;
        sc->nv = nv;
        const double den = (double)a.S * ad2 + P.alpha * nv;
        sc->c = (den > 0) ? dd / den : 0.0;
    }
}

// Finalize op `op` of job j of unit U (every thread of the block calls it).
template <typename T>
__device__ __forceinline__ void fin_job(const EngArgs<T> &a, const UG<T> &U, int j, int op, int s, double *lds) {
    const int js = U.job0 + j;
    if (!job_live(a, js)) return;
    GL Scalars *sc = &jsc(a, js);
    const GL JobParams &P = jpar(a, js);
    const int64_t npx = cdiv(U.npix_t, kPixBlock);
    switch (op) {
        case EF_BEGIN:
            if (threadIdx.x == 0) {
                sc->df_pass = 0.0;
                sc->edge_done = 0;
            }
            break;
        case EF_EDGE_C: {
            if (sc->edge_done || s >= P.iter_edge) return;
            const GL double *const p[3] = {part_at(a, U, SL_E0, j), part_at(a, U, SL_E1, j),
                                           part_at(a, U, SL_E2, j)};
            const int64_t n[3] = {npx, npx, npx};
            double r[3];
            fin_reduce_n<3>(p, n, r, lds);
            const double nd = r[0], nd2 = r[1], nnd = r[2];
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

// One block per job slot (units listed in m; m.nunits * kMaxK blocks).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_eng_fin(EngArgs<T> a, int op, int s) {
    __shared__ double lds[4 * (kBlock / kWave)];
    const int ui = blockIdx.x / kMaxK, j = blockIdx.x % kMaxK;
    const UG<T> U(a.units[ui]);
    if (j >= U.K) return;
    if (op == EF_STEP) fin_step(a, U, j, s, lds);
    else fin_job(a, U, j, op, s, lds);
}

enum PixOp { PX_AXPY = 0, PX_MEDIAN, PX_COPY_WU, PX_EDGE_GRAD, PX_EDGE_NABLA, PX_EDGE_UPDATE, PX_NV };

template <typename T, int KMAX, int OP>
__global__ __launch_bounds__(kBlock) void k_eng_pix(EngArgs<T> a, BlockMap m, int it) {
    __shared__ double lds[kBlock / kWave * kMaxK];
    // strips of the arrays the stencil reads around (u and v for the edge
    // gradient, one array for the median, nV and the gradient of the edge
    // direction)
    constexpr int kStrips = OP == PX_EDGE_GRAD ? 2 : (OP == PX_MEDIAN || OP == PX_EDGE_NABLA || OP == PX_NV ? 1 : 0);
    __shared__ __align__(16) T L[kStrips ? kStrips * kHaloN * KMAX : 1];
    const int b = (int)xcd_block();
    if (b >= m.start[m.nunits]) return;
    const int ui = find_unit(m, b);
    const UG<T> U(a.units[ui]);
    if (unit_idle(a, U)) return;
    const int lb = b - m.start[ui];
    if constexpr (OP == PX_AXPY) { MI_KSWITCH(KMAX, (axpy_unit<T, K>(a, U, lb))) }
    else if constexpr (OP == PX_MEDIAN) { MI_KSWITCH(KMAX, (median_unit<T, K>(a, U, lb, L))) }
    else if constexpr (OP == PX_NV) { MI_KSWITCH(KMAX, (nv_unit<T, K>(a, U, lb, lds, L))) }
    else if constexpr (OP == PX_COPY_WU) { MI_KSWITCH(KMAX, (copy_wu_unit<T, K>(U, lb))) }
    else if constexpr (OP == PX_EDGE_GRAD) { MI_KSWITCH(KMAX, (edge_grad_unit<T, K>(a, U, lb, it, lds, L))) }
    else if constexpr (OP == PX_EDGE_NABLA) { MI_KSWITCH(KMAX, (edge_nabla_unit<T, K>(a, U, lb, it, lds, L))) }
    else { MI_KSWITCH(KMAX, (edge_update_unit<T, K>(a, U, lb, it, lds))) }
}

template <typename T>
__global__ void k_put_rows(T *dst, int K, int slot, int64_t n, const int64_t *perm, const T *src) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i * K + slot] = src[perm[i]];
}
// Row-major image <-> slot of a tiled interleaved image.
template <typename T>
__global__ void k_put_img(T *dst, int K, int slot, int nx, int ny, int tny, const T *src, T fillv) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)nx * ny) return;
    const int64_t q = tile_index((int)(p / ny), (int)(p % ny), tny);
    dst[q * K + slot] = src ? src[p] : fillv;
}
template <typename T>
__global__ void k_get_img(T *dst, const T *src, int K, int slot, int nx, int ny, int tny) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)nx * ny) return;
    dst[p] = src[tile_index((int)(p / ny), (int)(p % ny), tny) * K + slot];
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
void eng_fp_resid(hipStream_t st, const EngArgs<T> &a, const BlockMap &m, int s, int kmax, bool all_walk) {
    const unsigned grid = grid8(m.start[m.nunits]);
    if (all_walk) {
        if (kmax == 1) launch(k_eng_fp<T, 1, true, 1>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, m, s);
        else if (kmax <= 4) launch(k_eng_fp<T, 4, true, 1>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, m, s);
        else launch(k_eng_fp<T, kMaxK, true, 1>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, m, s);
    } else {
        if (kmax == 1) launch(k_eng_fp<T, 1, true, 3>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, m, s);
        else if (kmax <= 4) launch(k_eng_fp<T, 4, true, 3>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, m, s);
        else launch(k_eng_fp<T, kMaxK, true, 3>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, m, s);
    }
    MI_HIP(hipGetLastError());
}

// Dynamic LDS of a back-projection launch.  The u / v squares of the
// regulariser either take a region of their own behind the residual slots,
// loaded while the slots are (the kernel's SQ form, squares at element sq_off),
// or reuse the slots' LDS after the walk.  The staged form takes its own
// region when both fit the launch's LDS and that kernel keeps as many blocks
// per CU (its registers and LDS, by the occupancy calculator, up to the
// blocks per CU the grid has); the global-gather form holds only the squares.
template <typename T, int KM, bool LDS>
void bp_launch(hipStream_t st, unsigned grid, const EngArgs<T> &a, const BlockMap &m, int s, size_t slot_bytes,
               int sq_mode) {
    const size_t sq = (size_t)2 * kSqN * KM * sizeof(T), at = (slot_bytes + 15) & ~(size_t)15;
    const uint32_t own = (uint32_t)(at + sq), shared = (uint32_t)std::max(slot_bytes, sq);
    const int sq_off = (int)(at / sizeof(T));
    if constexpr (!LDS) {
        launch(k_eng_bp<T, KM, false, true>, dim3(grid), dim3(kBlock), own, st, a, m, s, sq_off);
    } else {
        bool sep = false;
        if (at + sq > kBpLdsLaunch) {
            sep = false;
        } else if (sq_mode >= 0) {
            sep = sq_mode == 1;
        } else {
            int nb_own = 0, nb_shared = 0;
            MI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_own, (const void *)k_eng_bp<T, KM, true, true>,
                                                                kBlock, own));
            MI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb_shared, (const void *)k_eng_bp<T, KM, true, false>, kBlock, shared));
            int dev = 0, ncu = 1;
            MI_HIP(hipGetDevice(&dev));
            MI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            const int used = (int)((grid + ncu - 1) / ncu);   // blocks per CU the launch can place
            sep = std::min(nb_own, used) >= std::min(nb_shared, used);
        }
        if (sep) launch(k_eng_bp<T, KM, true, true>, dim3(grid), dim3(kBlock), own, st, a, m, s, sq_off);
        else launch(k_eng_bp<T, KM, true, false>, dim3(grid), dim3(kBlock), shared, st, a, m, s, 0);
    }
}

template <typename T>
void eng_bp_grad(hipStream_t st, const EngArgs<T> &a, const BlockMap &m, int s, int kmax, size_t lds_bytes,
                 int sq_mode) {
    const unsigned grid = grid8(m.start[m.nunits]);
    if (lds_bytes) {
        if (kmax == 1) bp_launch<T, 1, true>(st, grid, a, m, s, lds_bytes, sq_mode);
        else if (kmax <= 4) bp_launch<T, 4, true>(st, grid, a, m, s, lds_bytes, sq_mode);
        else bp_launch<T, kMaxK, true>(st, grid, a, m, s, lds_bytes, sq_mode);
    } else {
        if (kmax == 1) bp_launch<T, 1, false>(st, grid, a, m, s, 0, sq_mode);
        else if (kmax <= 4) bp_launch<T, 4, false>(st, grid, a, m, s, 0, sq_mode);
        else bp_launch<T, kMaxK, false>(st, grid, a, m, s, 0, sq_mode);
    }
    MI_HIP(hipGetLastError());
}

template <typename T>
void eng_fp_sq(hipStream_t st, const EngArgs<T> &a, const BlockMap &rows, int s, int kmax, bool all_walk) {
    const unsigned grid = grid8(rows.start[rows.nunits]);
    if (all_walk) {
        if (kmax == 1) launch(k_eng_fp<T, 1, false, 1>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, rows, s);
        else if (kmax <= 4) launch(k_eng_fp<T, 4, false, 1>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, rows, s);
        else launch(k_eng_fp<T, kMaxK, false, 1>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, rows, s);
    } else {
        if (kmax == 1) launch(k_eng_fp<T, 1, false, 3>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, rows, s);
        else if (kmax <= 4) launch(k_eng_fp<T, 4, false, 3>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, rows, s);
        else launch(k_eng_fp<T, kMaxK, false, 3>, dim3(grid), dim3(kBlock), (uint32_t)(0), st, a, rows, s);
    }
    MI_HIP(hipGetLastError());
}

template <typename T, int OP>
static void launch_pix(hipStream_t st, const EngArgs<T> &a, const BlockMap &m, int it, int kmax) {
    const unsigned grid = grid8(m.start[m.nunits]);
    if (kmax == 1) k_eng_pix<T, 1, OP><<<grid, kBlock, 0, st>>>(a, m, it);
    else if (kmax <= 4) k_eng_pix<T, 4, OP><<<grid, kBlock, 0, st>>>(a, m, it);
    else k_eng_pix<T, kMaxK, OP><<<grid, kBlock, 0, st>>>(a, m, it);
    MI_HIP(hipGetLastError());
}

template <typename T>
void eng_nv(hipStream_t st, const EngArgs<T> &a, const BlockMap &pix, int kmax) {
    launch_pix<T, PX_NV>(st, a, pix, 0, kmax);
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
void eng_put_img(hipStream_t st, T *dst, int K, int slot, int nx, int ny, int tny, const T *src, T fillv) {
    const int64_t n = (int64_t)nx * ny;
    if (n <= 0) return;
    k_put_img<T><<<grid_of(n), kBlock, 0, st>>>(dst, K, slot, nx, ny, tny, src, fillv);
    MI_HIP(hipGetLastError());
}
template <typename T>
void eng_get_img(hipStream_t st, T *dst, const T *src, int K, int slot, int nx, int ny, int tny) {
    const int64_t n = (int64_t)nx * ny;
    if (n <= 0) return;
    k_get_img<T><<<grid_of(n), kBlock, 0, st>>>(dst, src, K, slot, nx, ny, tny);
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
    template void eng_fp_resid<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int, int, bool);       \
    template void eng_bp_grad<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int, int, size_t, int); \
    template void eng_fp_sq<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int, int, bool);         \
    template void eng_nv<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int);                        \
    template void eng_axpy<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int);                      \
    template void eng_median<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int);                    \
    template void eng_edge_iter<T>(hipStream_t, const EngArgs<T> &, const BlockMap &, int, int);            \
    template void eng_finalize<T>(hipStream_t, const EngArgs<T> &, int, int);                               \
    template void eng_put_rows<T>(hipStream_t, T *, int, int, int64_t, const int64_t *, const T *);         \
    template void eng_put_img<T>(hipStream_t, T *, int, int, int, int, int, const T *, T);                  \
    template void eng_get_img<T>(hipStream_t, T *, const T *, int, int, int, int, int);                     \
    template void eng_move<T>(hipStream_t, T *, int, int, const T *, int, int, int64_t);

MI_ENG_INST(float)
MI_ENG_INST(double)

}  // namespace mi
