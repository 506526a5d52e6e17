// C ABI of libmedimage_amd.so (include/medimage_amd.h): argument checks,
// error translation, context and geometry cache, instrumentation.
#include "batch.hpp"
#include "cone.hpp"
#include "ordered.hpp"
#include "solver.hpp"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

namespace mi {

namespace {
thread_local std::string g_last_error;

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return MI_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return MI_E_NOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return MI_E_STATE;
    }
}
}  // namespace

void fail(int code, const std::string &msg) { throw Error(code, msg); }

void hip_check(hipError_t e, const char *what, const char *file, int line) {
    if (e == hipSuccess) return;
    std::string m = std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")";
    throw Error(e == hipErrorOutOfMemory ? MI_E_NOMEM : MI_E_DEVICE, m);
}

std::vector<double> view_table(const mi_geometry &g) {
    if (!g.view_deg) return {};
    return std::vector<double>(g.view_deg, g.view_deg + g.n_views);
}

void validate_geometry(const mi_geometry &g) {
    MI_REQUIRE(g.nx > 0 && g.ny > 0 && g.n_views > 0 && g.n_bins > 0, MI_E_ARG, "geometry sizes must be positive");
    MI_REQUIRE(!g.view_deg || g.beam == MI_BEAM_FAN_ARC, MI_E_ARG, "per-view angles are for the arc beam only");
    if (g.view_deg)
        for (int v = 0; v < g.n_views; ++v)
            MI_REQUIRE(std::isfinite(g.view_deg[v]), MI_E_ARG, "per-view angles must be finite");
    MI_REQUIRE((int64_t)g.nx * g.ny < (int64_t)1 << 31, MI_E_LIMIT, "image too large");
    switch (g.beam) {
        case MI_BEAM_FAN_FLAT:
            MI_REQUIRE(g.pixel_mm > 0 && g.bin_mm > 0 && g.sod_mm > 0 && g.sdd_mm > g.sod_mm, MI_E_ARG,
                       "fan geometry needs pixel_mm, bin_mm > 0 and sdd > sod > 0");
            MI_REQUIRE(std::isfinite(g.span_deg) && g.span_deg > 0, MI_E_ARG, "span_deg must be positive");
            break;
        case MI_BEAM_FAN_ARC:
            MI_REQUIRE(g.pixel_mm > 0 && g.bin_mm > 0 && g.sod_mm > 0 && g.sdd_mm > g.sod_mm, MI_E_ARG,
                       "arc geometry needs pixel_mm, bin_mm (degrees) > 0 and sdd > sod > 0");
            MI_REQUIRE(g.nx == g.ny && g.n_bins % 2 == 0, MI_E_ARG, "arc geometry needs a square image and even n_bins");
            MI_REQUIRE(std::isfinite(g.span_deg) && g.span_deg > 0 && std::isfinite(g.det_offset), MI_E_ARG,
                       "span_deg must be positive, det_offset finite");
            MI_REQUIRE(g.bin_mm * g.n_bins < 180.0, MI_E_ARG, "the arc detector must span less than 180 degrees");
            break;
        case MI_BEAM_PARALLEL:
            MI_REQUIRE(g.nx == g.ny, MI_E_ARG, "parallel beam needs a square image");
            MI_REQUIRE(g.flavour == 0 || g.flavour == 1, MI_E_ARG, "flavour must be 0 or 1");
            break;
        case MI_BEAM_GOLDENC:
            MI_REQUIRE(g.nx == 512 && g.ny == 512 && g.n_views == 180 && g.n_bins == 767, MI_E_ARG,
                       "GoldenC geometry is 512 x 512, 180 views, 767 bins");
            break;
        default:
            fail(MI_E_ARG, "unknown beam kind");
    }
}

GeomKey make_key(const mi_geometry &g, int S, int layout) {
    GeomKey k;
    k.i = {g.beam, g.nx, g.ny, g.n_views, g.n_bins, g.beam == MI_BEAM_PARALLEL ? g.flavour : 0, S, layout};
    if (g.beam == MI_BEAM_FAN_FLAT || g.beam == MI_BEAM_FAN_ARC)
        k.d = {g.pixel_mm, g.view_deg ? 0.0 : g.span_deg, g.bin_mm, g.sod_mm, g.sdd_mm, g.det_offset};
    else k.d = {0, 0, 0, 0, 0, 0};
    k.views = view_table(g);
    return k;
}

std::shared_ptr<Matrix> get_matrix(mi_context *ctx, const mi_geometry &g, int S, bool want64) {
    const GeomKey key = make_key(g, S);
    auto it = ctx->matrices.find(key);
    if (it != ctx->matrices.end() && (!want64 || it->second->has64)) {
        it->second->last_use = ++ctx->tick;
        return it->second;
    }
    if (it != ctx->matrices.end()) ctx->matrices.erase(it);
    std::shared_ptr<Matrix> m(build_matrix(g, S, want64, ctx->stream).release());
    m->last_use = ++ctx->tick;
    ctx->matrices[key] = m;
    return m;
}

std::shared_ptr<Matrix> find_engine_matrix(mi_context *ctx, const mi_geometry &g, int S, bool want64) {
    auto it = ctx->matrices.find(make_key(g, S, 1));
    if (it == ctx->matrices.end() || (want64 && !it->second->has64)) return nullptr;
    return it->second;
}

std::shared_ptr<Matrix> get_engine_matrix(mi_context *ctx, const mi_geometry &g, int S, bool want64) {
    const GeomKey key = make_key(g, S, 1);
    auto it = ctx->matrices.find(key);
    if (it != ctx->matrices.end() && (!want64 || it->second->has64)) {
        it->second->last_use = ++ctx->tick;
        return it->second;
    }
    if (it != ctx->matrices.end()) ctx->matrices.erase(it);
    std::shared_ptr<Matrix> m;
    for (int attempt = 0;; ++attempt) {
        try {
            m.reset(build_matrix(g, S, want64, ctx->stream, true).release());
            build_sell(*m, ctx->stream, true);
            break;
        } catch (const Error &e) {
            // Out of device memory: drop idle cached matrices (least recently
            // used first) and build once more.
            (void)hipGetLastError();
            m.reset();
            if (e.code != MI_E_NOMEM || attempt > 0 || evict_idle(ctx, -1) <= 0) throw;
        }
    }
    m->last_use = ++ctx->tick;
    ctx->matrices[key] = m;
    return m;
}

int64_t evict_idle(mi_context *ctx, int64_t bytes) {
    int64_t freed = 0;
    while (bytes < 0 || freed < bytes) {
        auto victim = ctx->matrices.end();
        for (auto it = ctx->matrices.begin(); it != ctx->matrices.end(); ++it)
            if (it->second.use_count() == 1 &&
                (victim == ctx->matrices.end() || it->second->last_use < victim->second->last_use))
                victim = it;
        if (victim == ctx->matrices.end()) break;
        MI_HIP(hipStreamSynchronize(ctx->stream));
        freed += (int64_t)victim->second->bytes();
        ctx->matrices.erase(victim);
    }
    return freed;
}

thread_local DispatchEvents *tl_dispatch = nullptr;

LaunchTimer::LaunchTimer(mi_context *c, int k, double by, bool disp) : ctx(c), kernel(k), bytes(by) {
    if (!ctx->stats.on || !((ctx->stats.mask >> k) & 1u)) return;
    auto &pool = ctx->stats.pool;
    auto take = [&]() {
        hipEvent_t e;
        // timing only, read after a stream synchronize: no system-scope
        // fence (cache write-back) when the event is recorded
        if (!pool.empty()) { e = pool.back(); pool.pop_back(); }
        else MI_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        return e;
    };
    a = take();
    b = take();
    if (disp) {
        dispatch = true;
        de.a = a;
        de.b = b;
        prev = tl_dispatch;
        tl_dispatch = &de;
    } else {
        MI_HIP(hipEventRecord(a, ctx->stream));
    }
}

LaunchTimer::~LaunchTimer() {
    if (!a) return;
    if (dispatch) {
        tl_dispatch = prev;
        if (de.launches == 0) {          // nothing was launched in the scope
            ctx->stats.pool.push_back(a);
            ctx->stats.pool.push_back(b);
            return;
        }
        if (de.launches > 1 && hipEventRecord(b, ctx->stream) != hipSuccess) return;
    } else if (hipEventRecord(b, ctx->stream) != hipSuccess) {
        return;
    }
    ctx->stats.pending.push_back({kernel, a, b, bytes});
}

void stats_collect(mi_context *ctx) {
    auto &S = ctx->stats;
    if (S.pending.empty()) return;
    MI_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &p : S.pending) {
        float ms = 0.f;
        MI_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        auto &k = S.k[p.kernel];
        k.launches += 1;
        k.total_ms += ms;
        k.bytes += p.bytes;
        S.pool.push_back(p.a);
        S.pool.push_back(p.b);
    }
    S.pending.clear();
}

}  // namespace mi

using namespace mi;

// The single-job API is a batch holding one job that is never retired, so
// it runs the same engine code as the batched form.
struct mi_job {
    mi_context *ctx = nullptr;
    int precision = MI_FP32;
    std::unique_ptr<Batch<float>> b32;
    std::unique_ptr<Batch<double>> b64;
    int64_t id = 0;
    double rel_tol = 0;
};

struct mi_batch {
    mi_context *ctx = nullptr;
    int precision = MI_FP32;
    std::unique_ptr<Batch<float>> b32;
    std::unique_ptr<Batch<double>> b64;
};

namespace {

mi_context *check_ctx(mi_context *ctx) {
    MI_REQUIRE(ctx != nullptr, MI_E_ARG, "null context");
    MI_HIP(hipSetDevice(ctx->device));
    return ctx;
}

template <typename T>
int64_t single_job(Batch<T> &b, const mi_geometry &g, mi_at_params p, const void *sino, const void *u0,
                   const void *v0) {
    b.auto_retire = false;
    const int64_t id = b.submit(g, p, (const T *)sino, (const T *)u0, (const T *)v0);
    b.admit(true);
    return id;
}

// Queues passes in chunks of `every` (all at once when 0) until the job's
// device flag says done or max_passes were queued; returns passes run.
template <typename T>
int run_job(mi_context *ctx, Batch<T> &b, int64_t id, int max_passes, int every) {
    const Scalars h0 = b.scalars(id);
    Scalars h = h0;
    int queued = 0;
    while (queued < max_passes && !h.done) {
        const int chunk = std::min(every > 0 ? every : max_passes, max_passes - queued);
        b.enqueue(chunk);
        queued += chunk;
        h = b.scalars(id);
    }
    stats_collect(ctx);
    return h.iters - h0.iters;
}

template <typename T>
void fill_report(Batch<T> &b, int64_t id, mi_report *rep) {
    const Scalars h = b.scalars(id);
    std::memset(rep, 0, sizeof(*rep));
    rep->iterations = h.iters;
    rep->converged = h.converged;
    rep->data_fidelity = h.df_prev;
    rep->rel_change = h.rel;
    rep->nnz = b.job(id).m ? b.job(id).m->nnz : 0;
}

template <typename T>
void at_solve(mi_context *ctx, const mi_geometry &g, const mi_at_params &p, const T *sino, T *u, T *v,
              mi_report *rep) {
    MI_REQUIRE(p.n_subsets >= 1 && p.max_iters >= 0 && p.iter_edge >= 0, MI_E_ARG, "bad AT parameters");
    MI_REQUIRE((p.precision == MI_FP64) == (sizeof(T) == 8), MI_E_ARG, "precision does not match the buffers");
    Batch<T> b(ctx, p.n_subsets);
    const int64_t id = single_job<T>(b, g, p, sino, u, v);
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    MI_HIP(hipEventRecord(e0, ctx->stream));
    run_job<T>(ctx, b, id, p.max_iters, p.rel_tol > 0 ? 4 : 0);
    MI_HIP(hipEventRecord(e1, ctx->stream));
    MI_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    MI_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b.read(id, u, v);
    if (rep) {
        fill_report<T>(b, id, rep);
        rep->elapsed_ms = ms;
    }
}

template <typename T>
void forward_impl(mi_context *ctx, const mi_geometry *g, const T *img, T *sino) {
    MI_REQUIRE(g && img && sino, MI_E_ARG, "null argument");
    const bool f64 = sizeof(T) == 8;
    Matrix &m = *get_matrix(ctx, *g, 1, f64);
    const hipStream_t st = ctx->stream;
    DevBuf<T> x(m.npix), rows(m.nrows), user(m.nrows);
    MI_HIP(hipMemcpyAsync(x.p, img, m.npix * sizeof(T), hipMemcpyHostToDevice, st));
    if constexpr (sizeof(T) == 8) {
        // GoldenC's A in the reference's summation order (bitwise its result).
        if (g->beam == MI_BEAM_GOLDENC) {
            const SeqMatrix &q = get_seq(st, m);
            LaunchTimer lt(ctx, MI_K_FP, fp_bytes(m, m.nrows, sizeof(T), FP_WRITE));
            seq_forward<double>(st, q, x.p, rows.p);
        }
    }
    if (!(f64 && g->beam == MI_BEAM_GOLDENC)) {
        LaunchTimer lt(ctx, MI_K_FP, fp_bytes(m, m.nrows, sizeof(T), FP_WRITE));
        launch_fp<T>(st, FP_WRITE, csr_view(m, f64), 0, m.nrows, x.p, nullptr, nullptr, rows.p, nullptr);
    }
    launch_scatter_rows<T>(st, m.nrows, m.perm.p, rows.p, user.p);
    MI_HIP(hipMemcpyAsync(sino, user.p, m.nrows * sizeof(T), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    stats_collect(ctx);
}

template <typename T>
void backproject_impl(mi_context *ctx, const mi_geometry *g, const T *sino, T *img) {
    MI_REQUIRE(g && img && sino, MI_E_ARG, "null argument");
    const bool f64 = sizeof(T) == 8;
    Matrix &m = *get_matrix(ctx, *g, 1, f64);
    const hipStream_t st = ctx->stream;
    DevBuf<T> x(m.npix), rows(m.nrows);
    load_sino<T>(st, m, sino, rows.p);
    if constexpr (sizeof(T) == 8) {
        // GoldenC's AStar in the reference's summation order.
        if (g->beam == MI_BEAM_GOLDENC) {
            const SeqMatrix &q = get_seq(st, m);
            LaunchTimer lt(ctx, MI_K_BP, bp_bytes(m, m.nrows, sizeof(T), false));
            seq_backproject<double>(st, q, g->nx, g->ny, rows.p, nullptr, nullptr, 0.0, 0.0, x.p);
        }
    }
    if (!(f64 && g->beam == MI_BEAM_GOLDENC)) {
        LaunchTimer lt(ctx, MI_K_BP, bp_bytes(m, m.nrows, sizeof(T), false));
        launch_bp_grad<T>(st, csc_view(m, 0, f64), g->nx, g->ny, rows.p, 1.0, nullptr, nullptr, 0.0, 0.0, x.p,
                          nullptr, bp_lanes(m));
    }
    MI_HIP(hipMemcpyAsync(img, x.p, m.npix * sizeof(T), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    stats_collect(ctx);
}

template <typename T>
void helical_slices(mi_context *ctx, const char *prefix, int first, int count, const mi_geometry &g0,
                    const mi_at_params &p, T *u_stack, T *v_stack, mi_report *reps) {
    MI_REQUIRE(g0.beam == MI_BEAM_FAN_ARC, MI_E_ARG, "the helical slices use the arc beam");
    MI_REQUIRE(p.n_subsets >= 1 && p.max_iters >= 1 && p.iter_edge >= 0, MI_E_ARG, "bad AT parameters");
    const int64_t npix = (int64_t)g0.nx * g0.ny;
    Batch<T> b(ctx, p.n_subsets);
    std::vector<int64_t> ids(count);
    // The slices' files are read on host threads (up to 16 at once), then the
    // slices are submitted in order; the first slice (in order) whose files
    // do not read gives the error.
    struct SliceIn {
        std::vector<double> angle;
        std::vector<T> sino;
        std::string err;
    };
    std::vector<SliceIn> in(count);
    std::atomic<int> next{0};
    auto load = [&] {
        std::vector<double> stage;
        for (int k; (k = next.fetch_add(1)) < count;) {
            SliceIn &S = in[k];
            try {
                const std::string base = std::string(prefix) + std::to_string(first + k);
                // the slice's table fixes its view count and angles (load_table)
                int nv = 0;
                if (mi_read_pos((base + "_pos").c_str(), nullptr, nullptr, 0, &nv) != MI_OK) {
                    S.err = mi_last_error();
                    continue;
                }
                if (nv < p.n_subsets) {
                    S.err = base + "_pos: fewer views than subsets";
                    continue;
                }
                S.angle.resize(nv);
                if (mi_read_pos((base + "_pos").c_str(), nullptr, S.angle.data(), nv, &nv) != MI_OK) {
                    S.err = mi_last_error();
                    continue;
                }
                stage.resize((size_t)nv * g0.n_bins);
                if (mi_read_raw((base + "_raw").c_str(), nv, g0.n_bins, stage.data()) != MI_OK) {
                    S.err = mi_last_error();
                    continue;
                }
                S.sino.assign(stage.begin(), stage.end());
            } catch (const std::exception &e) {
                S.err = e.what();
            }
        }
    };
    const int nt = std::max(1, std::min({count, 16, (int)std::thread::hardware_concurrency()}));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(load);
    load();
    for (auto &t : pool) t.join();
    for (int k = 0; k < count; ++k) {
        MI_REQUIRE(in[k].err.empty(), MI_E_ARG, in[k].err);
        mi_geometry g = g0;
        g.n_views = (int)in[k].angle.size();
        g.view_deg = in[k].angle.data();
        ids[k] = b.submit(g, p, in[k].sino.data(), nullptr, nullptr);
        std::vector<T>().swap(in[k].sino);   // the batch holds its own copy
    }
    // Every slice advances in the same passes; finished slices leave (their
    // result goes to host memory) and pending ones join as device memory
    // frees up.
    for (;;) {
        b.step(1);
        bool all_done = true;
        for (int k = 0; k < count; ++k) {
            const BatchJob<T> &j = b.job(ids[k]);
            MI_REQUIRE(j.state != BJ_FAILED, MI_E_NOMEM, "slice " + std::to_string(first + k) + ": " + j.error);
            if (j.state != BJ_DONE) all_done = false;
        }
        if (all_done) break;
    }
    for (int k = 0; k < count; ++k) {
        b.read(ids[k], u_stack + (size_t)k * npix, v_stack + (size_t)k * npix);
        if (reps) fill_report<T>(b, ids[k], &reps[k]);
    }
    stats_collect(ctx);
}

// Projections in the reference's summation order (mi_forward_ordered).
template <typename T>
void ordered_impl(mi_context *ctx, const mi_geometry *g, const T *in, T *out, bool forward) {
    MI_REQUIRE(g && in && out, MI_E_ARG, "null argument");
    // The parallel beam's tracer works in float; the others' lengths are
    // double and the lists keep them so.
    Matrix &m = *get_matrix(ctx, *g, 1, g->beam != MI_BEAM_PARALLEL);
    const hipStream_t st = ctx->stream;
    const SeqMatrix &q = get_seq(st, m);
    DevBuf<T> x(m.npix), rows(m.nrows);
    if (forward) {
        MI_HIP(hipMemcpyAsync(x.p, in, m.npix * sizeof(T), hipMemcpyHostToDevice, st));
        {
            LaunchTimer lt(ctx, MI_K_FP, fp_bytes(m, m.nrows, sizeof(T), FP_WRITE));
            seq_forward<T>(st, q, x.p, rows.p);
        }
        // one subset: rows are already in [view][bin] order
        MI_HIP(hipMemcpyAsync(out, rows.p, m.nrows * sizeof(T), hipMemcpyDeviceToHost, st));
    } else {
        MI_HIP(hipMemcpyAsync(rows.p, in, m.nrows * sizeof(T), hipMemcpyHostToDevice, st));
        {
            LaunchTimer lt(ctx, MI_K_BP, bp_bytes(m, m.nrows, sizeof(T), false));
            seq_backproject<T>(st, q, g->nx, g->ny, rows.p, nullptr, nullptr, 0.0, 0.0, x.p);
        }
        MI_HIP(hipMemcpyAsync(out, x.p, m.npix * sizeof(T), hipMemcpyDeviceToHost, st));
    }
    MI_HIP(hipStreamSynchronize(st));
    stats_collect(ctx);
}

// Algorithmic bytes of a cone projection when its kernel is timed (the
// entry count is traced once per geometry, only then).
double cone_timed_bytes(mi_context *ctx, ConeTables &t, int kernel) {
    if (!ctx->stats.on || !((ctx->stats.mask >> kernel) & 1u)) return 0.0;
    cone_nnz(ctx->stream, t);
    return kernel == MI_K_CONE_FP ? cone_fp_bytes(t) : cone_bp_bytes(t);
}

void cone_forward_dev(mi_context *ctx, const mi_cone_geometry *g, const float *vol, float *sino) {
    MI_REQUIRE(g && vol && sino, MI_E_ARG, "null argument");
    ConeTables &t = cone_tables(ctx, *g);
    {
        LaunchTimer lt(ctx, MI_K_CONE_FP, cone_timed_bytes(ctx, t, MI_K_CONE_FP), true);
        cone_forward(ctx->stream, t, vol, sino);
    }
    MI_HIP(hipStreamSynchronize(ctx->stream));
    stats_collect(ctx);
}

void cone_backproject_dev(mi_context *ctx, const mi_cone_geometry *g, const float *sino, float *vol) {
    MI_REQUIRE(g && vol && sino, MI_E_ARG, "null argument");
    ConeTables &t = cone_tables(ctx, *g);
    {
        LaunchTimer lt(ctx, MI_K_CONE_BP, cone_timed_bytes(ctx, t, MI_K_CONE_BP), true);
        cone_backproject(ctx->stream, t, sino, vol);
    }
    MI_HIP(hipStreamSynchronize(ctx->stream));
    stats_collect(ctx);
}

// The same back projection with no device allocation and no wait: the
// geometry's tables must be cached already, the carried sums use the
// device's fixed scratch array.
void cone_backproject_dev_noalloc(mi_context *ctx, const mi_cone_geometry *g, const float *sino, float *vol) {
    MI_REQUIRE(g && vol && sino, MI_E_ARG, "null argument");
    validate_cone(*g);
    MI_REQUIRE(ctx->cone && std::memcmp(&ctx->cone->g, g, sizeof(*g)) == 0, MI_E_STATE,
               "the geometry's tables are not on the device: make one cone call with this geometry first");
    cone_bp_views(ctx->stream, *ctx->cone, ctx->cone->all.p, ctx->cone->d.nviews, sino, vol, ctx->device);
}

size_t cone_voxels(const mi_cone_geometry &g) { return (size_t)g.nx * g.ny * g.nz; }
size_t cone_rays(const mi_cone_geometry &g) { return (size_t)g.n_views * g.n_rows * g.n_cols; }

}  // namespace

extern "C" {

const char *mi_last_error(void) { return g_last_error.c_str(); }

int mi_context_create(int device, mi_context **out) {
    return guarded([&] {
        MI_REQUIRE(out != nullptr, MI_E_ARG, "null output pointer");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            (void)hipGetLastError();
            fail(MI_E_DEVICE, "no HIP device available");
        }
        MI_REQUIRE(device >= 0 && device < n, MI_E_ARG, "device index out of range");
        MI_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        MI_HIP(hipGetDeviceProperties(&prop, device));
        MI_REQUIRE(std::string(prop.gcnArchName).rfind("gfx950", 0) == 0, MI_E_DEVICE,
                   std::string("library is built for gfx950, device is ") + prop.gcnArchName);
        // the device's address of the fixed scratch array (once per device)
        cone_scratch_init(device);
        auto c = new mi_context();
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            MI_HIP(e);
        }
        *out = c;
    });
}

int mi_context_destroy(mi_context *ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->matrices.clear();
        for (auto &p : ctx->stats.pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : ctx->stats.pool) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int mi_context_mem(mi_context *ctx, int64_t *bytes) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(bytes, MI_E_ARG, "null output");
        int64_t b = 0;
        for (auto &kv : ctx->matrices) b += (int64_t)kv.second->bytes();
        if (ctx->cone) {
            const ConeTables &t = *ctx->cone;
            b += (int64_t)(t.cs.bytes() + t.rows.bytes() + t.all.bytes() + t.vinv.bytes() + t.vl2.bytes() +
                           t.vsrc.bytes() + t.vpos.bytes());
        }
        *bytes = b;
    });
}

int mi_context_trim(mi_context *ctx) {
    return guarded([&] {
        check_ctx(ctx);
        MI_HIP(hipStreamSynchronize(ctx->stream));
        ctx->matrices.clear();
        ctx->cone.reset();
    });
}

int mi_geometry_prepare(mi_context *ctx, const mi_geometry *g, int n_subsets, int64_t *nnz) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && n_subsets >= 1, MI_E_ARG, "bad arguments");
        Matrix &m = *get_engine_matrix(ctx, *g, n_subsets, false);
        if (nnz) *nnz = m.nnz;
    });
}

int mi_autotune(mi_context *ctx, const mi_geometry *g, int n_subsets, mi_tuning *out) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g, MI_E_ARG, "null geometry");
        auto m = get_matrix(ctx, *g, n_subsets, false);
        autotune(ctx, *m);
        if (out) *out = m->tune;
    });
}

int mi_set_tuning(mi_context *ctx, const mi_geometry *g, int n_subsets, const mi_tuning *t) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && t, MI_E_ARG, "null argument");
        MI_REQUIRE(t->fp_vec == 0 || t->fp_vec == 1, MI_E_ARG, "fp_vec must be 0 or 1");
        MI_REQUIRE(t->bp_lanes == 4 || t->bp_lanes == 8 || t->bp_lanes == 16 || t->bp_lanes == 32 ||
                       t->bp_lanes == 64,
                   MI_E_ARG, "bp_lanes must be 4, 8, 16, 32 or 64");
        MI_REQUIRE(t->fp_grid >= 0 && t->fp_grid <= kMaxGrid && t->bp_grid >= 0 && t->bp_grid <= kMaxGrid, MI_E_ARG,
                   "grid caps must be in [0, 8192]");
        get_matrix(ctx, *g, n_subsets, false)->tune = *t;
    });
}

int mi_engine_autotune(mi_context *ctx, const mi_geometry *g, int n_subsets, int precision, mi_engine_tuning *out) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && out && n_subsets >= 1, MI_E_ARG, "bad arguments");
        MI_REQUIRE(precision == MI_FP32 || precision == MI_FP64, MI_E_ARG, "bad precision");
        *out = precision == MI_FP64 ? engine_autotune<double>(ctx, *g, n_subsets)
                                    : engine_autotune<float>(ctx, *g, n_subsets);
    });
}

int mi_engine_set_tuning(mi_context *ctx, const mi_geometry *g, int n_subsets, int precision,
                         const mi_engine_tuning *t) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && t && n_subsets >= 1, MI_E_ARG, "bad arguments");
        MI_REQUIRE(t->k_cap == 0 || t->k_cap == 1 || t->k_cap == 2 || t->k_cap == 4 || t->k_cap == 8, MI_E_ARG,
                   "k_cap must be 0, 1, 2, 4 or 8");
        MI_REQUIRE(t->bp_lds >= -1 && t->bp_lds <= 1, MI_E_ARG, "bp_lds must be -1, 0 or 1");
        Matrix &m = *get_engine_matrix(ctx, *g, n_subsets, precision == MI_FP64);
        m.etune = Matrix::EngineTune{t->k_cap, t->bp_lds, t->us_per_job_pass};
    });
}

int mi_engine_get_tuning(mi_context *ctx, const mi_geometry *g, int n_subsets, int precision, mi_engine_tuning *out) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && out && n_subsets >= 1, MI_E_ARG, "bad arguments");
        const Matrix &m = *get_engine_matrix(ctx, *g, n_subsets, precision == MI_FP64);
        *out = mi_engine_tuning{m.etune.k_cap, m.etune.bp_lds, m.etune.us_per_job_pass};
    });
}

int mi_get_tuning(mi_context *ctx, const mi_geometry *g, int n_subsets, mi_tuning *out) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && out, MI_E_ARG, "null argument");
        auto m = get_matrix(ctx, *g, n_subsets, false);
        *out = m->tune;
        if (!out->bp_lanes) out->bp_lanes = bp_lanes(*m);
    });
}

int mi_trace_ray(mi_context *ctx, const mi_geometry *g, int view, int bin, int32_t *ind, double *len, int cap,
                 int *numb) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && numb && cap >= 0 && (cap == 0 || (ind && len)), MI_E_ARG, "bad arguments");
        MI_REQUIRE(view >= 0 && view < g->n_views && bin >= 0 && bin < g->n_bins, MI_E_ARG, "ray out of range");
        Matrix &m = *get_matrix(ctx, *g, 1, true);
        const int64_t r = (int64_t)view * g->n_bins + bin;  // one subset: rows in view order
        int64_t rp[2];
        MI_HIP(hipMemcpy(rp, m.rowptr.p + r, 2 * sizeof(int64_t), hipMemcpyDeviceToHost));
        const int n = (int)(rp[1] - rp[0]);
        const int take = n < cap ? n : cap;
        if (take > 0) {
            std::vector<uint32_t> ix(take);
            MI_HIP(hipMemcpy(ix.data(), m.idx.p + rp[0], take * sizeof(uint32_t), hipMemcpyDeviceToHost));
            MI_HIP(hipMemcpy(len, m.len64.p + rp[0], take * sizeof(double), hipMemcpyDeviceToHost));
            for (int i = 0; i < take; ++i) ind[i] = (int32_t)ix[i];
        }
        *numb = n;
    });
}

int mi_matrix_export(mi_context *ctx, const mi_geometry *g, int64_t *rowptr, int32_t *ind, double *len,
                     int64_t nnz) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && rowptr && (nnz == 0 || (ind && len)), MI_E_ARG, "null argument");
        Matrix &m = *get_matrix(ctx, *g, 1, true);
        MI_REQUIRE(nnz == m.nnz, MI_E_ARG, "nnz does not match the geometry's matrix");
        MI_HIP(hipMemcpy(rowptr, m.rowptr.p, (m.nrows + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (nnz > 0) {
            static_assert(sizeof(uint32_t) == sizeof(int32_t), "index width");
            MI_HIP(hipMemcpy(ind, m.idx.p, nnz * sizeof(uint32_t), hipMemcpyDeviceToHost));
            MI_HIP(hipMemcpy(len, m.len64.p, nnz * sizeof(double), hipMemcpyDeviceToHost));
        }
    });
}

int mi_forward(mi_context *ctx, const mi_geometry *g, const float *img, float *sino) {
    return guarded([&] { forward_impl<float>(check_ctx(ctx), g, img, sino); });
}
int mi_backproject(mi_context *ctx, const mi_geometry *g, const float *sino, float *img) {
    return guarded([&] { backproject_impl<float>(check_ctx(ctx), g, sino, img); });
}
int mi_forward_f64(mi_context *ctx, const mi_geometry *g, const double *img, double *sino) {
    return guarded([&] { forward_impl<double>(check_ctx(ctx), g, img, sino); });
}
int mi_backproject_f64(mi_context *ctx, const mi_geometry *g, const double *sino, double *img) {
    return guarded([&] { backproject_impl<double>(check_ctx(ctx), g, sino, img); });
}

int mi_forward_ordered(mi_context *ctx, const mi_geometry *g, int precision, const void *img, void *sino) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(precision == MI_FP32 || precision == MI_FP64, MI_E_ARG, "bad precision");
        if (precision == MI_FP64) ordered_impl<double>(ctx, g, (const double *)img, (double *)sino, true);
        else ordered_impl<float>(ctx, g, (const float *)img, (float *)sino, true);
    });
}
int mi_backproject_ordered(mi_context *ctx, const mi_geometry *g, int precision, const void *sino, void *img) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(precision == MI_FP32 || precision == MI_FP64, MI_E_ARG, "bad precision");
        if (precision == MI_FP64) ordered_impl<double>(ctx, g, (const double *)sino, (double *)img, false);
        else ordered_impl<float>(ctx, g, (const float *)sino, (float *)img, false);
    });
}

int mi_image_step_f64(mi_context *ctx, const mi_geometry *g, double *f_out, const double *sino, const double *f,
                      const double *v, double alpha, double epsilon, int iter_image) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && f_out && sino && f && v && iter_image >= 1, MI_E_ARG, "bad arguments");
        Matrix &m = *get_matrix(ctx, *g, 1, true);
        JobState<double> js(m);
        const hipStream_t st = ctx->stream;
        load_sino<double>(st, m, sino, js.g.p);
        MI_HIP(hipMemcpyAsync(js.u.p, f, js.npix * sizeof(double), hipMemcpyHostToDevice, st));
        MI_HIP(hipMemcpyAsync(js.v.p, v, js.npix * sizeof(double), hipMemcpyHostToDevice, st));
        gc_image_step_ordered(ctx, m, js, alpha, epsilon, iter_image, nullptr);
        MI_HIP(hipMemcpyAsync(f_out, js.u.p, js.npix * sizeof(double), hipMemcpyDeviceToHost, st));
        MI_HIP(hipStreamSynchronize(st));
    });
}

int mi_edge_step_f64(mi_context *ctx, int nx, int ny, double *v_out, const double *f, const double *v,
                     double alpha, double beta, double epsilon, int iter_edge) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(nx > 0 && ny > 0 && v_out && f && v && iter_edge >= 1, MI_E_ARG, "bad arguments");
        // An edge step needs no projector; a 1-view dummy matrix is not built.
        Matrix dummy;
        dummy.g.nx = nx;
        dummy.g.ny = ny;
        dummy.npix = nx * ny;
        dummy.nrows = 1;
        JobState<double> js(dummy);
        const hipStream_t st = ctx->stream;
        MI_HIP(hipMemcpyAsync(js.u.p, f, js.npix * sizeof(double), hipMemcpyHostToDevice, st));
        MI_HIP(hipMemcpyAsync(js.v.p, v, js.npix * sizeof(double), hipMemcpyHostToDevice, st));
        gc_edge_step_ordered(ctx, js, alpha, beta, epsilon, iter_edge);
        MI_HIP(hipMemcpyAsync(v_out, js.v.p, js.npix * sizeof(double), hipMemcpyDeviceToHost, st));
        MI_HIP(hipStreamSynchronize(st));
    });
}

int mi_medfilt2_f64(mi_context *ctx, int nx, int ny, double *f) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(nx > 0 && ny > 0 && f, MI_E_ARG, "bad arguments");
        const int64_t n = (int64_t)nx * ny;
        DevBuf<double> a(n), b(n);
        const hipStream_t st = ctx->stream;
        MI_HIP(hipMemcpyAsync(a.p, f, n * sizeof(double), hipMemcpyHostToDevice, st));
        launch_median<double>(st, nx, ny, a.p, b.p);
        MI_HIP(hipMemcpyAsync(f, b.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
        MI_HIP(hipStreamSynchronize(st));
    });
}

int mi_mumford_shah(mi_context *ctx, const mi_geometry *g, const mi_at_params *p, const double *sino, double *f,
                    double *v, mi_report *rep) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && p && sino && f && v, MI_E_ARG, "null argument");
        MI_REQUIRE(p->iter_alt >= 0 && p->iter_image >= 1 && p->iter_edge >= 1, MI_E_ARG, "bad iteration counts");
        MI_REQUIRE(p->order == MI_ORDER_REFERENCE || p->order == MI_ORDER_PARALLEL, MI_E_ARG, "bad order");
        const bool ref = p->order == MI_ORDER_REFERENCE;
        Matrix &m = *get_matrix(ctx, *g, 1, true);
        if (ref) get_seq(ctx->stream, m);
        JobState<double> js(m);
        const hipStream_t st = ctx->stream;
        load_sino<double>(st, m, sino, js.g.p);
        fill<double>(st, js.u.p, js.npix, 0.0);
        fill<double>(st, js.v.p, js.npix, 1.0);
        hipEvent_t e0, e1;
        MI_HIP(hipEventCreate(&e0));
        MI_HIP(hipEventCreate(&e1));
        MI_HIP(hipEventRecord(e0, st));
        for (int i = 1; i <= p->iter_alt; ++i) {
            if (ref) {
                gc_image_step_ordered(ctx, m, js, p->alpha, p->epsilon, p->iter_image, nullptr);
                gc_edge_step_ordered(ctx, js, p->alpha, p->beta, p->epsilon, p->iter_edge);
            } else {
                gc_image_step<double>(ctx, m, js, p->alpha, p->epsilon, p->iter_image, nullptr);
                gc_edge_step<double>(ctx, js, p->alpha, p->beta, p->epsilon, p->iter_edge);
            }
        }
        MI_HIP(hipEventRecord(e1, st));
        MI_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        MI_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        MI_HIP(hipMemcpyAsync(f, js.u.p, js.npix * sizeof(double), hipMemcpyDeviceToHost, st));
        MI_HIP(hipMemcpyAsync(v, js.v.p, js.npix * sizeof(double), hipMemcpyDeviceToHost, st));
        MI_HIP(hipStreamSynchronize(st));
        if (rep) {
            std::memset(rep, 0, sizeof(*rep));
            rep->iterations = p->iter_alt;
            rep->elapsed_ms = ms;
            rep->nnz = m.nnz;
        }
    });
}

int mi_at_reconstruct(mi_context *ctx, const mi_geometry *g, const mi_at_params *p, const float *sino, float *u,
                      float *v, mi_report *rep) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && p && sino && u && v, MI_E_ARG, "null argument");
        at_solve<float>(ctx, *g, *p, sino, u, v, rep);
    });
}

int mi_at_reconstruct_f64(mi_context *ctx, const mi_geometry *g, const mi_at_params *p, const double *sino,
                          double *u, double *v, mi_report *rep) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && p && sino && u && v, MI_E_ARG, "null argument");
        at_solve<double>(ctx, *g, *p, sino, u, v, rep);
    });
}

int mi_helical_reconstruct(mi_context *ctx, const char *prefix, int first, int count, const mi_geometry *g,
                           const mi_at_params *p, void *u_stack, void *v_stack, mi_report *reps) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(prefix && g && p && u_stack && v_stack && count >= 1, MI_E_ARG, "bad arguments");
        if (p->precision == MI_FP64)
            helical_slices<double>(ctx, prefix, first, count, *g, *p, (double *)u_stack, (double *)v_stack, reps);
        else
            helical_slices<float>(ctx, prefix, first, count, *g, *p, (float *)u_stack, (float *)v_stack, reps);
    });
}

int mi_job_create(mi_context *ctx, const mi_geometry *g, const mi_at_params *p, const void *sino, const void *u0,
                  const void *v0, mi_job **out) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && p && sino && out, MI_E_ARG, "null argument");
        MI_REQUIRE(p->n_subsets >= 1 && p->iter_edge >= 0, MI_E_ARG, "bad AT parameters");
        MI_REQUIRE(p->precision == MI_FP32 || p->precision == MI_FP64, MI_E_ARG, "bad precision");
        *out = nullptr;
        auto j = std::make_unique<mi_job>();
        j->ctx = ctx;
        j->precision = p->precision;
        j->rel_tol = p->rel_tol;
        mi_at_params q = *p;
        q.max_iters = 0;  // passes are counted by the caller
        if (p->precision == MI_FP64) {
            j->b64 = std::make_unique<Batch<double>>(ctx, p->n_subsets);
            j->id = single_job<double>(*j->b64, *g, q, sino, u0, v0);
        } else {
            j->b32 = std::make_unique<Batch<float>>(ctx, p->n_subsets);
            j->id = single_job<float>(*j->b32, *g, q, sino, u0, v0);
        }
        *out = j.release();
    });
}

int mi_job_run(mi_job *job, int max_passes, int *passes_run) {
    return guarded([&] {
        MI_REQUIRE(job && max_passes >= 0, MI_E_ARG, "bad arguments");
        check_ctx(job->ctx);
        const int every = job->rel_tol > 0 ? 4 : 0;
        const int n = job->b32 ? run_job<float>(job->ctx, *job->b32, job->id, max_passes, every)
                               : run_job<double>(job->ctx, *job->b64, job->id, max_passes, every);
        if (passes_run) *passes_run = n;
    });
}

int mi_job_enqueue(mi_job *job, int passes) {
    return guarded([&] {
        MI_REQUIRE(job && passes >= 0, MI_E_ARG, "bad arguments");
        check_ctx(job->ctx);
        if (job->b32) job->b32->enqueue(passes);
        else job->b64->enqueue(passes);
    });
}

int mi_job_synchronize(mi_job *job) {
    return guarded([&] {
        MI_REQUIRE(job, MI_E_ARG, "null job");
        check_ctx(job->ctx);
        MI_HIP(hipStreamSynchronize(job->ctx->stream));
        stats_collect(job->ctx);
    });
}

int mi_job_report(mi_job *job, mi_report *rep) {
    return guarded([&] {
        MI_REQUIRE(job && rep, MI_E_ARG, "null argument");
        check_ctx(job->ctx);
        if (job->b32) fill_report<float>(*job->b32, job->id, rep);
        else fill_report<double>(*job->b64, job->id, rep);
    });
}

int mi_job_read(mi_job *job, void *u, void *v) {
    return guarded([&] {
        MI_REQUIRE(job, MI_E_ARG, "null job");
        check_ctx(job->ctx);
        if (job->b32) job->b32->read(job->id, (float *)u, (float *)v);
        else job->b64->read(job->id, (double *)u, (double *)v);
    });
}

int mi_job_destroy(mi_job *job) {
    return guarded([&] {
        if (!job) return;
        (void)hipSetDevice(job->ctx->device);
        (void)hipStreamSynchronize(job->ctx->stream);
        delete job;
    });
}

// ---- batches -------------------------------------------------------------

int mi_batch_create(mi_context *ctx, int n_subsets, int precision, mi_batch **out) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(out && n_subsets >= 1, MI_E_ARG, "bad arguments");
        MI_REQUIRE(precision == MI_FP32 || precision == MI_FP64, MI_E_ARG, "bad precision");
        *out = nullptr;
        auto b = std::make_unique<mi_batch>();
        b->ctx = ctx;
        b->precision = precision;
        if (precision == MI_FP64) b->b64 = std::make_unique<Batch<double>>(ctx, n_subsets);
        else b->b32 = std::make_unique<Batch<float>>(ctx, n_subsets);
        *out = b.release();
    });
}

int mi_batch_destroy(mi_batch *b) {
    return guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->ctx->device);
        delete b;
    });
}

int mi_batch_set(mi_batch *b, int option, int64_t value) {
    return guarded([&] {
        MI_REQUIRE(b, MI_E_ARG, "null batch");
        auto apply = [&](auto &B) {
            switch (option) {
                case MI_BATCH_GRAPH:
                    B.use_graph = value != 0;
                    B.drop_graph();
                    break;
                case MI_BATCH_MIN_K:
                    MI_REQUIRE(value == 1 || value == 2 || value == 4 || value == 8, MI_E_ARG, "min K must be 1, 2, 4 or 8");
                    B.min_k = (int)value;
                    break;
                case MI_BATCH_RESERVE_MB:
                    MI_REQUIRE(value >= 0, MI_E_ARG, "negative reserve");
                    B.reserve_bytes = value << 20;
                    break;
                case MI_BATCH_BP_LDS:
                    B.use_bp_lds = value != 0;
                    B.dirty = true;
                    break;
                case MI_BATCH_BP_SQ:
                    MI_REQUIRE(value >= -1 && value <= 1, MI_E_ARG, "BP_SQ must be -1, 0 or 1");
                    B.bp_sq = (int)value;
                    B.dirty = true;
                    break;
                case MI_BATCH_MAX_ENTRIES:
                    MI_REQUIRE(value >= 0, MI_E_ARG, "negative entry budget");
                    B.max_entries = value;
                    break;
                case MI_BATCH_MAX_K:
                    MI_REQUIRE(value == 1 || value == 2 || value == 4 || value == 8, MI_E_ARG, "max K must be 1, 2, 4 or 8");
                    B.max_k = (int)value;
                    break;
                default:
                    fail(MI_E_ARG, "unknown batch option");
            }
        };
        if (b->b32) apply(*b->b32);
        else apply(*b->b64);
    });
}

int mi_batch_submit(mi_batch *b, const mi_geometry *g, const mi_at_params *p, const void *sino, const void *u0,
                    const void *v0, int64_t *job) {
    return guarded([&] {
        MI_REQUIRE(b && g && p && sino && job, MI_E_ARG, "null argument");
        check_ctx(b->ctx);
        *job = b->b32 ? b->b32->submit(*g, *p, (const float *)sino, (const float *)u0, (const float *)v0)
                      : b->b64->submit(*g, *p, (const double *)sino, (const double *)u0, (const double *)v0);
    });
}

int mi_batch_step(mi_batch *b, int passes, int *n_active, int *n_pending) {
    return guarded([&] {
        MI_REQUIRE(b && passes >= 0, MI_E_ARG, "bad arguments");
        check_ctx(b->ctx);
        const int a = b->b32 ? b->b32->step(passes) : b->b64->step(passes);
        if (n_active) *n_active = a;
        if (n_pending) *n_pending = b->b32 ? b->b32->n_pending() : b->b64->n_pending();
    });
}

int mi_batch_enqueue(mi_batch *b, int passes) {
    return guarded([&] {
        MI_REQUIRE(b && passes >= 0, MI_E_ARG, "bad arguments");
        check_ctx(b->ctx);
        if (b->b32) b->b32->enqueue(passes);
        else b->b64->enqueue(passes);
    });
}

int mi_batch_poll(mi_batch *b, int *n_active) {
    return guarded([&] {
        MI_REQUIRE(b, MI_E_ARG, "null batch");
        check_ctx(b->ctx);
        const int a = b->b32 ? b->b32->poll() : b->b64->poll();
        if (n_active) *n_active = a;
    });
}

int mi_batch_status(mi_batch *b, int64_t job, mi_job_status *out) {
    return guarded([&] {
        MI_REQUIRE(b && out, MI_E_ARG, "null argument");
        check_ctx(b->ctx);
        auto fillst = [&](auto &B) {
            auto &j = B.job(job);
            const Scalars h = B.scalars(job);
            std::memset(out, 0, sizeof(*out));
            out->state = j.state;
            out->iterations = h.iters;
            out->converged = h.converged;
            out->data_fidelity = h.df_prev;
            out->rel_change = h.rel;
            out->t_submit = j.t_submit;
            out->t_admit = j.t_admit;
            out->t_done = j.t_done;
            out->nnz = j.m ? j.m->nnz : 0;
            out->unit_k = j.state == BJ_ACTIVE ? B.units[j.uid]->K : 0;
        };
        if (b->b32) fillst(*b->b32);
        else fillst(*b->b64);
    });
}

int mi_batch_read(mi_batch *b, int64_t job, void *u, void *v) {
    return guarded([&] {
        MI_REQUIRE(b, MI_E_ARG, "null batch");
        check_ctx(b->ctx);
        if (b->b32) b->b32->read(job, (float *)u, (float *)v);
        else b->b64->read(job, (double *)u, (double *)v);
    });
}

int mi_batch_release(mi_batch *b, int64_t job) {
    return guarded([&] {
        MI_REQUIRE(b, MI_E_ARG, "null batch");
        check_ctx(b->ctx);
        if (b->b32) b->b32->release(job);
        else b->b64->release(job);
    });
}

int mi_batch_retired(mi_batch *b, int64_t *ids, int cap, int *n) {
    return guarded([&] {
        MI_REQUIRE(b && n && cap >= 0 && (cap == 0 || ids), MI_E_ARG, "bad arguments");
        auto drain = [&](auto &B) {
            auto &log = B.retired_log;
            const int take = std::min<int>(cap, (int)log.size());
            for (int i = 0; i < take; ++i) ids[i] = log[i];
            log.erase(log.begin(), log.begin() + take);
            *n = take;
        };
        if (b->b32) drain(*b->b32);
        else drain(*b->b64);
    });
}

int mi_batch_units(mi_batch *b, int32_t *k, int cap, int *n) {
    return guarded([&] {
        MI_REQUIRE(b && n && cap >= 0 && (cap == 0 || k), MI_E_ARG, "bad arguments");
        auto list = [&](auto &B) {
            *n = (int)B.live.size();
            for (int i = 0; i < *n && i < cap; ++i) k[i] = B.units[B.live[i]]->K;
        };
        if (b->b32) list(*b->b32);
        else list(*b->b64);
    });
}

int mi_device_copy(mi_context *ctx, int64_t bytes, int reps, double *ms) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(bytes > 0 && bytes % 16 == 0 && reps >= 1, MI_E_ARG, "bad copy size");
        DevBuf<unsigned char> a((size_t)bytes), b((size_t)bytes);
        const hipStream_t st = ctx->stream;
        MI_HIP(hipMemsetAsync(a.p, 1, bytes, st));
        hipEvent_t e0, e1;
        MI_HIP(hipEventCreate(&e0));
        MI_HIP(hipEventCreate(&e1));
        MI_HIP(hipEventRecord(e0, st));
        for (int i = 0; i < reps; ++i) eng_copy(st, b.p, a.p, bytes);
        MI_HIP(hipEventRecord(e1, st));
        MI_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        MI_HIP(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms) *ms = t / reps;
    });
}

int mi_stats_enable(mi_context *ctx, int on) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->stats.on = on != 0;
        ctx->stats.mask = ~0u;
    });
}

int mi_stats_select(mi_context *ctx, uint32_t mask) {
    return guarded([&] {
        check_ctx(ctx);
        ctx->stats.on = mask != 0;
        ctx->stats.mask = mask;
    });
}

int mi_stats_get(mi_context *ctx, int kernel, mi_kernel_stats *out) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(kernel >= 0 && kernel < MI_K_COUNT && out, MI_E_ARG, "bad kernel id");
        stats_collect(ctx);
        *out = ctx->stats.k[kernel];
    });
}

int mi_stats_reset(mi_context *ctx) {
    return guarded([&] {
        check_ctx(ctx);
        stats_collect(ctx);
        for (auto &k : ctx->stats.k) k = mi_kernel_stats{};
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reference configuration and data formats on the host side of the C ABI:
// the flat JSON configuration objects of the reference's drivers and the
// industry3d projection file (.dr).  No device work.

namespace mi {
namespace {

// Top-level members of a JSON object: numbers (as double) and strings;
// nested objects / arrays and literals are skipped.
struct JsonObject {
    std::map<std::string, double> num;
    std::map<std::string, std::string> str;
};

struct JsonReader {
    const char *p;
    void ws() {
        while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') ++p;
    }
    void expect(char c) {
        ws();
        MI_REQUIRE(*p == c, MI_E_ARG, std::string("configuration JSON: expected '") + c + "'");
        ++p;
    }
    std::string string() {
        expect('"');
        std::string s;
        while (*p && *p != '"') {
            if (*p == '\\') {
                ++p;
                MI_REQUIRE(*p, MI_E_ARG, "configuration JSON: unterminated string");
                const char e = *p;
                s += e == 'n' ? '\n' : e == 't' ? '\t' : e;   // \uXXXX is not used by these files
            } else {
                s += *p;
            }
            ++p;
        }
        MI_REQUIRE(*p == '"', MI_E_ARG, "configuration JSON: unterminated string");
        ++p;
        return s;
    }
    void skip_value() {
        ws();
        if (*p == '"') {
            string();
        } else if (*p == '{' || *p == '[') {
            const char open = *p, close = open == '{' ? '}' : ']';
            int depth = 0;
            do {
                if (*p == '"') {
                    string();
                    continue;
                }
                if (*p == open) ++depth;
                else if (*p == close) --depth;
                MI_REQUIRE(*p, MI_E_ARG, "configuration JSON: unterminated object or array");
                ++p;
            } while (depth > 0);
        } else {
            while (*p && *p != ',' && *p != '}' && *p != ']' && *p != ' ' && *p != '\n' && *p != '\r' && *p != '\t')
                ++p;
        }
    }
    JsonObject object() {
        JsonObject o;
        expect('{');
        ws();
        if (*p == '}') {
            ++p;
            return o;
        }
        for (;;) {
            const std::string key = string();
            expect(':');
            ws();
            if (*p == '"') {
                o.str[key] = string();
            } else if (*p == '-' || (*p >= '0' && *p <= '9')) {
                char *end = nullptr;
                const double v = std::strtod(p, &end);
                MI_REQUIRE(end != p, MI_E_ARG, "configuration JSON: bad number for " + key);
                o.num[key] = v;
                p = end;
            } else {
                skip_value();
            }
            ws();
            if (*p == ',') {
                ++p;
                continue;
            }
            expect('}');
            return o;
        }
    }
};

double need(const JsonObject &o, const char *k) {
    auto it = o.num.find(k);
    MI_REQUIRE(it != o.num.end(), MI_E_ARG, std::string("configuration JSON lacks the number ") + k);
    return it->second;
}
double opt(const JsonObject &o, const char *k, double d) {
    auto it = o.num.find(k);
    return it == o.num.end() ? d : it->second;
}

}  // namespace
}  // namespace mi


int mi_config_parse(const char *json, mi_geometry *g, mi_at_params *p) {
    return guarded([&] {
        MI_REQUIRE(json && g && p, MI_E_ARG, "null argument");
        JsonReader r{json};
        const JsonObject o = r.object();
        *g = mi_geometry{};
        *p = mi_at_params{};
        p->iter_edge = 1;
        p->n_subsets = 1;
        p->lower = -INFINITY;
        p->upper = INFINITY;
        p->precision = MI_FP32;
        if (o.num.count("NDETECTOR")) {
            // helical/rebin/src/main.cpp:85-106: a slice of the arc-detector
            // scanner; vx is read from "vy" and vy from "vx" there
            const double vx = need(o, "vy"), vy = need(o, "vx");
            MI_REQUIRE(vx == vy, MI_E_ARG, "the arc geometry needs square pixels (vx == vy)");
            g->beam = MI_BEAM_FAN_ARC;
            g->nx = (int)need(o, "NX");
            g->ny = (int)need(o, "NY");
            g->pixel_mm = vx;
            g->n_views = (int)need(o, "NPROJ");
            g->span_deg = 360.0;
            g->n_bins = (int)need(o, "NDETECTOR");
            g->bin_mm = need(o, "LENGTH_PER_DET");
            g->sod_mm = need(o, "SOURCE_TO_ISO");
            g->sdd_mm = need(o, "SOURCE_TO_DET");
            g->det_offset = need(o, "OFF_CENTER");
            p->alpha = need(o, "ALPHA");
            p->beta = need(o, "BETA");
            p->epsilon = need(o, "EPSILON");
            p->max_iters = p->iter_alt = (int)need(o, "MS_ITERATIONS");
        } else {
            // industry3d Parameter::parse_config, industry3d/src/ct3d.cpp:20-93;
            // the 2D path takes the mid-plane of the cone ("Cone") or the
            // parallel beam ("Parallel")
            auto b = o.str.find("BEAM");
            MI_REQUIRE(b != o.str.end(), MI_E_ARG, "configuration JSON lacks BEAM");
            const std::string beam = b->second;
            g->nx = (int)need(o, "NX");
            g->ny = (int)need(o, "NY");
            g->n_views = (int)need(o, "NPROJ");
            g->n_bins = (int)need(o, "NDX");
            if (beam == "Cone" || beam == "cone") {
                g->beam = MI_BEAM_FAN_FLAT;
                g->pixel_mm = need(o, "SAMPLESIZE");
                g->bin_mm = need(o, "PIXELSIZE");
                g->sod_mm = need(o, "SOD");
                g->sdd_mm = need(o, "SDD");
                g->span_deg = 360.0;
                g->det_offset = opt(o, "OFF_CENTER", 0.0);
            } else if (beam == "Parallel" || beam == "parallel") {
                g->beam = MI_BEAM_PARALLEL;
                g->pixel_mm = 1.0;
                g->bin_mm = 1.0;
                g->span_deg = 180.0;
            } else {
                fail(MI_E_ARG, "unknown BEAM " + beam);
            }
            p->alpha = need(o, "ALPHA");
            p->beta = need(o, "BETA");
            p->epsilon = need(o, "EPSILON");
            p->max_iters = p->iter_alt = (int)need(o, "ITERATIONS");
            p->lower = opt(o, "LOWER_X", -INFINITY);
            p->upper = opt(o, "UPPER_X", INFINITY);
        }
        validate_geometry(*g);
    });
}

int mi_read_pos(const char *path, double *pos_mm, double *angle_deg, int cap, int *n) {
    return guarded([&] {
        MI_REQUIRE(path && n && cap >= 0, MI_E_ARG, "bad arguments");
        FILE *fp = std::fopen(path, "r");
        MI_REQUIRE(fp, MI_E_ARG, std::string("cannot open table file ") + path);
        // load_table: fscanf "%lf,%lf\n" until a line does not parse
        double pos = 0, angle = 0;
        int rows = 0;
        while (std::fscanf(fp, "%lf,%lf\n", &pos, &angle) == 2) {
            if (rows < cap) {
                if (pos_mm) pos_mm[rows] = pos;
                if (angle_deg) angle_deg[rows] = angle;
            }
            ++rows;
        }
        std::fclose(fp);
        *n = rows;
    });
}

int mi_read_raw(const char *path, int64_t n_views, int64_t n_bins, double *sino) {
    return guarded([&] {
        MI_REQUIRE(path && sino && n_views > 0 && n_bins > 0, MI_E_ARG, "bad arguments");
        FILE *fp = std::fopen(path, "rb");
        MI_REQUIRE(fp, MI_E_ARG, std::string("cannot open sinogram file ") + path);
        // ms2d.cpp:270-274: whitespace-separated values, [view][bin].  The
        // file is read whole and parsed with from_chars (correctly rounded,
        // the values fscanf's %lf gives, several times faster); a token it
        // does not take (a leading '+', a value out of range) goes through
        // strtod, as fscanf would.
        std::vector<char> buf;
        {
            char chunk[1 << 16];
            size_t k;
            while ((k = std::fread(chunk, 1, sizeof chunk, fp)) > 0) buf.insert(buf.end(), chunk, chunk + k);
            std::fclose(fp);
        }
        buf.push_back('\0');
        const char *q = buf.data(), *end = buf.data() + buf.size() - 1;
        const int64_t want = n_views * n_bins;
        int64_t got = 0;
        while (got < want) {
            while (q < end && std::isspace((unsigned char)*q)) ++q;
            if (q == end) break;
            double x;
            const auto r = std::from_chars(q, end, x);
            if (r.ec == std::errc()) {
                q = r.ptr;
            } else {
                char *e = nullptr;
                x = std::strtod(q, &e);
                if (e == q) break;   // not a number
                q = e;
            }
            sino[got++] = x;
        }
        MI_REQUIRE(got == want, MI_E_ARG, std::string(path) + ": " + std::to_string(got) + " values, expected " +
                                              std::to_string(want));
    });
}

int mi_read_dr(const char *path, int64_t nproj, int64_t ndx, int64_t ndy, uint16_t *proj, float *flat) {
    return guarded([&] {
        MI_REQUIRE(path && proj && nproj > 0 && ndx > 0 && ndy > 0, MI_E_ARG, "bad .dr arguments");
        FILE *f = std::fopen(path, "rb");
        MI_REQUIRE(f, MI_E_ARG, std::string("cannot open ") + path);
        struct Closer {
            FILE *f;
            ~Closer() { std::fclose(f); }
        } closer{f};
        const int64_t n = nproj * ndx * ndy, nf = ndx * ndy;
        MI_REQUIRE(std::fseek(f, 0, SEEK_END) == 0, MI_E_ARG, "cannot size the .dr file");
        const int64_t size = (int64_t)std::ftell(f);
        const int64_t need_bytes = 1024 + n * 2 + (flat ? nf * 4 : 0);
        MI_REQUIRE(size >= need_bytes, MI_E_ARG,
                   ".dr file shorter than its 1024-byte header, the projections and the flat field asked for");
        // industry3d/src/ct3d.cpp:199-214: projections after a 1024-byte header
        MI_REQUIRE(std::fseek(f, 1024, SEEK_SET) == 0, MI_E_ARG, "cannot seek past the .dr header");
        MI_REQUIRE((int64_t)std::fread(proj, 2, (size_t)n, f) == n, MI_E_ARG, "short read of .dr projections");
        if (flat)
            MI_REQUIRE((int64_t)std::fread(flat, 4, (size_t)nf, f) == nf, MI_E_ARG, "short read of .dr flat field");
    });
}

int mi_cone_trace_ray(mi_context *ctx, const mi_cone_geometry *g, int view, int row, int col, int64_t *ind,
                      float *len, int cap, int *numb) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && numb && cap >= 0 && (cap == 0 || (ind && len)), MI_E_ARG, "null argument");
        ConeTables &t = cone_tables(ctx, *g);
        MI_REQUIRE(view >= 0 && view < g->n_views && row >= 0 && row < g->n_rows && col >= 0 && col < g->n_cols,
                   MI_E_ARG, "ray outside the geometry");
        DevBuf<int64_t> di(cap > 0 ? cap : 1);
        DevBuf<float> dl(cap > 0 ? cap : 1);
        DevBuf<int> dn(1);
        cone_trace(ctx->stream, t, view, row, col, di.p, dl.p, cap, dn.p);
        int n = 0;
        MI_HIP(hipMemcpyAsync(&n, dn.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        MI_HIP(hipStreamSynchronize(ctx->stream));
        const int k = n < cap ? n : cap;
        if (k > 0) {
            MI_HIP(hipMemcpyAsync(ind, di.p, k * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
            MI_HIP(hipMemcpyAsync(len, dl.p, k * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            MI_HIP(hipStreamSynchronize(ctx->stream));
        }
        *numb = n;
    });
}

int mi_cone_forward(mi_context *ctx, const mi_cone_geometry *g, const float *vol, float *sino) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && vol && sino, MI_E_ARG, "null argument");
        validate_cone(*g);
        DevBuf<float> dv(cone_voxels(*g)), ds(cone_rays(*g));
        MI_HIP(hipMemcpyAsync(dv.p, vol, dv.bytes(), hipMemcpyHostToDevice, ctx->stream));
        cone_forward_dev(ctx, g, dv.p, ds.p);
        MI_HIP(hipMemcpyAsync(sino, ds.p, ds.bytes(), hipMemcpyDeviceToHost, ctx->stream));
        MI_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int mi_cone_forward_dev(mi_context *ctx, const mi_cone_geometry *g, const float *vol, float *sino) {
    return guarded([&] { cone_forward_dev(check_ctx(ctx), g, vol, sino); });
}

int mi_cone_backproject(mi_context *ctx, const mi_cone_geometry *g, const float *sino, float *vol) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && vol && sino, MI_E_ARG, "null argument");
        validate_cone(*g);
        DevBuf<float> dv(cone_voxels(*g)), ds(cone_rays(*g));
        MI_HIP(hipMemcpyAsync(ds.p, sino, ds.bytes(), hipMemcpyHostToDevice, ctx->stream));
        cone_backproject_dev(ctx, g, ds.p, dv.p);
        MI_HIP(hipMemcpyAsync(vol, dv.p, dv.bytes(), hipMemcpyDeviceToHost, ctx->stream));
        MI_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int mi_cone_backproject_dev(mi_context *ctx, const mi_cone_geometry *g, const float *sino, float *vol) {
    return guarded([&] { cone_backproject_dev(check_ctx(ctx), g, sino, vol); });
}

int mi_cone_backproject_dev_noalloc(mi_context *ctx, const mi_cone_geometry *g, const float *sino, float *vol) {
    return guarded([&] { cone_backproject_dev_noalloc(check_ctx(ctx), g, sino, vol); });
}


int mi_cone_at_reconstruct(mi_context *ctx, const mi_cone_geometry *g, const mi_at_params *p, const float *sino,
                           float *u, float *v, mi_report *rep) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && p, MI_E_ARG, "null argument");
        cone_at_solve(ctx, *g, *p, sino, u, v, rep);
    });
}

int mi_cone_at_reconstruct_dev(mi_context *ctx, const mi_cone_geometry *g, const mi_at_params *p,
                               const float *sino, float *u, float *v, mi_report *rep) {
    return guarded([&] {
        check_ctx(ctx);
        MI_REQUIRE(g && p, MI_E_ARG, "null argument");
        cone_at_solve(ctx, *g, *p, sino, u, v, rep, true);
    });
}

int mi_cone_config_parse(const char *json, mi_cone_geometry *g, mi_at_params *p) {
    return guarded([&] {
        MI_REQUIRE(json && g && p, MI_E_ARG, "null argument");
        JsonReader r{json};
        const JsonObject o = r.object();
        auto b = o.str.find("BEAM");
        MI_REQUIRE(b != o.str.end(), MI_E_ARG, "configuration JSON lacks BEAM");
        const bool cone = b->second == "Cone" || b->second == "cone";
        const bool par = b->second == "Parallel" || b->second == "parallel";
        MI_REQUIRE(cone || par, MI_E_ARG, "unknown BEAM \"" + b->second + "\"");
        *g = mi_cone_geometry{};
        g->beam = cone ? MI_CONE_BEAM : MI_CONE_PARALLEL;
        g->nx = (int)need(o, "NX");
        g->ny = (int)need(o, "NY");
        g->nz = (int)need(o, "NZ");
        g->n_views = (int)need(o, "NPROJ");
        g->n_rows = (int)need(o, "NDX");
        g->n_cols = (int)need(o, "NDY");
        g->sample_mm = (float)need(o, "SAMPLESIZE");
        g->pixel_mm = (float)need(o, "PIXELSIZE");
        g->sod_mm = (float)need(o, "SOD");
        g->sdd_mm = (float)need(o, "SDD");
        *p = mi_at_params{};
        p->alpha = need(o, "ALPHA");
        p->beta = need(o, "BETA");
        p->epsilon = need(o, "EPSILON");
        p->max_iters = p->iter_alt = (int)need(o, "ITERATIONS");
        p->iter_edge = 1;
        p->n_subsets = 1;
        p->lower = opt(o, "LOWER_X", -INFINITY);
        p->upper = opt(o, "UPPER_X", INFINITY);
        p->precision = MI_FP32;
        validate_cone(*g);
    });
}
