// Batch runtime of the ordered-subset AT engine: unit packing, admission by
// free device memory, pass recording / hipGraph replay, retirement.
#include "batch.hpp"
#include "solver.hpp"

#include <algorithm>
#include <cstring>

namespace mi {

namespace {

int pow2ceil(int n) {
    int k = 1;
    while (k < n) k <<= 1;
    return k;
}

template <typename T>
int64_t unit_state_bytes(const Matrix &m, int K) {
    return (int64_t)K * (2 * m.nrows + 4 * m.npix_t) * (int64_t)sizeof(T);
}

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Blocks of the projector launches: four slices each.
int64_t max_slices(const Sell &q) {
    int64_t r = 0;
    for (size_t s = 0; s + 1 < q.slice0.size(); ++s) r = std::max(r, q.slice0[s + 1] - q.slice0[s]);
    return r;
}
int64_t fp_blocks(const Matrix &m) {
    if (m.has_otf) {
        int64_t r = 0;
        for (int s = 0; s < m.S; ++s) r = std::max(r, m.otf.slice0[s + 1] - m.otf.slice0[s]);
        return cdiv(r, kBlock / kWave);
    }
    return cdiv(max_slices(m.fs), kBlock / kWave);
}
int64_t bp_blocks(const Matrix &m) { return cdiv(max_slices(m.bs), kBlock / kWave); }
int64_t px_blocks(const Matrix &m) { return cdiv(m.npix_t, kPixBlock); }

// Partials per job per slot of a unit over `m`.
int64_t part_cap(const Matrix &m) {
    return std::max({fp_blocks(m), bp_blocks(m), px_blocks(m), (int64_t)1});
}

int64_t free_device_bytes() {
    size_t fr = 0, tot = 0;
    MI_HIP(hipMemGetInfo(&fr, &tot));
    return (int64_t)fr;
}

}  // namespace

template <typename T>
Batch<T>::Batch(mi_context *c, int n_subsets) : ctx(c), S(n_subsets) {
    MI_REQUIRE(S >= 1, MI_E_ARG, "n_subsets must be >= 1");
    units.resize(kMaxUnits);
    d_units.alloc(kMaxUnits);
    d_jp.alloc((size_t)kMaxUnits * kMaxK);
    d_sc.alloc((size_t)kMaxUnits * kMaxK);
    h_jp.assign((size_t)kMaxUnits * kMaxK, JobParams{});
    h_sc.assign((size_t)kMaxUnits * kMaxK, Scalars{});
    MI_HIP(hipMemsetAsync(d_sc.p, 0, d_sc.bytes(), ctx->stream));
    MI_HIP(hipMemcpyAsync(d_jp.p, h_jp.data(), d_jp.bytes(), hipMemcpyHostToDevice, ctx->stream));
    MI_HIP(hipStreamSynchronize(ctx->stream));
    MI_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    MI_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    MI_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    t0 = std::chrono::steady_clock::now();
}

template <typename T>
Batch<T>::~Batch() {
    (void)hipStreamSynchronize(ctx->stream);
    drop_graph();
    if (side) {
        (void)hipStreamSynchronize(side);
        (void)hipStreamDestroy(side);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
}

template <typename T>
double Batch<T>::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

template <typename T>
int64_t Batch<T>::submit(const mi_geometry &g, const mi_at_params &p, const T *sino, const T *u0, const T *v0) {
    validate_geometry(g);
    MI_REQUIRE(sino != nullptr, MI_E_ARG, "null sinogram");
    MI_REQUIRE(p.n_subsets == 0 || p.n_subsets == S, MI_E_ARG,
               "job n_subsets must equal the batch's (" + std::to_string(S) + ")");
    MI_REQUIRE(S <= g.n_views, MI_E_ARG, "more subsets than views");
    MI_REQUIRE(p.iter_edge >= 0 && p.max_iters >= 0 && p.median_every >= 0, MI_E_ARG, "bad AT parameters");
    MI_REQUIRE((p.precision == MI_FP64) == (sizeof(T) == 8), MI_E_ARG, "job precision differs from the batch's");
    MI_REQUIRE(p.epsilon > 0, MI_E_ARG, "epsilon must be positive");
    BatchJob<T> j;
    j.id = next_id++;
    j.geom = g;
    j.views = view_table(g);
    j.geom.view_deg = nullptr;
    j.p = p;
    j.p.n_subsets = S;
    const int64_t nrows = (int64_t)g.n_views * g.n_bins, npix = (int64_t)g.nx * g.ny;
    j.sino.assign(sino, sino + nrows);
    if (u0) j.u0.assign(u0, u0 + npix);
    if (v0) j.v0.assign(v0, v0 + npix);
    j.t_submit = now();
    const int64_t id = j.id;
    jobs.emplace(id, std::move(j));
    pending.push_back(id);
    return id;
}

template <typename T>
BatchJob<T> &Batch<T>::job(int64_t id) {
    auto it = jobs.find(id);
    MI_REQUIRE(it != jobs.end(), MI_E_ARG, "unknown job id " + std::to_string(id));
    return it->second;
}

template <typename T>
int Batch<T>::n_active() const {
    int n = 0;
    for (int uid : live) n += units[uid]->used();
    return n;
}

template <typename T>
int64_t Batch<T>::active_entries() const {
    // one stream of a unit's matrix serves all of the unit's jobs
    int64_t e = 0;
    for (int uid : live) e += units[uid]->m->nnz;
    return e;
}

template <typename T>
bool Batch<T>::has_free_slot(const Matrix *m) const {
    for (int uid : live)
        if (units[uid]->m.get() == m && units[uid]->used() < units[uid]->K) return true;
    return false;
}

template <typename T>
void Batch<T>::ensure_stage(int64_t n) {
    if ((int64_t)stage.n >= n) return;
    MI_HIP(hipStreamSynchronize(ctx->stream));
    stage.alloc((size_t)n);
}

template <typename T>
Unit<T> *Batch<T>::place(BatchJob<T> &j, int group) {
    for (int uid : live) {
        Unit<T> *U = units[uid].get();
        if (U->m != j.m || U->used() >= U->K) continue;
        for (int k = 0; k < U->K; ++k)
            if (U->jobs[k] < 0) {
                U->jobs[k] = j.id;
                j.uid = uid;
                j.slot = k;
                return U;
            }
    }
    int uid = -1;
    for (int i = 0; i < kMaxUnits; ++i)
        if (!units[i]) { uid = i; break; }
    MI_REQUIRE(uid >= 0, MI_E_LIMIT, "batch holds the maximum number of units");
    const int K = std::min(k_cap(*j.m), pow2ceil(std::min(kMaxK, std::max(std::max(group, min_k), 1))));
    const Matrix &m = *j.m;
    const int64_t need = unit_state_bytes<T>(m, K);
    if (free_device_bytes() - need < reserve_bytes) evict_idle(ctx, need + reserve_bytes - free_device_bytes());
    MI_REQUIRE(free_device_bytes() - need >= reserve_bytes, MI_E_NOMEM, "not enough free device memory for a unit");
    auto U = std::make_unique<Unit<T>>();
    U->m = j.m;
    U->uid = uid;
    U->K = K;
    for (auto &x : U->jobs) x = -1;
    const hipStream_t st = ctx->stream;
    U->g.alloc((size_t)m.nrows * K);
    U->r.alloc((size_t)m.nrows * K);
    U->u.alloc((size_t)m.npix_t * K);
    U->v.alloc((size_t)m.npix_t * K);
    U->d.alloc((size_t)m.npix_t * K);
    U->w.alloc((size_t)m.npix_t * K);
    for (auto *b : {&U->g, &U->r, &U->u, &U->v, &U->d, &U->w}) MI_HIP(hipMemsetAsync(b->p, 0, b->bytes(), st));
    U->jobs[0] = j.id;
    j.uid = uid;
    j.slot = 0;
    units[uid] = std::move(U);
    live.push_back(uid);
    std::sort(live.begin(), live.end());
    return units[uid].get();
}

template <typename T>
void Batch<T>::load_job(BatchJob<T> &j) {
    Unit<T> &U = *units[j.uid];
    const Matrix &m = *U.m;
    const hipStream_t st = ctx->stream;
    const int js = j.uid * kMaxK + j.slot;
    ensure_stage(std::max<int64_t>(m.nrows, m.npix));
    MI_HIP(hipMemcpyAsync(stage.p, j.sino.data(), m.nrows * sizeof(T), hipMemcpyHostToDevice, st));
    eng_put_rows<T>(st, U.g.p, U.K, j.slot, m.nrows, m.perm.p, stage.p);
    MI_HIP(hipStreamSynchronize(st));
    const int nx = m.g.nx, ny = m.g.ny;
    if (!j.u0.empty()) {
        MI_HIP(hipMemcpyAsync(stage.p, j.u0.data(), m.npix * sizeof(T), hipMemcpyHostToDevice, st));
        eng_put_img<T>(st, U.u.p, U.K, j.slot, nx, ny, m.tny, stage.p, T(0));
        MI_HIP(hipStreamSynchronize(st));
    } else {
        eng_put_img<T>(st, U.u.p, U.K, j.slot, nx, ny, m.tny, nullptr, T(0));
    }
    if (!j.v0.empty()) {
        MI_HIP(hipMemcpyAsync(stage.p, j.v0.data(), m.npix * sizeof(T), hipMemcpyHostToDevice, st));
        eng_put_img<T>(st, U.v.p, U.K, j.slot, nx, ny, m.tny, stage.p, T(1));
        MI_HIP(hipStreamSynchronize(st));
    } else {
        eng_put_img<T>(st, U.v.p, U.K, j.slot, nx, ny, m.tny, nullptr, T(1));
    }
    JobParams &P = h_jp[js];
    P.alpha = j.p.alpha;
    P.beta = j.p.beta;
    P.eps = j.p.epsilon;
    P.lower = j.p.lower;
    P.upper = j.p.upper;
    P.rel_tol = j.p.rel_tol;
    P.iter_edge = j.p.iter_edge;
    P.median_every = j.p.median_every;
    P.max_iters = j.p.max_iters;
    P.active = 1;
    MI_HIP(hipMemsetAsync(d_sc.p + js, 0, sizeof(Scalars), st));
    MI_HIP(hipStreamSynchronize(st));
    std::vector<T>().swap(j.sino);
    std::vector<T>().swap(j.u0);
    std::vector<T>().swap(j.v0);
    j.state = BJ_ACTIVE;
    j.t_admit = now();
    dirty = true;
}

template <typename T>
int Batch<T>::admit(bool require) {
    int n = 0;
    // Arrival order, until a job that needs a new unit does not fit (the
    // entry budget or memory); after that only jobs that fill a free slot
    // of a unit already on the device join: they share its matrix stream
    // and its allocated state, so they cost the others little.
    bool blocked = false;
    for (auto it = pending.begin(); it != pending.end();) {
        const int64_t id = *it;
        BatchJob<T> &j = job(id);
        if (blocked) {
            if (!j.m) j.m = find_engine_matrix(ctx, j.geometry(), S, sizeof(T) == 8);
            if (!j.m || !has_free_slot(j.m.get())) {
                ++it;
                continue;
            }
        }
        try {
            if (!j.m) {
                try {
                    j.m = get_engine_matrix(ctx, j.geometry(), S, sizeof(T) == 8);
                } catch (const Error &e) {
                    if (e.code != MI_E_NOMEM) throw;
                    (void)hipGetLastError();
                    if (evict_idle(ctx, -1) <= 0) throw;
                    j.m = get_engine_matrix(ctx, j.geometry(), S, sizeof(T) == 8);
                }
            }
            j.m->last_use = ++ctx->tick;
            // the budget keeps the device as full as it needs to be and lets
            // later jobs wait in arrival order instead of sharing it
            if (!has_free_slot(j.m.get()) && max_entries > 0 && n_active() > 0 &&
                active_entries() + j.m->nnz > max_entries) {
                blocked = true;
                ++it;
                continue;
            }
            const GeomKey key = make_key(j.geometry(), S);
            int group = 0;
            for (int64_t q : pending) {
                const GeomKey kq = make_key(jobs.at(q).geometry(), S);
                if (!(kq < key) && !(key < kq)) ++group;
            }
            place(j, group);
            load_job(j);
        } catch (const Error &e) {
            (void)hipGetLastError();
            if (j.uid >= 0) free_slot(j);
            const bool wait = !require && (e.code == MI_E_NOMEM || e.code == MI_E_LIMIT) && n_active() > 0;
            if (wait) {
                blocked = true;
                ++it;
                continue;
            }
            if (require) {
                pending.erase(it);
                jobs.erase(id);
                throw;
            }
            j.state = BJ_FAILED;
            j.error = e.what();
            it = pending.erase(it);
            continue;
        }
        it = pending.erase(it);
        ++n;
    }
    if (n) compact();
    return n;
}

// Repacks the jobs of each matrix into units of 8 plus one unit of the
// next power of two above the remainder, when that needs fewer slots than
// the current units.  Job state (g, u, v, scalars) moves with the job; r, d
// and w are scratch that every pass rewrites.
template <typename T>
void Batch<T>::compact() {
    std::map<const Matrix *, std::vector<int>> by_m;
    for (int uid : live) by_m[units[uid]->m.get()].push_back(uid);
    const hipStream_t st = ctx->stream;
    for (auto &kv : by_m) {
        auto &uids = kv.second;
        int n = 0, cap = 0;
        for (int uid : uids) {
            n += units[uid]->used();
            cap += units[uid]->K;
        }
        if (n == 0) continue;
        std::vector<int> target;
        const int kc = k_cap(*kv.first);
        for (int left = n; left > 0; left -= kc) target.push_back(pow2ceil(std::min(left, kc)));
        int tcap = 0;
        for (int k : target) tcap += std::min(kc, std::max(k, min_k));
        if (tcap >= cap && target.size() >= uids.size()) continue;
        // Collect the jobs in slot order, then rebuild the units.
        std::vector<int64_t> ids;
        for (int uid : uids)
            for (int k = 0; k < units[uid]->K; ++k)
                if (units[uid]->jobs[k] >= 0) ids.push_back(units[uid]->jobs[k]);
        std::vector<std::unique_ptr<Unit<T>>> old;
        for (int uid : uids) {
            old.push_back(std::move(units[uid]));
            live.erase(std::find(live.begin(), live.end(), uid));
        }
        std::shared_ptr<Matrix> mp = old.front()->m;
        size_t next = 0;
        std::vector<JobParams> jp_old = h_jp;
        // Scalars are copied from a snapshot: a new unit may reuse an old uid.
        DevBuf<Scalars> sc_old(d_sc.n);
        MI_HIP(hipMemcpyAsync(sc_old.p, d_sc.p, d_sc.bytes(), hipMemcpyDeviceToDevice, st));
        for (int K : target) {
            K = std::min(kc, pow2ceil(std::max(K, min_k)));
            int uid = -1;
            for (int i = 0; i < kMaxUnits; ++i)
                if (!units[i]) { uid = i; break; }
            auto U = std::make_unique<Unit<T>>();
            U->m = mp;
            U->uid = uid;
            U->K = K;
            for (auto &x : U->jobs) x = -1;
            U->g.alloc((size_t)mp->nrows * K);
            U->r.alloc((size_t)mp->nrows * K);
            U->u.alloc((size_t)mp->npix_t * K);
            U->v.alloc((size_t)mp->npix_t * K);
            U->d.alloc((size_t)mp->npix_t * K);
            U->w.alloc((size_t)mp->npix_t * K);
            for (auto *b : {&U->g, &U->r, &U->u, &U->v, &U->d, &U->w}) MI_HIP(hipMemsetAsync(b->p, 0, b->bytes(), st));
            for (int k = 0; k < K && next < ids.size(); ++k, ++next) {
                BatchJob<T> &j = job(ids[next]);
                Unit<T> *src = nullptr;
                for (auto &o : old)
                    if (o->uid == j.uid) src = o.get();
                const int from = j.uid * kMaxK + j.slot, to = uid * kMaxK + k;
                eng_move<T>(st, U->g.p, K, k, src->g.p, src->K, j.slot, mp->nrows);
                eng_move<T>(st, U->u.p, K, k, src->u.p, src->K, j.slot, mp->npix_t);
                eng_move<T>(st, U->v.p, K, k, src->v.p, src->K, j.slot, mp->npix_t);
                MI_HIP(hipMemcpyAsync(d_sc.p + to, sc_old.p + from, sizeof(Scalars), hipMemcpyDeviceToDevice, st));
                h_jp[to] = jp_old[from];
                U->jobs[k] = j.id;
                j.uid = uid;
                j.slot = k;
            }
            units[uid] = std::move(U);
            live.push_back(uid);
        }
        MI_HIP(hipStreamSynchronize(st));
        old.clear();
        std::sort(live.begin(), live.end());
    }
    // A slot is active exactly when a unit holds a job in it.
    for (int uid = 0; uid < kMaxUnits; ++uid)
        for (int k = 0; k < kMaxK; ++k)
            if (!units[uid] || k >= units[uid]->K || units[uid]->jobs[k] < 0) h_jp[uid * kMaxK + k].active = 0;
    dirty = true;
}

template <typename T>
void Batch<T>::drop_graph() {
    // no replay of the graph may still be running when it is destroyed
    if (gexec) (void)hipStreamSynchronize(ctx->stream);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
}

template <typename T>
void Batch<T>::rebuild() {
    const hipStream_t st = ctx->stream;
    drop_graph();
    MI_REQUIRE((int)live.size() <= kMaxUnits, MI_E_LIMIT, "too many units");
    std::vector<UnitDesc<T>> desc(live.size());
    int64_t off = 0;
    kmax = 1;
    all_walk = true;
    edge_iters = 0;
    any_median = false;
    pix.nunits = (int)live.size();
    pix.start[0] = 0;
    for (Launch *L : {&rows[0], &rows[1], &bpl[0][0], &bpl[0][1], &bpl[1][0], &bpl[1][1], &bpg[0], &bpg[1]}) {
        *L = Launch{};
        L->map.nunits = (int)live.size();
        L->map.start[0] = 0;
    }
    const bool f64 = sizeof(T) == 8;
    for (size_t i = 0; i < live.size(); ++i) {
        Unit<T> &U = *units[live[i]];
        const Matrix &m = *U.m;
        UnitDesc<T> &D = desc[i];
        std::memset(&D, 0, sizeof(D));
        D.fs_slice0 = m.has_otf ? m.otf.d_slice0.p : m.fs.d_slice0.p;
        if (m.has_otf) {
            D.ot_lam = m.otf.lam.p;
            D.ot_ax = m.otf.ax.p;
            D.ot_ay = m.otf.ay.p;
            D.ot_stx = m.otf.stx.p;
            D.ot_sty = m.otf.sty.p;
            D.ot_cx = m.otf.cx.p;
            D.ot_cy = m.otf.cy.p;
            D.ot_sg = m.otf.sg.p;
            D.ot_cnt = m.otf.cnt.p;
        }
        D.fs_base = m.fs.base.p;
        D.fs_T = m.fs.T.p;
        D.fs_idx = m.fs.idx.p;
        D.fs_first = m.fs.first.p;
        D.fs_d16 = m.fs.d16.p;
        D.fs_len = f64 ? (const T *)m.fs.l64.p : (const T *)m.fs.l32.p;
        D.bs_slice0 = m.bs.d_slice0.p;
        D.bs_base = m.bs.base.p;
        D.bs_T = m.bs.T.p;
        D.bs_idx = m.bs.idx.p;
        D.bs_first = m.bs.first.p;
        D.bs_d16 = m.bs.d16.p;
        D.bs_len = f64 ? (const T *)m.bs.l64.p : (const T *)m.bs.l32.p;
        D.bs_slot16 = m.bs.slot16.p;
        D.bs_vlo = m.bs.vlo.p;
        D.bs_voff = m.bs.voff.p;
        D.bs_vtab0 = m.bs.d_vtab0.p;
        D.nbins = m.g.n_bins;
        const size_t need = m.bs.slot16.p ? (size_t)m.bs.max_slots * U.K * sizeof(T) : kBpLdsHw + 1;
        const bool in_lds = bp_in_lds(m, need);
        const int rc = U.K > 1 ? 1 : 0, bc = U.K > 4 ? 1 : 0;   // the unit's classes
        const int lc = need > kBpLdsSplit ? 1 : 0;
        Launch &B = in_lds ? bpl[bc][lc] : bpg[bc];
        B.kmax = std::max(B.kmax, U.K);
        if (in_lds) B.lds = std::max(B.lds, need);
        rows[rc].kmax = std::max(rows[rc].kmax, U.K);
        D.sub_row0 = m.d_sub_row0.p;
        D.g = U.g.p;
        D.r = U.r.p;
        D.u = U.u.p;
        D.v = U.v.p;
        D.d = U.d.p;
        D.w = U.w.p;
        D.npix_t = m.npix_t;
        D.nrows = m.nrows;
        D.cap = part_cap(m);
        D.part_off = off;
        off += (int64_t)SL_COUNT * U.K * D.cap;
        D.nx = m.g.nx;
        D.ny = m.g.ny;
        D.tny = m.tny;
        D.K = U.K;
        D.fs_P = m.fs.P;
        D.bs_P = m.bs.P;
        D.job0 = U.uid * kMaxK;
        MI_REQUIRE(m.has_sell && (!f64 || m.has64), MI_E_STATE, "matrix lacks the engine's sliced layout");
        for (int c = 0; c < 2; ++c) {
            rows[c].map.start[i + 1] = rows[c].map.start[i] + (c == rc ? (int)fp_blocks(m) : 0);
            for (int l = 0; l < 2; ++l)
                bpl[c][l].map.start[i + 1] =
                    bpl[c][l].map.start[i] + (in_lds && c == bc && l == lc ? (int)bp_blocks(m) : 0);
            bpg[c].map.start[i + 1] = bpg[c].map.start[i] + (!in_lds && c == bc ? (int)bp_blocks(m) : 0);
        }
        pix.start[i + 1] = pix.start[i] + (int)px_blocks(m);
        kmax = std::max(kmax, U.K);
        all_walk = all_walk && m.has_otf;
        for (int k = 0; k < U.K; ++k) {
            if (U.jobs[k] < 0) continue;
            const JobParams &P = h_jp[D.job0 + k];
            edge_iters = std::max(edge_iters, (int)P.iter_edge);
            any_median = any_median || P.median_every > 0;
        }
    }
    MI_HIP(hipStreamSynchronize(st));
    if ((int64_t)d_part.n < off) d_part.alloc((size_t)std::max<int64_t>(off, 1));
    if (!desc.empty())
        MI_HIP(hipMemcpyAsync(d_units.p, desc.data(), desc.size() * sizeof(UnitDesc<T>), hipMemcpyHostToDevice, st));
    MI_HIP(hipMemcpyAsync(d_jp.p, h_jp.data(), d_jp.bytes(), hipMemcpyHostToDevice, st));
    MI_HIP(hipStreamSynchronize(st));
    args.units = d_units.p;
    args.jp = d_jp.p;
    args.sc = d_sc.p;
    args.part = d_part.p;
    args.nunits = (int)live.size();
    args.S = S;
    dirty = false;
}

// Algorithmic bytes of one launch summed over the live units (instrumented
// runs only).
template <typename T>
void Batch<T>::record_pass(bool timed) {
    const hipStream_t st = ctx->stream;
    const double sz = (double)sizeof(T);
    auto sum = [&](auto f) {
        double b = 0;
        if (!timed) return b;
        for (int uid : live) b += f(*units[uid]);
        return b;
    };
    // The two classes of a phase: class 1 on the side stream, class 0 on
    // the context's stream, joined before the next phase.  Timed passes run
    // on one stream (the per-kernel events time whole phases).
    auto branches = [&](auto &&launch0, auto &&launch1, bool has0, bool has1) {
        if (has0 && has1 && !timed) {
            MI_HIP(hipEventRecord(ev_fork, st));
            MI_HIP(hipStreamWaitEvent(side, ev_fork, 0));
            launch1(side);
            launch0(st);
            MI_HIP(hipEventRecord(ev_join, side));
            MI_HIP(hipStreamWaitEvent(st, ev_join, 0));
        } else {
            if (has0) launch0(st);
            if (has1) launch1(st);
        }
    };
    eng_finalize<T>(st, args, EF_BEGIN, 0);
    for (int s = 0; s < S; ++s) {
        auto nnz_s = [&](const Unit<T> &U) { return (double)U.m->sub_nnz[s]; };
        auto nr_s = [&](const Unit<T> &U) { return (double)(U.m->sub_row0[s + 1] - U.m->sub_row0[s]); };
        {
            LaunchTimer lt(ctx, MI_K_FP, sum([&](const Unit<T> &U) {
                return nnz_s(U) * (4 + sz) + U.m->npix * sz * U.K + 2 * nr_s(U) * sz * U.K;
            }), true);
            auto one = [&](int c) {
                return [&, c](hipStream_t q) { eng_fp_resid<T>(q, args, rows[c].map, s, rows[c].kmax, all_walk); };
            };
            branches(one(0), one(1), rows[0].any(), rows[1].any());
        }
        {
            LaunchTimer lt(ctx, MI_K_BP, sum([&](const Unit<T> &U) {
                return nnz_s(U) * (4 + sz) + nr_s(U) * sz * U.K + 3.0 * U.m->npix * sz * U.K;
            }), true);
            auto one = [&](int c) {
                return [&, c](hipStream_t q) {
                    for (const Launch &L : bpl[c])
                        if (L.any()) eng_bp_grad<T>(q, args, L.map, s, L.kmax, L.lds, bp_sq);
                    if (bpg[c].any()) eng_bp_grad<T>(q, args, bpg[c].map, s, bpg[c].kmax, 0);
                };
            };
            branches(one(0), one(1), bpl[0][0].any() || bpl[0][1].any() || bpg[0].any(),
                     bpl[1][0].any() || bpl[1][1].any() || bpg[1].any());
        }
        {
            LaunchTimer lt(ctx, MI_K_FPSQ, sum([&](const Unit<T> &U) {
                return nnz_s(U) * (4 + sz) + U.m->npix * sz * U.K;
            }), true);
            auto one = [&](int c) {
                return [&, c](hipStream_t q) { eng_fp_sq<T>(q, args, rows[c].map, s, rows[c].kmax, all_walk); };
            };
            branches(one(0), one(1), rows[0].any(), rows[1].any());
        }
        {
            LaunchTimer lt(ctx, MI_K_STEP, sum([&](const Unit<T> &U) { return 5.0 * U.m->npix * sz * U.K; }));
            eng_nv<T>(st, args, pix, kmax);
            eng_finalize<T>(st, args, EF_STEP, s);
            eng_axpy<T>(st, args, pix, kmax);
        }
    }
    if (any_median) eng_median<T>(st, args, pix, kmax);
    if (edge_iters > 0) {
        LaunchTimer lt(ctx, MI_K_EDGE, sum([&](const Unit<T> &U) { return 7.0 * U.m->npix * sz * U.K * edge_iters; }));
        for (int it = 0; it < edge_iters; ++it) eng_edge_iter<T>(st, args, pix, it, kmax);
    }
    eng_finalize<T>(st, args, EF_END, 0);
}

template <typename T>
void Batch<T>::enqueue(int passes) {
    MI_REQUIRE(passes >= 0, MI_E_ARG, "negative pass count");
    if (live.empty() || passes == 0) return;
    if (dirty) rebuild();
    const hipStream_t st = ctx->stream;
    const bool timed = ctx->stats.on;
    if (use_graph && !timed) {
        if (!gexec) {
            MI_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                record_pass(false);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            MI_HIP(hipStreamEndCapture(st, &graph));
            MI_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        }
        for (int i = 0; i < passes; ++i) MI_HIP(hipGraphLaunch(gexec, st));
    } else {
        for (int i = 0; i < passes; ++i) record_pass(timed);
    }
    passes_run += passes;
}

template <typename T>
void Batch<T>::retire(BatchJob<T> &j) {
    Unit<T> &U = *units[j.uid];
    const Matrix &m = *U.m;
    const hipStream_t st = ctx->stream;
    const int js = j.uid * kMaxK + j.slot;
    j.last = h_sc[js];
    j.u.resize(m.npix);
    j.v.resize(m.npix);
    ensure_stage(m.npix);
    eng_get_img<T>(st, stage.p, U.u.p, U.K, j.slot, m.g.nx, m.g.ny, m.tny);
    MI_HIP(hipMemcpyAsync(j.u.data(), stage.p, m.npix * sizeof(T), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    eng_get_img<T>(st, stage.p, U.v.p, U.K, j.slot, m.g.nx, m.g.ny, m.tny);
    MI_HIP(hipMemcpyAsync(j.v.data(), stage.p, m.npix * sizeof(T), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    j.state = BJ_DONE;
    j.t_done = now();
    retired_log.push_back(j.id);
    free_slot(j);
}

template <typename T>
void Batch<T>::free_slot(BatchJob<T> &j) {
    if (j.uid < 0) return;
    Unit<T> &U = *units[j.uid];
    const int js = j.uid * kMaxK + j.slot;
    U.jobs[j.slot] = -1;
    h_jp[js].active = 0;
    if (U.used() == 0) {
        MI_HIP(hipStreamSynchronize(ctx->stream));
        live.erase(std::find(live.begin(), live.end(), j.uid));
        units[j.uid].reset();
    }
    j.uid = -1;
    j.slot = -1;
    dirty = true;
}

template <typename T>
int Batch<T>::poll() {
    const hipStream_t st = ctx->stream;
    MI_HIP(hipStreamSynchronize(st));
    stats_collect(ctx);
    if (live.empty()) return 0;
    MI_HIP(hipMemcpyAsync(h_sc.data(), d_sc.p, d_sc.bytes(), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    if (!auto_retire) return n_active();
    std::vector<int64_t> fin;
    for (int uid : live) {
        Unit<T> &U = *units[uid];
        for (int k = 0; k < U.K; ++k)
            if (U.jobs[k] >= 0 && h_sc[uid * kMaxK + k].done) fin.push_back(U.jobs[k]);
    }
    for (int64_t id : fin) retire(job(id));
    if (!fin.empty()) compact();
    return n_active();
}

template <typename T>
int Batch<T>::step(int passes) {
    admit(false);
    enqueue(passes);
    return poll();
}

template <typename T>
Scalars Batch<T>::scalars(int64_t id) {
    BatchJob<T> &j = job(id);
    if (j.state != BJ_ACTIVE) return j.last;
    Scalars h;
    const hipStream_t st = ctx->stream;
    MI_HIP(hipMemcpyAsync(&h, d_sc.p + j.uid * kMaxK + j.slot, sizeof(Scalars), hipMemcpyDeviceToHost, st));
    MI_HIP(hipStreamSynchronize(st));
    return h;
}

template <typename T>
void Batch<T>::read(int64_t id, T *u, T *v) {
    BatchJob<T> &j = job(id);
    if (j.state == BJ_DONE) {
        if (u) std::memcpy(u, j.u.data(), j.u.size() * sizeof(T));
        if (v) std::memcpy(v, j.v.data(), j.v.size() * sizeof(T));
        return;
    }
    MI_REQUIRE(j.state == BJ_ACTIVE, MI_E_STATE, "job has no state on the device");
    Unit<T> &U = *units[j.uid];
    const Matrix &m = *U.m;
    const int64_t npix = m.npix;
    const hipStream_t st = ctx->stream;
    ensure_stage(npix);
    if (u) {
        eng_get_img<T>(st, stage.p, U.u.p, U.K, j.slot, m.g.nx, m.g.ny, m.tny);
        MI_HIP(hipMemcpyAsync(u, stage.p, npix * sizeof(T), hipMemcpyDeviceToHost, st));
        MI_HIP(hipStreamSynchronize(st));
    }
    if (v) {
        eng_get_img<T>(st, stage.p, U.v.p, U.K, j.slot, m.g.nx, m.g.ny, m.tny);
        MI_HIP(hipMemcpyAsync(v, stage.p, npix * sizeof(T), hipMemcpyDeviceToHost, st));
        MI_HIP(hipStreamSynchronize(st));
    }
}

template <typename T>
void Batch<T>::cancel(int64_t id) {
    BatchJob<T> &j = job(id);
    if (j.state == BJ_PENDING) {
        pending.erase(std::find(pending.begin(), pending.end(), id));
    } else if (j.state == BJ_ACTIVE) {
        MI_HIP(hipStreamSynchronize(ctx->stream));
        free_slot(j);
        compact();
    }
    jobs.erase(id);
}

template <typename T>
void Batch<T>::release(int64_t id) {
    cancel(id);
}

// Launch settings of the engine for one geometry: every (jobs per unit,
// residual staging) candidate runs passes of a unit of synthetic jobs (a
// constant sinogram; the kernels' work does not depend on the values),
// replayed from a hipGraph as in a batch, and the fastest time per job and
// pass is kept on the matrix.  None of the settings changes a job's
// arithmetic.
template <typename T>
mi_engine_tuning engine_autotune(mi_context *ctx, const mi_geometry &g, int S) {
    std::shared_ptr<Matrix> m = get_engine_matrix(ctx, g, S, sizeof(T) == 8);
    const hipStream_t st = ctx->stream;
    const double inf = std::numeric_limits<double>::infinity();
    std::vector<T> sino((size_t)m->nrows, T(0.5));
    mi_at_params p{};
    p.alpha = 1e-3;
    p.beta = 1e-4;
    p.epsilon = 1e-2;
    p.iter_edge = 1;
    p.n_subsets = S;
    p.max_iters = 0;   // no cap: every timed pass does its work
    p.lower = -inf;
    p.upper = inf;
    p.precision = sizeof(T) == 8 ? MI_FP64 : MI_FP32;
    mi_engine_tuning best{0, -1, inf};
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    const int reps = 3;
    try {
        for (int K = 1; K <= kMaxK; K *= 2)
            for (int lds = 0; lds <= 1; ++lds) {
                if (lds && (!m->bs.slot16.p || (size_t)m->bs.max_slots * K * sizeof(T) > kBpLdsHw)) continue;
                m->etune.k_cap = K;
                m->etune.bp_lds = lds;
                Batch<T> b(ctx, S);
                b.min_k = b.max_k = K;
                for (int k = 0; k < K; ++k) b.submit(g, p, sino.data(), nullptr, nullptr);
                b.admit(false);
                if (b.n_active() != K) break;   // does not fit: no larger K either
                b.enqueue(1);                   // warm, graph capture
                MI_HIP(hipEventRecord(e0, st));
                b.enqueue(reps);
                MI_HIP(hipEventRecord(e1, st));
                MI_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                MI_HIP(hipEventElapsedTime(&ms, e0, e1));
                const double us = 1e3 * ms / reps / K;
                if (us < best.us_per_job_pass) best = mi_engine_tuning{K, lds, us};
            }
    } catch (...) {
        m->etune = Matrix::EngineTune{};
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    MI_REQUIRE(best.k_cap > 0, MI_E_NOMEM, "no engine setting fits in device memory");
    m->etune = Matrix::EngineTune{best.k_cap, best.bp_lds, best.us_per_job_pass};
    return best;
}

template struct Batch<float>;
template struct Batch<double>;
template mi_engine_tuning engine_autotune<float>(mi_context *, const mi_geometry &, int);
template mi_engine_tuning engine_autotune<double>(mi_context *, const mi_geometry &, int);

}  // namespace mi
