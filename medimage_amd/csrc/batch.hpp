// Batches of device-resident ordered-subset AT jobs (batch.cpp).
//
// A batch owns units (K <= 8 jobs sharing one system matrix, state
// interleaved) and runs one pass -- every subset step of every unit, then
// the median and edge steps, then the convergence test -- as one sequence of
// engine launches, captured once per batch composition into a hipGraph and
// replayed.  Jobs join and leave only between passes (continuous batching):
// submit() queues a job with host copies of its inputs, admit() moves
// pending jobs onto the device while device memory allows, and jobs whose
// device flag says done are retired (results copied to the host, slot
// freed) by poll().
#pragma once

#include "engine.hpp"
#include "runtime.hpp"

#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace mi {

enum BatchJobState { BJ_PENDING = 0, BJ_ACTIVE = 1, BJ_DONE = 2, BJ_FAILED = 3 };

template <typename T>
struct Unit {
    std::shared_ptr<Matrix> m;
    int uid = -1;                  // stable: job slots uid * kMaxK + j
    int K = 1;
    int64_t jobs[kMaxK];           // job id per slot, -1 when free
    DevBuf<T> g, r, u, v, d, w;
    int used() const {
        int n = 0;
        for (int j = 0; j < K; ++j) n += jobs[j] >= 0;
        return n;
    }
};

template <typename T>
struct BatchJob {
    int64_t id = 0;
    mi_geometry geom{};           // geom.view_deg unused: see geometry()
    std::vector<double> views;    // owned copy of the caller's per-view angles
    mi_geometry geometry() const {
        mi_geometry g = geom;
        g.view_deg = views.empty() ? nullptr : views.data();
        return g;
    }
    mi_at_params p{};
    int state = BJ_PENDING;
    std::shared_ptr<Matrix> m;
    int uid = -1, slot = -1;
    std::vector<T> sino, u0, v0;   // inputs until admission
    std::vector<T> u, v;           // results after retirement
    Scalars last{};                // device scalars at retirement
    double t_submit = 0, t_admit = 0, t_done = 0;   // seconds on the batch clock
    std::string error;
};

template <typename T>
struct Batch {
    mi_context *ctx;
    int S;
    int min_k = 1;                 // smallest K of a new unit
    int max_k = kMaxK;             // largest K of a unit
    int64_t max_entries = 0;       // admission budget: matrix entries of the active jobs (0: none)
    // Largest K of a unit over m (at most max_k): the tuned value, else 8 for
    // images of up to 2^17 pixels and 4 above.  A K-wide FP32 gather is one
    // 16-byte load up to K = 4 and two at K = 8, so K = 4 costs a large image
    // little more per matrix entry than K = 1 while sharing the entry four
    // ways; small images also amortise the launches at K = 8.  Measured with
    // the engine tuner on the 20 C5 geometries (N = 256 .. 1024): K = 8 is
    // fastest per job for 256^2, K = 4 for 512^2 to 1024^2 (1024^2 x 2304
    // views: 5.1 ms per job pass against 14.1 ms at K = 1).
    int k_cap(const Matrix &m) const {
        if (m.etune.k_cap > 0) return std::min(m.etune.k_cap, max_k);
        return std::min(m.npix <= (int64_t)1 << 17 ? 8 : 4, max_k);
    }
    // Back projection staged in LDS for unit U of matrix m needing `need`
    // bytes of slots (MI_BATCH_BP_LDS off: never).
    bool bp_in_lds(const Matrix &m, size_t need) const {
        if (!use_bp_lds) return false;
        if (m.etune.bp_lds == 0) return false;
        return need <= (m.etune.bp_lds == 1 ? kBpLdsHw : kBpLdsMax);
    }
    bool use_graph = true;
    bool auto_retire = true;
    int64_t reserve_bytes = (int64_t)1 << 30;   // device memory kept free at admission
    int64_t next_id = 1;
    std::map<int64_t, BatchJob<T>> jobs;
    std::deque<int64_t> pending;
    std::vector<std::unique_ptr<Unit<T>>> units;   // by uid (kMaxUnits)

    DevBuf<UnitDesc<T>> d_units;
    DevBuf<JobParams> d_jp;
    DevBuf<Scalars> d_sc;
    DevBuf<double> d_part;
    DevBuf<T> stage;
    std::vector<JobParams> h_jp;
    std::vector<Scalars> h_sc;
    std::vector<int> live;         // uids in table order
    BlockMap pix{};
    // Projector launches, each split in two classes of units by K so that
    // small units do not run with the register budget (and occupancy) of
    // code compiled for K = 8: forward rows by K = 1 / K >= 2, back
    // projection by K <= 4 / K = 8.  Back projection: units whose residual
    // slots fit the staging limit in the LDS launches (bpl), the others in
    // the global-gather launches (bpg); a unit has blocks in exactly one map
    // of each phase.  The LDS launches are split once more by the units'
    // staging bytes (up to kBpLdsSplit, above), as a launch gives every block
    // the LDS of its largest unit and with it the blocks per CU.  The two classes of a phase run as concurrent branches
    // (second stream, forked and joined by events; captured as parallel
    // graph nodes).
    struct Launch {
        BlockMap map{};
        int kmax = 1;
        size_t lds = 0;
        bool any() const { return map.start[map.nunits] > 0; }
    };
    Launch rows[2], bpl[2][2], bpg[2];   // bpl[K class][LDS class]
    hipStream_t side = nullptr;   // the second class of a phase
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    EngArgs<T> args{};
    int kmax = 1;
    bool all_walk = true;   // every unit's forward projector is the checkpointed walk
    bool use_bp_lds = true; // MI_BATCH_BP_LDS
    int bp_sq = -1;         // MI_BATCH_BP_SQ
    int edge_iters = 0;            // max iter_edge over live jobs
    bool any_median = false;
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::chrono::steady_clock::time_point t0;
    int64_t passes_run = 0;
    std::vector<int64_t> retired_log;   // ids retired since the caller last drained it

    Batch(mi_context *c, int n_subsets);
    ~Batch();
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    double now() const;
    int64_t submit(const mi_geometry &g, const mi_at_params &p, const T *sino, const T *u0, const T *v0);
    // Admits pending jobs in arrival order while memory allows; returns the
    // number admitted.  `require` makes a job that cannot be placed an error
    // (used by the single-job API) instead of leaving it pending.
    int admit(bool require = false);
    // Queues `passes` passes of every live job (no wait).
    void enqueue(int passes);
    // Waits for the stream, reads every job's device scalars and retires
    // the finished ones (when auto_retire).  Returns the number still active.
    int poll();
    // admit + enqueue + poll.
    int step(int passes);
    void cancel(int64_t id);
    void release(int64_t id);
    BatchJob<T> &job(int64_t id);
    Scalars scalars(int64_t id);
    void read(int64_t id, T *u, T *v);
    int n_active() const;
    int64_t active_entries() const;   // sum over the units on the device of their matrix entries
    bool has_free_slot(const Matrix *m) const;   // a unit of m has an empty slot
    int n_pending() const { return (int)pending.size(); }
    int n_units() const { return (int)live.size(); }

    // Internals (also used by the launch-shape tuner).
    Unit<T> *place(BatchJob<T> &j, int group_size);
    void load_job(BatchJob<T> &j);
    void retire(BatchJob<T> &j);
    void free_slot(BatchJob<T> &j);
    void compact();
    void rebuild();
    void record_pass(bool timed);
    void drop_graph();
    void ensure_stage(int64_t n);
};

// Times the engine's launch settings for (g, S) and stores the fastest on
// the geometry's engine matrix (Matrix::etune).
template <typename T>
mi_engine_tuning engine_autotune(mi_context *ctx, const mi_geometry &g, int S);

}  // namespace mi
