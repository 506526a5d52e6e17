"""GPU tests of the batched engine (mi_batch): many jobs of mixed geometry
per launch, continuous batching, and the property the engine is built for:
a job's result is bit-for-bit the same alone and in any mix (every
reduction is cut into fixed chunks combined in a fixed order, and the
per-job arithmetic does not depend on K).  FP64 batches are also checked
against the CPU oracle (ordered-subset AT, oracle/solvers.c)."""
import numpy as np
import pytest

import medimage_amd as mi
from medimage_amd import configs
import oracle_py as O

pytestmark = pytest.mark.gpu

GEOMS = {
    "a": mi.fan_geometry(64, 4.0, 60, 96, 4.0),
    "b": mi.fan_geometry(96, 2.5, 72, 128, 3.0),
    "c": mi.fan_geometry(128, 2.0, 90, 128, 2.0),
}


def _jobs(ctx, precision=mi.MI_FP32):
    """Nine jobs over three geometries with differing parameters: a x3, b x1,
    c x5; pass caps 3..6, one with a median every 2 passes, one with two
    edge steps, one stopped by rel_tol."""
    dt = np.float64 if precision == mi.MI_FP64 else np.float32
    spec = [("a", 3, {}), ("a", 5, {"median_every": 2}), ("a", 4, {}),
            ("b", 6, {"iter_edge": 2}),
            ("c", 3, {}), ("c", 4, {"alpha": 2e-3}), ("c", 6, {"rel_tol": 5e-2}), ("c", 5, {}), ("c", 4, {})]
    out = []
    for i, (gk, iters, kw) in enumerate(spec):
        g = GEOMS[gk]
        _, sino = configs.make_case(ctx, g, 3e4, seed=100 + i)
        p = mi.at_params(kw.get("alpha", 1e-3), 1e-4, 1e-2, n_subsets=4, max_iters=iters,
                         iter_edge=kw.get("iter_edge", 1), rel_tol=kw.get("rel_tol", 0.0),
                         median_every=kw.get("median_every", 0), precision=precision)
        out.append((g, sino.astype(dt), p))
    return out


def _run_batch(ctx, jobs, precision=mi.MI_FP32, graph=True, waves=(0,), options=()):
    """Submits jobs in waves (wave w enters after waves[w] passes) and runs
    to completion.  Returns {index: (u, v, status)}."""
    b = ctx.batch(4, precision, graph=graph)
    for opt, val in options:
        b.set(opt, val)
    ids = {}
    order = np.array_split(np.arange(len(jobs)), len(waves))
    passes = 0
    for w, start in enumerate(waves):
        while passes < start:
            b.step(1)
            passes += 1
        for i in order[w]:
            g, sino, p = jobs[i]
            ids[int(i)] = b.submit(g, sino, p)
    b.run()
    out = {}
    for i, jid in ids.items():
        st = b.status(jid)
        assert st.state == mi.MI_JOB_DONE, f"job {i} state {st.state}"
        u, v = b.read(jid)
        out[i] = (u, v, st)
    b.close()
    return out


def _alone(ctx, jobs, precision=mi.MI_FP32):
    out = {}
    for i, (g, sino, p) in enumerate(jobs):
        out[i] = _run_batch(ctx, [(g, sino, p)], precision)[0]
    return out


def _same(x, y):
    return x.dtype == y.dtype and np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_batch_alone_equals_mixed_bitwise(ctx):
    jobs = _jobs(ctx)
    print(f"[batch mixed] {len(jobs)} jobs, 3 geometries", flush=True)
    mixed = _run_batch(ctx, jobs)
    alone = _alone(ctx, jobs)
    for i in range(len(jobs)):
        um, vm, sm = mixed[i]
        ua, va, sa = alone[i]
        assert sm.iterations == sa.iterations, i
        assert _same(um, ua) and _same(vm, va), f"job {i} differs alone vs mixed"
        assert sm.data_fidelity == sa.data_fidelity
    # the rel_tol job may stop early; the capped ones run exactly their cap
    assert mixed[0][2].iterations == 3 and mixed[3][2].iterations == 6


def test_batch_continuous_admission_bitwise(ctx):
    """Jobs joining at pass 0, 2 and 3 (continuous batching) end with the
    same bits as alone."""
    jobs = _jobs(ctx)
    staged = _run_batch(ctx, jobs, waves=(0, 2, 3))
    alone = _alone(ctx, jobs)
    for i in range(len(jobs)):
        assert _same(staged[i][0], alone[i][0]) and _same(staged[i][1], alone[i][1]), f"job {i}"


def test_batch_graph_equals_eager(ctx):
    jobs = _jobs(ctx)
    a = _run_batch(ctx, jobs, graph=True)
    b = _run_batch(ctx, jobs, graph=False)
    for i in range(len(jobs)):
        assert _same(a[i][0], b[i][0]) and _same(a[i][1], b[i][1]), f"job {i}"


def test_batch_graph_replays_repeat(ctx):
    """The captured pass replayed many times, with host synchronisation and
    eager work in between, gives the same result as eager launches."""
    g = GEOMS["c"]
    jobs = [(g, configs.make_case(ctx, g, 3e4, seed=300 + i)[1].astype(np.float32),
             mi.at_params(n_subsets=4, max_iters=8)) for i in range(3)]
    res = []
    for graph in (True, False):
        b = ctx.batch(4, graph=graph)
        ids = [b.submit(*j) for j in jobs]
        for _ in range(8):
            b.step(1)
            ctx.forward(g, np.ones((g.nx, g.ny), np.float32))   # eager work between replays
        res.append([b.read(i) for i in ids])
        b.close()
    for (ua, va), (ub, vb) in zip(*res):
        assert _same(ua, ub) and _same(va, vb)


def test_batch_packing(ctx):
    jobs = _jobs(ctx)
    b = ctx.batch(4)
    for g, sino, p in jobs:
        b.submit(g, sino, p)
    b.step(0)
    ks = sorted(b.units())
    print(f"[packing] units K = {ks}", flush=True)
    assert ks == [1, 4, 8]          # b x1, a x3 (K=4), c x5 (K=8)
    b.run()
    assert b.units() == []
    b.close()


def test_batch_fp64_matches_oracle(ctx):
    jobs = _jobs(ctx, mi.MI_FP64)
    res = _run_batch(ctx, jobs, mi.MI_FP64)
    for i in (0, 1, 3, 5):
        g, sino, p = jobs[i]
        ou, ov, n, hist = O.os_at(O.op_fan(O.fan_from(g)), sino, p.alpha, p.beta, p.epsilon, p.n_subsets,
                                  p.max_iters, iter_edge=p.iter_edge, rel_tol=p.rel_tol,
                                  median_every=p.median_every)
        u, v, st = res[i]
        err_u = float(np.linalg.norm(u - ou) / np.linalg.norm(ou))
        err_v = float(np.linalg.norm(v - ov) / np.linalg.norm(ov))
        print(f"[batch fp64 job {i}] iters {st.iterations}/{n} rel u {err_u:.2e} v {err_v:.2e}", flush=True)
        assert st.iterations == n
        assert err_u < 1e-10 and err_v < 1e-10


def test_batch_fp32_within_tolerance_of_oracle(ctx):
    jobs = _jobs(ctx)
    res = _run_batch(ctx, jobs)
    for i in (2, 7):
        g, sino, p = jobs[i]
        ou, ov, n, _ = O.os_at(O.op_fan(O.fan_from(g)), sino.astype(np.float64), p.alpha, p.beta, p.epsilon,
                               p.n_subsets, p.max_iters, iter_edge=p.iter_edge)
        u, v, _ = res[i]
        assert np.linalg.norm(u - ou) / np.linalg.norm(ou) < 1e-5
        assert np.abs(u - ou).max() / np.abs(ou).max() < 1e-5
        assert np.linalg.norm(v - ov) / np.linalg.norm(ov) < 1e-5


def test_batch_walk_and_sliced_units_mixed_bitwise(ctx):
    """Fan-beam jobs (checkpointed forward walk) and a parallel-beam job
    (stored sliced forward matrix) in one batch run the kernels that hold
    both forward forms; every job still ends with the bits it has alone."""
    jobs = _jobs(ctx)[:4]
    gp = mi.parallel_geometry(64, 60, 96, flavour=0)
    _, sp = configs.make_case(ctx, gp, 3e4, seed=123)
    jobs.append((gp, sp.astype(np.float32), mi.at_params(n_subsets=4, max_iters=4)))
    mixed = _run_batch(ctx, jobs)
    alone = _alone(ctx, jobs)
    for i in range(len(jobs)):
        assert _same(mixed[i][0], alone[i][0]) and _same(mixed[i][1], alone[i][1]), f"job {i}"


def test_batch_entry_budget_and_max_k(ctx):
    """MI_BATCH_MAX_ENTRIES: a job that needs a new unit waits while the
    units on the device already stream that many matrix entries, and the
    jobs behind it wait in arrival order, except those that fill a free slot
    of a unit already on the device (they share its matrix stream).
    MI_BATCH_MAX_K caps the unit size.  Results stay bitwise those of the
    job alone."""
    allj = _jobs(ctx)
    a0, b0, c0, c1, c2 = allj[0], allj[3], allj[4], allj[5], allj[7]
    for jobs, want_active in (([a0, b0, c0, c1], 1),      # a0 alone; b0 and then c need new units
                              ([c0, a0, c1, c2], 3)):     # c's unit (K=4) takes c1, c2 past a0
        alone = _alone(ctx, jobs)
        b = ctx.batch(4)
        b.set(mi.MI_BATCH_MAX_ENTRIES, 1)
        ids = [b.submit(*j) for j in jobs]
        a, q = b.step(1)
        assert (a, q) == (want_active, len(jobs) - want_active)
        b.run()
        for i, jid in enumerate(ids):
            u, v = b.read(jid)
            assert _same(u, alone[i][0]) and _same(v, alone[i][1]), f"budget job {i}"
        b.close()
    jobs = allj[4:8]                          # four jobs of geometry c
    alone = _alone(ctx, jobs)
    b = ctx.batch(4)
    b.set(mi.MI_BATCH_MAX_K, 2)
    ids = [b.submit(*j) for j in jobs]
    b.step(0)
    assert max(b.units()) <= 2
    b.run()
    for i, jid in enumerate(ids):
        u, v = b.read(jid)
        assert _same(u, alone[i][0]) and _same(v, alone[i][1]), f"max-K job {i}"
    b.close()
    b = ctx.batch(4)
    with pytest.raises(mi.MIError):
        b.set(mi.MI_BATCH_MAX_K, 3)
    b.close()


@pytest.mark.parametrize("precision", [mi.MI_FP32, mi.MI_FP64])
def test_batch_bp_lds_equals_global_gathers(ctx, precision):
    """The back projection that stages each block's residual values in LDS
    and the one that gathers them from global memory sum the same values in
    the same order: bitwise equal results.  So do the two places of the
    regulariser's u / v squares in the staged form (their own LDS region,
    loaded before the walk, or the residual's LDS after it), forced for
    every staged launch (units of K = 1, 4 and 8)."""
    jobs = _jobs(ctx, precision)
    a = _run_batch(ctx, jobs, precision)
    for opts in ([(mi.MI_BATCH_BP_LDS, 0)], [(mi.MI_BATCH_BP_SQ, 0)], [(mi.MI_BATCH_BP_SQ, 1)]):
        b = _run_batch(ctx, jobs, precision, options=opts)
        for i in range(len(jobs)):
            assert _same(a[i][0], b[i][0]) and _same(a[i][1], b[i][1]), f"job {i}, options {opts}"


@pytest.mark.parametrize("precision", [mi.MI_FP32, mi.MI_FP64])
def test_batch_bp_lds_many_views_per_subset(ctx, precision):
    """300 views per subset: each wave of a back-projection block stages its
    views' residual slots in more than one round of 64 views (a 16 x 16
    block of 2 mm pixels covers at most 21 of the 4 mm bins of a view, so
    the slots take at most 25 KiB per job in FP32 and stay staged for a unit
    of one job, and of two in FP32).  Staged and global gathers give the same
    bits, and FP64 matches the oracle."""
    g = mi.fan_geometry(48, 2.0, 600, 32, 4.0)
    dt = np.float64 if precision == mi.MI_FP64 else np.float32
    jobs = []
    for i in range(2):
        _, sino = configs.make_case(ctx, g, 3e4, seed=300 + i)
        jobs.append((g, sino.astype(dt), mi.at_params(1e-3, 1e-4, 1e-2, n_subsets=2, max_iters=3,
                                                        precision=precision)))
    for sel in (jobs[:1], jobs):
        runs = []
        for lds in (1, 0):
            b = ctx.batch(2, precision)
            b.set(mi.MI_BATCH_BP_LDS, lds)
            ids = [b.submit(*j) for j in sel]
            b.run()
            runs.append([b.read(i) for i in ids])
            b.close()
        for (ua, va), (ub, vb) in zip(*runs):
            assert _same(ua, ub) and _same(va, vb)
        if precision == mi.MI_FP64 and len(sel) == 1:
            gs, sino, p = sel[0]
            u, v, n, _ = O.os_at(O.op_fan(O.fan_from(gs)), sino, 1e-3, 1e-4, 1e-2, 2, 3)
            ud, vd = runs[0][0]
            assert np.linalg.norm(ud - u) <= 1e-10 * np.linalg.norm(u)
            assert np.linalg.norm(vd - v) <= 1e-10 * np.linalg.norm(v)


def test_batch_release_and_errors(ctx):
    g = GEOMS["a"]
    _, sino = configs.make_case(ctx, g, 3e4, seed=9)
    b = ctx.batch(4)
    j1 = b.submit(g, sino.astype(np.float32), mi.at_params(n_subsets=4, max_iters=50))
    j2 = b.submit(g, sino.astype(np.float32), mi.at_params(n_subsets=4, max_iters=2))
    b.step(1)
    b.release(j1)                     # cancel an active job
    a, q = b.step(1)
    assert b.status(j2).state == mi.MI_JOB_DONE and a == 0 and q == 0
    with pytest.raises(mi.MIError) as e:
        b.status(j1)
    assert e.value.code == mi.MI_E_ARG
    with pytest.raises(mi.MIError) as e:
        b.submit(g, sino, mi.at_params(n_subsets=3))
    assert e.value.code == mi.MI_E_ARG
    with pytest.raises(mi.MIError) as e:
        b.set(mi.MI_BATCH_BP_SQ, 2)
    assert e.value.code == mi.MI_E_ARG
    b.close()


def test_serving_stream_results_equal_alone(ctx):
    """A Poisson stream served by the request loop (continuous batching,
    jobs joining and leaving between passes): every job's result equals the
    same job run alone, bit for bit; the plain one-at-a-time loop agrees too."""
    from medimage_amd import serve, phantom
    g = phantom.rng(77, 1)
    reqs = []
    t = 0.0
    for i in range(10):
        t += float(g.exponential(1 / 400.0))
        geo = GEOMS["abc"[i % 3]]
        _, sino = configs.make_case(ctx, geo, 3e4, seed=500 + i)
        p = mi.at_params(n_subsets=4, max_iters=40, rel_tol=1e-3)
        reqs.append(dict(t=t, geom=geo, sino=sino.astype(np.float32), params=p))
    bat = serve.run_stream(ctx, reqs, n_subsets=4, mode="batched")
    pla = serve.run_stream(ctx, reqs, n_subsets=4, mode="plain")
    print(f"[serve] batched p50 {bat['latency_p50_s']*1e3:.1f} ms p99 {bat['latency_p99_s']*1e3:.1f} ms; "
          f"plain p50 {pla['latency_p50_s']*1e3:.1f} ms p99 {pla['latency_p99_s']*1e3:.1f} ms", flush=True)
    for i, r in enumerate(reqs):
        alone = _run_batch(ctx, [(r["geom"], r["sino"], r["params"])])[0]
        for res in (bat, pla):
            u, v = res["results"][i]
            assert _same(u, alone[0]) and _same(v, alone[1]), f"job {i} ({res['mode']})"
        assert bat["iterations"][i] == alone[2].iterations


def test_batch_admission_refuses_without_memory(ctx):
    """With a reserve larger than the device, a job cannot be admitted and,
    with nothing else on the device to wait for, it fails with a message."""
    g = GEOMS["a"]
    _, sino = configs.make_case(ctx, g, 3e4, seed=9)
    b = ctx.batch(4)
    b.set(mi.MI_BATCH_RESERVE_MB, 1 << 30)
    jid = b.submit(g, sino.astype(np.float32), mi.at_params(n_subsets=4, max_iters=2))
    a, q = b.step(1)
    assert a == 0 and q == 0
    assert b.status(jid).state == mi.MI_JOB_FAILED
    b.close()


def test_engine_autotune_keeps_results_bitwise(ctx):
    """The engine tuner picks jobs per unit and the residual staging per
    geometry; every setting it may pick gives the default settings' results
    bit for bit, and batches pack the geometry's jobs by the chosen K."""
    jobs = [j for j in _jobs(ctx) if j[0] is GEOMS["c"]]
    g = GEOMS["c"]
    base = _run_batch(ctx, jobs)
    t = ctx.engine_autotune(g, 4)
    print(f"[engine tune] {g}: K {t.k_cap}, LDS {t.bp_lds}, {t.us_per_job_pass:.1f} us per job and pass",
          flush=True)
    assert t.k_cap in (1, 2, 4, 8) and t.bp_lds in (0, 1) and t.us_per_job_pass > 0
    got = ctx.engine_tuning(g, 4)
    assert (got.k_cap, got.bp_lds) == (t.k_cap, t.bp_lds)
    try:
        for k, lds in [(t.k_cap, t.bp_lds), (1, 0), (2, 1), (8, 0), (8, 1)]:
            ctx.engine_tuning(g, 4, set_to=mi.EngineTuning(k, lds, 0.0))
            b = ctx.batch(4)
            for gj, sino, p in jobs:
                b.submit(gj, sino, p)
            b.step(0)
            assert max(b.units()) <= k
            b.close()
            res = _run_batch(ctx, jobs)
            for i in base:
                assert _same(res[i][0], base[i][0]) and _same(res[i][1], base[i][1]), (k, lds, i)
    finally:
        ctx.engine_tuning(g, 4, set_to=mi.EngineTuning(0, -1, 0.0))
