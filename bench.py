"""Benchmark of the AT reconstruction loop on one MI355X.

Workload: the precision form (C2): 512 x 512 image, 0.7 mm pixels, 720 views
over 360 degrees, 736 detector bins of 1.0 mm, flat fan beam (SOD 541 mm,
SDD 949 mm), 12 ordered subsets, FP32 bulk with Kahan sums and FP64
reductions.  One step = one AT iteration (a pass over all 12 subsets: forward
projection, residual, gather back projection + image gradient, line search,
image update; then one edge step) of one job whose sinogram and state are
resident in HBM.  Synthetic input: a seeded 10-ellipse phantom with Poisson
noise at I0 = 2e4.

Prints one JSON line: AT iterations/s (value), rays/s, the roofline of the
dominant kernel from HIP events on the library's stream during the timed
region (the dispatch's own start/stop events), and the CPU oracle timed on a
bounded sample of the same workload on this host.  Extra blocks of the same
line: the 64-slice batched form (C3), the served request stream (C5) and the
3D cone beam (C4: 512^3 volume, 720 views of a 1024 x 768 panel; projector
kernel times and ordered-subset AT passes).
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "AT iters/sec & rays/sec (512², 720 views) plain vs batched; p50/p99 job latency"


def host_cpus():
    """CPUs this process may run on, the machine's count and the CPU model.
    A GPU host may hand each GPU's job a share of its cores (its affinity
    mask, OMP_NUM_THREADS); the CPU baseline uses that share."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    omp = int(os.environ.get("OMP_NUM_THREADS") or 0)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"threads": min(allowed, omp) if omp else allowed, "affinity_cpus": allowed,
            "machine_cpus": os.cpu_count(), "omp_num_threads": omp or None, "model": model}


def cpu_baseline(sino, passes, sino_c1=None, passes_c1=0):
    """The CPU port (oracle/cpu_at.c: the same AT iteration with the oracle's
    tracer, rays re-traced in every projection as the reference's OpenMP
    programs do, FP32 bulk, OpenMP over rays) on `passes` AT iterations of
    the C2 job, on every CPU this process may use; then the C1 job (256^2,
    360 views, 512 bins, one subset) for `passes_c1` iterations."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_py as O
    from medimage_amd import configs
    host = host_cpus()
    threads = host["threads"]
    g = configs.c2_geometry()
    op = O.op_fan(O.fan_from(g))
    t0 = time.perf_counter()
    O.cpu_os_at_f32(op, sino.astype(np.float32), configs.ALPHA, configs.BETA, configs.EPS, 12, passes,
                    iter_edge=1, threads=threads)
    dt = time.perf_counter() - t0
    out = {"value": passes / dt, "unit": "AT iterations/s", "cores": threads, "kind": "port",
           "sample": f"{passes} AT iterations of the same 512x512 / 720-view / 736-bin / 12-subset job, "
                     f"FP32 CPU port with OpenMP over rays, {threads} threads, {dt:.1f} s",
           "rays_per_s": passes * g.n_views * g.n_bins / dt, "host": host}
    if sino_c1 is not None and passes_c1 > 0:
        g1 = configs.c1_geometry()
        op1 = O.op_fan(O.fan_from(g1))
        t0 = time.perf_counter()
        O.cpu_os_at_f32(op1, sino_c1.astype(np.float32), configs.ALPHA, configs.BETA, configs.EPS, 1, passes_c1,
                        iter_edge=1, threads=threads)
        dt1 = time.perf_counter() - t0
        out["c1"] = {"value": passes_c1 / dt1, "unit": "AT iterations/s",
                     "rays_per_s": passes_c1 * g1.n_views * g1.n_bins / dt1,
                     "sample": f"{passes_c1} AT iterations of the C1 job (256x256, 1.0 mm, 360 views, 512 bins, "
                               f"one subset), {threads} threads, {dt1:.1f} s"}
    return out


def c1_gpu(ctx, iters):
    """C1 (256^2 / 360 views / 512 bins, 50 AT iterations, one subset) as a
    plain per-job GPU solve: returns (AT iterations/s, sinogram)."""
    from medimage_amd import configs
    g = configs.c1_geometry()
    _, sino = configs.make_case(ctx, g, 1e5, seed=1)
    p = configs.c1_params(max_iters=iters)
    ctx.at_reconstruct(g, sino.astype(np.float32), configs.c1_params(max_iters=2))  # warm (matrix, code)
    u, v, rep = ctx.at_reconstruct(g, sino.astype(np.float32), p)
    return {"workload": "C1: 256x256 (1.0 mm), 360 views, 512 bins (1.0 mm), one subset, "
                        f"{rep.iterations} AT iterations, FP32 (FP64 reductions)",
            "at_iterations_per_s": rep.iterations / (rep.elapsed_ms * 1e-3),
            "rays_per_s": rep.iterations * g.n_views * g.n_bins / (rep.elapsed_ms * 1e-3),
            "elapsed_ms": rep.elapsed_ms}, sino


def msbeam_gpu(ctx, iters):
    """The parallel beam at msbeam's own configuration (msbeam/CPU_OpenMP/src/
    proto.h: 512 x 512, 180 views over 180 degrees, 768 rays; its tracer,
    bit-exact): the plain per-job ordered-subset AT solve with one subset
    (every ray each iteration, as msbeam's sweeps), FP32 (FP64 reductions)."""
    import medimage_amd as mi
    from medimage_amd import configs
    g = mi.parallel_geometry(512, 180, 768, flavour=0)
    _, sino = configs.make_case(ctx, g, 1e5, seed=6)
    p = mi.at_params(configs.ALPHA, configs.BETA, configs.EPS, n_subsets=1, max_iters=iters)
    ctx.at_reconstruct(g, sino.astype(np.float32), mi.at_params(configs.ALPHA, configs.BETA, configs.EPS,
                                                                n_subsets=1, max_iters=2))  # warm
    u, v, rep = ctx.at_reconstruct(g, sino.astype(np.float32), p)
    return {"workload": f"msbeam parallel beam: 512x512, 180 views over 180 degrees, 768 rays, one subset, "
                        f"{rep.iterations} AT iterations, FP32 (FP64 reductions)",
            "at_iterations_per_s": rep.iterations / (rep.elapsed_ms * 1e-3),
            "rays_per_s": rep.iterations * g.n_views * g.n_bins / (rep.elapsed_ms * 1e-3),
            "elapsed_ms": rep.elapsed_ms}


def bench_goldenc(ctx, log, cpu=True):
    """The GoldenC drop-in at the reference program's own configuration
    (origin/GoldenC/src/config.h: 512 x 512, 180 views x 767 bins,
    MumfordShah(1000, 0.002, 1e-4) with 10 x (10 CG + 10 SD)), in the
    reference's summation order: device time of the whole run, and whether
    its f and v equal the reference program's (tests/golden/goldenc_ms.npz)
    bit for bit.  CPU leg: one alternation of the oracle's restatement on
    this host (the reference's single-threaded loop order)."""
    import medimage_amd as mi
    gold = np.load(ROOT / "tests" / "golden" / "goldenc_ms.npz")
    f0 = np.load(ROOT / "tests" / "golden" / "blockrecon_std.npz")["f"].astype(np.float64)
    g = mi.goldenc_geometry()
    sino = ctx.forward(g, f0, np.float64)
    warm = mi.at_params(1000.0, 0.002, 1e-4, iter_alt=1, iter_image=2, iter_edge=2, precision=mi.MI_FP64)
    ctx.mumford_shah(g, sino, warm)        # reference-order lists, code
    p = mi.at_params(1000.0, 0.002, 1e-4, iter_alt=10, iter_image=10, iter_edge=10, precision=mi.MI_FP64)
    t0 = time.perf_counter()
    f, v, rep = ctx.mumford_shah(g, sino, p)
    wall = time.perf_counter() - t0
    same = bool(np.array_equal(f.view(np.uint64), gold["f"].view(np.uint64)) and
                np.array_equal(v.view(np.uint64), gold["v"].view(np.uint64)))
    log(f"[goldenc] 10 alternations: {rep.elapsed_ms:.0f} ms on the device, bitwise = reference: {same}")
    out = {"workload": "GoldenC MumfordShah(1000, 0.002, 1e-4), 512x512, 180 views x 767 bins, 10 x (10 CG + 10 SD), "
                       "FP64, reference summation order",
           "device_ms": rep.elapsed_ms, "wall_s": wall, "alternations_per_s": 10.0 / (rep.elapsed_ms * 1e-3),
           "bitwise_equal_reference_output": same,
           "sino_bitwise_equal_reference_gDelta": bool(np.array_equal(sino.view(np.uint64), gold["g"].view(np.uint64)))}
    if cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_py as O
        t0 = time.perf_counter()
        O.mumford_shah(O.op_goldenc(), sino, 1000.0, 0.002, 1e-4, 1, 10, 10)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "alternations/s", "cores": 1, "kind": "port",
                               "sample": f"one alternation (10 CG + 10 SD) of the oracle's restatement, {dt:.1f} s"}
    return out


def bench_c3(ctx, n_jobs, iters, log, plain_pass=True, tune=True):
    """Batched form: n_jobs mixed fan-beam slices (sizes, views, bins and I0
    drawn from the seed), 8 subsets, `iters` AT iterations each, run (a) one
    at a time through the plain per-job form and (b) all together in one
    batch.  Returns throughput of both and the per-job agreement."""
    import medimage_amd as mi
    from medimage_amd import configs
    spec = configs.c3_jobs(n_jobs, seed=3)
    t = time.perf_counter()
    jobs = []
    geoms = {}
    for j in spec:
        g = configs.mixed_geometry(j["n"], j["views"], j["bins"])
        geoms[g.key()] = g
        _, sino = configs.make_case(ctx, g, j["i0"], j["seed"])
        jobs.append((g, sino.astype(np.float32), j))
    ctx.trim()
    nnz = {}
    for k, g in geoms.items():
        nnz[k] = ctx.prepare(g, 8)
    log(f"[c3] {n_jobs} jobs over {len(geoms)} geometries prepared in {time.perf_counter() - t:.1f} s")
    # launch settings per geometry from the engine tuner (both passes use them)
    t = time.perf_counter()
    tuned = {}
    if tune:
        for k, g in geoms.items():
            tt = ctx.engine_autotune(g, 8)
            tuned[f"{g.nx}x{g.n_views}x{g.n_bins}"] = {"k_cap": tt.k_cap, "bp_lds": tt.bp_lds,
                                                       "us_per_job_pass": tt.us_per_job_pass}
        log(f"[c3] tuned {len(geoms)} geometries in {time.perf_counter() - t:.1f} s")
    p = configs.at_params(configs.ALPHA, configs.BETA, configs.EPS, n_subsets=8, max_iters=iters)
    rays = sum(g.n_views * g.n_bins for g, _, _ in jobs)

    # (a) plain: one job at a time (each a batch of one; graph replay).
    plain, t_plain = None, None
    if plain_pass:
        ctx.at_reconstruct(jobs[0][0], jobs[0][1], configs.at_params(n_subsets=8, max_iters=2))  # warm
        t0 = time.perf_counter()
        plain = [ctx.at_reconstruct(g, s, p) for g, s, _ in jobs]
        t_plain = time.perf_counter() - t0
        log(f"[c3] plain: {n_jobs} x {iters} iterations in {t_plain:.2f} s")

    # (b) batched: every job in one batch.
    b = ctx.batch(8)
    ids = [b.submit(g, s, p) for g, s, _ in jobs]
    b.step(0)                       # admission (device copies, unit packing)
    units = b.units()
    t0 = time.perf_counter()
    b.run()
    t_batch = time.perf_counter() - t0
    log(f"[c3] batched: {len(units)} units K={sorted(units)}; {t_batch:.2f} s")
    same = 0
    worst = 0.0
    for (u0, v0, r0), jid in zip(plain or [], ids):
        u1, v1 = b.read(jid)
        if np.array_equal(u0.view(np.uint32), u1.view(np.uint32)) and np.array_equal(v0.view(np.uint32), v1.view(np.uint32)):
            same += 1
        worst = max(worst, float(np.abs(u1 - u0).max() / max(np.abs(u0).max(), 1e-30)))
    b.close()
    work = n_jobs * iters
    # Matrix-sharing bound of the mix: a unit walks its matrix once for its K
    # jobs, so at equal per-entry rates the batch gains at most
    # sum_jobs nnz / sum_units nnz over running the jobs one at a time (one
    # unit per ceil(jobs / K cap) of a geometry; with the tuned caps), and
    # sum_jobs nnz / sum_geometries nnz with every geometry's jobs in one unit.
    per_geom = {}
    for g, _, _ in jobs:
        per_geom[g.key()] = per_geom.get(g.key(), 0) + 1
    job_nnz = sum(nnz[k] * c for k, c in per_geom.items())
    bound = {"sum_jobs_nnz": job_nnz, "one_unit_per_geometry": job_nnz / sum(nnz.values())}
    if tune:
        caps = {g.key(): tuned[f"{g.nx}x{g.n_views}x{g.n_bins}"]["k_cap"] for g in geoms.values()}
        unit_nnz = sum(nnz[k] * -(-c // caps[k]) for k, c in per_geom.items())
        bound["tuned_units"] = job_nnz / unit_nnz
    return {
        "workload": f"{n_jobs} fan-beam slices, N in {{256,384,512}}, views in {{360,720,1152}}, bins in "
                    f"{{512,736}}, I0 U[1e4,1e5], 8 subsets, {iters} AT iterations each",
        "geometries": len(geoms),
        "units_K": sorted(units),
        "plain_iters_per_s": work / t_plain if plain_pass else None,
        "batched_iters_per_s": work / t_batch,
        "speedup": t_plain / t_batch if plain_pass else None,
        "plain_rays_per_s": rays * iters / t_plain if plain_pass else None,
        "batched_rays_per_s": rays * iters / t_batch,
        "jobs_bitwise_identical": same if plain_pass else None,
        "max_rel_diff": worst,
        "tuning": tuned,
        "sharing_bound": bound,
    }


def bench_serving(ctx, n_jobs, rate, log):
    """Serving form: a Poisson stream (rate jobs/s) of fan-beam jobs with N in
    {256..1024}, views in {360..2304}, bins 1.5 N, I0 log-uniform, each run
    until its relative data-fidelity change is below 1e-4 or 200 AT
    iterations (8 subsets), served (a) one job at a time and (b) by the
    continuous-batching request loop.  Sinograms: exact ellipse line
    integrals with Poisson noise.  The system matrices of the stream's
    geometries are built before the clock starts (a warm matrix cache)."""
    from medimage_amd import configs, serve
    spec = configs.c5_requests(n_jobs, seed=5, rate=rate)
    t = time.perf_counter()
    reqs = []
    geoms = {}
    for r in spec:
        g = configs.mixed_geometry(r["n"], r["views"], r["bins"])
        geoms.setdefault(g.key(), g)
        sino = configs.analytic_case(g, r["i0"], r["seed"]).astype(np.float32)
        p = configs.at_params(configs.ALPHA, configs.BETA, configs.EPS, n_subsets=8, max_iters=200, rel_tol=1e-4)
        reqs.append(dict(t=r["t"], geom=g, sino=sino, params=p))
    nnz = sum(ctx.prepare(g, 8) for g in geoms.values())
    for g in geoms.values():        # launch settings per geometry, as the full stream uses them
        ctx.engine_autotune(g, 8)
    log(f"[serve] {n_jobs} requests over {len(geoms)} geometries ({nnz / 1e9:.2f}e9 matrix entries) "
        f"prepared and tuned in {time.perf_counter() - t:.1f} s")
    out = {"workload": f"the first {n_jobs} requests of the C5 stream (not C5 itself: the whole 2000-request "
                       f"stream is bench.py --serve-full): Poisson arrivals at {rate:g} jobs/s, N in "
                       f"{{256,384,512,768,1024}}, views in {{360,720,1152,2304}}, bins 1.5N, I0 log-uniform "
                       f"[1e4,1e5], 8 subsets, rel. data-fidelity change < 1e-4 or 200 AT iterations",
           "geometries": len(geoms)}
    res = {}
    for mode in ("plain", "batched"):
        r = serve.run_stream(ctx, reqs, n_subsets=8, mode=mode, log=log)
        res[mode] = r
        log(f"[serve] {mode}: {r['makespan_s']:.1f} s, {r['iters_per_s']:.0f} AT iter/s, "
            f"p50 {r['latency_p50_s']:.2f} s, p99 {r['latency_p99_s']:.2f} s")
        out[mode] = {k: r[k] for k in ("makespan_s", "jobs_per_s", "iters_per_s", "latency_p50_s",
                                       "latency_p99_s", "latency_mean_s", "mean_iterations")}
    same = sum(np.array_equal(res["plain"]["results"][i][0].view(np.uint32),
                              res["batched"]["results"][i][0].view(np.uint32)) for i in range(n_jobs))
    out["jobs_bitwise_identical"] = int(same)
    out["throughput_ratio"] = res["batched"]["iters_per_s"] / res["plain"]["iters_per_s"]
    return out


def _c5_sino(args):
    """Worker: one request's noisy analytic sinogram (float32)."""
    from medimage_amd import configs
    n, views, bins, i0, seed = args
    g = configs.mixed_geometry(n, views, bins)
    return configs.analytic_case(g, i0, seed).astype(np.float32)


def bench_serving_full(ctx, n_jobs, rate, log, tune=True, workers=16):
    """The whole C5 stream (n_jobs Poisson arrivals at `rate` jobs/s) served
    live by the continuous-batching request loop.  The plain one-job-at-a-
    time form of the same stream is not run (it would take as long as its
    total work); its first-come-first-served latencies are derived from each
    geometry's measured per-pass time of the plain form (one job alone, the
    same kernels) times each job's iteration count, which is the same alone
    and in the mix (results are bitwise identical), queued in arrival
    order.  Geometries are tuned first (mi_engine_autotune) and their
    matrices built before the clock starts; sinograms are exact ellipse line
    integrals with Poisson noise, made on `workers` host processes."""
    from concurrent.futures import ProcessPoolExecutor
    import medimage_amd as mi
    from medimage_amd import configs, serve
    spec = configs.c5_requests(n_jobs, seed=5, rate=rate)
    t = time.perf_counter()
    geoms = {}
    for r in spec:
        g = configs.mixed_geometry(r["n"], r["views"], r["bins"])
        geoms.setdefault(g.key(), g)
    nnz = sum(ctx.prepare(g, 8) for g in geoms.values())
    log(f"[serve-full] {len(geoms)} geometries ({nnz / 1e9:.2f}e9 entries) prepared in {time.perf_counter() - t:.1f} s")
    tuned = {}
    if tune:
        for k, g in geoms.items():
            tt = ctx.engine_autotune(g, 8)
            tuned[k] = (tt.k_cap, tt.bp_lds, tt.us_per_job_pass)
        log(f"[serve-full] tuned: " + ", ".join(f"{g.nx}^2/{g.n_views}: K{v[0]} L{v[1]}"
                                               for k, g in geoms.items() for v in [tuned[k]]))
    # plain per-pass time of each geometry: one job alone, 3 passes
    plain_pass_ms = {}
    for k, g in geoms.items():
        s0 = np.zeros((g.n_views, g.n_bins), np.float32) + np.float32(0.5)
        _, _, rep = ctx.at_reconstruct(g, s0, configs.at_params(configs.ALPHA, configs.BETA, configs.EPS,
                                                                n_subsets=8, max_iters=3))
        plain_pass_ms[k] = rep.elapsed_ms / 3.0
    t = time.perf_counter()
    with ProcessPoolExecutor(max_workers=workers) as ex:
        sinos = list(ex.map(_c5_sino, [(r["n"], r["views"], r["bins"], r["i0"], r["seed"]) for r in spec],
                            chunksize=8))
    log(f"[serve-full] {n_jobs} sinograms in {time.perf_counter() - t:.1f} s")
    reqs = []
    for r, sino in zip(spec, sinos):
        g = geoms[configs.mixed_geometry(r["n"], r["views"], r["bins"]).key()]
        p = configs.at_params(configs.ALPHA, configs.BETA, configs.EPS, n_subsets=8, max_iters=200, rel_tol=1e-4)
        reqs.append(dict(t=r["t"], geom=g, sino=sino, params=p))
    del sinos
    res = serve.run_stream(ctx, reqs, n_subsets=8, mode="batched", log=log, keep_results=False)
    # plain FCFS derived: service = iterations x the geometry's plain pass time
    svc = np.array([res["iterations"][i] * plain_pass_ms[reqs[i]["geom"].key()] * 1e-3 for i in range(n_jobs)])
    arr = np.array([r["t"] for r in reqs])
    order = np.argsort(arr, kind="stable")
    free, lat = 0.0, np.zeros(n_jobs)
    for i in order:
        start = max(free, arr[i])
        free = start + svc[i]
        lat[i] = free - arr[i]
    iters = float(res["iterations"].sum())
    out = {"workload": f"the C5 stream: {n_jobs} Poisson arrivals at {rate:g} jobs/s, N in {{256,384,512,768,1024}}, "
                       "views in {360,720,1152,2304}, bins 1.5N, I0 log-uniform [1e4,1e5], 8 subsets, rel. "
                       "data-fidelity change < 1e-4 or 200 AT iterations",
           "geometries": len(geoms), "matrix_entries": nnz,
           "tuning": {f"{g.nx}x{g.n_views}x{g.n_bins}": {"k_cap": v[0], "bp_lds": v[1], "us_per_job_pass": v[2]}
                      for k, g in geoms.items() for v in [tuned.get(k, (0, -1, 0.0))]},
           "batched": {k: res[k] for k in ("makespan_s", "jobs_per_s", "iters_per_s", "latency_p50_s", "latency_p99_s",
                                          "latency_mean_s", "mean_iterations")},
           "plain_derived": {"makespan_s": float(free), "jobs_per_s": n_jobs / float(free),
                             "iters_per_s": iters / float(svc.sum()),
                             "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
                             "latency_mean_s": float(lat.mean()),
                             "how": "first come first served over the measured per-pass time of each geometry's "
                                    "plain form times each job's iteration count"},
           "plain_pass_ms": {f"{g.nx}x{g.n_views}x{g.n_bins}": plain_pass_ms[k] for k, g in geoms.items()}}
    out["throughput_ratio"] = out["batched"]["iters_per_s"] / out["plain_derived"]["iters_per_s"]
    return out


def bench_helical(ctx, n_slices, iters, log, cpu_passes=20):
    """The helical per-slice path at the reference's clinical geometry
    (helical/rebin/json/UCLA_PHANTOM/SCAN1_1.5mm.json: 874 x 874 at 0.5859 mm,
    736 arc detector elements of 0.06786 degrees, SOD 595 mm, SDD 1085.6 mm,
    off-centre 1.75 elements; 1152 views per slice): n_slices rebinned slices,
    each with its own uneven source-angle table, written as <prefix>i_pos /
    <prefix>i_raw, then (a) every slice through mi_at_reconstruct one at a
    time and (b) all of them through the per-slice driver
    (mi_helical_reconstruct, slices batched).  Matrices are built before the
    clock starts.  Each slice's matrix is its own (its angle table), so the
    batch shares no matrix: (b) gains only from running the slices together."""
    import tempfile
    import medimage_amd as mi
    from medimage_amd import phantom
    views, n = 1152, 874
    g0 = mi.arc_geometry(n, 0.5859, 1, 736, 0.06786112, sod=595.0, sdd=1085.6, off_center=1.75)
    p = mi.at_params(0.1, 0.01, 1e-3, n_subsets=8, max_iters=iters)
    rng = np.random.default_rng(11)
    tmp = tempfile.mkdtemp(prefix="helical_")
    prefix = os.path.join(tmp, "slice")
    t = time.perf_counter()
    slices = []
    nnz = 0
    for i in range(n_slices):
        steps = rng.uniform(0.6, 1.4, views) * (360.0 / views)
        angles = np.cumsum(steps) + rng.uniform(0.0, 360.0)
        g = mi.arc_geometry(n, 0.5859, views, 736, 0.06786112, sod=595.0, sdd=1085.6,
                            off_center=1.75).with_views(angles)
        ell = phantom.random_ellipses(10, n * 0.5859, 500 + i, mu_max=0.02)
        img = phantom.rasterize(ell, n, 0.5859)
        sino = phantom.noisy_sinogram(ctx.forward(g, img, np.float32), 5e4, 500 + i).astype(np.float32)
        # the files hold exactly these values: the angles and the float32
        # sinogram as the doubles they are (17 significant digits)
        with open(f"{prefix}{i}_pos", "w") as f:
            for v, a in enumerate(angles):
                f.write(f"{float(v) * 0.5!r},{float(a)!r}\n")
        np.savetxt(f"{prefix}{i}_raw", sino.astype(np.float64), fmt="%.17g")
        ctx.trim()   # the projector's matrix of this slice
        slices.append((g, sino))
        log(f"[helical] slice {i}: {views} views written ({time.perf_counter() - t:.1f} s)")
    for g, _ in slices:
        nnz += ctx.prepare(g, 8)
    log(f"[helical] engine matrices of {n_slices} slices ready ({time.perf_counter() - t:.1f} s)")
    ctx.at_reconstruct(slices[0][0], slices[0][1], mi.at_params(0.1, 0.01, 1e-3, n_subsets=8, max_iters=1))  # warm
    t0 = time.perf_counter()
    alone = [ctx.at_reconstruct(g, s, p) for g, s in slices]
    t_plain = time.perf_counter() - t0
    log(f"[helical] one at a time: {n_slices} x {iters} iterations in {t_plain:.2f} s")
    # the files read one after another with the same readers (the driver
    # reads them on host threads)
    t0 = time.perf_counter()
    for i in range(n_slices):
        _, ang = mi.read_pos(f"{prefix}{i}_pos")
        mi.read_raw(f"{prefix}{i}_raw", len(ang), 736)
    t_read = time.perf_counter() - t0
    t0 = time.perf_counter()
    u, v, reps = ctx.helical_reconstruct(prefix, 0, n_slices, g0, p)
    t_drv = time.perf_counter() - t0
    log(f"[helical] per-slice driver (files read, slices batched): {t_drv:.2f} s, of which reading the files "
        f"alone takes {t_read:.2f} s")
    same = sum(1 for k, (ua, va, _) in enumerate(alone)
               if np.array_equal(u[k].view(np.uint32), ua.view(np.uint32))
               and np.array_equal(v[k].view(np.uint32), va.view(np.uint32)))
    for i in range(n_slices):
        os.remove(f"{prefix}{i}_pos")
        os.remove(f"{prefix}{i}_raw")
    os.rmdir(tmp)
    work = n_slices * iters
    out = {"workload": f"{n_slices} rebinned helical slices at the clinical geometry: 874x874 (0.5859 mm), "
                       f"1152 views each with its own source-angle table, 736 arc elements (0.06786 deg), "
                       f"SOD 595 mm, SDD 1085.6 mm, 8 subsets, {iters} AT iterations, alpha 0.1, beta 0.01, eps 1e-3",
           "matrix_entries_per_slice": nnz / n_slices,
           "plain_iters_per_s": work / t_plain, "driver_iters_per_s": work / t_drv,
           "driver_wall_s": t_drv, "read_files_one_after_another_s": t_read,
           "slices_bitwise_identical": same,
           "driver_includes": "reading the slice files (text, on host threads) and submitting every slice"}
    if cpu_passes > 0:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_py as O
        host = host_cpus()
        g, s = slices[0]
        t0 = time.perf_counter()
        O.cpu_os_at_f32(O.op_fan(O.fan_from(g)), s, 0.1, 0.01, 1e-3, 8, cpu_passes, iter_edge=1,
                        threads=host["threads"])
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_passes / dt, "unit": "AT iterations/s", "cores": host["threads"],
                               "kind": "port", "sample": f"{cpu_passes} AT iteration(s) of slice 0, FP32 CPU port, "
                                                         f"{host['threads']} threads, {dt:.1f} s"}
    return out


def bench_k_scaling(ctx, log):
    """Jobs per unit: for geometries of the C5 mix, the time per job and
    pass of a unit of K identical jobs (K = 1, 2, 4, 8, set through the
    engine's tuning setting, residual staged in LDS where it fits) and the
    per-phase times of an instrumented pass (HIP events per phase)."""
    import medimage_amd as mi
    from medimage_amd import configs
    out = {}
    for (n, v, b) in [(256, 720, 384), (512, 720, 768), (768, 1152, 1152), (1024, 2304, 1536)]:
        g = configs.mixed_geometry(n, v, b)
        ctx.prepare(g, 8)
        sino = np.full((g.n_views, g.n_bins), 0.5, np.float32)
        p = configs.at_params(configs.ALPHA, configs.BETA, configs.EPS, n_subsets=8, max_iters=0)
        rows = {}
        for K in (1, 2, 4, 8):
            ctx.engine_tuning(g, 8, set_to=mi.EngineTuning(K, -1, 0.0))
            bt = ctx.batch(8, min_k=K)
            for _ in range(K):
                bt.submit(g, sino, p)
            bt.step(2)
            t0 = time.perf_counter()
            bt.step(4)
            t1 = time.perf_counter()
            ctx.stats_reset()
            ctx.stats_enable(True)
            bt.step(2)
            ctx.stats_enable(False)
            ph = {name: ctx.stats(k).total_ms / 2 for name, k in
                  (("forward_walk", mi.MI_K_FP), ("back_projection", mi.MI_K_BP), ("direction_walk", mi.MI_K_FPSQ),
                   ("nv_step_update", mi.MI_K_STEP), ("edge", mi.MI_K_EDGE))}
            rows[K] = {"ms_per_job_pass": (t1 - t0) * 1e3 / 4 / K, "units": bt.units(), "phase_ms_per_pass": ph}
            log(f"[k-scaling] {n}^2 x {v} x {b} K={K}: {rows[K]['ms_per_job_pass']:.3f} ms per job pass")
            bt.close()
        ctx.engine_tuning(g, 8, set_to=mi.EngineTuning(0, -1, 0.0))
        out[f"{n}x{v}x{b}"] = rows
        ctx.trim()
    return out


def bench_c4(ctx, reps, at_iters, log):
    """C4 cone beam (512^3 volume, 720 views of the 1024 x 768 panel): the
    forward and gather back projection kernels (HIP-event time; host copies
    are outside the timed launches), then ordered-subset AT passes (20
    subsets) on a noisy sinogram of the 20-ellipsoid phantom."""
    import medimage_amd as mi
    from medimage_amd import configs, phantom
    g = configs.c4_geometry()
    t0 = time.perf_counter()
    ell = phantom.random_ellipsoids(20, g.nx * g.sample_mm, seed=4, mu_max=0.05)
    vol = phantom.rasterize3d(ell, g.nz, g.nx, g.ny, g.sample_mm)
    ks = [mi.MI_K_CONE_FP, mi.MI_K_CONE_BP]
    ctx.stats_select(ks)
    log(f"[bench] C4: phantom {time.perf_counter() - t0:.1f} s; first projection (counts the traced entries once)")
    sino = ctx.cone_forward(g, vol)
    ctx.cone_backproject(g, sino)
    ctx.stats_reset()
    for i in range(reps):
        sino = ctx.cone_forward(g, vol)
        ctx.cone_backproject(g, sino)
        log(f"[bench] C4 projector rep {i + 1}/{reps}")
    st = {k: ctx.stats(k) for k in ks}
    rays = g.n_views * g.n_rows * g.n_cols
    vox = g.nx * g.ny * g.nz
    out = {"workload": "C4: 512^3 volume (0.5 mm), 720 views over 360 degrees, flat panel 1024 x 768 "
                       "(0.4 mm), SOD 600 mm, SDD 1000 mm; 20-ellipsoid phantom (0-0.05 / mm)",
           "rays": rays, "voxels": vox}
    for k, name in ((mi.MI_K_CONE_FP, "forward"), (mi.MI_K_CONE_BP, "backproject")):
        s = st[k]
        ms = s.total_ms / max(s.launches, 1)
        b = s.bytes / max(s.launches, 1)
        # The projectors are matrix-free: the algorithmic bytes are those of
        # the equivalent stored matrix (8 B per traced entry), so the rate
        # can exceed what HBM could stream.
        out[name] = {"launches": s.launches, "avg_ms": ms, "rays_per_s": rays / (ms * 1e-3),
                     "voxel_views_per_s": vox * g.n_views / (ms * 1e-3),
                     "algorithmic_bytes": b, "algorithmic_GBps": b / (ms * 1e-3) / 1e9,
                     "stored_matrix_time_at_peak_ms": b / (HBM_PEAK_GBS * 1e9) * 1e3}
    out["nnz"] = int(round((out["forward"]["algorithmic_bytes"] - 4.0 * (rays + vox)) / 8.0))
    out["traffic_source"] = "profiles/c4_cone_kernel_stats.md"
    if at_iters > 0:
        t0 = time.perf_counter()
        noisy = np.empty_like(sino)
        for v in range(g.n_views):   # per view, so the float64 temporaries stay small
            noisy[v] = phantom.noisy_sinogram(sino[v], 5e4, seed=4, stream=v)
        del sino
        log(f"[bench] C4: noisy sinogram (I0 5e4) {time.perf_counter() - t0:.1f} s; {at_iters} AT passes")
        p = mi.at_params(1e-3, 1e-4, 1e-2, n_subsets=20, max_iters=at_iters)
        ctx.stats_reset()
        u, v, rep = ctx.cone_at_reconstruct(g, noisy, p)
        st = {k: ctx.stats(k) for k in ks}
        sec = rep.elapsed_ms * 1e-3
        out["at"] = {"subsets": 20, "passes": rep.iterations, "elapsed_ms": rep.elapsed_ms,
                     "at_iterations_per_s": rep.iterations / sec, "rays_per_s": rays * rep.iterations / sec,
                     "data_fidelity": rep.data_fidelity,
                     "kernel_ms": {"forward": st[mi.MI_K_CONE_FP].total_ms, "backproject": st[mi.MI_K_CONE_BP].total_ms}}
    ctx.stats_select([])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-passes", type=int, default=80)      # about 10 s of the CPU port at C2
    ap.add_argument("--cpu-passes-c1", type=int, default=500)  # about 10 s at C1
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-goldenc", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="C3: keep the engine's default launch settings")
    ap.add_argument("--c3-jobs", type=int, default=64)
    ap.add_argument("--c3-iters", type=int, default=50)
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--c3-batched-only", action="store_true",
                    help="run the batched form of the 64-job workload without the one-at-a-time pass (profiling)")
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-full", action="store_true",
                    help="only the whole C5 stream (--serve-jobs 2000 by default here): batched live, plain derived")
    ap.add_argument("--helical", action="store_true",
                    help="the helical per-slice path at the clinical arc geometry (one JSON line of its own)")
    ap.add_argument("--helical-slices", type=int, default=8)
    ap.add_argument("--helical-iters", type=int, default=10)
    ap.add_argument("--k-scaling", action="store_true",
                    help="only the jobs-per-unit scaling of four C5 geometries (K = 1, 2, 4, 8)")
    ap.add_argument("--c4-reps", type=int, default=2)
    ap.add_argument("--c4-iters", type=int, default=2, help="AT passes of the C4 cone-beam solve (0: projectors only)")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="also run a plain device copy of this many GiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(independent jobs shard by GPU without any exchange)", file=sys.stderr)
        sys.exit(2)

    import medimage_amd as mi
    from medimage_amd import configs

    ctx = mi.Context(0)
    if args.k_scaling:
        print(json.dumps({"k_scaling": bench_k_scaling(ctx, lambda m: print(m, file=sys.stderr, flush=True))}))
        ctx.close()
        return
    if args.helical:
        out = bench_helical(ctx, args.helical_slices, args.helical_iters,
                            lambda m: print(m, file=sys.stderr, flush=True), 0 if args.no_cpu else 20)
        print(json.dumps({"helical_slices": out}))
        ctx.close()
        return
    if args.serve_full:
        n = args.serve_jobs if args.serve_jobs != 48 else 2000
        out = bench_serving_full(ctx, n, args.serve_rate, lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps({"c5_full_stream": out}))
        ctx.close()
        return
    g = configs.c2_geometry()
    print(f"[bench] tracing C2 geometry and generating the phantom", file=sys.stderr, flush=True)
    t = time.perf_counter()
    _, sino = configs.make_case(ctx, g, 2e4, seed=2)
    ctx.trim()
    nnz = ctx.prepare(g, 12)
    print(f"[bench] nnz {nnz}, setup {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    # every pass below runs (warm-up, instrumented, timed): none is skipped
    # by the iteration cap
    p = configs.c2_params(max_iters=args.warmup + max(args.warmup, 5) + args.steps)
    job = ctx.job(g, sino.astype(np.float32), p)
    job.enqueue(args.warmup)
    job.synchronize()

    names = {mi.MI_K_FP: "fp_residual", mi.MI_K_BP: "bp_gather_grad", mi.MI_K_FPSQ: "fp_direction_norm"}
    kernel_symbol = {mi.MI_K_FP: "k_eng_fp<float, 1, true, 1>", mi.MI_K_BP: "k_eng_bp<float, 1, true, true>",
                     mi.MI_K_FPSQ: "k_eng_fp<float, 1, false, 1>"}

    def per_kernel_stats():
        ks = {k: ctx.stats(k) for k in names}
        return ks, {names[k]: {"launches": ks[k].launches,
                               "avg_us": 1e3 * ks[k].total_ms / max(ks[k].launches, 1),
                               "GBps": (ks[k].bytes / max(ks[k].launches, 1)) /
                                       (max(ks[k].total_ms, 1e-9) / max(ks[k].launches, 1) * 1e-3) / 1e9}
                    for k in names}

    # Instrumented passes (an event pair around every projector launch) give
    # the per-kernel table and pick the dominant kernel; the timed region then
    # records events around the dominant kernel's launches only, so the
    # instrumentation does not slow the run it measures.
    ctx.stats_reset()
    ctx.stats_enable(True)
    job.enqueue(max(args.warmup, 5))
    job.synchronize()
    ctx.stats_enable(False)
    ks, per_kernel = per_kernel_stats()
    dom = max(ks, key=lambda k: ks[k].total_ms)

    ctx.stats_reset()
    ctx.stats_select([dom])
    job.synchronize()
    t0 = time.perf_counter()
    job.enqueue(args.steps)
    job.synchronize()
    t1 = time.perf_counter()
    ctx.stats_select([])
    elapsed = t1 - t0
    rep = job.report()
    print(f"[bench] {args.steps} AT iterations in {elapsed * 1e3:.1f} ms; data fidelity {rep.data_fidelity:.6e}",
          file=sys.stderr, flush=True)
    total_passes = rep.iterations
    u32, v32 = job.read()
    s = ctx.stats(dom)
    avg_ms = s.total_ms / max(s.launches, 1)
    bytes_per = s.bytes / max(s.launches, 1)
    achieved = bytes_per / (avg_ms * 1e-3) / 1e9

    # HBM bytes per launch of the same kernel on the same workload, from the
    # rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated against a plain copy
    # kernel) recorded next to the kernel statistics.
    traffic = None
    tfile = ROOT / "profiles" / "c2_plain_traffic.json"
    if tfile.exists():
        tk = json.loads(tfile.read_text())["kernels"].get(kernel_symbol[dom])
        if tk:
            traffic = tk["traffic_bytes"]
    rays = g.n_views * g.n_bins
    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "AT iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 10-ellipse phantom, Poisson noise I0=2e4, log transform",
        "config": {"workload": "C2 plain per-job form: 512x512 fan beam (0.7 mm), 720 views, 736 bins (1.0 mm), "
                               "SOD 541 mm, SDD 949 mm, 12 ordered subsets, alpha 1e-3, beta 1e-4, eps 1e-2",
                   "rays": rays, "nnz": nnz, "subsets": 12},
        "rays_per_s": rays * args.steps / elapsed,
        "roofline": {"bound": "hbm", "kernel": kernel_symbol[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/c2_plain_traffic.json" if traffic else None,
                     "bytes_per_launch": bytes_per, "avg_launch_us": avg_ms * 1e3,
                     "launches_timed": s.launches},
        "kernels": per_kernel,
    }
    # The benched result's error: the same job, the same number of passes,
    # run in FP64 throughout on the device (the full-FP64 GPU oracle of the
    # C2 form; its agreement with the CPU FP64 oracle is a GPU test).
    job.close()
    p64 = configs.c2_params(max_iters=total_passes, precision=mi.MI_FP64)
    u64, v64, rep64 = ctx.at_reconstruct(g, sino.astype(np.float64), p64)

    def rel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        return (float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)),
                float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)))
    eu, ev = rel(u32, u64), rel(v32, v64)
    out["accuracy"] = {"reference": f"the same job in FP64 on the device, {rep64.iterations} passes",
                       "u_rel_l2": eu[0], "u_rel_max": eu[1], "v_rel_l2": ev[0], "v_rel_max": ev[1],
                       "tolerance": 1e-5, "within": max(eu + ev) <= 1e-5}
    print(f"[bench] FP32 vs FP64 after {total_passes} passes: u {eu[0]:.2e} / {eu[1]:.2e}, "
          f"v {ev[0]:.2e} / {ev[1]:.2e}", file=sys.stderr, flush=True)
    sino_c1 = None
    if not args.no_c1:
        out["c1"], sino_c1 = c1_gpu(ctx, 50)
        out["msbeam_parallel"] = msbeam_gpu(ctx, 50)
    if not args.no_goldenc:
        out["goldenc"] = bench_goldenc(ctx, lambda m: print(m, file=sys.stderr, flush=True), not args.no_cpu)
    if args.copy_gib > 0:
        nb = int(args.copy_gib * (1 << 30)) // 16 * 16
        ms = ctx.device_copy(nb, 3)
        out["copy_reference"] = {"kernel": "k_copy16", "bytes_read": nb, "bytes_written": nb, "launches": 3,
                                 "ms": ms, "GBps": 2 * nb / (ms * 1e-3) / 1e9}
    if not args.no_c4:
        out["cone_c4"] = bench_c4(ctx, args.c4_reps, args.c4_iters, lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_c3:
        out["batched"] = bench_c3(ctx, args.c3_jobs, args.c3_iters,
                                  lambda m: print(m, file=sys.stderr, flush=True), not args.c3_batched_only,
                                  not args.no_tune)
    if not args.no_serve:
        out["serving"] = bench_serving(ctx, args.serve_jobs, args.serve_rate,
                                       lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_cpu and args.cpu_passes > 0:
        print(f"[bench] CPU baseline: {args.cpu_passes} AT iterations of the CPU port", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sino.astype(np.float32), args.cpu_passes, sino_c1, args.cpu_passes_c1)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
