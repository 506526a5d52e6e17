"""Request loop over the batched engine: jobs arrive on a clock, are
admitted into a batch between passes while device memory allows (the batch
evicts idle cached matrices first), advance together, and leave with their
result on the host.  Latency is arrival -> converged image on the host.

`mode="plain"` serves the same stream one job at a time (first come, first
served; a job starts when the previous one has left), the per-job form the
batched one is compared with.
"""
from __future__ import annotations

import time

import numpy as np

from . import MI_BATCH_MAX_ENTRIES, MI_FP32


def run_stream(ctx, requests, n_subsets=8, mode="batched", passes_per_step=1, log=None, max_entries=400_000_000,
               keep_results=True):
    """requests: list of dicts with keys t (arrival, seconds from start),
    geom, sino, params.  Returns a dict with per-job latency, iterations and
    the totals.  max_entries (batched mode): admission budget in matrix
    entries of the active jobs (MI_BATCH_MAX_ENTRIES; 0 admits whatever fits
    in memory).  The default, about one and a half 512^2 x 720-view jobs,
    keeps one MI355X busy; later arrivals wait in order instead of slowing
    every job down (measured on the 48-request stream of the bench: p50
    4.4 s against 7.3 s with no budget, at the same throughput)."""
    order = sorted(range(len(requests)), key=lambda i: requests[i]["t"])
    b = ctx.batch(n_subsets, MI_FP32)
    if mode == "batched" and max_entries:
        b.set(MI_BATCH_MAX_ENTRIES, int(max_entries))
    t0 = time.perf_counter()
    nxt = 0
    inflight = {}
    lat = np.zeros(len(requests))
    iters = np.zeros(len(requests), np.int64)
    results = {}
    finished = 0
    steps = 0
    while finished < len(requests):
        now = time.perf_counter() - t0
        while nxt < len(order) and requests[order[nxt]]["t"] <= now and (mode == "batched" or not inflight):
            r = requests[order[nxt]]
            inflight[b.submit(r["geom"], r["sino"], r["params"])] = order[nxt]
            nxt += 1
        if not inflight:
            time.sleep(max(0.0, requests[order[nxt]]["t"] - (time.perf_counter() - t0)))
            continue
        b.step(passes_per_step)
        steps += 1
        done_at = time.perf_counter() - t0
        for jid in b.retired():
            i = inflight.pop(jid)
            st = b.status(jid)
            lat[i] = done_at - requests[i]["t"]
            iters[i] = st.iterations
            if keep_results:
                results[i] = b.read(jid)
            b.release(jid)
            finished += 1
            if log and finished % 10 == 0:
                log(f"[serve {mode}] {finished}/{len(requests)} done at {done_at:.1f} s")
    makespan = time.perf_counter() - t0
    b.close()
    return {
        "mode": mode,
        "jobs": len(requests),
        "makespan_s": makespan,
        "jobs_per_s": len(requests) / makespan,
        "iters_per_s": float(iters.sum()) / makespan,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_mean_s": float(lat.mean()),
        "mean_iterations": float(iters.mean()),
        "steps": steps,
        "latency": lat,
        "iterations": iters,
        "results": results,
    }
