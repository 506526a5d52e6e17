"""The batched (C3) and served (C5) workloads themselves on the GPU, at
their real shapes (fewer AT iterations than the benchmark runs):

* C3: the 64 mixed fan-beam slices of configs.c3_jobs(64) (17 geometries,
  N in {256, 384, 512}, views in {360, 720, 1152}, bins in {512, 736}),
  8 subsets: every job bit for bit the same alone and in the one batch; the
  two smallest jobs against the CPU FP64 oracle within 1e-5;
* C5: a sample of the serving stream configs.c5_requests(2000) that
  includes its largest geometry (1024 x 1024, 2304 views, 1536 bins),
  served by the request loop (continuous batching, entry budget) and one
  at a time: every result bit for bit the job's result alone.
"""
import numpy as np
import pytest

import medimage_amd as mi
from medimage_amd import configs
import oracle_py as O

pytestmark = pytest.mark.gpu


def _same(x, y):
    return x.dtype == y.dtype and np.array_equal(x.view(np.uint8), y.view(np.uint8))


def _alone(ctx, g, sino, p, n_subsets):
    b = ctx.batch(n_subsets)
    jid = b.submit(g, sino, p)
    b.run()
    u, v = b.read(jid)
    st = b.status(jid)
    b.close()
    return u, v, st


def test_c3_batch_of_64_mixed_slices(ctx):
    spec = configs.c3_jobs(64, seed=3)
    jobs = []
    for j in spec:
        g = configs.mixed_geometry(j["n"], j["views"], j["bins"])
        _, sino = configs.make_case(ctx, g, j["i0"], j["seed"])
        jobs.append((g, sino, j))
    print(f"[c3] {len(jobs)} jobs over {len({g.key() for g, _, _ in jobs})} geometries", flush=True)
    iters = 3
    p = configs.at_params(configs.ALPHA, configs.BETA, configs.EPS, n_subsets=8, max_iters=iters)
    b = ctx.batch(8)
    ids = [b.submit(g, s.astype(np.float32), p) for g, s, _ in jobs]
    b.run()
    batched = [b.read(i) for i in ids]
    units = b.units()
    b.close()
    print(f"[c3] batch done", flush=True)
    for k, (g, s, j) in enumerate(jobs):
        u, v, st = _alone(ctx, g, s.astype(np.float32), p, 8)
        assert st.iterations == iters
        assert _same(u, batched[k][0]) and _same(v, batched[k][1]), f"job {k} {j}"
    print(f"[c3] 64 jobs bitwise alone == batched", flush=True)
    small = sorted(range(len(jobs)), key=lambda k: (jobs[k][2]["n"], jobs[k][2]["views"]))[:2]
    for k in small:
        g, s, j = jobs[k]
        s64 = s.astype(np.float32).astype(np.float64)
        ou, ov, n, _ = O.os_at(O.op_fan(O.fan_from(g)), s64, configs.ALPHA, configs.BETA, configs.EPS, 8, iters)
        u, v = batched[k]
        eu = float(np.linalg.norm(u - ou) / np.linalg.norm(ou))
        mu = float(np.abs(u - ou).max() / np.abs(ou).max())
        ev = float(np.linalg.norm(v - ov) / np.linalg.norm(ov))
        print(f"[c3 job {k} {j['n']}^2 x {j['views']} x {j['bins']}] rel L2 u {eu:.2e} max {mu:.2e} v {ev:.2e}",
              flush=True)
        assert eu < 1e-5 and mu < 1e-5 and ev < 1e-5


def test_c5_stream_sample_with_largest_geometry(ctx):
    from medimage_amd import serve
    spec = configs.c5_requests(2000, seed=5, rate=20.0)
    big = next(r for r in spec if r["n"] == 1024 and r["views"] == 2304)
    sample = spec[:7] + [big]
    t0 = sample[0]["t"]
    reqs = []
    for k, r in enumerate(sample):
        g = configs.mixed_geometry(r["n"], r["views"], r["bins"])
        sino = configs.analytic_case(g, r["i0"], r["seed"]).astype(np.float32)
        p = configs.at_params(configs.ALPHA, configs.BETA, configs.EPS, n_subsets=8, max_iters=4, rel_tol=1e-4)
        # arrivals 5 ms apart, so later jobs join while earlier ones run
        reqs.append(dict(t=0.005 * k, geom=g, sino=sino, params=p))
        print(f"[c5 sample {k}] {r['n']}^2 x {r['views']} views x {r['bins']} bins (t {r['t'] - t0:.2f} s)",
              flush=True)
    bat = serve.run_stream(ctx, reqs, n_subsets=8, mode="batched")
    pla = serve.run_stream(ctx, reqs, n_subsets=8, mode="plain")
    print(f"[c5] batched p50 {bat['latency_p50_s']:.2f} s p99 {bat['latency_p99_s']:.2f} s; "
          f"plain p50 {pla['latency_p50_s']:.2f} s p99 {pla['latency_p99_s']:.2f} s", flush=True)
    for k, r in enumerate(reqs):
        u, v, st = _alone(ctx, r["geom"], r["sino"], r["params"], 8)
        for res in (bat, pla):
            ru, rv = res["results"][k]
            assert _same(ru, u) and _same(rv, v), f"request {k} ({res['mode']})"
        assert bat["iterations"][k] == st.iterations
    ctx.trim()
