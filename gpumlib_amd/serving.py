"""Serving workload: a Poisson stream of NMF jobs of uneven sizes, run through the
continuous-batching server and, for comparison, one job at a time through the
single-problem path with the same stopping rule.

Job mix (the batched-serving configuration): Euclidean and KL jobs 50/50; m
log-uniform in [64, 4096], n log-uniform in [64, 2048], r from {8, 16, 32, 64};
V = W0 H0 + U[0, 0.01) noise (Euclidean) or Poisson counts with rate W0 H0
(KL), from the seeded generators; each job runs until the relative objective
change over one iteration drops below 1e-4, or 500 iterations.

Stopping rule, identical on both paths: iterations run in rounds of R; the FP64
objective is taken on entry to the first two iterations of a round, and the job
stops after the round when their relative difference is below the tolerance or
the iteration cap is reached.
"""
import ctypes
import time

import numpy as np

from . import capi
from . import nmf as G


def make_jobs(count, seed=0x5E7E, n_range=(64, 2048), m_range=(64, 4096), ranks=(8, 16, 32, 64)):
    """Job descriptions (rule, n, m, r, data seed), drawn from a seeded RNG."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(count):
        rule = G.EUCLIDEAN if rng.random() < 0.5 else G.KL
        n = int(round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        m = int(round(np.exp(rng.uniform(np.log(m_range[0]), np.log(m_range[1])))))
        r = int(rng.choice(ranks))
        jobs.append((rule, n, m, r, seed * 1000003 + k))
    return jobs


def job_data(job):
    """V, W, H of one job (flat column-major float32)."""
    rule, n, m, r, s = job
    V = G.gen_euclidean(n, m, r, s) if rule == G.EUCLIDEAN else G.gen_poisson(n, m, r, s)
    W, H = G.gen_init(n, m, r, s)
    return V, W, H


def iteration_work(rule, n, m, r):
    """Algorithmic work of one iteration of a job: FLOPs (Euclidean 4nmr + 4r^2(n+m), KL
    8nmr) and bytes (the FP32 V streamed twice, W and H each read and written once)."""
    flops = 4.0 * n * m * r + 4.0 * r * r * (n + m) if rule == G.EUCLIDEAN else 8.0 * n * m * r
    return flops, 8.0 * n * m + 8.0 * (n + m) * r


def _aggregate(jobs, iters, span):
    """Aggregate rates over the span from the jobs' logged iteration counts (the
    stopping rule's objective evaluations are not counted as work)."""
    flops = sum(iteration_work(j[0], j[1], j[2], j[3])[0] * it for j, it in zip(jobs, iters))
    nbytes = sum(iteration_work(j[0], j[1], j[2], j[3])[1] * it for j, it in zip(jobs, iters))
    return {"iterations_total": int(sum(iters)), "aggregate_tflops": flops / span / 1e12,
            "aggregate_gbs": nbytes / span / 1e9}


def _percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if xs else None


def run_server(lib, jobs, data, rate, tol=1e-4, max_iters=500, budget=32 << 30, max_active=1024, round_iters=10,
               use_graph=1, autotune=0):
    """Submits the jobs with exponential inter-arrival times (mean 1/rate) and waits for all.

    autotune > 0 lets the server run the batch autotuner (that many timed iterations per
    candidate) whenever the number of resident jobs doubles or halves.
    Returns throughput (jobs/s over first submit .. last finish), latency percentiles
    (finish - submit, seconds) and the server's counters."""
    cfg = capi.ServerConfig(budget, max_active, round_iters, use_graph, 0, autotune)
    h = ctypes.c_void_p()
    capi.check(lib.gpumlib_nmf_server_create(ctypes.byref(cfg), ctypes.byref(h)), "gpumlib_nmf_server_create")
    rng = np.random.default_rng(12345)
    gaps = rng.exponential(1.0 / rate, size=len(jobs)) if rate > 0 else np.zeros(len(jobs))
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    ids = []
    lat, iters, conv, first, last = [], [], 0, None, None
    st = capi.ServerStats()
    try:
        due = time.perf_counter()
        for job, (V, W, H), gap in zip(jobs, data, gaps):
            due += gap
            wait = due - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            rule, n, m, r, _ = job
            eps = G.REFERENCE_EPS if rule == G.EUCLIDEAN else 1e-9
            jid = ctypes.c_longlong()
            capi.check(lib.gpumlib_nmf_server_submit(h, rule, n, m, r, ptr(V), ptr(W), ptr(H), eps, tol, max_iters,
                                                     ctypes.byref(jid)), "gpumlib_nmf_server_submit")
            ids.append(jid.value)
        for jid in ids:
            res = capi.JobResult()
            capi.check(lib.gpumlib_nmf_server_wait(h, jid, None, None, ctypes.byref(res)), "gpumlib_nmf_server_wait")
            lat.append(res.finish_s - res.submit_s)
            iters.append(res.iterations)
            conv += res.converged
            first = res.submit_s if first is None else min(first, res.submit_s)
            last = res.finish_s if last is None else max(last, res.finish_s)
        capi.check(lib.gpumlib_nmf_server_stats_get(h, ctypes.byref(st)), "gpumlib_nmf_server_stats_get")
    finally:
        lib.gpumlib_nmf_server_destroy(h)  # finishes what was submitted, then releases the server
    span = last - first
    return {"jobs": len(ids), "jobs_per_s": len(ids) / span, "span_s": span,
            "latency_p50_s": _percentile(lat, 50), "latency_p99_s": _percentile(lat, 99),
            "iterations_mean": float(np.mean(iters)), **_aggregate(jobs, iters, span), "converged": conv,
            "rounds": st.rounds, "peak_active": st.peak_active, "pool_high_water_bytes": st.pool_high_water,
            "arrival_rate_per_s": rate, "round_iterations": round_iters, "graph": bool(use_graph),
            "autotune_trial_iterations": autotune, "autotunes": st.autotunes,
            "tuned": {"target_workgroups": st.tuned_target, "pass_grid": st.tuned_pass_grid,
                      "update_grid": st.tuned_update_grid} if st.autotunes else None}


def run_one_at_a_time(jobs, data, tol=1e-4, max_iters=500, round_iters=10, rate=0.0, progress=None):
    """The same jobs launched one after another through gpumlib_nmf_iterate (one
    problem, its own buffers), with the server's stopping rule and cadence.

    Latency is reported for the same Poisson arrival stream as the batched run
    (arrival times from the same seed): a job starts when it has arrived and the
    previous one has finished, and takes its measured service time."""
    lat, iters = [], []
    t0 = time.perf_counter()
    for job, (V, W, H) in zip(jobs, data):
        s0 = time.perf_counter()
        rule, n, m, r, _ = job
        eps = G.REFERENCE_EPS if rule == G.EUCLIDEAN else 1e-9
        kind = G.OBJ_FROBENIUS if rule == G.EUCLIDEAN else G.OBJ_KL
        d = G.DeviceNMF(rule, V, W, H, n, m, r, eps=eps)
        done = 0
        while done < max_iters:
            f0 = d.iterate(1, objective=kind)
            f1 = d.iterate(1, objective=kind)
            d.iterate(round_iters - 2)
            done += round_iters
            if f0 != 0 and abs(f0 - f1) < tol * abs(f0):
                break
        d.get_w()
        d.get_h()
        lat.append(time.perf_counter() - s0)
        iters.append(done)
        if progress and len(lat) % 256 == 0:
            progress(len(lat))
    span = time.perf_counter() - t0
    rng = np.random.default_rng(12345)
    arrive = np.cumsum(rng.exponential(1.0 / rate, size=len(jobs))) if rate > 0 else np.zeros(len(jobs))
    finish, queued = 0.0, []
    for a, svc in zip(arrive, lat):
        finish = max(finish, a) + svc
        queued.append(finish - a)
    return {"jobs": len(jobs), "jobs_per_s": len(jobs) / span, "span_s": span,
            "service_p50_s": _percentile(lat, 50), "service_p99_s": _percentile(lat, 99),
            "latency_p50_s": _percentile(queued, 50), "latency_p99_s": _percentile(queued, 99),
            "iterations_mean": float(np.mean(iters)), **_aggregate(jobs, iters, span), "arrival_rate_per_s": rate}
