"""Serving workload: a Poisson stream of NMF jobs of uneven sizes, run through the
continuous-batching server and, for comparison, through the same server's one-at-a-time
mode (each job alone through the single-problem iteration) with the same stopping rule.

Job mix (the batched-serving configuration): Euclidean and KL jobs 50/50; in GPUMLib's
naming (V is n x m: n rows, m columns, nmf_multiplicative_euclidian.h:47) m log-uniform in
[64, 4096] columns and n log-uniform in [64, 2048] rows, r from {8, 16, 32, 64};
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


def make_data(jobs, threads=16):
    """job_data of every job, generated on a thread pool (the generators are native and
    release the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(job_data, jobs))


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
               use_graph=1, autotune=0, split_k=0, sequential=0):
    """Submits the jobs with exponential inter-arrival times (mean 1/rate) and waits for all.

    autotune > 0 lets the server run the batch autotuner (that many timed iterations per
    candidate) whenever the number of resident jobs doubles or halves.  sequential = 1 runs
    the server's one-at-a-time mode (each job alone through the single-problem iteration,
    same rounds and stopping rule).
    Returns throughput (jobs/s over first submit .. last finish), latency percentiles
    (finish - submit, seconds), per-job service times (finish - admit) and the server's
    counters."""
    cfg = capi.ServerConfig(budget, max_active, round_iters, use_graph, 0, autotune, split_k, sequential)
    h = ctypes.c_void_p()
    capi.check(lib.gpumlib_nmf_server_create(ctypes.byref(cfg), ctypes.byref(h)), "gpumlib_nmf_server_create")
    rng = np.random.default_rng(12345)
    gaps = rng.exponential(1.0 / rate, size=len(jobs)) if rate > 0 else np.zeros(len(jobs))
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    ids = []
    lat, svc, iters, conv, first, last = [], [], [], 0, None, None
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
            svc.append(res.finish_s - res.admit_s)
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
            "service_p50_s": _percentile(svc, 50), "service_p99_s": _percentile(svc, 99),
            "_service_s": svc,
            "iterations_mean": float(np.mean(iters)), **_aggregate(jobs, iters, span), "converged": conv,
            "rounds": st.rounds, "peak_active": st.peak_active, "pool_high_water_bytes": st.pool_high_water,
            "arrival_rate_per_s": rate, "round_iterations": round_iters, "graph": bool(use_graph),
            "max_active": 1 if sequential else max_active,
            "engine": "one_at_a_time" if sequential else ("split_k" if split_k else "fused"),
            "autotune_trial_iterations": autotune, "autotunes": st.autotunes,
            "tuned": _tuned_report(st)
            if st.autotunes else None}


_BUCKETS = ("V<2^18", "V<2^21", "V>=2^21")


def _tuned_report(st):
    """The batch configuration the autotuner chose last: the engine, the grids, and per shape bucket
    with problems the pass tile shape -- K-split target (split-K engine), tile width, rank blocking."""
    fused = st.tuned_engine == capi.BATCH_FUSED
    buckets = {}
    for name, t, w, sp in zip(_BUCKETS, st.tuned_bucket_targets, st.tuned_bucket_tile_cols,
                              st.tuned_bucket_split_ranks):
        if not w:
            continue
        buckets[name] = {"k_split": "whole contraction per item" if fused else f"target {t} workgroups",
                         "tile_cols": w,
                         "rank_blocking": "r<=32 on 1-block kernels" if sp else "2-block kernels"}
    return {"engine": "fused" if fused else "split_k", "pass_grid": st.tuned_pass_grid,
            "update_grid": st.tuned_update_grid, "tile_shapes": buckets}


def queued_latency(service_s, rate, seed=12345):
    """Latency percentiles of one FIFO server with the given per-job service times (seconds,
    measured) under the Poisson arrival stream of run_server (same seed): job k starts when it
    has arrived and job k-1 has finished."""
    rng = np.random.default_rng(seed)
    arrive = np.cumsum(rng.exponential(1.0 / rate, size=len(service_s))) if rate > 0 else np.zeros(len(service_s))
    finish, lat = 0.0, []
    for a, s in zip(arrive, service_s):
        finish = max(finish, a) + s
        lat.append(finish - a)
    return {"latency_p50_s": _percentile(lat, 50), "latency_p99_s": _percentile(lat, 99),
            "span_s": finish - float(arrive[0]) if len(arrive) else 0.0}
