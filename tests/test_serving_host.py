"""Host-side checks of the serving workload's bookkeeping (no GPU)."""
import pytest

from gpumlib_amd import nmf as G
from gpumlib_amd import serving


@pytest.mark.parametrize("rule,n,m,r,gflop", [
    (G.EUCLIDEAN, 8192, 4096, 64, 8.79),    # 8192 x 4096, r = 64
    (G.KL, 4096, 4096, 32, 4.29),           # KL 4096 x 4096, r = 32
    (G.EUCLIDEAN, 16384, 8192, 64, 34.76),  # 16384 x 8192, r = 64
])
def test_iteration_work_matches_published_counts(rule, n, m, r, gflop):
    flops, nbytes = serving.iteration_work(rule, n, m, r)
    assert flops / 1e9 == pytest.approx(gflop, abs=0.005)
    # two FP32 passes over V dominate; W and H add (n + m) r terms only
    assert 8.0 * n * m <= nbytes <= 8.0 * n * m * 1.05


def test_aggregate_rates():
    jobs = [(G.EUCLIDEAN, 100, 50, 8, 0), (G.KL, 30, 20, 4, 1)]
    agg = serving._aggregate(jobs, [10, 30], 2.0)
    f0, b0 = serving.iteration_work(G.EUCLIDEAN, 100, 50, 8)
    f1, b1 = serving.iteration_work(G.KL, 30, 20, 4)
    assert agg["iterations_total"] == 40
    assert agg["aggregate_tflops"] == pytest.approx((10 * f0 + 30 * f1) / 2.0 / 1e12)
    assert agg["aggregate_gbs"] == pytest.approx((10 * b0 + 30 * b1) / 2.0 / 1e9)


def test_make_jobs_ranges_and_seed():
    jobs = serving.make_jobs(200)
    assert jobs == serving.make_jobs(200)  # seeded
    for rule, n, m, r, _ in jobs:
        assert rule in (G.EUCLIDEAN, G.KL)
        assert 64 <= n <= 2048 and 64 <= m <= 4096 and r in (8, 16, 32, 64)
    kinds = [j[0] for j in jobs]
    assert 0.35 < kinds.count(G.EUCLIDEAN) / len(kinds) < 0.65


def test_queued_latency_fifo():
    """One FIFO server: with every arrival after the previous finish the latency is the service
    time; with all arrivals at once it is the running sum of service times."""
    import numpy as np

    svc = [0.5, 0.25, 1.0, 0.125]
    burst = serving.queued_latency(svc, rate=0.0)  # all arrive at t = 0
    assert burst["latency_p50_s"] == pytest.approx(np.percentile(np.cumsum(svc), 50))
    assert burst["latency_p99_s"] == pytest.approx(np.percentile(np.cumsum(svc), 99))
    slow = serving.queued_latency([1e-6] * 50, rate=10.0)  # service far below the mean gap
    assert slow["latency_p99_s"] < 1e-3
    # the same arrival stream as run_server: exponential gaps from seed 12345
    rng = np.random.default_rng(12345)
    arrive = np.cumsum(rng.exponential(1.0 / 3.0, size=4))
    fin, lat = 0.0, []
    for a, s in zip(arrive, svc):
        fin = max(fin, a) + s
        lat.append(fin - a)
    assert serving.queued_latency(svc, rate=3.0)["latency_p50_s"] == pytest.approx(np.percentile(lat, 50))
