"""Grouped (batched) execution and the continuous-batching server.

* every problem of a mixed batch matches the FP64 oracle;
* a problem's W and H are bitwise the same alone and in a mix;
* hipGraph replays equal eager launches, replay after replay, with eager work
  and synchronisations in between;
* the server admits by HBM footprint, retires at convergence, and returns the
  same bits for a problem alone and in a mix of concurrent problems.
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_fro
from gpumlib_amd import capi
from gpumlib_amd import nmf as G

pytestmark = pytest.mark.gpu

JOBS = [  # rule, n, m, r
    (G.EUCLIDEAN, 300, 200, 16),
    (G.KL, 250, 130, 8),
    (G.EUCLIDEAN, 97, 61, 5),
    (G.KL, 512, 384, 32),
    (G.EUCLIDEAN, 1000, 700, 64),
    (G.KL, 64, 2048, 64),
]


def make_job(k, rule, n, m, r):
    V = G.gen_euclidean(n, m, r, 100 + k) if rule == G.EUCLIDEAN else G.gen_poisson(n, m, r, 100 + k)
    W, H = G.gen_init(n, m, r, 200 + k)
    return V, W, H


ENGINES = [capi.BATCH_FUSED, capi.BATCH_SPLIT_K]

EDGE_JOBS = [  # odd shapes for the fused engine's tiles: partial tiles, unaligned rows, long contractions
    (G.EUCLIDEAN, 1, 7, 1),
    (G.KL, 130, 257, 33),
    (G.EUCLIDEAN, 2048, 70, 64),
    (G.KL, 50, 4096, 17),
    (G.EUCLIDEAN, 333, 129, 63),
    (G.KL, 129, 1000, 64),
    (G.EUCLIDEAN, 90, 3000, 40),  # fused engine: the H gram in 3 K slices (the W gram of 2048 rows in 2)
]


class Batch:
    def __init__(self, hip, lib, max_jobs=16, engine=None, rank_split=None, tile_cols=None):
        self.hip, self.lib = hip, lib
        h = ctypes.c_void_p()
        capi.check(lib.gpumlib_nmf_batch_create(max_jobs, ctypes.byref(h)))
        self.h = h
        if engine is not None:
            capi.check(lib.gpumlib_nmf_batch_set_engine(h, engine))
            assert lib.gpumlib_nmf_batch_engine(h) == engine
        if rank_split is not None:
            capi.check(lib.gpumlib_nmf_batch_set_rank_split(h, rank_split))
            assert lib.gpumlib_nmf_batch_rank_split(h) == rank_split
        if tile_cols is not None:
            capi.check(lib.gpumlib_nmf_batch_set_tile_cols(h, tile_cols))
        self.keep = []
        self.stream = hip.Stream()

    def add(self, rule, n, m, r, V, W, H, eps=G.REFERENCE_EPS):
        dV, dW, dH = (self.hip.Buffer.from_array(np.ascontiguousarray(x)) for x in (V, W, H))
        wsb = int(self.lib.gpumlib_nmf_batch_workspace_bytes(self.h, rule, n, m, r))
        ws = self.hip.Buffer(wsb)
        slot = ctypes.c_int()
        capi.check(self.lib.gpumlib_nmf_batch_add(self.h, rule, n, m, r, dV.ptr, dW.ptr, dH.ptr, eps, ws.ptr, wsb,
                                                  ctypes.byref(slot)))
        self.keep.append((dV, dW, dH, ws))
        return slot.value, dW, dH

    def iterate(self, iters, objectives=0, graph=False):
        fn = self.lib.gpumlib_nmf_batch_graph_launch if graph else self.lib.gpumlib_nmf_batch_iterate
        capi.check(fn(self.h, iters, objectives, self.stream.handle))
        self.stream.synchronize()

    def objectives(self, max_jobs=16):
        out = (ctypes.c_double * (2 * max_jobs))()
        capi.check(self.lib.gpumlib_nmf_batch_objectives(self.h, out, self.stream.handle))
        return np.array(out)

    def close(self):
        self.hip.synchronize()
        self.lib.gpumlib_nmf_batch_destroy(self.h)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("jobs", ["mix", "edge"])
@pytest.mark.parametrize("rank_split", [0, 1])
@pytest.mark.parametrize("tile_cols", [128, 64])
def test_batch_matches_oracle(native, orc, cuda, engine, jobs, rank_split, tile_cols):
    b = Batch(cuda, native, engine=engine, rank_split=rank_split, tile_cols=tile_cols)
    outs = []
    for k, (rule, n, m, r) in enumerate(JOBS if jobs == "mix" else EDGE_JOBS):
        V, W, H = make_job(k, rule, n, m, r)
        slot, dW, dH = b.add(rule, n, m, r, V, W, H)
        outs.append((rule, n, m, r, V, W, H, dW, dH))
    b.iterate(25)
    for rule, n, m, r, V, W, H, dW, dH in outs:
        v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W, H))
        (orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter)(n, m, r, v, w, h, G.REFERENCE_EPS, iters=25)
        print(f"case rule={rule} n={n} m={m} r={r}", flush=True)
        assert rel_fro(dW.download(np.float32), w) < 1e-4
        assert rel_fro(dH.download(np.float32), h) < 1e-4
    b.close()


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("k", [3, 4])
def test_alone_equals_mix(native, cuda, engine, k):
    rule, n, m, r = JOBS[k]
    V, W, H = make_job(k, rule, n, m, r)
    b1 = Batch(cuda, native, engine=engine)
    _, w1, h1 = b1.add(rule, n, m, r, V, W, H)
    b1.iterate(15)
    b2 = Batch(cuda, native, engine=engine)
    for kk, job in enumerate(JOBS):
        if kk == k:
            _, w2, h2 = b2.add(rule, n, m, r, V, W, H)
        else:
            b2.add(*job, *make_job(kk, *job))
    b2.iterate(15)
    assert np.array_equal(w1.download(np.float32), w2.download(np.float32))
    assert np.array_equal(h1.download(np.float32), h2.download(np.float32))
    b1.close()
    b2.close()


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("tile_cols", [128, 64])
def test_rank_split_is_bit_identical(native, cuda, engine, tile_cols):
    """Problems with r <= 32 on the one-rank-block kernels give the bits of the two-block
    kernels (the same products in the same order), eager and graph-replayed, at either tile
    width."""
    outs = []
    for split in (0, 1):
        b = Batch(cuda, native, engine=engine, rank_split=split, tile_cols=tile_cols)
        bufs = []
        for k, job in enumerate(JOBS + EDGE_JOBS):
            _, dW, dH = b.add(*job, *make_job(k, *job))
            bufs.append((dW, dH))
        b.iterate(6, objectives=2)
        b.iterate(6, objectives=1, graph=True)
        outs.append(([(w.download(np.float32), h.download(np.float32)) for w, h in bufs],
                     b.objectives()[:2 * len(bufs)]))
        b.close()
    for (w0, h0), (w1, h1) in zip(outs[0][0], outs[1][0]):
        assert np.array_equal(w0, w1) and np.array_equal(h0, h1)
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("tile_cols", [128, 64])
def test_graph_replays_equal_eager(native, cuda, engine, tile_cols):
    eager = Batch(cuda, native, engine=engine, tile_cols=tile_cols)
    graphed = Batch(cuda, native, engine=engine, tile_cols=tile_cols)
    pairs = []
    for k, job in enumerate(JOBS[:4]):
        V, W, H = make_job(k, *job)
        _, we, he = eager.add(*job, V, W, H)
        _, wg, hg = graphed.add(*job, V, W, H)
        pairs.append((we, he, wg, hg))
    for replay in range(4):
        eager.iterate(5, objectives=2)
        graphed.iterate(5, objectives=2, graph=True)
        # eager work and synchronisations between replays
        x = cuda.Buffer.from_array(np.full(1000, replay, np.float32))
        y = cuda.Buffer(4000)
        y.copy_from(x)
        cuda.synchronize()
        assert float(y.download(np.float32).sum()) == 1000.0 * replay
        for we, he, wg, hg in pairs:
            assert np.array_equal(we.download(np.float32), wg.download(np.float32)), replay
            assert np.array_equal(he.download(np.float32), hg.download(np.float32)), replay
        assert np.array_equal(eager.objectives()[:8], graphed.objectives()[:8])
    eager.close()
    graphed.close()


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("tile_cols", [128, 64])
def test_batch_objectives_match_single(native, orc, cuda, engine, tile_cols):
    b = Batch(cuda, native, engine=engine, tile_cols=tile_cols)
    slots = []
    for k, job in enumerate(JOBS[:3]):
        V, W, H = make_job(k, *job)
        slots.append((b.add(*job, V, W, H)[0], job, V, W, H))
    b.iterate(2, objectives=2)
    obj = b.objectives()
    for slot, (rule, n, m, r), V, W, H in slots:
        v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W, H))
        f = orc.frob if rule == G.EUCLIDEAN else orc.kl_div
        f0 = f(n, m, r, v, w, h)
        (orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter)(n, m, r, v, w, h, G.REFERENCE_EPS)
        f1 = f(n, m, r, v, w, h)
        assert abs(obj[2 * slot] - f0) / f0 < 1e-5
        assert abs(obj[2 * slot + 1] - f1) / f1 < 1e-5
    b.close()


# --------------------------------------------------------------- autotuner --
TUNE_TARGETS, TUNE_PASS, TUNE_UPD = {4, 8, 16, 32, 64, 128}, {512, 1024, 2048}, {256, 512, 1024}


def test_autotune_keeps_state_and_parity(native, orc, cuda):
    """The autotuner times candidate K-split depths and grid sizes on the live mix; W and H
    are restored after every trial (bitwise unchanged by the call), the choice is cached
    per mix, and the tuned configuration still matches the FP64 oracle."""
    b = Batch(cuda, native)
    outs = []
    for k, (rule, n, m, r) in enumerate(JOBS):
        V, W, H = make_job(k, rule, n, m, r)
        slot, dW, dH = b.add(rule, n, m, r, V, W, H)
        outs.append((rule, n, m, r, V, W, H, dW, dH))
    b.iterate(3)
    before = [(dW.download(np.float32), dH.download(np.float32)) for *_, dW, dH in outs]
    t = capi.BatchTuning()
    capi.check(native.gpumlib_nmf_batch_autotune(b.h, 2, b.stream.handle, ctypes.byref(t)))
    print(f"tuned engine={t.engine} target={t.target_workgroups} pass_grid={t.pass_grid} "
          f"update_grid={t.update_grid} ms {t.ms_per_iteration_before:.3f} -> {t.ms_per_iteration:.3f} "
          f"trials={t.trials} buckets={list(t.bucket_targets)} tiles={list(t.bucket_tile_cols)} "
          f"split={list(t.bucket_split_ranks)}", flush=True)
    assert t.cached == 0 and t.trials >= 3
    assert t.engine in (capi.BATCH_FUSED, capi.BATCH_SPLIT_K) and native.gpumlib_nmf_batch_engine(b.h) == t.engine
    assert t.split_ranks == int(any(t.bucket_split_ranks)) and native.gpumlib_nmf_batch_rank_split(b.h) == t.split_ranks
    assert t.target_workgroups in TUNE_TARGETS | {32} and t.pass_grid in TUNE_PASS and t.update_grid in TUNE_UPD
    # the mix fills the buckets of fewer than 2^18 and fewer than 2^21 elements of V, not the third;
    # the single target reported is that of the largest bucket present
    assert t.bucket_targets[0] in TUNE_TARGETS and t.bucket_targets[1] in TUNE_TARGETS
    assert t.bucket_targets[2] == 0 and t.target_workgroups == t.bucket_targets[1]
    assert t.bucket_tile_cols[0] in (64, 128) and t.bucket_tile_cols[1] in (64, 128) and t.bucket_tile_cols[2] == 0
    assert all(x in (0, 1) for x in t.bucket_split_ranks) and t.bucket_split_ranks[2] == 0
    assert 0 < t.ms_per_iteration <= t.ms_per_iteration_before
    for (w0, h0), (*_, dW, dH) in zip(before, outs):
        assert np.array_equal(dW.download(np.float32), w0) and np.array_equal(dH.download(np.float32), h0)
    t2 = capi.BatchTuning()
    capi.check(native.gpumlib_nmf_batch_autotune(b.h, 2, b.stream.handle, ctypes.byref(t2)))
    assert t2.cached == 1 and t2.trials == 0
    assert (t2.engine, t2.target_workgroups, t2.pass_grid, t2.update_grid, t2.split_ranks) == \
        (t.engine, t.target_workgroups, t.pass_grid, t.update_grid, t.split_ranks)
    assert list(t2.bucket_targets) == list(t.bucket_targets)
    assert list(t2.bucket_tile_cols) == list(t.bucket_tile_cols)
    assert list(t2.bucket_split_ranks) == list(t.bucket_split_ranks)
    b.iterate(10)
    b.iterate(12, graph=True)  # graphs captured after tuning use the tuned grids
    for rule, n, m, r, V, W, H, dW, dH in outs:
        v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W, H))
        (orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter)(n, m, r, v, w, h, G.REFERENCE_EPS, iters=25)
        assert rel_fro(dW.download(np.float32), w) < 1e-4
        assert rel_fro(dH.download(np.float32), h) < 1e-4
    b.close()


def test_autotune_rejects_empty_batch(native, cuda):
    b = Batch(cuda, native)
    t = capi.BatchTuning()
    assert native.gpumlib_nmf_batch_autotune(b.h, 2, b.stream.handle, ctypes.byref(t)) != 0
    b.close()


# ------------------------------------------------------------------ server --
class Server:
    def __init__(self, lib, budget, max_active=64, round_iters=10, graph=1, autotune=0, sequential=0):
        self.lib = lib
        cfg = capi.ServerConfig(budget, max_active, round_iters, graph, 0, autotune, 0, sequential)
        h = ctypes.c_void_p()
        capi.check(lib.gpumlib_nmf_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def submit(self, rule, n, m, r, V, W, H, tol=1e-4, max_iters=200):
        jid = ctypes.c_longlong()
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        capi.check(self.lib.gpumlib_nmf_server_submit(self.h, rule, n, m, r, ptr(V), ptr(W), ptr(H), G.REFERENCE_EPS,
                                                      tol, max_iters, ctypes.byref(jid)))
        return jid.value

    def wait(self, jid, n, m, r):
        W = np.empty(n * r, np.float32)
        H = np.empty(r * m, np.float32)
        res = capi.JobResult()
        capi.check(self.lib.gpumlib_nmf_server_wait(self.h, jid, W.ctypes.data_as(ctypes.c_void_p),
                                                    H.ctypes.data_as(ctypes.c_void_p), ctypes.byref(res)))
        return W, H, res

    def stats(self):
        st = capi.ServerStats()
        capi.check(self.lib.gpumlib_nmf_server_stats_get(self.h, ctypes.byref(st)))
        return st

    def close(self):
        self.lib.gpumlib_nmf_server_destroy(self.h)


def test_server_matches_oracle_and_retires(native, orc, cuda):
    s = Server(native, budget=256 << 20)
    ids = []
    for k, job in enumerate(JOBS):
        V, W, H = make_job(k, *job)
        ids.append((s.submit(*job, V, W, H, tol=1e-4, max_iters=100), job, V, W, H))
    for jid, (rule, n, m, r), V, W, H in ids:
        gw, gh, res = s.wait(jid, n, m, r)
        assert res.iterations % 10 == 0 and 10 <= res.iterations <= 100
        assert res.converged or res.iterations == 100
        assert res.finish_s >= res.admit_s >= res.submit_s
        v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W, H))
        (orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter)(n, m, r, v, w, h, G.REFERENCE_EPS,
                                                               iters=res.iterations)
        assert rel_fro(gw, w) < 1e-4 and rel_fro(gh, h) < 1e-4
    st = s.stats()
    assert st.completed == len(JOBS)
    s.close()


def test_server_with_autotuning_matches_oracle(native, orc, cuda):
    """Autotuning inside the server (when the resident count doubles) keeps every job
    within the FP32 tolerance of the oracle."""
    s = Server(native, budget=512 << 20, autotune=2)
    # all inputs are made before the first submission and every job runs 8 rounds, so
    # that at least 4 are resident at some round boundary however fast a round is
    iters = 80
    data = [(job, *make_job(10 * k + j, *job)) for k in range(3) for j, job in enumerate(JOBS)]
    ids = [(s.submit(*job, V, W, H, tol=0.0, max_iters=iters), job, V, W, H) for job, V, W, H in data]
    for jid, (rule, n, m, r), V, W, H in ids:
        gw, gh, res = s.wait(jid, n, m, r)
        assert res.iterations == iters
        v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W, H))
        (orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter)(n, m, r, v, w, h, G.REFERENCE_EPS, iters=iters)
        assert rel_fro(gw, w) < 1e-4 and rel_fro(gh, h) < 1e-4
    st = s.stats()
    print(f"autotunes={st.autotunes} target={st.tuned_target} grids={st.tuned_pass_grid}/{st.tuned_update_grid}")
    assert st.autotunes >= 1 and st.tuned_target in TUNE_TARGETS
    assert all(t in TUNE_TARGETS | {0} for t in st.tuned_bucket_targets)
    assert st.tuned_target in list(st.tuned_bucket_targets)
    assert all(t in (0, 64, 128) for t in st.tuned_bucket_tile_cols)
    s.close()


def test_server_alone_equals_mix_and_admission(native, cuda):
    target = (G.KL, 512, 384, 32)
    V, W, H = make_job(3, *target)
    s1 = Server(native, budget=64 << 20)
    a = s1.wait(s1.submit(*target, V, W, H), 512, 384, 32)
    s1.close()
    # a pool that holds only a few problems at once: the rest wait for admission
    rng = np.random.default_rng(5)
    s2 = Server(native, budget=8 << 20, max_active=8)
    others = []
    for k in range(12):
        rule = int(rng.integers(0, 2))
        n, m, r = int(rng.integers(64, 400)), int(rng.integers(64, 400)), int(rng.choice([8, 16, 32]))
        Vk, Wk, Hk = make_job(50 + k, rule, n, m, r)
        others.append((s2.submit(rule, n, m, r, Vk, Wk, Hk), n, m, r))
        if k == 5:
            tid = s2.submit(*target, V, W, H)
    b = s2.wait(tid, 512, 384, 32)
    for jid, n, m, r in others:
        s2.wait(jid, n, m, r)
    st = s2.stats()
    assert st.completed == 13 and st.pool_high_water <= 8 << 20 and st.peak_active <= 8
    s2.close()
    assert a[2].iterations == b[2].iterations
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_serving_stream_error_still_destroys_server(native, cuda):
    """A submit that fails part-way through a serving stream (here a rank above 64)
    raises, and the server is destroyed on the way out: the jobs already submitted
    finish and a new server runs normally afterwards."""
    from gpumlib_amd import serving

    jobs = serving.make_jobs(6, seed=77)
    jobs[3] = (G.EUCLIDEAN, 80, 90, 65, 1)
    data = [serving.job_data(j if j[3] <= 64 else (j[0], j[1], j[2], 8, j[4])) for j in jobs]
    with pytest.raises(RuntimeError, match="gpumlib_nmf_server_submit"):
        serving.run_server(native, jobs, data, rate=0.0, max_iters=20)
    ok = serving.run_server(native, jobs[:3], data[:3], rate=0.0, max_iters=20)
    assert ok["jobs"] == 3


_EXIT_CHILD = r"""
import ctypes, sys
sys.path.insert(0, {root!r})
from gpumlib_amd import capi
from gpumlib_amd import nmf as G
lib = capi.load()
cfg = capi.ServerConfig(64 << 20, 16, 10, 1, 0, 0)
h = ctypes.c_void_p()
capi.check(lib.gpumlib_nmf_server_create(ctypes.byref(cfg), ctypes.byref(h)))
ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
keep = []
for k in range(12):
    V = G.gen_euclidean(400, 300, 16, k)
    W, H = G.gen_init(400, 300, 16, k)
    keep.append((V, W, H))
    jid = ctypes.c_longlong()
    capi.check(lib.gpumlib_nmf_server_submit(h, G.EUCLIDEAN, 400, 300, 16, ptr(V), ptr(W), ptr(H),
                                             G.REFERENCE_EPS, 0.0, 500, ctypes.byref(jid)))
print("submitted; leaving without destroying the server", flush=True)
"""


def test_server_left_running_at_process_exit(native, cuda):
    """A process that ends while its server still has work (never destroyed) exits
    cleanly: the library stops the engine after its round in flight and waits for its
    stream before the HIP runtime goes away."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, "-c", _EXIT_CHILD.format(root=root)], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr[-2000:]
    assert "submitted" in p.stdout


def test_server_cap_is_reached_at_a_round_end(native, orc, cuda):
    """Jobs advance in whole rounds: a cap of 15 with rounds of 10 stops after the second
    round, 20 iterations, and the result says so (and matches the oracle at 20)."""
    s = Server(native, budget=64 << 20)
    job = (G.EUCLIDEAN, 300, 200, 16)
    V, W, H = make_job(91, *job)
    gw, gh, res = s.wait(s.submit(*job, V, W, H, tol=0.0, max_iters=15), 300, 200, 16)
    s.close()
    assert res.iterations == 20 and not res.converged
    v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W, H))
    orc.euc_iter(300, 200, 16, v, w, h, G.REFERENCE_EPS, iters=20)
    assert rel_fro(gw, w) < 1e-4 and rel_fro(gh, h) < 1e-4


def test_server_one_at_a_time_matches_oracle(native, orc, cuda):
    """The server's one-at-a-time mode (the unbatched comparison point of the serving
    benchmark): each job alone through the single-problem iteration, the same rounds and
    stopping rule, results within the FP32 tolerance of the oracle, one job resident at a time."""
    s = Server(native, budget=256 << 20, sequential=1)
    ids = []
    for k, job in enumerate(JOBS):
        V, W, H = make_job(k, *job)
        ids.append((s.submit(*job, V, W, H, tol=1e-4, max_iters=60), job, V, W, H))
    for jid, (rule, n, m, r), V, W, H in ids:
        gw, gh, res = s.wait(jid, n, m, r)
        assert res.iterations % 10 == 0 and 10 <= res.iterations <= 60
        assert res.finish_s >= res.admit_s >= res.submit_s
        v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W, H))
        (orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter)(n, m, r, v, w, h, G.REFERENCE_EPS,
                                                               iters=res.iterations)
        assert rel_fro(gw, w) < 1e-4 and rel_fro(gh, h) < 1e-4
    st = s.stats()
    assert st.completed == len(JOBS) and st.peak_active == 1
    s.close()
