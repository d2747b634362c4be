"""Parity of the fused gfx950 iteration (gpumlib_nmf_iterate) with the FP64 oracle.

Gate: after a fixed number of iterations from the same seeded
inputs, W, H and the objective match the FP64 restatement of the reference
update rules within 1e-4 relative Frobenius error.
"""
import numpy as np
import pytest

from conftest import rel_fro
from gpumlib_amd import nmf as G

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _oracle_run(orc, rule, n, m, r, V, W, H, eps, iters, update_w=True):
    v, w, h = _f64(V), _f64(W), _f64(H)
    fn = orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter
    fn(n, m, r, v, w, h, eps, update_w=update_w, iters=iters)
    return v, w, h


def _problem(rule, n, m, r, seed, r0=None):
    r0 = r0 or max(1, r)
    V = G.gen_euclidean(n, m, r0, seed) if rule == G.EUCLIDEAN else G.gen_poisson(n, m, r0, seed)
    W, H = G.gen_init(n, m, r, seed)
    return V, W, H


@pytest.mark.parametrize("rule", [G.EUCLIDEAN, G.KL])
def test_c1_200_iterations(native, orc, cuda, rule):
    """C1: 256 x 128, r = 16, 200 iterations (KL on Poisson counts)."""
    n, m, r, seed = 256, 128, 16, 0x5EED + 1
    print(f"case rule={rule} n={n} m={m} r={r}", flush=True)
    V, W, H = _problem(rule, n, m, r, seed)
    eps = G.REFERENCE_EPS
    d = G.DeviceNMF(rule, V, W, H, n, m, r, eps=eps)
    d.iterate(200)
    v, w, h = _oracle_run(orc, rule, n, m, r, V, W, H, eps, 200)
    assert rel_fro(d.get_w(), w) < TOL
    assert rel_fro(d.get_h(), h) < TOL
    kind = G.OBJ_FROBENIUS if rule == G.EUCLIDEAN else G.OBJ_KL
    ref = orc.frob(n, m, r, v, w, h) if rule == G.EUCLIDEAN else orc.kl_div(n, m, r, v, w, h)
    got = d.objective(kind)
    assert abs(got - ref) / abs(ref) < TOL


SHAPES = [
    (97, 61, 5),     # nothing aligned: scalar loads, partial tiles
    (130, 257, 33),  # r just above one MFMA block
    (64, 64, 64),
    (300, 200, 100),  # four rank blocks
    (128, 384, 128),  # largest fused rank
    (1, 7, 1),
    (4100, 36, 12),   # tall: many K slices in the H pass
    (36, 4100, 12),   # wide: many K slices in the W pass
    (16500, 90, 64),  # n >= 16384: the 32-row update workgroups, a partial last one
    (16390, 70, 20),  # the same for one rank block, n not a multiple of 4
]


@pytest.mark.parametrize("rule", [G.EUCLIDEAN, G.KL])
@pytest.mark.parametrize("shape", SHAPES)
def test_shapes_vs_oracle(native, orc, cuda, rule, shape):
    n, m, r = shape
    print(f"case rule={rule} shape={shape}", flush=True)
    V, W, H = _problem(rule, n, m, r, seed=1234 + n + m + r, r0=min(r, 8))
    eps = G.REFERENCE_EPS
    iters = 20
    d = G.DeviceNMF(rule, V, W, H, n, m, r, eps=eps)
    d.iterate(iters)
    _, w, h = _oracle_run(orc, rule, n, m, r, V, W, H, eps, iters)
    assert rel_fro(d.get_w(), w) < TOL
    assert rel_fro(d.get_h(), h) < TOL


@pytest.mark.parametrize("rule", [G.EUCLIDEAN, G.KL])
def test_update_h_only(native, orc, cuda, rule):
    """DoIteration(false): H changes, W stays bit-identical."""
    n, m, r = 200, 150, 24
    V, W, H = _problem(rule, n, m, r, seed=77)
    d = G.DeviceNMF(rule, V, W, H, n, m, r)
    d.iterate(5, update_w=False)
    _, w, h = _oracle_run(orc, rule, n, m, r, V, W, H, G.REFERENCE_EPS, 5, update_w=False)
    assert np.array_equal(d.get_w(), W)
    assert rel_fro(d.get_h(), h) < TOL


@pytest.mark.parametrize("rule", [G.EUCLIDEAN, G.KL])
def test_objectives_match_oracle(native, orc, cuda, rule):
    """FP64 objectives of the entry state, fused into the first pass and standalone."""
    n, m, r = 333, 190, 20
    V, W, H = _problem(rule, n, m, r, seed=5)
    v, w, h = _f64(V), _f64(W), _f64(H)
    d = G.DeviceNMF(rule, V, W, H, n, m, r)
    frob = orc.frob(n, m, r, v, w, h)
    kl = orc.kl_div(n, m, r, v, w, h)
    wh = np.zeros(n * m)
    orc.gemm("N", "N", n, m, r, 1.0, w, n, h, r, 0.0, wh, n)
    qual = orc.quality(v, wh, G.REFERENCE_EPS)
    assert abs(d.objective(G.OBJ_FROBENIUS) - frob) / frob < 1e-5
    assert abs(d.objective(G.OBJ_KL) - kl) / abs(kl) < 1e-5
    assert abs(d.objective(G.OBJ_QUALITY) - qual) / abs(qual) < 1e-5
    kind = G.OBJ_FROBENIUS if rule == G.EUCLIDEAN else G.OBJ_KL
    ref = frob if rule == G.EUCLIDEAN else kl
    fused = d.iterate(1, objective=kind)  # objective of the state on entry
    assert abs(fused - ref) / abs(ref) < 1e-5


@pytest.mark.parametrize("rule", [G.EUCLIDEAN, G.KL])
def test_deterministic(native, cuda, rule):
    """Fixed reduction order: two runs give identical bits."""
    n, m, r = 1000, 700, 32
    V, W, H = _problem(rule, n, m, r, seed=9)
    outs = []
    for _ in range(2):
        d = G.DeviceNMF(rule, V, W, H, n, m, r)
        d.iterate(10)
        outs.append((d.get_w(), d.get_h()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("rule", [G.EUCLIDEAN, G.KL])
def test_objective_monotone(native, cuda, rule):
    """Lee & Seung: the objective does not increase under multiplicative updates."""
    n, m, r = 512, 384, 16
    V, W, H = _problem(rule, n, m, r, seed=21)
    d = G.DeviceNMF(rule, V, W, H, n, m, r)
    kind = G.OBJ_FROBENIUS if rule == G.EUCLIDEAN else G.OBJ_KL
    vals = [d.iterate(1, objective=kind) for _ in range(30)]
    for a, b in zip(vals, vals[1:]):
        assert b <= a * (1 + 1e-6)
    assert vals[-1] < vals[0]


@pytest.mark.timeout(600)
def test_c2_full_euclidean(native, orc, cuda):
    """C2: 8192 x 4096, r = 64, 500 iterations, FP32 GPU vs FP64 oracle (numpy)."""
    n, m, r, seed = 8192, 4096, 64, 0x5EED + 2
    print("case C2 full", flush=True)
    V, W, H = _problem(G.EUCLIDEAN, n, m, r, seed)
    d = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r)
    d.iterate(500)
    gw, gh = d.get_w(), d.get_h()
    gobj = d.objective(G.OBJ_FROBENIUS)
    v = G.as_matrix(V, n, m).astype(np.float64)
    w = G.as_matrix(W, n, r).astype(np.float64)
    h = G.as_matrix(H, r, m).astype(np.float64)
    for _ in range(500):
        w, h = orc.np_euc_iter(v, w, h, G.REFERENCE_EPS)
    assert rel_fro(G.as_matrix(gw, n, r), w) < TOL
    assert rel_fro(G.as_matrix(gh, r, m), h) < TOL
    ref = orc.np_frob(v, w, h)
    assert abs(gobj - ref) / ref < TOL


@pytest.mark.timeout(600)
def test_c3_full_kl(native, orc, cuda):
    """C3: KL, 4096 x 4096 Poisson counts, r = 32, eps = 1e-9, 300 iterations."""
    n, m, r, seed = 4096, 4096, 32, 0x5EED + 3
    eps = 1e-9
    print("case C3 full", flush=True)
    V, W, H = _problem(G.KL, n, m, r, seed)
    d = G.DeviceNMF(G.KL, V, W, H, n, m, r, eps=eps)
    d.iterate(300)
    gw, gh = d.get_w(), d.get_h()
    gobj = d.objective(G.OBJ_KL)
    v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W, H))
    for k in range(6):  # fused OpenMP FP64 oracle, 50 iterations at a time
        orc.kl_iter(n, m, r, v, w, h, eps, iters=50, omp=True)
        print(f"oracle {50 * (k + 1)} iterations", flush=True)
    assert rel_fro(gw, w) < TOL
    assert rel_fro(gh, h) < TOL
    ref = orc.np_kl(G.as_matrix(v, n, m), G.as_matrix(w, n, r), G.as_matrix(h, r, m))
    assert abs(gobj - ref) / abs(ref) < TOL
