"""Objective pass on split factors: the FP64 sum of the per-element objective terms of
(V, W H) over V stored in FP32, BF16 or FP8, with W H formed on the BF16 matrix cores
from a three-way BF16 split of W and H (six products, two FP32 accumulators).

Checked against the FP64 oracle on the stored V (the dequantised BF16 / FP8 store):
* Frobenius at an exact fit (V = fl32(W H), and W H + noise), where the objective is a
  residual many orders below |V|^2 and any loss of accuracy in W H shows: the difference
  must stay inside the bound of an FP32 r-term dot product per element;
* KL (0 ln 0 = 0) and the reference quality (nmf_kernels_quality.cu:35) on random states;
* shapes with ragged rows, columns and ranks (padding of the 128-row blocks, 32-column
  blocks and 16-wide rank steps).
"""
import numpy as np
import pytest

from conftest import objective_allowance
from gpumlib_amd import nmf as G

pytestmark = pytest.mark.gpu

SHAPES = [(300, 77, 5), (257, 300, 33), (128, 64, 64), (513, 129, 100), (200, 96, 128), (1000, 700, 64)]
PRECISIONS = {"fp32": G.FP32, "bf16": G.BF16, "fp8": G.FP8}


def _stored(orc, V, n, m, precision):
    """V as the precision's store holds it, FP64 n x m."""
    Vm = G.as_matrix(V, n, m)
    if precision == G.BF16:
        return orc.bf16_round(Vm).astype(np.float64)
    if precision == G.FP8:
        return orc.fp8_tile_round(Vm).astype(np.float64)
    return Vm.astype(np.float64)


def _skip_unplanned(precision, r):
    if precision == G.FP8 and r > 64:
        pytest.skip("the FP8 storage path plans ranks up to 64")


def _fit(n, m, r, seed, noise):
    """W, H ~ U[0, 1) and V = fl32(W H) + noise (flat column-major float32)."""
    rng = np.random.default_rng(seed)
    w = rng.random((n, r), dtype=np.float32)
    h = rng.random((r, m), dtype=np.float32)
    v = (w.astype(np.float64) @ h.astype(np.float64)).astype(np.float32)
    if noise:
        v = (v + noise * rng.random((n, m), dtype=np.float32)).astype(np.float32)
    flat = lambda a: np.asfortranarray(a).ravel(order="F").astype(np.float32)  # noqa: E731
    return flat(v), flat(w), flat(h)


@pytest.mark.parametrize("prec", list(PRECISIONS))
@pytest.mark.parametrize("nmr", SHAPES)
@pytest.mark.parametrize("noise", [0.0, 0.01])
def test_frobenius_at_a_fit(native, orc, cuda, prec, nmr, noise):
    n, m, r = nmr
    precision = PRECISIONS[prec]
    _skip_unplanned(precision, r)
    V, W, H = _fit(n, m, r, seed=n * 7 + m + r, noise=noise)
    d = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r, precision=precision)
    got = d.objective(G.OBJ_FROBENIUS)
    vs = _stored(orc, V, n, m, precision)
    w, h = G.as_matrix(W, n, r).astype(np.float64), G.as_matrix(H, r, m).astype(np.float64)
    ref = orc.np_frob(vs, w, h)
    allow = objective_allowance(ref, float(np.sum(vs * vs)), r, 0.0)
    assert abs(got - ref) <= allow, (got, ref, allow)


@pytest.mark.parametrize("prec", list(PRECISIONS))
@pytest.mark.parametrize("nmr", SHAPES)
def test_kl_and_quality(native, orc, cuda, prec, nmr):
    n, m, r = nmr
    precision = PRECISIONS[prec]
    _skip_unplanned(precision, r)
    V = G.gen_poisson(n, m, 8, seed=n + 3 * m)
    W, H = G.gen_init(n, m, r, seed=r)
    w, h = G.as_matrix(W, n, r).astype(np.float64), G.as_matrix(H, r, m).astype(np.float64)
    vs = _stored(orc, V, n, m, precision)
    if precision == G.FP32:  # the KL rules run FP32 storage only
        d = G.DeviceNMF(G.KL, V, W, H, n, m, r, eps=1e-9)
        got = d.objective(G.OBJ_KL)
        ref = orc.np_kl(vs, w, h)
        assert abs(got - ref) / ref < 1e-6, (got, ref)
    d = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r, precision=precision)
    got = d.objective(G.OBJ_QUALITY)
    ref = orc.np_quality(vs, w, h, G.REFERENCE_EPS)
    assert abs(got - ref) / abs(ref) < 1e-5, (got, ref)


def test_objective_with_the_iteration(native, orc, cuda):
    """The Euclidean FP32 iteration asked for its entry objective runs this pass before the
    H pass: the objective is that of the entry state and the step itself is unchanged."""
    n, m, r = 700, 500, 48
    V = G.gen_euclidean(n, m, r, seed=5)
    W, H = G.gen_init(n, m, r, seed=6)
    a = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r)
    b = G.DeviceNMF(G.EUCLIDEAN, V, W, H, n, m, r)
    f0 = a.iterate(1, objective=G.OBJ_FROBENIUS)
    q0 = b.iterate(1, objective=G.OBJ_QUALITY)
    w, h = G.as_matrix(W, n, r).astype(np.float64), G.as_matrix(H, r, m).astype(np.float64)
    vs = G.as_matrix(V, n, m).astype(np.float64)
    ref = orc.np_frob(vs, w, h)
    assert abs(f0 - ref) / ref < 1e-6
    assert abs(q0 - orc.np_quality(vs, w, h, G.REFERENCE_EPS)) / abs(q0) < 1e-5
    assert np.array_equal(a.get_w(), b.get_w()) and np.array_equal(a.get_h(), b.get_h())
