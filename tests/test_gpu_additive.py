"""Additive-update variants on the GPU (NMF_AdditiveEuclidian, NMF_AdditiveDivergence)
against the FP64 oracle.

The bar is the multiplicative path's: W, H and the objective within 1e-4 relative
Frobenius of the FP64 restatement after a fixed number of iterations from the same
seeded inputs; H bitwise unchanged under GPUMLIB_NMF_ADDITIVE_EUCLIDEAN (the
reference's H step writes only scratch); the reference-named launchers checked
element by element.
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_fro
from gpumlib_amd import capi
from gpumlib_amd import nmf as G

pytestmark = pytest.mark.gpu

TOL = 1e-4
RULES = [G.ADDITIVE_EUCLIDEAN, G.ADDITIVE_EUCLIDEAN_WITH_H, G.ADDITIVE_KL]
NAMES = {G.ADDITIVE_EUCLIDEAN: "add-euc", G.ADDITIVE_EUCLIDEAN_WITH_H: "add-euc-h", G.ADDITIVE_KL: "add-kl"}


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _oracle_run(orc, rule, n, m, r, V, W, H, eps, iters, update_w=True):
    v, w, h = _f64(V), _f64(W), _f64(H)
    if rule == G.ADDITIVE_KL:
        orc.add_kl_iter(n, m, r, v, w, h, eps, update_w=update_w, iters=iters)
    else:
        orc.add_euc_iter(n, m, r, v, w, h, update_w=update_w, fix_h=(rule == G.ADDITIVE_EUCLIDEAN_WITH_H),
                         iters=iters)
    return v, w, h


def _problem(rule, n, m, r, seed, r0=None):
    r0 = r0 or max(1, r)
    V = G.gen_poisson(n, m, r0, seed) if rule == G.ADDITIVE_KL else G.gen_euclidean(n, m, r0, seed)
    W, H = G.gen_init(n, m, r, seed)
    return V, W, H


def _objective(orc, rule, n, m, r, v, w, h):
    return orc.kl_div(n, m, r, v, w, h) if rule == G.ADDITIVE_KL else orc.frob(n, m, r, v, w, h)


@pytest.mark.parametrize("rule", RULES, ids=lambda x: NAMES[x])
def test_c1_200_iterations(native, orc, cuda, rule):
    n, m, r, seed = 256, 128, 16, 0x5EED + 1
    print(f"case {NAMES[rule]} C1", flush=True)
    V, W, H = _problem(rule, n, m, r, seed)
    eps = G.REFERENCE_EPS
    d = G.DeviceNMF(rule, V, W, H, n, m, r, eps=eps)
    d.iterate(200)
    v, w, h = _oracle_run(orc, rule, n, m, r, V, W, H, eps, 200)
    assert rel_fro(d.get_w(), w) < TOL
    assert rel_fro(d.get_h(), h) < TOL
    kind = G.OBJ_KL if rule == G.ADDITIVE_KL else G.OBJ_FROBENIUS
    ref = _objective(orc, rule, n, m, r, v, w, h)
    assert abs(d.objective(kind) - ref) / abs(ref) < TOL


SHAPES = [(97, 61, 5), (130, 257, 33), (300, 200, 100), (128, 384, 128), (1, 7, 1), (4100, 36, 12),
          (36, 4100, 12), (16500, 90, 64)]


@pytest.mark.parametrize("rule", RULES, ids=lambda x: NAMES[x])
@pytest.mark.parametrize("shape", SHAPES)
def test_shapes_vs_oracle(native, orc, cuda, rule, shape):
    n, m, r = shape
    print(f"case {NAMES[rule]} shape={shape}", flush=True)
    V, W, H = _problem(rule, n, m, r, seed=4321 + n + m + r, r0=min(r, 8))
    eps = G.REFERENCE_EPS
    d = G.DeviceNMF(rule, V, W, H, n, m, r, eps=eps)
    d.iterate(20)
    _, w, h = _oracle_run(orc, rule, n, m, r, V, W, H, eps, 20)
    assert rel_fro(d.get_w(), w) < TOL
    assert rel_fro(d.get_h(), h) < TOL


def test_additive_euclidean_keeps_h_bitwise(native, cuda):
    n, m, r = 700, 500, 24
    V, W, H = _problem(G.ADDITIVE_EUCLIDEAN, n, m, r, 77)
    d = G.DeviceNMF(G.ADDITIVE_EUCLIDEAN, V, W, H, n, m, r)
    d.iterate(50)
    assert np.array_equal(d.get_h(), H)
    assert not np.array_equal(d.get_w(), W)
    # update_w = 0: nothing changes at all
    d2 = G.DeviceNMF(G.ADDITIVE_EUCLIDEAN, V, W, H, n, m, r)
    d2.iterate(5, update_w=False)
    assert np.array_equal(d2.get_w(), W) and np.array_equal(d2.get_h(), H)


@pytest.mark.parametrize("rule", [G.ADDITIVE_EUCLIDEAN_WITH_H, G.ADDITIVE_KL], ids=lambda x: NAMES[x])
def test_update_h_only(native, orc, cuda, rule):
    n, m, r = 333, 222, 20
    V, W, H = _problem(rule, n, m, r, 78)
    d = G.DeviceNMF(rule, V, W, H, n, m, r)
    d.iterate(3, update_w=False)
    _, w, h = _oracle_run(orc, rule, n, m, r, V, W, H, G.REFERENCE_EPS, 3, update_w=False)
    assert np.array_equal(d.get_w(), W)
    assert rel_fro(d.get_h(), h) < TOL


@pytest.mark.parametrize("rule", RULES, ids=lambda x: NAMES[x])
def test_objective_on_entry(native, orc, cuda, rule):
    """The objective asked for with an iteration is the one of W, H on entry (the H
    pass computes it; GPUMLIB_NMF_ADDITIVE_EUCLIDEAN runs the objective pass instead)."""
    n, m, r = 500, 300, 16
    V, W, H = _problem(rule, n, m, r, 79)
    kind = G.OBJ_KL if rule == G.ADDITIVE_KL else G.OBJ_FROBENIUS
    d = G.DeviceNMF(rule, V, W, H, n, m, r)
    f0 = d.iterate(1, objective=kind)
    ref = _objective(orc, rule, n, m, r, _f64(V), _f64(W), _f64(H))
    assert abs(f0 - ref) / ref < 1e-6
    q = d.iterate(1, objective=G.OBJ_QUALITY)
    v, w, h = _oracle_run(orc, rule, n, m, r, V, W, H, G.REFERENCE_EPS, 1)
    WH = np.zeros(n * m)
    orc.gemm("N", "N", n, m, r, 1.0, w, n, h, r, 0.0, WH, n)
    assert abs(q - orc.quality(v, WH, G.REFERENCE_EPS)) / abs(q) < 1e-5


# ------------------------------------------------------- reference launchers --
def test_additive_euclidean_launcher(native, orc, cuda):
    """gpumlib_cuda_nmf_additive_euclidean_update_matrix (nmf_kernels.cu:31-39)."""
    from test_gpu_kernels import dev, get, stream

    rng = np.random.default_rng(41)
    E = 100_003
    d1, d2, X = (rng.random(E, dtype=np.float32) + np.float32(0.01) for _ in range(3))
    d1[:10] = d2[:10]      # unchanged exactly
    d1[10:20] = 0.0        # zero, negative roundings clamped
    X[20:25] = 0.0
    d2[20:25] = 0.0        # 0 / 0: NaN kept
    ref = X.astype(np.float64)
    orc.lib().orc_add_update_f64(orc._p64(d1.astype(np.float64)), orc._p64(d2.astype(np.float64)), orc._p64(ref), E)
    dX, dd1, dd2 = dev(cuda, X), dev(cuda, d1), dev(cuda, d2)
    capi.check(native.gpumlib_nmf_additive_update(dX.ptr, dd1.ptr, dd2.ptr, E, stream(cuda)))
    got = get(dX)
    ok = ~np.isnan(ref)
    assert np.array_equal(np.isnan(got), ~ok)
    # FP32 rounding scales with X: where d1 = 0 the exact result is 0 and x - (x/d2) d2 leaves a few ulp of x
    assert np.all(np.abs(got[ok] - ref[ok]) <= 1e-6 * np.maximum(ref[ok], X[ok]))
    assert np.array_equal(got[:10], X[:10])
    assert np.all(got[10:20] >= 0) and np.all(got[10:20] < 1e-6)
    dX2 = dev(cuda, X)
    cuda.synchronize()
    capi.check(native.gpumlib_cuda_nmf_additive_euclidean_update_matrix(1, 512, dX2.ptr, dd1.ptr, dd2.ptr, E))
    cuda.synchronize()
    assert np.array_equal(get(dX2)[ok], got[ok])


@pytest.mark.parametrize("nmr", [(96, 80, 16), (257, 131, 37)])
def test_additive_kl_launchers_with_materialised_wh(native, orc, cuda, nmr):
    """nmf_kernels_additive_divergence.cu:57-193 with WH given, against the FP64 kernels."""
    from test_gpu_kernels import dev, get, stream

    n, m, r = nmr
    V = G.gen_poisson(n, m, 8, seed=n + 1)
    W, H = G.gen_init(n, m, r, seed=m + 1)
    W[:n] = 0.0  # one all-zero column of W: clamped sum, that row of H kept
    WH = np.zeros(n * m, np.float32)
    orc.gemm("N", "N", n, m, r, 1.0, W, n, H, r, 0.0, WH, n)
    sw = np.zeros(r, np.float32)
    sh = np.zeros(r, np.float32)
    orc.lib().orc_sum_w_f32(n, r, orc._p32(W), np.float32(1e-8), orc._p32(sw))
    orc.lib().orc_sum_h_f32(r, m, orc._p32(H), np.float32(1e-8), orc._p32(sh))
    h_ref, w_ref = H.astype(np.float64), W.astype(np.float64)
    args = [orc._p64(x.astype(np.float64)) for x in (V, WH)]
    orc.lib().orc_add_kl_update_h_f64(orc._p64(h_ref), orc._p64(W.astype(np.float64)), *args,
                                      orc._p64(sw.astype(np.float64)), n, m, r, 1e-8)
    orc.lib().orc_add_kl_update_w_f64(orc._p64(w_ref), orc._p64(H.astype(np.float64)), *args,
                                      orc._p64(sh.astype(np.float64)), n, m, r, 1e-8)
    dV, dW, dH, dWH = dev(cuda, V), dev(cuda, W), dev(cuda, H), dev(cuda, WH)
    dsw, dsh = dev(cuda, sw), dev(cuda, sh)
    dH2, dW2 = dev(cuda, H), dev(cuda, W)
    capi.check(native.gpumlib_nmf_additive_kl_update_h(dH2.ptr, dW.ptr, dV.ptr, dWH.ptr, dsw.ptr, n, m, r, 1e-8,
                                                       stream(cuda)))
    capi.check(native.gpumlib_nmf_additive_kl_update_w(dW2.ptr, dH.ptr, dV.ptr, dWH.ptr, dsh.ptr, n, m, r, 1e-8,
                                                       stream(cuda)))
    gh, gw = get(dH2), get(dW2)
    assert rel_fro(gh, h_ref) < 1e-5 and rel_fro(gw, w_ref) < 1e-5
    assert np.array_equal(gh[0::r], H[0::r])  # row 0 of H: W(:,0) == 0
    g = capi.Dim3(1, 1, 1)
    dH3, dW3 = dev(cuda, H), dev(cuda, W)
    cuda.synchronize()
    capi.check(native.gpumlib_cuda_nmf_additive_divergence_update_matrix_h(g, g, dH3.ptr, dW.ptr, dV.ptr, dWH.ptr,
                                                                           dsw.ptr, n, m, r))
    capi.check(native.gpumlib_cuda_nmf_additive_divergence_update_matrix_w(g, g, dW3.ptr, dH.ptr, dV.ptr, dWH.ptr,
                                                                           dsh.ptr, n, m, r))
    cuda.synchronize()
    assert np.array_equal(get(dH3), gh) and np.array_equal(get(dW3), gw)


# --------------------------------------------------------------- object API --
@pytest.mark.parametrize("rule", RULES, ids=lambda x: NAMES[x])
def test_object_api(native, orc, cuda, rule):
    """gpumlib_nmf_object_* over NMF_AdditiveEuclidian / NMF_AdditiveDivergence."""
    n, m, r = 400, 250, 12
    V, W, H = _problem(rule, n, m, r, 80)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    h = ctypes.c_void_p()
    capi.check(native.gpumlib_nmf_object_create(rule, G.FP32, n, m, r, ptr(V), ptr(W), ptr(H), ctypes.byref(h)))
    try:
        capi.check(native.gpumlib_nmf_object_do_iterations(h, 30, 1))
        gw, gh = np.empty(n * r, np.float32), np.empty(r * m, np.float32)
        capi.check(native.gpumlib_nmf_object_get_w(h, ptr(gw)))
        capi.check(native.gpumlib_nmf_object_get_h(h, ptr(gh)))
    finally:
        native.gpumlib_nmf_object_destroy(h)
    _, w, hh = _oracle_run(orc, rule, n, m, r, V, W, H, G.REFERENCE_EPS, 30)
    assert rel_fro(gw, w) < TOL and rel_fro(gh, hh) < TOL
    if rule == G.ADDITIVE_EUCLIDEAN:
        assert np.array_equal(gh, H)


def test_batch_rejects_additive_rules(native, cuda):
    b = ctypes.c_void_p()
    capi.check(native.gpumlib_nmf_batch_create(4, ctypes.byref(b)))
    try:
        for rule in RULES:
            assert native.gpumlib_nmf_batch_workspace_bytes(b, rule, 100, 100, 8) == 0
        assert native.gpumlib_nmf_batch_workspace_bytes(b, G.EUCLIDEAN, 100, 100, 8) > 0
    finally:
        native.gpumlib_nmf_batch_destroy(b)


# ---------------------------------------------------------------- full size --
@pytest.mark.timeout(600)
@pytest.mark.parametrize("rule,iters", [(G.ADDITIVE_EUCLIDEAN, 500), (G.ADDITIVE_EUCLIDEAN_WITH_H, 200)],
                         ids=["add-euc-500", "add-euc-h-200"])
def test_c2_full_additive_euclidean(native, orc, cuda, rule, iters):
    """C2 shape (8192 x 4096, r = 64) against the numpy FP64 restatement."""
    n, m, r, seed = 8192, 4096, 64, 0x5EED + 2
    print(f"case {NAMES[rule]} C2 full, {iters} iterations", flush=True)
    V, W, H = _problem(rule, n, m, r, seed)
    d = G.DeviceNMF(rule, V, W, H, n, m, r)
    d.iterate(iters)
    gw, gh = d.get_w(), d.get_h()
    gobj = d.objective(G.OBJ_FROBENIUS)
    v = G.as_matrix(V, n, m).astype(np.float64)
    w = G.as_matrix(W, n, r).astype(np.float64)
    h = G.as_matrix(H, r, m).astype(np.float64)
    for k in range(iters):
        w, h = orc.np_add_euc_iter(v, w, h, fix_h=(rule == G.ADDITIVE_EUCLIDEAN_WITH_H))
        if (k + 1) % 100 == 0:
            print(f"oracle {k + 1} iterations", flush=True)
    assert rel_fro(G.as_matrix(gw, n, r), w) < TOL
    assert rel_fro(G.as_matrix(gh, r, m), h) < TOL
    ref = orc.np_frob(v, w, h)
    assert abs(gobj - ref) / ref < TOL


@pytest.mark.timeout(600)
def test_c3_full_additive_kl(native, orc, cuda):
    """C3 shape (KL, 4096 x 4096 Poisson counts, r = 32, eps = 1e-9), 300 iterations."""
    n, m, r, seed = 4096, 4096, 32, 0x5EED + 3
    eps = 1e-9
    print("case add-kl C3 full", flush=True)
    V, W, H = _problem(G.ADDITIVE_KL, n, m, r, seed)
    d = G.DeviceNMF(G.ADDITIVE_KL, V, W, H, n, m, r, eps=eps)
    d.iterate(300)
    gw, gh = d.get_w(), d.get_h()
    gobj = d.objective(G.OBJ_KL)
    v, w, h = (np.ascontiguousarray(x, dtype=np.float64) for x in (V, W, H))
    for k in range(6):
        orc.add_kl_iter(n, m, r, v, w, h, eps, iters=50, omp=True)
        print(f"oracle {50 * (k + 1)} iterations", flush=True)
    assert rel_fro(gw, w) < TOL
    assert rel_fro(gh, h) < TOL
    ref = orc.np_kl(G.as_matrix(v, n, m), G.as_matrix(w, n, r), G.as_matrix(h, r, m))
    assert abs(gobj - ref) / abs(ref) < TOL
