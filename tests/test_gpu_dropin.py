"""Parity of the drop-in surface beyond the fused iteration, against the FP64 oracle:

* ranks above 128 (the unfused GEMM sequence of nmf_multiplicative_euclidian.cpp:29-67,
  nmf_multiplicative_divergence.cpp:29-50 and the additive variants), through the C-ABI
  and through the C++ classes (the object handle);
* KernelSumH (nmf_kernels.cu:149-233) at any rank; the quality written as a float
  (KernelNMFquality, nmf_kernels_quality.cu:29-109);
* NMF::GetWH (nmf_base.h:135-141) and the QualityImprovement() bookkeeping
  (nmf_base.h:70-89, 148-150) at the deterministic quality cadence.
"""
import ctypes

import numpy as np
import pytest

from conftest import rel_fro
from gpumlib_amd import capi
from gpumlib_amd import nmf as G

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _problem(rule, n, m, r, seed, r0=8):
    kl = rule in (G.KL, G.ADDITIVE_KL)
    V = G.gen_poisson(n, m, r0, seed) if kl else G.gen_euclidean(n, m, r0, seed)
    W, H = G.gen_init(n, m, r, seed)
    return V, W, H


def _oracle(orc, rule, n, m, r, V, W, H, eps, iters, update_w=True):
    v, w, h = _f64(V), _f64(W), _f64(H)
    if rule == G.EUCLIDEAN:
        orc.euc_iter(n, m, r, v, w, h, eps, update_w=update_w, iters=iters)
    elif rule == G.KL:
        orc.kl_iter(n, m, r, v, w, h, eps, update_w=update_w, iters=iters)
    elif rule == G.ADDITIVE_KL:
        orc.add_kl_iter(n, m, r, v, w, h, eps, update_w=update_w, iters=iters)
    else:
        orc.add_euc_iter(n, m, r, v, w, h, update_w=update_w, fix_h=(rule == G.ADDITIVE_EUCLIDEAN_WITH_H),
                         iters=iters)
    return v, w, h


UNFUSED = [(300, 200, 150), (160, 140, 129), (97, 400, 257)]


@pytest.mark.parametrize("rule", [G.EUCLIDEAN, G.KL, G.ADDITIVE_KL, G.ADDITIVE_EUCLIDEAN,
                                  G.ADDITIVE_EUCLIDEAN_WITH_H])
@pytest.mark.parametrize("shape", UNFUSED)
def test_unfused_rank_vs_oracle(native, orc, cuda, rule, shape):
    """Ranks above the fused passes' 128 run the reference's GEMM sequence."""
    n, m, r = shape
    print(f"case unfused rule={rule} shape={shape}", flush=True)
    V, W, H = _problem(rule, n, m, r, seed=31 + n + r)
    eps = 1e-9 if rule in (G.KL, G.ADDITIVE_KL) else G.REFERENCE_EPS
    iters = 12
    d = G.DeviceNMF(rule, V, W, H, n, m, r, eps=eps)
    kind = G.OBJ_KL if rule in (G.KL, G.ADDITIVE_KL) else G.OBJ_FROBENIUS
    f0 = d.iterate(iters, objective=kind)
    v, w, h = _oracle(orc, rule, n, m, r, V, W, H, eps, iters)
    assert rel_fro(d.get_w(), w) < TOL
    assert rel_fro(d.get_h(), h) < TOL
    if rule == G.ADDITIVE_EUCLIDEAN:
        assert np.array_equal(d.get_h(), H)  # the reference's additive H step leaves H as it is
    v0, w0, h0 = _f64(V), _f64(W), _f64(H)
    ref0 = orc.frob(n, m, r, v0, w0, h0) if kind == G.OBJ_FROBENIUS else orc.kl_div(n, m, r, v0, w0, h0)
    assert abs(f0 - ref0) / abs(ref0) < 1e-6
    ref1 = orc.frob(n, m, r, v, w, h) if kind == G.OBJ_FROBENIUS else orc.kl_div(n, m, r, v, w, h)
    assert abs(d.objective(kind) - ref1) / abs(ref1) < TOL
    q = d.objective(G.OBJ_QUALITY)
    qref = orc.np_quality(G.as_matrix(V, n, m).astype(np.float64), G.as_matrix(w, n, r), G.as_matrix(h, r, m), eps)
    assert abs(q - qref) / abs(qref) < TOL


@pytest.mark.parametrize("rule", [G.EUCLIDEAN, G.KL])
def test_unfused_update_h_only(native, orc, cuda, rule):
    n, m, r = 180, 150, 140
    V, W, H = _problem(rule, n, m, r, seed=9)
    d = G.DeviceNMF(rule, V, W, H, n, m, r)
    d.iterate(4, update_w=False)
    _, w, h = _oracle(orc, rule, n, m, r, V, W, H, G.REFERENCE_EPS, 4, update_w=False)
    assert np.array_equal(d.get_w(), W)
    assert rel_fro(d.get_h(), h) < TOL


def _object(native, rule, n, m, r, V, W, H):
    h = ctypes.c_void_p()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    capi.check(native.gpumlib_nmf_object_create(rule, G.FP32, n, m, r, p(V), p(W), p(H), ctypes.byref(h)),
               "gpumlib_nmf_object_create")
    return h


def _get(native, fn, h, count):
    out = np.empty(count, np.float32)
    capi.check(fn(h, out.ctypes.data_as(ctypes.c_void_p)))
    return out


@pytest.mark.parametrize("rule", [G.EUCLIDEAN, G.KL])
def test_class_api_above_rank_128(native, orc, cuda, rule):
    """NMF_Multiplicative* at r = 150 iterate (the classes no longer stop at the fused passes' limit)."""
    n, m, r = 220, 170, 150
    V, W, H = _problem(rule, n, m, r, seed=12)
    h = _object(native, rule, n, m, r, V, W, H)
    try:
        capi.check(native.gpumlib_nmf_object_do_iterations(h, 6, 1))
        _, w, hh = _oracle(orc, rule, n, m, r, V, W, H, G.REFERENCE_EPS, 6)
        assert rel_fro(_get(native, native.gpumlib_nmf_object_get_w, h, n * r), w) < TOL
        assert rel_fro(_get(native, native.gpumlib_nmf_object_get_h, h, r * m), hh) < TOL
        wh = _get(native, native.gpumlib_nmf_object_get_wh, h, n * m)
        assert rel_fro(wh, (G.as_matrix(w, n, r) @ G.as_matrix(hh, r, m)).T.ravel()) < TOL
    finally:
        native.gpumlib_nmf_object_destroy(h)


@pytest.mark.parametrize("rm", [(1, 5000), (20, 3001), (64, 8192), (129, 1000), (300, 2000), (1000, 70)])
def test_sum_h_any_rank(native, orc, cuda, rm):
    """KernelSumH semantics at ranks the reference allows (one workgroup per row there)."""
    r, m = rm
    rng = np.random.default_rng(r + m)
    H = rng.random(r * m, dtype=np.float32)
    H[(r // 2)::r] = 0.0  # one row sums to 0 -> eps
    ref = np.zeros(r)
    orc.lib().orc_sum_h_f64(r, m, orc._p64(H.astype(np.float64)), 1e-8, orc._p64(ref))
    dH, out = cuda.Buffer.from_array(H), cuda.Buffer(4 * r)
    capi.check(native.gpumlib_nmf_sum_h(dH.ptr, r, m, 1e-8, out.ptr, None))
    cuda.synchronize()
    got = out.download(np.float32)
    assert np.allclose(got, ref, rtol=2e-6)
    assert got[r // 2] == np.float32(1e-8)


def test_quality_f32_is_the_rounded_fp64_quality(native, orc, cuda):
    n = 123_457
    rng = np.random.default_rng(3)
    V = rng.random(n, dtype=np.float32) * 3
    WH = rng.random(n, dtype=np.float32) * 3
    dV, dWH = cuda.Buffer.from_array(V), cuda.Buffer.from_array(WH)
    wsb = int(native.gpumlib_nmf_quality_workspace_bytes(n))
    ws, o64, o32 = cuda.Buffer(wsb), cuda.Buffer(8), cuda.Buffer(4)
    capi.check(native.gpumlib_nmf_quality(dV.ptr, dWH.ptr, n, 1e-8, ws.ptr, wsb, o64.ptr, None))
    capi.check(native.gpumlib_nmf_quality_f32(dV.ptr, dWH.ptr, n, 1e-8, ws.ptr, wsb, o32.ptr, None))
    cuda.synchronize()
    q64 = o64.download(np.float64)[0]
    assert o32.download(np.float32)[0] == np.float32(q64)
    ref = orc.quality(V.astype(np.float64), WH.astype(np.float64), 1e-8)
    assert abs(q64 - ref) / abs(ref) < 1e-7  # eps added in float before log10, as the reference does


@pytest.mark.parametrize("rule", [G.EUCLIDEAN, G.KL])
def test_get_wh(native, orc, cuda, rule):
    """GetWH() = W H of the current factors (nmf_base.h:135-141), column-major n x m."""
    n, m, r = 333, 290, 24
    V, W, H = _problem(rule, n, m, r, seed=40)
    h = _object(native, rule, n, m, r, V, W, H)
    try:
        wh0 = _get(native, native.gpumlib_nmf_object_get_wh, h, n * m)
        assert rel_fro(wh0, (G.as_matrix(W, n, r).astype(np.float64) @ G.as_matrix(H, r, m)).T.ravel()) < 1e-6
        capi.check(native.gpumlib_nmf_object_do_iterations(h, 7, 1))
        gw = _get(native, native.gpumlib_nmf_object_get_w, h, n * r)
        gh = _get(native, native.gpumlib_nmf_object_get_h, h, r * m)
        wh = _get(native, native.gpumlib_nmf_object_get_wh, h, n * m)
        # the product of the returned factors (FP32 GEMM of r terms)
        assert rel_fro(wh, (G.as_matrix(gw, n, r).astype(np.float64) @ G.as_matrix(gh, r, m)).T.ravel()) < 1e-6
        _, w, hh = _oracle(orc, rule, n, m, r, V, W, H, G.REFERENCE_EPS, 7)
        assert rel_fro(wh, (G.as_matrix(w, n, r) @ G.as_matrix(hh, r, m)).T.ravel()) < TOL
        # GetWH leaves the state alone: iterating on gives the oracle's next states
        capi.check(native.gpumlib_nmf_object_do_iterations(h, 3, 1))
        _, w, hh = _oracle(orc, rule, n, m, r, V, W, H, G.REFERENCE_EPS, 10)
        assert rel_fro(_get(native, native.gpumlib_nmf_object_get_w, h, n * r), w) < TOL
    finally:
        native.gpumlib_nmf_object_destroy(h)


@pytest.mark.parametrize("rule,interval", [(G.EUCLIDEAN, 1), (G.KL, 1), (G.EUCLIDEAN, 3)])
def test_quality_improvement_sequence(native, orc, cuda, rule, interval):
    """QualityImprovement() after each DoIteration, against the oracle's quality of the same
    states.  Reference bookkeeping (nmf_base.h:70-89): at each quality point the quality of the
    previous quality point is read (0 at the first), and from the third reading on the
    improvement is that reading minus the one before, in float.  Here a quality point is every
    `interval` iterations (the reference's is whenever its side stream is idle), and its quality
    is that of the state on entry to the iteration."""
    n, m, r, iters = 256, 200, 12, 14
    V, W, H = _problem(rule, n, m, r, seed=71)
    h = _object(native, rule, n, m, r, V, W, H)
    try:
        capi.check(native.gpumlib_nmf_object_set_quality_interval(h, interval))
        got = []
        for _ in range(iters):
            capi.check(native.gpumlib_nmf_object_do_iteration(h, 1))
            got.append(native.gpumlib_nmf_object_quality_improvement(h))
    finally:
        native.gpumlib_nmf_object_destroy(h)
    # oracle quality of the state on entry to every iteration
    Vm = G.as_matrix(V, n, m).astype(np.float64)
    w, hh = _f64(W), _f64(H)
    q = []
    for t in range(iters):
        q.append(orc.np_quality(Vm, G.as_matrix(w, n, r), G.as_matrix(hh, r, m), G.REFERENCE_EPS))
        _oracle_step = orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter
        _oracle_step(n, m, r, _f64(V), w, hh, G.REFERENCE_EPS)
    # restatement of the bookkeeping over the quality points
    times, last, delta, value, expected = 0, np.float32(-1.0), np.float32(1.0), np.float32(0.0), []
    for t in range(iters):
        if t % interval == 0:
            reading = value
            if times > 2:
                delta = np.float32(reading - last)
            else:
                times += 1
            times += 1
            last = reading
            value = np.float32(q[t])  # computed at this point, read at the next one
        expected.append(delta)
    scale = max(abs(x) for x in q)
    for t, (a, b) in enumerate(zip(got, expected)):
        if b == np.float32(1.0):
            assert a == 1.0, (t, a)  # no improvement figure before the third reading
        else:
            assert abs(a - b) <= 1e-5 * scale, (t, a, b)
    assert any(b != np.float32(1.0) for b in expected)
