"""Pins the oracle's additive-update variants (NMF_AdditiveEuclidian,
NMF_AdditiveDivergence) without the reference.

* Element rule X + (X/d2)(d1 - d2) (nmf_kernels.cu:31-39) equals X d1 / d2, i.e. the
  multiplicative element rule with eps = 0, which is pinned on its own; special
  cases: d1 = d2 leaves X exact, d1 = 0 gives 0 (clamped), 0/0 stays NaN.
* The additive Euclidean iteration leaves H unchanged, as the reference does
  (nmf_additive_euclidian.cpp:42 hands H to the multiplicative launcher as its
  numerator, so only the scratch deltaH2 is written); its W half equals the
  multiplicative W half at eps = 0 (through transpose duality with the H half).
* The additive KL iteration equals the multiplicative KL iteration whenever the
  clamped sums are the true sums (sum W (ratio - 1) = sum W ratio - sum W); a
  zero column of W (sum clamped at eps) tells them apart: the additive rule keeps
  that row of H, the multiplicative rule zeroes it.
* The numpy FP64 restatement used at large sizes agrees with the C loops; the
  FP32 restatement tracks FP64.
"""
import numpy as np

from conftest import rel_fro
from test_oracle import cm, mat, poisson_problem, rand_problem


def add_run(orc, rule, V, W, H, eps=1e-8, iters=1, update_w=True, fix_h=False):
    n, m = V.shape
    r = W.shape[1]
    v, w, h = cm(V), cm(W), cm(H)
    if rule == "euc":
        orc.add_euc_iter(n, m, r, v, w, h, update_w=update_w, fix_h=fix_h, iters=iters)
    else:
        orc.add_kl_iter(n, m, r, v, w, h, eps, update_w=update_w, iters=iters)
    return mat(w, n, r), mat(h, r, m)


def mu_run(orc, rule, V, W, H, eps, update_w=True):
    n, m = V.shape
    r = W.shape[1]
    v, w, h = cm(V), cm(W), cm(H)
    (orc.euc_iter if rule == "euc" else orc.kl_iter)(n, m, r, v, w, h, eps, update_w=update_w)
    return mat(w, n, r), mat(h, r, m)


def test_add_update_equals_mu_rule_at_eps0(orc):
    rng = np.random.default_rng(30)
    d1, d2, X = rng.random(500) + 0.05, rng.random(500) + 0.05, rng.random(500) + 0.05
    a = X.copy()
    orc.lib().orc_add_update_f64(orc._p64(d1), orc._p64(d2), orc._p64(a), a.size)
    b = X.copy()
    orc.lib().orc_mu_update_f64(orc._p64(d1), orc._p64(d2), orc._p64(b), b.size, 0.0)
    assert np.max(np.abs(a - b) / b) < 1e-13


def test_add_update_special_cases(orc):
    X = np.array([0.7, 0.7, 0.0, 2.0])
    d1 = np.array([0.3, 0.0, 0.0, 5.0])
    d2 = np.array([0.3, 0.9, 0.0, 5.0])
    orc.lib().orc_add_update_f64(orc._p64(d1), orc._p64(d2), orc._p64(X), X.size)
    assert X[0] == 0.7 and X[3] == 2.0  # d1 == d2: unchanged, exactly
    assert 0.0 <= X[1] < 1e-15          # d1 == 0: zero (a negative rounding is clamped)
    assert np.isnan(X[2])               # 0/0: NaN is kept, as `v < 0` is false for it
    Xf = np.array([0.7, 0.7], np.float32)
    orc.lib().orc_add_update_f32(orc._p32(np.array([0.0, 0.3], np.float32)), orc._p32(np.array([0.9, 0.3], np.float32)),
                                 orc._p32(Xf), 2)
    assert 0.0 <= Xf[0] < 1e-7 and Xf[1] == np.float32(0.7)


def test_add_euclidean_keeps_h_and_w_half_is_mu_at_eps0(orc):
    V, W, H, _, _ = rand_problem(37, 29, 5, 31)
    w, h = add_run(orc, "euc", V, W, H)
    assert np.array_equal(h, H)
    # W half on (V, W, H) == H half on (V^T, H^T, W^T), multiplicative rule at eps = 0
    _, wt = mu_run(orc, "euc", V.T, H.T, W.T, 0.0, update_w=False)
    assert rel_fro(w, wt.T) < 1e-12
    w0, h0 = add_run(orc, "euc", V, W, H, update_w=False)
    assert np.array_equal(w0, W) and np.array_equal(h0, H)


def test_add_euclidean_fixed_h_variant(orc):
    V, W, H, _, _ = rand_problem(33, 26, 4, 32)
    w, h = add_run(orc, "euc", V, W, H, fix_h=True)
    wm, hm = mu_run(orc, "euc", V, W, H, 0.0)
    assert rel_fro(h, hm) < 1e-12 and rel_fro(w, wm) < 1e-12


def test_add_euclidean_objective_non_increasing(orc):
    """The W half is Lee & Seung's W step, so with H fixed the objective cannot rise."""
    V, W, H, _, _ = rand_problem(40, 30, 4, 33)
    n, m = V.shape
    prev, w = np.inf, W
    for _ in range(40):
        f = orc.frob(n, m, 4, cm(V), cm(w), cm(H))
        assert f <= prev * (1 + 1e-12)
        prev = f
        w, _ = add_run(orc, "euc", V, w, H)


def test_add_kl_equals_mu_when_sums_unclamped(orc):
    V, W, H = poisson_problem(41, 27, 5, 34)
    w, h = add_run(orc, "kl", V, W, H, iters=1)
    wm, hm = mu_run(orc, "kl", V, W, H, 1e-8)
    assert rel_fro(h, hm) < 1e-12 and rel_fro(w, wm) < 1e-12


def test_add_kl_clamped_column_keeps_h_row(orc):
    V, W, H = poisson_problem(20, 15, 3, 35)
    W[:, 1] = 0.0
    _, h = add_run(orc, "kl", V, W, H, update_w=False)
    _, hm = mu_run(orc, "kl", V, W, H, 1e-8, update_w=False)
    assert np.array_equal(h[1], H[1])  # sum_i W(i,1)(ratio - 1) == 0 exactly
    assert np.all(hm[1] == 0.0)        # multiplicative: 0 / eps
    assert rel_fro(np.delete(h, 1, 0), np.delete(hm, 1, 0)) < 1e-12


def test_add_kl_update_kernels_closed_form(orc):
    """Kernel-level rules against numpy on materialised WH (both halves)."""
    V, W, H = poisson_problem(17, 13, 3, 36)
    WH = W @ H + 0.05
    eps = 1e-8
    sumW = np.maximum(W.sum(axis=0), eps)
    h = cm(H)
    orc.lib().orc_add_kl_update_h_f64(orc._p64(h), orc._p64(cm(W)), orc._p64(cm(V)), orc._p64(cm(WH)),
                                      orc._p64(sumW), 17, 13, 3, eps)
    R = V / (WH + eps) - 1.0
    expect = H + H / sumW[:, None] * (W.T @ R)
    assert rel_fro(mat(h, 3, 13), np.maximum(expect, 0)) < 1e-13
    sumH = np.maximum(H.sum(axis=1), eps)
    w = cm(W)
    orc.lib().orc_add_kl_update_w_f64(orc._p64(w), orc._p64(cm(H)), orc._p64(cm(V)), orc._p64(cm(WH)),
                                      orc._p64(sumH), 17, 13, 3, eps)
    expect = W + W / sumH[None, :] * (R @ H.T)
    assert rel_fro(mat(w, 17, 3), np.maximum(expect, 0)) < 1e-13


def test_add_numpy_restatements_match_c(orc):
    V, W, H, _, _ = rand_problem(36, 28, 4, 37)
    w1, h1 = add_run(orc, "euc", V, W, H, iters=3)
    w2, h2 = W, H
    for _ in range(3):
        w2, h2 = orc.np_add_euc_iter(V, w2, h2)
    assert rel_fro(w1, w2) < 1e-12 and np.array_equal(h1, h2)
    w1, h1 = add_run(orc, "euc", V, W, H, iters=2, fix_h=True)
    w2, h2 = W, H
    for _ in range(2):
        w2, h2 = orc.np_add_euc_iter(V, w2, h2, fix_h=True)
    assert rel_fro(w1, w2) < 1e-12 and rel_fro(h1, h2) < 1e-12
    V, W, H = poisson_problem(36, 28, 4, 38)
    w1, h1 = add_run(orc, "kl", V, W, H, iters=3)
    w2, h2 = W, H
    for _ in range(3):
        w2, h2 = orc.np_add_kl_iter(V, w2, h2, 1e-8)
    assert rel_fro(w1, w2) < 1e-12 and rel_fro(h1, h2) < 1e-12


def test_add_f32_tracks_f64(orc):
    V, W, H, _, _ = rand_problem(64, 48, 6, 39)
    n, m = V.shape
    for rule in ("euc", "kl"):
        a = [cm(x) for x in (V, W, H)]
        b = [x.astype(np.float32) for x in a]
        if rule == "euc":
            orc.add_euc_iter(n, m, 6, *a, iters=10)
            orc.add_euc_iter(n, m, 6, *b, iters=10)
        else:
            orc.add_kl_iter(n, m, 6, *a, 1e-8, iters=10)
            orc.add_kl_iter(n, m, 6, *b, 1e-8, iters=10)
        assert rel_fro(b[1].astype(np.float64), a[1]) < 1e-5 and rel_fro(b[2].astype(np.float64), a[2]) < 1e-5


def test_add_kl_f64_fused_omp_matches_textbook(orc):
    V, W, H = poisson_problem(130, 90, 7, 40)
    n, m = V.shape
    a = [cm(x) for x in (V, W, H)]
    b = [x.copy() for x in a]
    orc.add_kl_iter(n, m, 7, *a, 1e-9, iters=5)
    orc.add_kl_iter(n, m, 7, *b, 1e-9, iters=5, omp=True)
    assert rel_fro(a[1], b[1]) < 1e-12 and rel_fro(a[2], b[2]) < 1e-12
