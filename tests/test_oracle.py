"""Pins the CPU oracle (oracle/nmf_oracle.c, oracle/oracle.py) without the reference.

The reference has no tests, golden vectors or fixtures for this path, and it
cannot be built here (nvcc/cuBLAS absent; HostNMF.h does not compile with g++).
So each oracle function is pinned by what the mathematics fixes: library
routines (numpy matmul), closed forms, the Lee & Seung monotonicity theorem,
fixed points, transpose duality, permutation equivariance, the rank-1 answer of
the SVD, and an independent second implementation (numpy) of the same rules.
Each check is chosen so that a plausible slip (a dropped eps, a transposed
operand, the wrong update order, a missing clamp) fails it.
"""
import numpy as np
import pytest

from conftest import rel_fro


def cm(x):
    """2-D array -> flat column-major float64 (copy)."""
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).T).ravel()


def mat(flat, n, m):
    return flat.reshape(m, n).T


def rand_problem(n, m, r, seed, exact=False):
    rng = np.random.default_rng(seed)
    W0, H0 = rng.random((n, r)), rng.random((r, m))
    V = W0 @ H0 + (0.0 if exact else 0.01 * rng.random((n, m)))
    W, H = rng.random((n, r)) + 0.1, rng.random((r, m)) + 0.1
    return V, W, H, W0, H0


def run(orc, rule, V, W, H, eps=1e-8, iters=1, update_w=True):
    n, m = V.shape
    r = W.shape[1]
    v, w, h = cm(V), cm(W), cm(H)
    (orc.euc_iter if rule == "euc" else orc.kl_iter)(n, m, r, v, w, h, eps, update_w=update_w, iters=iters)
    return mat(w, n, r), mat(h, r, m)


# ---------------------------------------------------------------- GEMM -----
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
def test_gemm_matches_numpy(orc, ta, tb):
    M, N, K = 7, 5, 9
    rng = np.random.default_rng(1)
    A = rng.random((M, K) if ta == "N" else (K, M))
    B = rng.random((K, N) if tb == "N" else (N, K))
    C = rng.random((M, N))
    c = cm(C)
    orc.gemm(ta, tb, M, N, K, 2.0, cm(A), A.shape[0], cm(B), B.shape[0], 0.5, c, M)
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    assert np.allclose(mat(c, M, N), 2.0 * opA @ opB + 0.5 * C, rtol=1e-13, atol=0)


def test_gemm_beta_zero_ignores_c(orc):
    c = np.full(4, np.nan)
    orc.gemm("N", "N", 2, 2, 1, 1.0, np.ones(2), 2, np.ones(2), 1, 0.0, c, 2)
    assert np.all(c == 1.0)


# ------------------------------------------------------- element-wise rule -
def test_mu_update_closed_forms(orc):
    eps = np.float32(1e-8)
    m = np.array([1.0, 2.0, 3.0, 0.5], np.float32)
    nm = np.array([2.0, 0.0, 4.0, 5.0], np.float32)
    dm = np.array([2.0, 1.0, 0.0, 5.0], np.float32)
    out = m.copy()
    orc.lib().orc_mu_update_f32(orc._p32(nm), orc._p32(dm), orc._p32(out), 4, eps)
    expect = m * (nm / (dm + eps))  # numpy float32: IEEE ops in the same order
    assert np.array_equal(out, expect)
    assert out[1] == 0.0                           # zero numerator
    assert out[2] == np.float32(3.0) * (np.float32(4.0) / eps)  # dm = 0 -> m nm / eps
    assert abs(out[0] - 1.0) < 1e-7 and abs(out[3] - 0.5) < 1e-7  # nm = dm -> ~m


# ---------------------------------------------------------- Euclidean MU ---
def test_euclidean_matches_numpy_restatement(orc):
    V, W, H, _, _ = rand_problem(40, 30, 6, 2)
    w1, h1 = run(orc, "euc", V, W, H, iters=3)
    w2, h2 = W, H
    for _ in range(3):
        w2, h2 = orc.np_euc_iter(V, w2, h2, 1e-8)
    assert rel_fro(w1, w2) < 1e-12 and rel_fro(h1, h2) < 1e-12


def test_euclidean_monotone(orc):
    V, W, H, _, _ = rand_problem(50, 40, 5, 3)
    n, m = V.shape
    prev = np.inf
    w, h = W, H
    for _ in range(60):
        f = orc.frob(n, m, 5, cm(V), cm(w), cm(h))
        assert f <= prev * (1 + 1e-12)
        prev = f
        w, h = run(orc, "euc", V, w, h)


def test_euclidean_fixed_point(orc):
    V, _, _, W0, H0 = rand_problem(30, 20, 4, 4, exact=True)
    w, h = run(orc, "euc", V, W0, H0, iters=5)
    assert rel_fro(w, W0) < 1e-7 and rel_fro(h, H0) < 1e-7


def test_euclidean_transpose_duality(orc):
    """U_W(V, W, H) == U_H(V^T, H^T, W^T)^T: catches swapped operands in either half."""
    V, W, H, _, _ = rand_problem(25, 35, 4, 5)
    w2, h2 = run(orc, "euc", V, W, H)                      # h2 = U_H(V,W,H), w2 = U_W(V,W,h2)
    _, wt = run(orc, "euc", V.T, h2.T, W.T, update_w=False)  # U_H(V^T, h2^T, W^T)
    assert rel_fro(wt.T, w2) < 1e-12


def test_euclidean_update_order(orc):
    """H is updated first from the old W, then W from the new H."""
    V, W, H, _, _ = rand_problem(20, 15, 3, 6)
    w2, h2 = run(orc, "euc", V, W, H)
    _, h_only = run(orc, "euc", V, W, H, update_w=False)
    assert np.array_equal(h2, h_only)
    w_alt = W * ((V @ H.T) / (W @ (H @ H.T) + 1e-8))  # W from the OLD H: must differ
    assert rel_fro(w2, w_alt) > 1e-6


def test_euclidean_permutation_equivariance(orc):
    V, W, H, _, _ = rand_problem(30, 22, 5, 7)
    p = np.array([3, 0, 4, 1, 2])
    w1, h1 = run(orc, "euc", V, W, H, iters=4)
    w2, h2 = run(orc, "euc", V, W[:, p], H[p, :], iters=4)
    assert rel_fro(w2, w1[:, p]) < 1e-12 and rel_fro(h2, h1[p, :]) < 1e-12


def test_euclidean_rank1_converges_to_svd(orc):
    """r = 1 on positive V: W H -> sigma_1 u_1 v_1^T (Perron-Frobenius + Eckart-Young)."""
    rng = np.random.default_rng(8)
    V = rng.random((12, 9)) + 0.5
    w, h = run(orc, "euc", V, rng.random((12, 1)) + 0.1, rng.random((1, 9)) + 0.1, eps=0.0, iters=3000)
    U, s, Vt = np.linalg.svd(V)
    best = s[0] * np.outer(U[:, 0], Vt[0])
    assert rel_fro(w @ h, best) < 1e-9


def test_euclidean_f32_tracks_f64_c1(orc):
    """FP32 restatement vs FP64 after 200 iterations at 256 x 128, r = 16."""
    n, m, r = 256, 128, 16
    V, W, H, _, _ = rand_problem(n, m, r, 9)
    v32, w32, h32 = (cm(x).astype(np.float32) for x in (V, W, H))
    v64, w64, h64 = (x.astype(np.float64) for x in (v32, w32, h32))
    orc.euc_iter(n, m, r, v32, w32, h32, 1e-8, iters=200)
    orc.euc_iter(n, m, r, v64, w64, h64, 1e-8, iters=200)
    assert rel_fro(w32, w64) < 1e-4 and rel_fro(h32, h64) < 1e-4


def test_euclidean_omp_baseline_matches(orc):
    n, m, r = 300, 200, 12
    V, W, H, _, _ = rand_problem(n, m, r, 10)
    a = [cm(x).astype(np.float32) for x in (V, W, H)]
    b = [x.copy() for x in a]
    orc.euc_iter(n, m, r, *a, 1e-8, iters=10)
    orc.euc_iter(n, m, r, *b, 1e-8, iters=10, omp=True)
    assert rel_fro(a[1], b[1]) < 1e-5 and rel_fro(a[2], b[2]) < 1e-5


# ---------------------------------------------------------------- KL MU ---
def poisson_problem(n, m, r, seed):
    rng = np.random.default_rng(seed)
    W0, H0 = rng.exponential(size=(n, r)), rng.exponential(size=(r, m))
    V = rng.poisson(W0 @ H0).astype(np.float64)
    return V, rng.random((n, r)) + 0.1, rng.random((r, m)) + 0.1


def test_kl_matches_numpy_restatement(orc):
    V, W, H = poisson_problem(40, 30, 5, 11)
    w1, h1 = run(orc, "kl", V, W, H, iters=3)
    w2, h2 = W, H
    for _ in range(3):
        w2, h2 = orc.np_kl_iter(V, w2, h2, 1e-8)
    assert rel_fro(w1, w2) < 1e-12 and rel_fro(h1, h2) < 1e-12


def test_kl_f64_fused_omp_matches_textbook(orc):
    """The fused OpenMP FP64 KL iteration (used at full size) equals the textbook loops."""
    V, W, H = poisson_problem(130, 90, 7, 17)
    n, m = V.shape
    a = [cm(x) for x in (V, W, H)]
    b = [x.copy() for x in a]
    orc.kl_iter(n, m, 7, *a, 1e-9, iters=5)
    orc.kl_iter(n, m, 7, *b, 1e-9, iters=5, omp=True)
    assert rel_fro(a[1], b[1]) < 1e-12 and rel_fro(a[2], b[2]) < 1e-12
    c = [cm(x) for x in (V, W, H)]
    orc.kl_iter(n, m, 7, *c, 1e-9, iters=1, update_w=False, omp=True)
    assert np.array_equal(c[1], cm(W))  # W untouched


def test_kl_monotone(orc):
    V, W, H = poisson_problem(45, 35, 4, 12)
    n, m = V.shape
    prev = np.inf
    w, h = W, H
    for _ in range(60):
        d = orc.kl_div(n, m, 4, cm(V), cm(w), cm(h))
        assert d <= prev * (1 + 1e-12)
        prev = d
        w, h = run(orc, "kl", V, w, h)


def test_kl_fixed_point(orc):
    _, _, _, W0, H0 = rand_problem(25, 18, 3, 13, exact=True)
    V = W0 @ H0
    w, h = run(orc, "kl", V, W0, H0, iters=5)
    assert rel_fro(w, W0) < 1e-7 and rel_fro(h, H0) < 1e-7


def test_kl_transpose_duality(orc):
    V, W, H = poisson_problem(28, 19, 3, 14)
    w2, h2 = run(orc, "kl", V, W, H)
    _, wt = run(orc, "kl", V.T, h2.T, W.T, update_w=False)
    assert rel_fro(wt.T, w2) < 1e-12


def test_kl_rank1_closed_form(orc):
    """r = 1: one H update gives H(u) = colsum(V)(u) / sum(W); then W(i) = rowsum(V)(i) / sum(H)."""
    V, _, _ = poisson_problem(16, 11, 2, 15)
    V += 1.0
    rng = np.random.default_rng(16)
    W, H = rng.random((16, 1)) + 0.2, rng.random((1, 11)) + 0.2
    w, h = run(orc, "kl", V, W, H, eps=0.0)
    assert np.allclose(h[0], V.sum(axis=0) / W.sum(), rtol=1e-12)
    assert np.allclose(w[:, 0], V.sum(axis=1) / h.sum(), rtol=1e-12)


def test_kl_sums_are_clamped(orc):
    """A zero column of W gives sumW = eps (nmf_kernels.cu:104), not 0 as in HostNMF."""
    n, r = 10, 3
    W = np.ones((n, r))
    W[:, 1] = 0.0
    out = np.zeros(r)
    orc.lib().orc_sum_w_f64(n, r, orc._p64(cm(W)), 1e-8, orc._p64(out))
    assert out[1] == 1e-8 and out[0] == n
    H = np.ones((r, 7))
    H[2] = 0.0
    out = np.zeros(r)
    orc.lib().orc_sum_h_f64(r, 7, orc._p64(cm(H)), 1e-8, orc._p64(out))
    assert out[2] == 1e-8 and out[0] == 7


# ------------------------------------------------------------ objectives ---
def test_objectives_closed_forms(orc):
    V = np.array([[1.0, 0.0], [2.0, 3.0]])
    W = np.array([[1.0], [2.0]])
    H = np.array([[1.0, 1.0]])
    WH = W @ H  # [[1,1],[2,2]]
    frob = orc.frob(2, 2, 1, cm(V), cm(W), cm(H))
    assert frob == pytest.approx(0.0 + 1.0 + 0.0 + 1.0)
    kl = orc.kl_div(2, 2, 1, cm(V), cm(W), cm(H))
    # 0 ln 0 = 0 for V = 0: that entry contributes WH = 1
    expect = (0 - 1 + 1) + (0 - 0 + 1) + (0 - 2 + 2) + (3 * np.log(3 / 2) - 3 + 2)
    assert kl == pytest.approx(expect, rel=1e-14)
    q = orc.quality(cm(V), cm(WH), 0.0)
    assert q == pytest.approx(sum(V.ravel() * np.log10(WH.ravel()) - WH.ravel()), rel=1e-14)
    assert orc.np_frob(V, W, H) == pytest.approx(frob)
    assert orc.np_kl(V, W, H) == pytest.approx(kl)


# ------------------------------------------ low-precision operand restatement -
def _e4m3_values():
    """All non-negative finite OCP E4M3 values, by code (0x00 .. 0x7e)."""
    out = []
    for q in range(0x7F):
        E, mnt = q >> 3, q & 7
        out.append(mnt / 8 * 2.0 ** -6 if E == 0 else (1 + mnt / 8) * 2.0 ** (E - 7))
    return np.array(out)


def test_bf16_round_matches_torch(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(21)
    x = np.concatenate([rng.random(5000) * 100, rng.random(1000) * 1e-30, [0.0, 1.0, 3.0]]).astype(np.float32)
    # exact midpoints between neighbouring bf16 values: ties go to the even value
    b = (x.view(np.uint32) & 0xFFFF0000).view(np.float32)
    mid = (b.view(np.uint32) | 0x8000).view(np.float32)
    x = np.concatenate([x, mid])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(orc.bf16_round(x), ref)


def test_e4m3_round_brute_force(orc):
    vals = _e4m3_values()
    rng = np.random.default_rng(22)
    ys = np.concatenate([rng.random(2000) * 500, rng.random(500) * 0.02, vals, (vals[1:] + vals[:-1]) / 2,
                         [0.0, 448.0, 449.0, 464.0, 500.0, 2.0 ** -10, 3 * 2.0 ** -10]])
    got = orc.e4m3_round(ys)
    for y, g in zip(ys, got):
        d = np.abs(vals - y)
        ties = np.nonzero(d == d.min())[0]
        code = ties[ties % 2 == 0][0] if np.any(ties % 2 == 0) else ties[0]
        assert g == vals[code], (y, g, vals[code])


@pytest.mark.parametrize("amax", [1e-3, 0.7, 1.0, 447.0, 448.0, 449.0, 3e4, 2.5e7])
def test_e8m0_smallest_fitting_power(orc, amax):
    s = int(orc.e8m0_for_amax(amax))
    assert amax / 2.0 ** (s - 127) <= 448.0 < amax / 2.0 ** (s - 128)


def test_fp8_tile_and_block_rounding(orc):
    """Every 32 x 32 tile (or 32-block) is rounded with its own smallest fitting power-of-two scale."""
    rng = np.random.default_rng(23)
    V = (rng.random((70, 45)) * np.exp(rng.normal(0, 3, (70, 45)))).astype(np.float32)
    q = orc.fp8_tile_round(V)
    vals = _e4m3_values()
    for i0 in (0, 32, 64):
        for u0 in (0, 32):
            blk = V[i0:i0 + 32, u0:u0 + 32].astype(np.float64)
            s = int(orc.e8m0_for_amax(blk.max()))
            unit = 2.0 ** (s - 127)
            y = (blk / unit).ravel()
            near = vals[np.abs(vals[None, :] - y[:, None]).argmin(axis=1)]
            # nearest value (ties are measure-zero for random data)
            assert np.array_equal(q[i0:i0 + 32, u0:u0 + 32].ravel(), near * unit)
    W = rng.random((70, 3)).astype(np.float32)
    qb = orc.fp8_block_round(W, 0)
    for a in range(3):
        for i0 in (0, 32, 64):
            blk = W[i0:i0 + 32, a].astype(np.float64)
            unit = 2.0 ** (int(orc.e8m0_for_amax(blk.max())) - 127)
            assert np.array_equal(qb[i0:i0 + 32, a], orc.e4m3_round(blk / unit) * unit)
    assert np.array_equal(orc.fp8_block_round(W.T, 1), qb.T)


def test_shadow_iteration_order(orc):
    """W is shadowed before the H half, the new H before the W half; identity shadows give np_euc_iter."""
    V, W, H, _, _ = rand_problem(20, 15, 3, 24)
    seen = []
    ident_w = lambda X: (seen.append(("w", X.copy())), X)[1]
    ident_h = lambda X: (seen.append(("h", X.copy())), X)[1]
    w1, h1 = orc.np_euc_iter_shadow(V, W, H, 1e-8, ident_w, ident_h)
    w2, h2 = orc.np_euc_iter(V, W, H, 1e-8)
    assert np.array_equal(w1, w2) and np.array_equal(h1, h2)
    assert [k for k, _ in seen] == ["w", "h"]
    assert np.array_equal(seen[0][1], W) and np.array_equal(seen[1][1], h2)


def test_kl_shadow_identity_is_np_kl_iter(orc):
    """Identity shadows and ratio rounding give np_kl_iter bit for bit; the shadow is taken of
    the old W, the old H and then the new H, and the W half keeps the old W's shadow."""
    V, W, H = poisson_problem(30, 22, 4, 31)
    seen = []
    ident = lambda X: (seen.append(X.copy()), X)[1]
    w1, h1 = orc.np_kl_iter_shadow(V, W, H, 1e-9, ident, lambda R: R)
    w2, h2 = orc.np_kl_iter(V, W, H, 1e-9)
    assert np.array_equal(w1, w2) and np.array_equal(h1, h2)
    assert len(seen) == 3
    assert np.array_equal(seen[0], W) and np.array_equal(seen[1], H) and np.array_equal(seen[2], h2)


def test_kl_shadow_transpose_duality(orc):
    """With BF16 shadows too, the W half on (V, W, H) is the H half on (V^T, H^T, W^T)."""
    V, W, H = poisson_problem(26, 33, 5, 32)
    w2, h2 = orc.np_kl_iter_shadow(V, W, H, 1e-9, orc.bf16_round, orc.bf16_round)
    _, wt = orc.np_kl_iter_shadow(V.T, h2.T, W.T, 1e-9, orc.bf16_round, orc.bf16_round, update_w=False)
    assert rel_fro(wt.T, w2) < 1e-13


def test_kl_shadow_fixed_point_uses_shadow_sums(orc):
    """V = bf16(W0) bf16(H0) exactly, masters W0, H0 not representable in BF16: the ratio rounds to
    exactly 1, so the factors stay put only if the numerator and the clamped sums both come from
    the shadows (sums of the FP32 masters would move every entry by ~2^-9)."""
    rng = np.random.default_rng(33)
    W0 = rng.random((24, 3)) + 0.5
    H0 = rng.random((3, 17)) + 0.5
    Ws, Hs = orc.bf16_round(W0), orc.bf16_round(H0)
    assert not np.array_equal(Ws, W0)
    V = Ws @ Hs
    w, h = orc.np_kl_iter_shadow(V, W0, H0, 1e-9, orc.bf16_round, orc.bf16_round)
    assert rel_fro(h, H0) < 1e-14 and rel_fro(w, W0) < 1e-14
    sums_from_masters = H0 * ((Ws.T @ np.ones_like(V)) / W0.sum(axis=0)[:, None])
    assert rel_fro(sums_from_masters, H0) > 1e-4


def test_kl_shadow_tracks_exact_iteration(orc):
    """BF16 operands move a Poisson problem's objective by far less than the iterations do."""
    V, W, H = poisson_problem(60, 50, 4, 34)
    n, m = V.shape
    w, h, we, he = W, H, W, H
    for _ in range(30):
        w, h = orc.np_kl_iter(V, w, h, 1e-9)
        we, he = orc.np_kl_iter_shadow(V, we, he, 1e-9, orc.bf16_round, orc.bf16_round)
    d0 = orc.kl_div(n, m, 4, cm(V), cm(W), cm(H))
    d, de = orc.kl_div(n, m, 4, cm(V), cm(w), cm(h)), orc.kl_div(n, m, 4, cm(V), cm(we), cm(he))
    assert abs(de - d) / d < 1e-3 and d < 0.5 * d0
