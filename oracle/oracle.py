"""Python access to the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker or as the timed CPU restatement; the product never does.

Two implementations of the same rules:
  * liborcnmf.so (nmf_oracle.c): textbook loops in FP64 / FP32, plus the
    OpenMP host-cores FP32 baseline;
  * the numpy FP64 functions below, for sizes where the textbook loops are too
    slow (full-size parity runs).  They follow the same reference lines:
      Euclidean  src/algorithms/nmf/nmf_multiplicative_euclidian.cpp:29-67,
                 src/cuda/nmf/nmf_kernels.cu:43-47
      KL         src/algorithms/nmf/nmf_multiplicative_divergence.cpp:29-50,
                 src/cuda/nmf/nmf_kernels.cu:63-233, 247-341
All matrices are column-major flat arrays (element (i, u) of an n x m matrix at
u*n + i), i.e. numpy arrays X of logical shape (n, m) are passed as X.T.ravel().
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborcnmf.so")

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i, _d, _f, _l, _c = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_long, ctypes.c_char

_SIGS = {
    "orc_gemm_f64": (None, [_c, _c, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]),
    "orc_gemm_f32": (None, [_c, _c, _i, _i, _i, _f, _fp, _i, _fp, _i, _f, _fp, _i]),
    "orc_mu_update_f32": (None, [_fp, _fp, _fp, _l, _f]),
    "orc_mu_update_f64": (None, [_dp, _dp, _dp, _l, _d]),
    "orc_sum_w_f64": (None, [_i, _i, _dp, _d, _dp]),
    "orc_sum_h_f64": (None, [_i, _i, _dp, _d, _dp]),
    "orc_sum_w_f32": (None, [_i, _i, _fp, _f, _fp]),
    "orc_sum_h_f32": (None, [_i, _i, _fp, _f, _fp]),
    "orc_kl_update_h_f64": (None, [_dp, _dp, _dp, _dp, _dp, _i, _i, _i, _d]),
    "orc_kl_update_w_f64": (None, [_dp, _dp, _dp, _dp, _dp, _i, _i, _i, _d]),
    "orc_kl_update_h_f32": (None, [_fp, _fp, _fp, _fp, _fp, _i, _i, _i, _f]),
    "orc_kl_update_w_f32": (None, [_fp, _fp, _fp, _fp, _fp, _i, _i, _i, _f]),
    "orc_euc_iter_f64": (None, [_i, _i, _i, _dp, _dp, _dp, _d, _i]),
    "orc_euc_iter_f32": (None, [_i, _i, _i, _fp, _fp, _fp, _f, _i]),
    "orc_kl_iter_f64": (None, [_i, _i, _i, _dp, _dp, _dp, _d, _i]),
    "orc_kl_iter_f32": (None, [_i, _i, _i, _fp, _fp, _fp, _f, _i]),
    "orc_euc_iter_f32_omp": (None, [_i, _i, _i, _fp, _fp, _fp, _f, _i]),
    "orc_kl_iter_f32_omp": (None, [_i, _i, _i, _fp, _fp, _fp, _f, _i]),
    "orc_kl_iter_f64_omp": (None, [_i, _i, _i, _dp, _dp, _dp, _d, _i]),
    "orc_add_update_f64": (None, [_dp, _dp, _dp, _l]),
    "orc_add_update_f32": (None, [_fp, _fp, _fp, _l]),
    "orc_add_kl_update_h_f64": (None, [_dp, _dp, _dp, _dp, _dp, _i, _i, _i, _d]),
    "orc_add_kl_update_w_f64": (None, [_dp, _dp, _dp, _dp, _dp, _i, _i, _i, _d]),
    "orc_add_kl_update_h_f32": (None, [_fp, _fp, _fp, _fp, _fp, _i, _i, _i, _f]),
    "orc_add_kl_update_w_f32": (None, [_fp, _fp, _fp, _fp, _fp, _i, _i, _i, _f]),
    "orc_add_euc_iter_f64": (None, [_i, _i, _i, _dp, _dp, _dp, _i, _i]),
    "orc_add_euc_iter_f32": (None, [_i, _i, _i, _fp, _fp, _fp, _i, _i]),
    "orc_add_kl_iter_f64": (None, [_i, _i, _i, _dp, _dp, _dp, _d, _i]),
    "orc_add_kl_iter_f32": (None, [_i, _i, _i, _fp, _fp, _fp, _f, _i]),
    "orc_add_kl_iter_f64_omp": (None, [_i, _i, _i, _dp, _dp, _dp, _d, _i]),
    "orc_reduce_sum": (_d, [_fp, _l, _d]),
    "orc_reduce_extreme": (None, [_fp, _l, _i, _fp, ctypes.POINTER(_i)]),
    "orc_frob_f64": (_d, [_i, _i, _i, _dp, _dp, _dp]),
    "orc_kl_f64": (_d, [_i, _i, _i, _dp, _dp, _dp]),
    "orc_quality_f64": (_d, [_l, _dp, _dp, _d]),
    "orc_omp_threads": (_i, []),
}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _p64(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _p32(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


# ---------------------------------------------------------------------------
# Wrappers of the C oracle (arrays are flat, column-major; updated in place).
def gemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc):
    if C.dtype == np.float64:
        lib().orc_gemm_f64(ta.encode(), tb.encode(), M, N, K, alpha, _p64(A), lda, _p64(B), ldb, beta, _p64(C), ldc)
    else:
        lib().orc_gemm_f32(ta.encode(), tb.encode(), M, N, K, alpha, _p32(A), lda, _p32(B), ldb, beta, _p32(C), ldc)
    return C


def euc_iter(n, m, r, V, W, H, eps, update_w=True, iters=1, omp=False):
    L = lib()
    if V.dtype == np.float64:
        fn, p, e = L.orc_euc_iter_f64, _p64, float(eps)
    else:
        fn, p, e = (L.orc_euc_iter_f32_omp if omp else L.orc_euc_iter_f32), _p32, float(np.float32(eps))
    for _ in range(iters):
        fn(n, m, r, p(V), p(W), p(H), e, 1 if update_w else 0)
    return W, H


def kl_iter(n, m, r, V, W, H, eps, update_w=True, iters=1, omp=False):
    L = lib()
    if V.dtype == np.float64:
        fn, p, e = (L.orc_kl_iter_f64_omp if omp else L.orc_kl_iter_f64), _p64, float(eps)
    else:
        fn, p, e = (L.orc_kl_iter_f32_omp if omp else L.orc_kl_iter_f32), _p32, float(np.float32(eps))
    for _ in range(iters):
        fn(n, m, r, p(V), p(W), p(H), e, 1 if update_w else 0)
    return W, H


def add_euc_iter(n, m, r, V, W, H, update_w=True, fix_h=False, iters=1):
    """NMF_AdditiveEuclidian iterations (H is left unchanged unless fix_h)."""
    L = lib()
    fn, p = (L.orc_add_euc_iter_f64, _p64) if V.dtype == np.float64 else (L.orc_add_euc_iter_f32, _p32)
    for _ in range(iters):
        fn(n, m, r, p(V), p(W), p(H), 1 if update_w else 0, 1 if fix_h else 0)
    return W, H


def add_kl_iter(n, m, r, V, W, H, eps, update_w=True, iters=1, omp=False):
    """NMF_AdditiveDivergence iterations (omp: the fused OpenMP FP64 loops)."""
    L = lib()
    if V.dtype == np.float64:
        fn, p, e = (L.orc_add_kl_iter_f64_omp if omp else L.orc_add_kl_iter_f64), _p64, float(eps)
    else:
        fn, p, e = L.orc_add_kl_iter_f32, _p32, float(np.float32(eps))
    for _ in range(iters):
        fn(n, m, r, p(V), p(W), p(H), e, 1 if update_w else 0)
    return W, H


def reduce_sum(x, factor=1.0):
    """factor * sum of a float32 array, accumulated in double (Reduction::Sum / Average)."""
    return lib().orc_reduce_sum(_p32(x), x.size, float(factor))


def reduce_extreme(x, is_max):
    """(value, index) of Reduction::Min/MinIndex (is_max False) or Max/MaxIndex."""
    v, i = ctypes.c_float(), ctypes.c_int()
    lib().orc_reduce_extreme(_p32(x), x.size, 1 if is_max else 0, ctypes.byref(v), ctypes.byref(i))
    return v.value, i.value


def frob(n, m, r, V, W, H):
    return lib().orc_frob_f64(n, m, r, _p64(V), _p64(W), _p64(H))


def kl_div(n, m, r, V, W, H):
    return lib().orc_kl_f64(n, m, r, _p64(V), _p64(W), _p64(H))


def quality(V, WH, eps):
    return lib().orc_quality_f64(V.size, _p64(V), _p64(WH), float(eps))


def omp_threads():
    return lib().orc_omp_threads()


# ---------------------------------------------------------------------------
# numpy FP64 restatement of the same rules for large sizes.  Inputs are n x m
# (V), n x r (W), r x m (H) 2-D arrays in float64; returns new arrays.
def np_euc_iter(V, W, H, eps, update_w=True):
    WtV = W.T @ V
    WtW = W.T @ W
    H = H * (WtV / (WtW @ H + eps))
    if update_w:
        VHt = V @ H.T
        HHt = H @ H.T
        W = W * (VHt / (W @ HHt + eps))
    return W, H


def np_kl_iter(V, W, H, eps, update_w=True):
    WH = W @ H
    sumW = np.maximum(W.sum(axis=0), eps)
    H = H * ((W.T @ (V / (WH + eps))) / sumW[:, None])
    if update_w:
        WH = W @ H
        sumH = np.maximum(H.sum(axis=1), eps)
        W = W * (((V / (WH + eps)) @ H.T) / sumH[None, :])
    return W, H


def _np_add(X, d1, d2):
    # nmf_kernels.cu:31-39: X + (X / d2) * (d1 - d2), negative results set to 0 (NaN kept)
    v = X + (X / d2) * (d1 - d2)
    return np.where(v < 0, 0.0, v)


def np_add_euc_iter(V, W, H, update_w=True, fix_h=False):
    """nmf_additive_euclidian.cpp:31-56 (H unchanged unless fix_h)."""
    if fix_h:
        H = _np_add(H, W.T @ V, (W.T @ W) @ H)
    if update_w:
        W = _np_add(W, V @ H.T, W @ (H @ H.T))
    return W, H


def np_add_kl_iter(V, W, H, eps, update_w=True):
    """nmf_additive_divergence.cpp:30-51 with the kernels of
    nmf_kernels_additive_divergence.cu:57-193 (sums of W * (ratio - 1))."""
    WH = W @ H
    sumW = np.maximum(W.sum(axis=0), eps)
    v = H + (H / sumW[:, None]) * (W.T @ (V / (WH + eps) - 1.0))
    H = np.where(v < 0, 0.0, v)
    if update_w:
        WH = W @ H
        sumH = np.maximum(H.sum(axis=1), eps)
        v = W + (W / sumH[None, :]) * ((V / (WH + eps) - 1.0) @ H.T)
        W = np.where(v < 0, 0.0, v)
    return W, H


# ---------------------------------------------------------------------------
# Low-precision operand restatement.  The BF16 and FP8 storage paths of the
# library are additions with no counterpart in the reference: they keep V in a
# narrow format and feed the MFMA passes narrow "shadow" copies of W and H,
# while the updates run on FP32 masters.  The functions below restate that
# arithmetic in FP64 (exact products and sums of the rounded operands), so that
# a test can separate the rounding the format implies (stated as an objective
# gap to the FP64 oracle) from the accumulation error of the kernels.
def bf16_round(x):
    """float32 -> nearest-even bfloat16 value, as float64 (finite inputs)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def e8m0_for_amax(amax):
    """Smallest s with amax / 2^(s-127) <= 448 (E4M3 maximum), s in [0, 254]; 127 for amax = 0."""
    amax = np.asarray(amax, dtype=np.float64)
    mant, ex = np.frexp(np.where(amax > 0, amax, 1.0))  # amax = mant * 2^ex, mant in [0.5, 1)
    s = (ex - 1) - 8 + (mant * 2 > 1.75) + 127
    return np.where(amax > 0, np.clip(s, 0, 254), 127).astype(np.int64)


def e4m3_round(y):
    """Non-negative finite y -> nearest OCP E4M3 value (ties to even), saturating at 448."""
    y = np.asarray(y, dtype=np.float64)
    _, ex = np.frexp(np.where(y > 0, y, 1.0))
    e = np.maximum(ex - 1, -6)  # binade exponent; below 2^-6 the spacing stays 2^-9
    step = np.ldexp(1.0, e - 3)
    return np.minimum(np.round(y / step) * step, 448.0)


def fp8_block_round(X, axis):
    """Dequantised E4M3 copy of X (2-D, >= 0) with one E8M0 scale per 32 entries along axis."""
    X = np.asarray(X, dtype=np.float32).astype(np.float64)
    Xt = X if axis == 0 else X.T
    out = np.empty_like(Xt)
    for k0 in range(0, Xt.shape[0], 32):
        blk = Xt[k0:k0 + 32]
        unit = np.ldexp(1.0, e8m0_for_amax(np.abs(blk).max(axis=0)) - 127)
        out[k0:k0 + 32] = e4m3_round(blk / unit) * unit
    return out if axis == 0 else out.T


def fp8_tile_round(V):
    """Dequantised E4M3 copy of V (n x m) with one E8M0 scale per 32 x 32 tile."""
    V = np.asarray(V, dtype=np.float32).astype(np.float64)
    out = np.empty_like(V)
    for i0 in range(0, V.shape[0], 32):
        for u0 in range(0, V.shape[1], 32):
            blk = V[i0:i0 + 32, u0:u0 + 32]
            unit = np.ldexp(1.0, int(e8m0_for_amax(np.abs(blk).max())) - 127)
            out[i0:i0 + 32, u0:u0 + 32] = e4m3_round(blk / unit) * unit
    return out


def np_euc_iter_shadow(Vq, W, H, eps, shadow_w, shadow_h, update_w=True):
    """Euclidean MU with the passes fed shadow operands: W^T V and W^T W from
    shadow_w(W), V H^T and H H^T from shadow_h(H_new); ratios and the masters
    in full precision (the same update rules as np_euc_iter)."""
    Ws = shadow_w(W)
    H = H * ((Ws.T @ Vq) / ((Ws.T @ Ws) @ H + eps))
    if update_w:
        Hs = shadow_h(H)
        W = W * ((Vq @ Hs.T) / (W @ (Hs @ Hs.T) + eps))
    return W, H


def np_kl_iter_shadow(Vq, W, H, eps, shadow, ratio_round, update_w=True):
    """KL MU (the rules of np_kl_iter, nmf_kernels.cu:63-341) with the passes fed shadow
    operands, as the BF16 storage path runs it: W H from shadow(W) and shadow(H); the ratio
    V / (W H + eps) rounded by ratio_round (bf16_round: to BF16 through FP32) before the second product;
    numerator and the clamped column (row) sums from the same shadows; the FP32 masters
    updated.  The W half takes shadow(new H) and the shadow of the old W."""
    Ws, Hs = shadow(W), shadow(H)
    R = ratio_round(Vq / (Ws @ Hs + eps))
    sumW = np.maximum(Ws.sum(axis=0), eps)
    H = H * ((Ws.T @ R) / sumW[:, None])
    if update_w:
        Hs = shadow(H)
        R = ratio_round(Vq / (Ws @ Hs + eps))
        sumH = np.maximum(Hs.sum(axis=1), eps)
        W = W * ((R @ Hs.T) / sumH[None, :])
    return W, H


def np_frob(V, W, H):
    D = V - W @ H
    return float(np.sum(D * D))


def np_kl(V, W, H):
    WH = W @ H
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(V > 0, V * np.log(np.where(V > 0, V, 1.0) / WH), 0.0)
    return float(np.sum(t - V + WH))


def np_quality(V, W, H, eps):
    WH = W @ H
    # the reference adds eps in float before log10 (nmf_kernels_quality.cu:34-35)
    whe = (WH.astype(np.float32) + np.float32(eps)).astype(np.float64)
    return float(np.sum(V * np.log10(whe) - WH))


# ---------------------------------------------------------------------------
# FacesNMF image loading (src/examples/FacesNMF/FacesNMFDlg.cpp:409-472).
_WS = b" \t\n\v\f\r"


def pgm_column(raw, equalize, standard=False, width=19, height=19):
    """The V column LoadImage builds from the bytes of one P5 file (float32, width*height).

    The header is read as `imageFile >> s >> width >> height >> maxGrayValue` reads it
    (each token after skipping whitespace, the last one stopping before the whitespace
    that follows it); the raster is read from there (:441), so the whitespace byte that
    ends the header is pixel 0.  standard=True skips that byte first (Netpbm).  Scaling
    (:464-469): p / 255 in float.  Histogram equalisation (:443-462): float counts,
    running sums, the first non-zero running sum as cdfmin, then (h - cdfmin) /
    (pixels - cdfmin) in float, looked up per pixel.  Returns None where LoadImage fails."""
    pos = 0

    def token():
        nonlocal pos
        while pos < len(raw) and raw[pos] in _WS:
            pos += 1
        start = pos
        while pos < len(raw) and raw[pos] not in _WS:
            pos += 1
        return raw[start:pos]

    try:
        magic, w, h, maxv = token(), int(token()), int(token()), int(token())
    except ValueError:
        return None
    if magic != b"P5" or w != width or h != height or maxv != 255:
        return None
    if standard:
        pos += 1
    pixels = width * height
    buf = np.frombuffer(raw[pos:pos + pixels], np.uint8)
    if buf.size < pixels:
        return None
    if not equalize:
        return buf.astype(np.float32) / np.float32(255.0)
    hist = np.cumsum(np.bincount(buf, minlength=256)).astype(np.float32)  # exact: counts <= pixels
    i = int(np.argmax(hist != 0))
    cdfmin = hist[i]
    with np.errstate(divide="ignore", invalid="ignore"):
        hist[i:] = (hist[i:] - cdfmin) / (np.float32(pixels) - cdfmin)
    return hist[buf]
