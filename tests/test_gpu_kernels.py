"""Kernel-level parity: the reference-shaped entry points against the CPU oracle."""
import ctypes

import numpy as np
import pytest

from conftest import rel_fro
from gpumlib_amd import capi
from gpumlib_amd import nmf as G

pytestmark = pytest.mark.gpu


def dev(hip, a):
    return hip.Buffer.from_array(np.ascontiguousarray(a))


_STREAM = None


def stream(hip):
    global _STREAM
    if _STREAM is None:
        _STREAM = hip.Stream()
    return _STREAM.handle


def get(buf, dtype=np.float32):
    if _STREAM is not None:
        _STREAM.synchronize()
    return buf.download(dtype)


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
# (256, 384, 320): whole 128 x 128 tiles and whole stages of K, 16-byte loads; (300, 260, 1000):
# interior tiles on the vector path beside edge tiles, and a partial last stage of K
@pytest.mark.parametrize("mnk", [(1, 1, 1), (33, 17, 5), (64, 64, 64), (130, 257, 95), (1000, 3, 700),
                                 (256, 384, 320), (300, 260, 1000)])
def test_sgemm(native, orc, cuda, ta, tb, mnk):
    M, N, K = mnk
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    Ar, Ac = (M, K) if ta == "N" else (K, M)
    Br, Bc = (K, N) if tb == "N" else (N, K)
    A = rng.random(Ar * Ac, dtype=np.float32)
    B = rng.random(Br * Bc, dtype=np.float32)
    C = rng.random(M * N, dtype=np.float32)
    alpha, beta = 1.5, 0.25
    dA, dB, dC = dev(cuda, A), dev(cuda, B), dev(cuda, C)
    capi.check(native.gpumlib_nmf_sgemm(ta.encode(), tb.encode(), M, N, K, alpha, dA.ptr, Ar,
                                        dB.ptr, Br, beta, dC.ptr, M, stream(cuda)))
    ref = C.astype(np.float64)
    orc.gemm(ta, tb, M, N, K, alpha, A.astype(np.float64), Ar, B.astype(np.float64), Br, beta, ref, M)
    assert rel_fro(get(dC), ref) < 1e-5


def test_sgemm_beta_zero_ignores_c(native, cuda):
    """beta == 0 never reads C (cuBLAS contract): NaN in C does not leak."""
    M = N = K = 40
    A = np.ones(M * K, np.float32)
    B = np.ones(K * N, np.float32)
    C = np.full(M * N, np.nan, np.float32)
    dA, dB, dC = dev(cuda, A), dev(cuda, B), dev(cuda, C)
    capi.check(native.gpumlib_nmf_sgemm(b"N", b"N", M, N, K, 1.0, dA.ptr, M, dB.ptr, K, 0.0,
                                        dC.ptr, M, stream(cuda)))
    assert np.all(get(dC) == K)


def test_sgemm_rejects_bad_flags(native, cuda):
    assert native.gpumlib_nmf_sgemm(b"X", b"N", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4, None) != 0


def bf16_value(bits):
    """BF16 bit patterns -> their values (float64)."""
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def e4m3_value(q):
    """OCP E4M3 bytes -> their values (float64), from the format's definition: sign, 4-bit
    exponent (bias 7), 3-bit mantissa, subnormals at exponent 0 (the NaN codes are not used)."""
    q = q.astype(np.int64)
    sign = np.where(q & 0x80, -1.0, 1.0)
    e, mt = (q >> 3) & 0xF, q & 7
    mag = np.where(e == 0, mt / 8.0 * 2.0 ** -6, (1.0 + mt / 8.0) * np.ldexp(1.0, e - 7))
    return sign * mag


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (33, 17, 5), (130, 257, 95), (1000, 3, 700), (256, 384, 320),
                                 (264, 136, 1000)])
def test_gemm_bf16(native, orc, cuda, ta, tb, mnk):
    """BF16 operands, FP32 accumulation: the FP64 GEMM of the stored values within FP32
    accumulation error (the products are exact in FP32)."""
    M, N, K = mnk
    rng = np.random.default_rng(M * 5 + N * 11 + K)
    Ar, Ac = (M, K) if ta == "N" else (K, M)
    Br, Bc = (K, N) if tb == "N" else (N, K)
    A = (rng.standard_normal(Ar * Ac).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    B = (rng.standard_normal(Br * Bc).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    C = rng.random(M * N, dtype=np.float32)
    alpha, beta = 0.75, -0.5
    dA, dB, dC = dev(cuda, A), dev(cuda, B), dev(cuda, C)
    capi.check(native.gpumlib_nmf_gemm_bf16(ta.encode(), tb.encode(), M, N, K, alpha, dA.ptr, Ar, dB.ptr, Br, beta,
                                            dC.ptr, M, stream(cuda)))
    ref = C.astype(np.float64)
    orc.gemm(ta, tb, M, N, K, alpha, bf16_value(A), Ar, bf16_value(B), Br, beta, ref, M)
    assert rel_fro(get(dC), ref) < 1e-5


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("mnk", [(33, 17, 5), (130, 257, 95), (64, 64, 512), (256, 384, 320), (272, 144, 1000)])
def test_gemm_fp8(native, orc, cuda, ta, tb, mnk):
    """E4M3 operands with one E8M0 scale each, on the FP8 matrix cores: the FP64 GEMM of the
    scaled stored values within 2e-4.  The products are exact, but the matrix cores sum each
    64-term block of K at less than FP32 accuracy (measured on these operands: 3.8e-5 relative
    at K = 5, 5.3e-5 at K = 64, e.g. 143808 for an exact 143819.4), so the bound is that of
    the hardware's block sums, not of FP32 accumulation."""
    M, N, K = mnk
    rng = np.random.default_rng(M * 13 + N + K * 3)
    Ar, Ac = (M, K) if ta == "N" else (K, M)
    Br, Bc = (K, N) if tb == "N" else (N, K)
    codes = np.array([c for c in range(256) if (c & 0x7F) != 0x7F], np.uint8)  # no NaN codes
    A = rng.choice(codes, Ar * Ac)
    B = rng.choice(codes, Br * Bc)
    sa, sb = 121, 130  # 2^-6 and 2^3
    C = rng.random(M * N, dtype=np.float32)
    dA, dB, dC = dev(cuda, A), dev(cuda, B), dev(cuda, C)
    capi.check(native.gpumlib_nmf_gemm_fp8(ta.encode(), tb.encode(), M, N, K, 1.25, dA.ptr, Ar, sa, dB.ptr, Br, sb,
                                           0.5, dC.ptr, M, stream(cuda)))
    ref = C.astype(np.float64)
    orc.gemm(ta, tb, M, N, K, 1.25, e4m3_value(A) * 2.0 ** (sa - 127), Ar, e4m3_value(B) * 2.0 ** (sb - 127), Br,
             0.5, ref, M)
    assert rel_fro(get(dC), ref) < 2e-4
    # the decoding is the library's own host decoder's
    assert all(native.gpumlib_nmf_fp8_decode(int(c), 127) == e4m3_value(np.array([c]))[0] for c in codes)


def test_mu_update_bit_exact(native, orc, cuda):
    rng = np.random.default_rng(3)
    E = 100_003
    nm, dm, mm = (rng.random(E, dtype=np.float32) for _ in range(3))
    dm[:10] = 0.0          # m * nm / eps
    nm[10:20] = dm[10:20]  # m * dm / (dm + eps)
    ref = mm.copy()
    orc.lib().orc_mu_update_f32(orc._p32(nm), orc._p32(dm), orc._p32(ref), E, np.float32(1e-8))
    dnm, ddm, dmm = dev(cuda, nm), dev(cuda, dm), dev(cuda, mm)
    capi.check(native.gpumlib_nmf_mu_update(dnm.ptr, ddm.ptr, dmm.ptr, E, 1e-8, stream(cuda)))
    assert np.array_equal(get(dmm), ref)
    # the reference-named launcher (grid/block ignored, eps = 1e-8f, null stream)
    dmm2 = dev(cuda, mm)
    cuda.synchronize()
    capi.check(native.gpumlib_cuda_nmf_multiplicative_euclidean_update_matrix(1, 512, dnm.ptr, ddm.ptr,
                                                                              dmm2.ptr, E))
    cuda.synchronize()
    assert np.array_equal(get(dmm2), ref)


def test_sums_clamped(native, orc, cuda):
    n, m, r = 1000, 777, 40
    rng = np.random.default_rng(11)
    W = rng.random(n * r, dtype=np.float32)
    H = rng.random(r * m, dtype=np.float32)
    W[3 * n:4 * n] = 0.0            # column 3 sums to 0 -> eps
    H[5::r] = 0.0                   # row 5 sums to 0 -> eps
    sw = np.zeros(r)
    sh = np.zeros(r)
    orc.lib().orc_sum_w_f64(n, r, orc._p64(W.astype(np.float64)), 1e-8, orc._p64(sw))
    orc.lib().orc_sum_h_f64(r, m, orc._p64(H.astype(np.float64)), 1e-8, orc._p64(sh))
    dW, dH = dev(cuda, W), dev(cuda, H)
    osw = cuda.Buffer(4 * r)
    osh = cuda.Buffer(4 * r)
    capi.check(native.gpumlib_nmf_sum_w(dW.ptr, n, r, 1e-8, osw.ptr, stream(cuda)))
    capi.check(native.gpumlib_nmf_sum_h(dH.ptr, r, m, 1e-8, osh.ptr, stream(cuda)))
    assert np.allclose(get(osw), sw, rtol=1e-5)
    assert np.allclose(get(osh), sh, rtol=1e-5)
    assert get(osw)[3] == np.float32(1e-8) and get(osh)[5] == np.float32(1e-8)


def test_quality(native, orc, cuda):
    n = 300_001
    rng = np.random.default_rng(12)
    V = rng.random(n, dtype=np.float32) * 5
    WH = rng.random(n, dtype=np.float32) * 5
    WH[:7] = 0.0
    ref = orc.quality(V.astype(np.float64), WH.astype(np.float64), 1e-8)
    dV, dWH = dev(cuda, V), dev(cuda, WH)
    wsb = int(native.gpumlib_nmf_quality_workspace_bytes(n))
    ws = cuda.Buffer(wsb)
    out = cuda.Buffer(8)
    capi.check(native.gpumlib_nmf_quality(dV.ptr, dWH.ptr, n, 1e-8, ws.ptr, wsb,
                                          out.ptr, stream(cuda)))
    assert abs(float(get(out, np.float64)[0]) - ref) / abs(ref) < 1e-6


@pytest.mark.parametrize("nmr", [(96, 80, 16), (257, 131, 37)])
def test_kl_updates_with_materialised_wh(native, orc, cuda, nmr):
    """Reference kernel contract (nmf_kernels.cu:247-341) with WH given."""
    n, m, r = nmr
    V = G.gen_poisson(n, m, 8, seed=n)
    W, H = G.gen_init(n, m, r, seed=m)
    WH = np.zeros(n * m, np.float32)
    orc.gemm("N", "N", n, m, r, 1.0, W, n, H, r, 0.0, WH, n)
    sw = np.zeros(r, np.float32)
    sh = np.zeros(r, np.float32)
    orc.lib().orc_sum_w_f32(n, r, orc._p32(W), np.float32(1e-8), orc._p32(sw))
    orc.lib().orc_sum_h_f32(r, m, orc._p32(H), np.float32(1e-8), orc._p32(sh))
    h_ref, w_ref = H.astype(np.float64), W.astype(np.float64)
    orc.lib().orc_kl_update_h_f64(orc._p64(h_ref), orc._p64(W.astype(np.float64)), orc._p64(V.astype(np.float64)),
                                  orc._p64(WH.astype(np.float64)), orc._p64(sw.astype(np.float64)), n, m, r, 1e-8)
    orc.lib().orc_kl_update_w_f64(orc._p64(w_ref), orc._p64(H.astype(np.float64)), orc._p64(V.astype(np.float64)),
                                  orc._p64(WH.astype(np.float64)), orc._p64(sh.astype(np.float64)), n, m, r, 1e-8)
    dV, dW, dH, dWH = dev(cuda, V), dev(cuda, W), dev(cuda, H), dev(cuda, WH)
    dsw, dsh = dev(cuda, sw), dev(cuda, sh)
    dH2 = dev(cuda, H)
    dW2 = dev(cuda, W)
    capi.check(native.gpumlib_nmf_kl_update_h(dH2.ptr, dW.ptr, dV.ptr, dWH.ptr,
                                              dsw.ptr, n, m, r, 1e-8, stream(cuda)))
    capi.check(native.gpumlib_nmf_kl_update_w(dW2.ptr, dH.ptr, dV.ptr, dWH.ptr,
                                              dsh.ptr, n, m, r, 1e-8, stream(cuda)))
    assert rel_fro(get(dH2), h_ref) < 1e-5
    assert rel_fro(get(dW2), w_ref) < 1e-5
    # reference-named launchers (dim3 arguments ignored; null stream)
    g = capi.Dim3(1, 1, 1)
    dH3, dW3 = dev(cuda, H), dev(cuda, W)
    cuda.synchronize()
    capi.check(native.gpumlib_cuda_nmf_multiplicative_divergence_update_matrix_h(
        g, g, dH3.ptr, dW.ptr, dV.ptr, dWH.ptr, dsw.ptr, n, m, r))
    capi.check(native.gpumlib_cuda_nmf_multiplicative_divergence_update_matrix_w(
        g, g, dW3.ptr, dH.ptr, dV.ptr, dWH.ptr, dsh.ptr, n, m, r))
    cuda.synchronize()
    assert np.array_equal(get(dH3), get(dH2))
    assert np.array_equal(get(dW3), get(dW2))
