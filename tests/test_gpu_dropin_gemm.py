"""The container GEMM surface of the drop-in C++ API (DeviceMatrix<float>::Multiply,
MultiplyBySelfTranspose: reference src/memory/DeviceMatrix.h:195-227) and the C++ helpers
KernelSumW / KernelSumH / KernelNMFquality (nmf_kernels.h:44-73), run by the native probe
tests/native/dropin_gemm_probe.cpp on the GPU and checked here against FP64 products.

The probe writes each case's operands in logical row-major order, so the checks are
independent of the storage orders the C++ code was given; they also check the
reference's contract that the result C comes back column-major."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "gpumlib_amd", "bin", "dropin_gemm_probe")


@pytest.fixture(scope="module")
def cases(native, tmp_path_factory):
    out = tmp_path_factory.mktemp("dropin_gemm")
    p = subprocess.run([PROBE, str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    res = []
    for f in sorted(out.glob("case_*.bin"), key=lambda x: int(x.stem.split("_")[1])):
        raw = f.read_bytes()
        hdr = np.frombuffer(raw[:32], np.int32)
        ab = np.frombuffer(raw[32:40], np.float32)
        res.append((hdr, float(ab[0]), float(ab[1]), np.frombuffer(raw[40:], np.float32)))
    return res


def _take(data, off, rows, cols):
    return data[off:off + rows * cols].reshape(rows, cols).astype(np.float64), off + rows * cols


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_probe_ran_every_case(cases):
    kinds = [int(h[0]) for h, *_ in cases]
    assert kinds.count(0) == 42 and kinds.count(1) == 20 and kinds.count(2) == 2 and kinds.count(3) == 1


def test_multiply(cases):
    seen = set()
    for hdr, alpha, beta, data in cases:
        if hdr[0] != 0:
            continue
        _, M, N, K, arow, brow, crow0, crow1 = (int(x) for x in hdr)
        A, o = _take(data, 0, M, K)
        B, o = _take(data, o, K, N)
        C0, o = _take(data, o, M, N)
        C, o = _take(data, o, M, N)
        assert o == data.size
        ref = alpha * (A @ B) + beta * C0
        # FP32 accumulation over K products of values in [0.25, 1.25)
        assert rel(C, ref) < 1e-5, (M, N, K, arow, brow, crow0)
        assert crow1 == 0  # C is returned column-major
        if beta != 0:
            assert crow0 == 0  # beta != 0 needs a column-major C (DeviceMatrix.h:218-223)
        seen.add((arow, brow, crow0))
    assert seen == {(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)}


def test_multiply_by_self_transpose(cases):
    seen = set()
    for hdr, alpha, beta, data in cases:
        if hdr[0] != 1:
            continue
        _, M, _, K, arow, _, crow0, crow1 = (int(x) for x in hdr)
        A, o = _take(data, 0, M, K)
        C0, o = _take(data, o, M, M)
        C, o = _take(data, o, M, M)
        ref = alpha * (A @ A.T) + beta * C0
        assert rel(C, ref) < 1e-5, (M, K, arow, crow0)
        assert crow1 == 0
        seen.add((arow, crow0))
    assert seen == {(0, 0), (0, 1), (1, 0), (1, 1)}


def test_kernel_sums(cases, orc):
    count = 0
    for hdr, _, _, data in cases:
        if hdr[0] != 2:
            continue
        _, n, m, r = (int(x) for x in hdr[:4])
        W, o = _take(data, 0, n, r)
        H, o = _take(data, o, r, m)
        sw, sh = data[o:o + r], data[o + r:o + 2 * r]
        ref_w = np.maximum(W.sum(axis=0), 1e-8)
        ref_h = np.maximum(H.sum(axis=1), 1e-8)
        assert np.allclose(sw, ref_w, rtol=2e-6) and np.allclose(sh, ref_h, rtol=2e-6)
        assert sw[r // 2] == np.float32(1e-8) and sh[r - 1] == np.float32(1e-8)  # clamped at eps
        count += 1
    assert count == 2


def test_kernel_nmf_quality(cases, orc):
    hdr, _, _, data = next(c for c in cases if c[0][0] == 3)
    count = int(hdr[1])
    V, WH, q = data[:count], data[count:2 * count], data[2 * count]
    ref = orc.quality(V.astype(np.float64), WH.astype(np.float64), 1e-8)
    assert abs(float(q) - ref) <= 1e-6 * abs(ref)
