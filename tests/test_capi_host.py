"""The C-ABI library loads, exports every symbol the header declares, and rejects
bad arguments before touching a device (CPU only: no compute calls)."""
import os
import re

import pytest

from gpumlib_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_INVALID_VALUE = 1


def header_functions():
    inc = os.path.join(ROOT, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(gpumlib_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = header_functions()
    assert "gpumlib_nmf_iterate" in names and "gpumlib_cuda_nmf_multiplicative_euclidean_update_matrix" in names
    assert len(names) >= 30


def test_every_declared_symbol_is_exported(native):
    missing = [n for n in header_functions() if not hasattr(native, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(capi.SIGNATURES) == header_functions()


def test_workspace_bytes(native):
    ok = native.gpumlib_nmf_workspace_bytes(capi.EUCLIDEAN, capi.FP32, 16384, 8192, 64)
    assert ok > 0 and ok < 64 << 20  # partials only: far below V itself (512 MiB)
    assert native.gpumlib_nmf_workspace_bytes(capi.KL, capi.FP32, 1, 1, 1) > 0
    for bad in [(capi.EUCLIDEAN, capi.FP32, 10, 10, 0), (7, capi.FP32, 10, 10, 4), (capi.EUCLIDEAN, 9, 10, 10, 4),
                (capi.KL, capi.FP32, 0, 10, 4), (capi.EUCLIDEAN, capi.BF16, 10, 10, 129),
                (capi.EUCLIDEAN, capi.FP8, 10, 10, 65)]:
        assert native.gpumlib_nmf_workspace_bytes(*bad) == 0
    # ranks above 128 plan the unfused iteration (FP32 storage only): its workspace holds an
    # n x m buffer (W H, or the KL ratio) beside the products
    for rule in (capi.EUCLIDEAN, capi.KL, capi.ADDITIVE_KL, capi.ADDITIVE_EUCLIDEAN):
        big = native.gpumlib_nmf_workspace_bytes(rule, capi.FP32, 1000, 700, 129)
        assert big >= 4 * 1000 * 700 + 2 * 4 * 129 * 1000
    assert native.gpumlib_nmf_workspace_bytes(capi.EUCLIDEAN, capi.FP32, 10, 10, 300) > 0
    # the largest rank of each precision plans: 128 (FP32, BF16), 64 (FP8)
    for ok in [(capi.EUCLIDEAN, capi.BF16, 10, 10, 128), (capi.EUCLIDEAN, capi.BF16, 10, 10, 65),
               (capi.EUCLIDEAN, capi.FP8, 10, 10, 64), (capi.KL, capi.FP32, 10, 10, 128)]:
        assert native.gpumlib_nmf_workspace_bytes(*ok) > 0


def test_kl_bf16_plan(native):
    """KL with BF16 storage plans ranks up to 128 (the multiplicative rule only); its workspace
    adds the rank-contiguous BF16 shadows of W and H (rows of r rounded up to 32, 64 or 128) to
    what the Euclidean BF16 path holds; eps must be positive (refused before any device work)."""
    import ctypes

    n, m = 1000, 700
    for r in (1, 17, 32, 33, 64, 65, 100, 128):
        kl = native.gpumlib_nmf_workspace_bytes(capi.KL, capi.BF16, n, m, r)
        rps = 32 if r <= 32 else (64 if r <= 64 else 128)
        assert kl >= 2 * rps * (1024 + 768), r
    for bad in [(capi.KL, capi.BF16, n, m, 129), (capi.ADDITIVE_KL, capi.BF16, n, m, 8), (capi.KL, capi.FP8, n, m, 8),
                (capi.KL, capi.FP32X, n, m, 8)]:
        assert native.gpumlib_nmf_workspace_bytes(*bad) == 0
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: eps is checked first
    ws = native.gpumlib_nmf_workspace_bytes(capi.KL, capi.BF16, n, m, 16)
    for eps in (0.0, -1e-9):
        assert native.gpumlib_nmf_iterate(capi.KL, capi.BF16, n, m, 16, fake, fake, fake, eps, 1, 1, capi.OBJ_NONE,
                                          None, fake, ws, None) == HIP_INVALID_VALUE


def test_fp32x_plan_and_store(native):
    """FP32X storage: Euclidean only, ranks up to 128; the V store holds V and V^T in FP32,
    each padded to 128-deep chunks and 64-column tiles; the workspace holds three BF16
    piece planes of W and of H^T beside the partials."""
    n, m = 1000, 700
    ws32 = native.gpumlib_nmf_workspace_bytes(capi.EUCLIDEAN, capi.FP32, n, m, 64)
    wsx = native.gpumlib_nmf_workspace_bytes(capi.EUCLIDEAN, capi.FP32X, n, m, 64)
    assert wsx > 0 and wsx >= 3 * 2 * 64 * (1024 + 768)
    assert native.gpumlib_nmf_workspace_bytes(capi.EUCLIDEAN, capi.FP32X, n, m, 128) > 0
    for bad in [(capi.EUCLIDEAN, capi.FP32X, n, m, 129), (capi.KL, capi.FP32X, n, m, 8),
                (capi.ADDITIVE_EUCLIDEAN, capi.FP32X, n, m, 8), (capi.EUCLIDEAN, capi.FP32X, 0, m, 8)]:
        assert native.gpumlib_nmf_workspace_bytes(*bad) == 0
    assert ws32 > 0
    pad = lambda k: (k + 127) // 128 * 128  # noqa: E731
    tiles = lambda c: (c + 63) // 64  # noqa: E731
    a256 = lambda b: (b + 255) // 256 * 256  # noqa: E731
    for nn, mm in [(1000, 700), (1, 1), (128, 64), (4097, 130)]:
        want = a256(pad(nn) * 64 * tiles(mm) * 4) + a256(pad(mm) * 64 * tiles(nn) * 4)
        assert native.gpumlib_nmf_vstore_bytes(capi.FP32X, nn, mm) == want


def test_additive_rules_plan_fp32_only(native):
    """The additive rules run the FP32 passes (same workspace as their multiplicative
    counterparts); no narrow-precision store; the object API takes FP32 only."""
    import ctypes

    for add, mult in [(capi.ADDITIVE_EUCLIDEAN, capi.EUCLIDEAN), (capi.ADDITIVE_EUCLIDEAN_WITH_H, capi.EUCLIDEAN),
                      (capi.ADDITIVE_KL, capi.KL)]:
        a = native.gpumlib_nmf_workspace_bytes(add, capi.FP32, 1000, 700, 32)
        assert a > 0 and a == native.gpumlib_nmf_workspace_bytes(mult, capi.FP32, 1000, 700, 32)
        assert native.gpumlib_nmf_workspace_bytes(add, capi.BF16, 1000, 700, 32) == 0
        assert native.gpumlib_nmf_workspace_bytes(add, capi.FP8, 1000, 700, 32) == 0
    assert native.gpumlib_nmf_workspace_bytes(5, capi.FP32, 100, 100, 8) == 0
    assert native.gpumlib_nmf_additive_update(None, None, None, -1, None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_additive_kl_update_h(None, None, None, None, None, -1, 4, 4, 1e-8, None) \
        == HIP_INVALID_VALUE
    h = ctypes.c_void_p()
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: the precision check comes first
    assert native.gpumlib_nmf_object_create(capi.ADDITIVE_KL, capi.BF16, 8, 8, 2, fake, fake, fake,
                                            ctypes.byref(h)) == HIP_INVALID_VALUE


def test_iterate_rejects_bad_arguments(native):
    f = native.gpumlib_nmf_iterate
    ws = native.gpumlib_nmf_workspace_bytes(capi.EUCLIDEAN, capi.FP32, 64, 64, 8)
    fake = 0x1000  # never dereferenced: validation fails first
    # null V
    assert f(0, 0, 64, 64, 8, None, fake, fake, 1e-8, 1, 1, 0, None, fake, ws, None) == HIP_INVALID_VALUE
    # workspace too small
    assert f(0, 0, 64, 64, 8, fake, fake, fake, 1e-8, 1, 1, 0, None, fake, ws - 1, None) == HIP_INVALID_VALUE
    # objective requested without an output
    assert f(0, 0, 64, 64, 8, fake, fake, fake, 1e-8, 1, 1, 1, None, fake, ws, None) == HIP_INVALID_VALUE
    # unsupported rank: above 128 with narrow storage (FP32 storage plans any rank)
    assert f(0, 1, 64, 64, 200, fake, fake, fake, 1e-8, 1, 1, 0, None, fake, 1 << 30, None) == HIP_INVALID_VALUE
    assert f(0, 0, 64, 64, 0, fake, fake, fake, 1e-8, 1, 1, 0, None, fake, 1 << 30, None) == HIP_INVALID_VALUE
    # an unfused-rank workspace one byte short
    wsu = native.gpumlib_nmf_workspace_bytes(capi.KL, capi.FP32, 64, 64, 200)
    assert f(1, 0, 64, 64, 200, fake, fake, fake, 1e-8, 1, 1, 0, None, fake, wsu - 1, None) == HIP_INVALID_VALUE
    # negative iteration count
    assert f(0, 0, 64, 64, 8, fake, fake, fake, 1e-8, -1, 1, 0, None, fake, ws, None) == HIP_INVALID_VALUE
    # zero iterations, no objective: nothing to do
    assert f(0, 0, 64, 64, 8, fake, fake, fake, 1e-8, 0, 1, 0, None, fake, ws, None) == 0


def test_other_entry_points_reject_bad_arguments(native):
    assert native.gpumlib_nmf_sgemm(b"Q", b"N", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4, None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_sgemm(b"N", b"N", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 2, None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_gemm_bf16(b"N", b"Z", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4,
                                        None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_gemm_fp8(b"N", b"N", 4, 4, 4, 1.0, None, 4, 255, None, 4, 127, 0.0, None, 4,
                                       None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_gemm_fp8(b"N", b"N", 4, 4, 4, 1.0, None, 4, 127, None, 4, -1, 0.0, None, 4,
                                       None) == HIP_INVALID_VALUE
    # empty products launch nothing
    assert native.gpumlib_nmf_gemm_bf16(b"N", b"T", 0, 4, 4, 1.0, None, 1, None, 4, 0.0, None, 1, None) == 0
    assert native.gpumlib_nmf_quality(None, None, 10, 1e-8, None, 0, None, None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_sum_h(None, -1, 10, 1e-8, None, None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_sum_h(None, 300, -1, 1e-8, None, None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_quality_f32(None, None, 10, 1e-8, None, 0, None, None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_mu_update(None, None, None, -1, 1e-8, None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_objective(0, 0, 8, 8, 2, 0x10, 0x10, 0x10, 1e-8, 0, 0x10, 0x10, 1 << 20,
                                        None) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_pass(0, 0, 2, 8, 8, 2, 0x10, 0x10, 0x10, 1e-8, 0x10, 1 << 20, None) == HIP_INVALID_VALUE
    # the single-pass entry runs the fused passes only (r <= 128)
    assert native.gpumlib_nmf_pass(0, 0, 0, 8, 8, 130, 0x10, 0x10, 0x10, 1e-8, 0x10, 1 << 20, None) == HIP_INVALID_VALUE
    import ctypes

    h = ctypes.c_void_p()
    assert native.gpumlib_nmf_object_create(0, 0, 0, 8, 2, None, None, None, ctypes.byref(h)) == HIP_INVALID_VALUE
    assert native.gpumlib_nmf_object_do_iteration(None, 1) == HIP_INVALID_VALUE
    native.gpumlib_nmf_object_destroy(None)  # no-op


def test_reduction_rejects_bad_arguments(native):
    ws = native.gpumlib_reduce_workspace_bytes(1 << 20)
    assert ws > 0 and ws % 256 == 0 and ws <= 1024 * 8
    assert native.gpumlib_reduce_workspace_bytes(0) > 0
    fake = 0x1000  # never dereferenced: validation fails first
    assert native.gpumlib_reduce_sum(None, 10, 1.0, fake, fake, ws, None) == HIP_INVALID_VALUE
    assert native.gpumlib_reduce_sum(fake, -1, 1.0, fake, fake, ws, None) == HIP_INVALID_VALUE
    assert native.gpumlib_reduce_sum(fake, 10, 1.0, None, fake, ws, None) == HIP_INVALID_VALUE
    assert native.gpumlib_reduce_sum(fake, 1 << 30, 1.0, fake, fake, 8, None) == HIP_INVALID_VALUE
    assert native.gpumlib_reduce_min(fake, 10, None, None, fake, ws, None) == HIP_INVALID_VALUE
    assert native.gpumlib_reduce_max(fake, 1 << 32, fake, fake, fake, 1 << 20, None) == HIP_INVALID_VALUE


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(capi, "_lib", None)
    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(capi.NativeLibraryMissing):
        capi.load()
