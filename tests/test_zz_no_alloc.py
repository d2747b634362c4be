"""gpumlib_nmf_iterate_noalloc: the fused iteration with its workspace in the
library's static device array.  Its results must be bitwise those of
gpumlib_nmf_iterate with a caller workspace (the same kernels and plan), on two
calls in a row, the second with a smaller problem (the array keeps the first
call's contents), for every rule and storage precision the fused iteration runs.
"""
import ctypes

import numpy as np
import pytest

from gpumlib_amd import capi
from gpumlib_amd import nmf as G

SCRATCH_BYTES = 32 << 20
HIP_ERROR_OUT_OF_MEMORY = 2


def test_noalloc_rejects_oversize_and_bad_problems(native):
    """Refused on the host, before any device work: a problem whose workspace exceeds the
    32 MiB array (FP32 16384 x 8192 at r = 128 needs 32.6 MiB), and problems the fused
    iteration does not plan."""
    assert native.gpumlib_nmf_workspace_bytes(G.EUCLIDEAN, G.FP32, 16384, 8192, 128) > SCRATCH_BYTES
    args = (G.EUCLIDEAN, G.FP32, 16384, 8192, 128, None, None, None, 1e-8, 1, 1, G.OBJ_NONE, None, None)
    assert native.gpumlib_nmf_iterate_noalloc(*args) == HIP_ERROR_OUT_OF_MEMORY
    assert native.gpumlib_nmf_workspace_bytes(G.EUCLIDEAN, G.FP32, 16384, 8192, 64) <= SCRATCH_BYTES
    for bad in [(G.EUCLIDEAN, G.FP32, 10, 10, 0), (7, G.FP32, 10, 10, 4), (G.KL, G.FP8, 10, 10, 4),
                (G.KL, G.BF16, 10, 10, 129)]:
        a = (*bad, None, None, None, 1e-8, 1, 1, G.OBJ_NONE, None, None)
        assert native.gpumlib_nmf_iterate_noalloc(*a) == 1  # hipErrorInvalidValue


class _Problem:
    def __init__(self, rule, prec, n, m, r, seed):
        from gpumlib_amd import hip

        self.rule, self.prec, self.n, self.m, self.r = rule, prec, n, m, r
        kl = rule in (G.KL, G.ADDITIVE_KL)
        V = G.gen_poisson(n, m, min(r, 8), seed) if kl else G.gen_euclidean(n, m, min(r, 8), seed)
        self.W0, self.H0 = G.gen_init(n, m, r, seed)
        self.Vf = hip.Buffer.from_array(V)
        if prec == G.FP32:
            self.V = self.Vf
        else:
            vsb = int(capi.load().gpumlib_nmf_vstore_bytes(prec, n, m))
            self.V = hip.Buffer(vsb)
            capi.check(capi.load().gpumlib_nmf_vstore_pack(prec, n, m, self.Vf.ptr, self.V.ptr, vsb, None),
                       "gpumlib_nmf_vstore_pack")
        self.eps = 1e-9 if kl else G.REFERENCE_EPS
        self.obj_kind = G.OBJ_KL if kl else G.OBJ_FROBENIUS

    def run(self, lib, stream, noalloc, iters=7, update_w=1):
        from gpumlib_amd import hip

        W, H, obj = hip.Buffer.from_array(self.W0), hip.Buffer.from_array(self.H0), hip.Buffer(8)
        common = (self.rule, self.prec, self.n, self.m, self.r, self.V.ptr, W.ptr, H.ptr, self.eps, iters, update_w,
                  self.obj_kind, obj.ptr)
        if noalloc:
            err = lib.gpumlib_nmf_iterate_noalloc(*common, stream.handle)
        else:
            wsb = int(lib.gpumlib_nmf_workspace_bytes(self.rule, self.prec, self.n, self.m, self.r))
            ws = hip.Buffer(wsb)
            err = lib.gpumlib_nmf_iterate(*common, ws.ptr, wsb, stream.handle)
        capi.check(err, "noalloc" if noalloc else "iterate")
        stream.synchronize()
        return (W.download(np.float32, self.n * self.r), H.download(np.float32, self.r * self.m),
                obj.download(np.float64)[0])


CASES = [
    # (rule, precision, first (n, m, r), second, smaller (n, m, r))
    (G.EUCLIDEAN, G.FP32, (1500, 900, 48), (300, 200, 20)),
    (G.KL, G.FP32, (1200, 1000, 40), (257, 130, 17)),
    (G.EUCLIDEAN, G.BF16, (2048, 1536, 64), (600, 500, 32)),
    (G.EUCLIDEAN, G.FP8, (2048, 1024, 64), (700, 300, 16)),
    (G.KL, G.BF16, (1800, 1300, 64), (333, 250, 20)),
    (G.ADDITIVE_KL, G.FP32, (800, 600, 24), (100, 90, 8)),
    (G.ADDITIVE_EUCLIDEAN_WITH_H, G.FP32, (640, 480, 32), (64, 64, 4)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"rule{c[0]}-prec{c[1]}")
def test_noalloc_matches_iterate(native, cuda, case):
    rule, prec, big, small = case
    stream = cuda.Stream()
    for k, (n, m, r) in enumerate((big, small)):
        print(f"case noalloc rule={rule} prec={prec} call={k} n={n} m={m} r={r}", flush=True)
        p = _Problem(rule, prec, n, m, r, seed=900 + 17 * k + rule)
        w_ref, h_ref, o_ref = p.run(native, stream, noalloc=False)
        w, h, o = p.run(native, stream, noalloc=True)
        assert np.array_equal(w, w_ref) and np.array_equal(h, h_ref)
        assert o == o_ref
        assert np.all(np.isfinite(w)) and np.all(np.isfinite(h)) and o > 0


@pytest.mark.gpu
def test_noalloc_update_h_only_then_objective_only(native, cuda):
    """DoIteration(false) through the static workspace (W untouched), then zero
    iterations with an objective (the objective pass alone)."""
    stream = cuda.Stream()
    p = _Problem(G.EUCLIDEAN, G.FP32, 500, 400, 24, seed=4)
    w_ref, h_ref, _ = p.run(native, stream, noalloc=False, iters=3, update_w=0)
    w, h, _ = p.run(native, stream, noalloc=True, iters=3, update_w=0)
    assert np.array_equal(w, p.W0) and np.array_equal(w, w_ref) and np.array_equal(h, h_ref)
    _, _, o_ref = p.run(native, stream, noalloc=False, iters=0)
    _, _, o = p.run(native, stream, noalloc=True, iters=0)
    assert o == o_ref and o > 0
