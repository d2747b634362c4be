"""Python harness over the C-ABI: seeded inputs on the host and NMF problems
resident in HBM.  Device memory and streams come from the library's own HIP
runtime (gpumlib_amd.hip); all arithmetic runs in the native gfx950 library.

Matrices are column-major, stored as flat float32 arrays: an n x m matrix X is
held as X.T.ravel() (element (i, u) at index u*n + i), the layout of GPUMLib's
ColumnMajor containers.
"""
import ctypes

import numpy as np

from . import capi
from .capi import (ADDITIVE_EUCLIDEAN, ADDITIVE_EUCLIDEAN_WITH_H, ADDITIVE_KL, BF16, EUCLIDEAN, FP8, FP32, FP32X,
                   KL, OBJ_FROBENIUS, OBJ_KL, OBJ_NONE, OBJ_QUALITY, REFERENCE_EPS)

__all__ = ["colmajor", "as_matrix", "gen_euclidean", "gen_poisson", "gen_init", "DeviceNMF",
           "EUCLIDEAN", "KL", "ADDITIVE_EUCLIDEAN", "ADDITIVE_KL", "ADDITIVE_EUCLIDEAN_WITH_H",
           "FP32", "BF16", "FP8", "FP32X", "OBJ_NONE", "OBJ_FROBENIUS", "OBJ_KL", "OBJ_QUALITY", "REFERENCE_EPS"]


def colmajor(x):
    """n x m array -> flat column-major float32."""
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).T).ravel()


def as_matrix(flat, n, m):
    """flat column-major -> n x m array view."""
    return np.asarray(flat).reshape(m, n).T


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def gen_euclidean(n, m, r0, seed, noise=0.01):
    """V = W0 H0 + noise U[0,1), W0, H0 ~ U[0,1) (flat column-major float32)."""
    lib = capi.load()
    v = np.empty(n * m, dtype=np.float32)
    lib.gpumlib_nmf_gen_euclidean_v(seed, n, m, r0, noise, _ptr(v))
    return v


def gen_poisson(n, m, r0, seed):
    """V ~ Poisson(W0 H0), W0, H0 ~ Exp(1) (flat column-major float32)."""
    lib = capi.load()
    v = np.empty(n * m, dtype=np.float32)
    lib.gpumlib_nmf_gen_poisson_v(seed, n, m, r0, _ptr(v))
    return v


def gen_init(n, m, r, seed):
    """Initial W (n x r) and H (r x m) ~ U[0,1), flat column-major float32."""
    lib = capi.load()
    w = np.empty(n * r, dtype=np.float32)
    h = np.empty(r * m, dtype=np.float32)
    lib.gpumlib_nmf_gen_init(seed, n, m, r, _ptr(w), _ptr(h))
    return w, h


class DeviceNMF:
    """One NMF problem resident on the GPU, driven through gpumlib_nmf_iterate.

    precision BF16/FP8/FP32X: V is packed into the V store of that precision on the device.
    setup_s: wall-clock seconds of the set-up phases, "upload" (device buffers allocated and
    V, W, H copied from pageable host memory) and "pack" (the V store built on the device)."""

    def __init__(self, rule, V, W, H, n, m, r, eps=REFERENCE_EPS, precision=FP32, stream=None):
        import time

        from . import hip

        self.hip = hip
        self.lib = capi.load()
        self.rule, self.precision, self.n, self.m, self.r, self.eps = rule, precision, n, m, r, float(eps)
        self.stream = stream or hip.Stream()
        t0 = time.perf_counter()
        self.Vf = hip.Buffer.from_array(np.asarray(V, dtype=np.float32))
        self.W = hip.Buffer.from_array(np.asarray(W, dtype=np.float32))
        self.H = hip.Buffer.from_array(np.asarray(H, dtype=np.float32))
        t1 = time.perf_counter()
        if precision == FP32:
            self.V = self.Vf
        else:
            vsb = int(self.lib.gpumlib_nmf_vstore_bytes(precision, n, m))
            if vsb == 0:
                raise ValueError("unsupported precision")
            self.V = hip.Buffer(vsb)
            capi.check(self.lib.gpumlib_nmf_vstore_pack(precision, n, m, self.Vf.ptr, self.V.ptr, vsb,
                                                        self.stream.handle), "gpumlib_nmf_vstore_pack")
            self.stream.synchronize()
        self.setup_s = {"upload": t1 - t0, "pack": time.perf_counter() - t1}
        self.ws_bytes = int(self.lib.gpumlib_nmf_workspace_bytes(rule, precision, n, m, r))
        if self.ws_bytes == 0:
            raise ValueError("unsupported problem shape / precision")
        self.ws = hip.Buffer(self.ws_bytes)
        self.obj = hip.Buffer(8)

    def launch(self, iterations, update_w=True, objective=OBJ_NONE):
        """Enqueues iterations on the problem's stream (no synchronisation)."""
        err = self.lib.gpumlib_nmf_iterate(
            self.rule, self.precision, self.n, self.m, self.r, self.V.ptr, self.W.ptr, self.H.ptr, self.eps,
            iterations, 1 if update_w else 0, objective, self.obj.ptr if objective else None, self.ws.ptr,
            self.ws_bytes, self.stream.handle)
        capi.check(err, "gpumlib_nmf_iterate")

    def iterate(self, iterations, update_w=True, objective=OBJ_NONE):
        """Runs iterations; returns the objective of the entry state when requested."""
        self.launch(iterations, update_w, objective)
        self.stream.synchronize()
        if objective:
            return float(self.obj.download(np.float64)[0])
        return None

    def run_pass(self, which):
        capi.check(self.lib.gpumlib_nmf_pass(self.rule, self.precision, which, self.n, self.m, self.r, self.V.ptr,
                                             self.W.ptr, self.H.ptr, self.eps, self.ws.ptr, self.ws_bytes,
                                             self.stream.handle), "gpumlib_nmf_pass")

    def objective(self, kind):
        err = self.lib.gpumlib_nmf_objective(
            self.rule, self.precision, self.n, self.m, self.r, self.V.ptr, self.W.ptr, self.H.ptr, self.eps, kind,
            self.obj.ptr, self.ws.ptr, self.ws_bytes, self.stream.handle)
        capi.check(err, "gpumlib_nmf_objective")
        self.stream.synchronize()
        return float(self.obj.download(np.float64)[0])

    def objective_fp32v(self, kind):
        """Objective of the current W, H against the FP32 V (whatever the V store's precision)."""
        if self.precision == FP32:
            return self.objective(kind)
        if getattr(self, "ws32", None) is None:
            self.ws32_bytes = int(self.lib.gpumlib_nmf_workspace_bytes(self.rule, FP32, self.n, self.m, self.r))
            self.ws32 = self.hip.Buffer(self.ws32_bytes)
        err = self.lib.gpumlib_nmf_objective(
            self.rule, FP32, self.n, self.m, self.r, self.Vf.ptr, self.W.ptr, self.H.ptr, self.eps, kind,
            self.obj.ptr, self.ws32.ptr, self.ws32_bytes, self.stream.handle)
        capi.check(err, "gpumlib_nmf_objective")
        self.stream.synchronize()
        return float(self.obj.download(np.float64)[0])

    def set_wh(self, W, H):
        """Replaces W and H (the V store stays)."""
        self.stream.synchronize()
        self.W.upload(np.asarray(W, dtype=np.float32))
        self.H.upload(np.asarray(H, dtype=np.float32))

    def get_w(self):
        self.stream.synchronize()
        return self.W.download(np.float32, self.n * self.r)

    def get_h(self):
        self.stream.synchronize()
        return self.H.download(np.float32, self.r * self.m)
