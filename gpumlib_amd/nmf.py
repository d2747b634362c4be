"""Python harness over the C-ABI: seeded inputs on the host and NMF problems
resident in HBM.  Device memory and streams come from PyTorch (plumbing only);
all arithmetic runs in the native gfx950 library.

Matrices are column-major, stored as flat float32 arrays: an n x m matrix X is
held as X.T.ravel() (element (i, u) at index u*n + i), the layout of GPUMLib's
ColumnMajor containers.
"""
import ctypes

import numpy as np

from . import capi
from .capi import BF16, EUCLIDEAN, FP8, FP32, KL, OBJ_FROBENIUS, OBJ_KL, OBJ_NONE, OBJ_QUALITY, REFERENCE_EPS

__all__ = ["colmajor", "as_matrix", "gen_euclidean", "gen_poisson", "gen_init", "DeviceNMF",
           "EUCLIDEAN", "KL", "FP32", "BF16", "FP8", "OBJ_NONE", "OBJ_FROBENIUS", "OBJ_KL", "OBJ_QUALITY",
           "REFERENCE_EPS"]


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
    """One NMF problem resident on the GPU, driven through gpumlib_nmf_iterate."""

    def __init__(self, rule, V, W, H, n, m, r, eps=REFERENCE_EPS, precision=FP32, device="cuda"):
        import torch

        self.torch = torch
        self.lib = capi.load()
        self.rule, self.precision, self.n, self.m, self.r, self.eps = rule, precision, n, m, r, float(eps)
        self.V = torch.from_numpy(np.ascontiguousarray(V, dtype=np.float32)).to(device)
        self.W = torch.from_numpy(np.ascontiguousarray(W, dtype=np.float32)).to(device)
        self.H = torch.from_numpy(np.ascontiguousarray(H, dtype=np.float32)).to(device)
        self.ws_bytes = int(self.lib.gpumlib_nmf_workspace_bytes(rule, precision, n, m, r))
        if self.ws_bytes == 0:
            raise ValueError("unsupported problem shape / precision")
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=device)
        self.obj = torch.zeros(1, dtype=torch.float64, device=device)

    def stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def iterate(self, iterations, update_w=True, objective=OBJ_NONE):
        """Runs iterations; returns the objective of the entry state when requested (syncs)."""
        err = self.lib.gpumlib_nmf_iterate(
            self.rule, self.precision, self.n, self.m, self.r, self.V.data_ptr(), self.W.data_ptr(),
            self.H.data_ptr(), self.eps, iterations, 1 if update_w else 0, objective,
            self.obj.data_ptr() if objective else None, self.ws.data_ptr(), self.ws_bytes, self.stream())
        capi.check(err, "gpumlib_nmf_iterate")
        if objective:
            return float(self.obj.cpu()[0])
        return None

    def objective(self, kind):
        err = self.lib.gpumlib_nmf_objective(
            self.rule, self.precision, self.n, self.m, self.r, self.V.data_ptr(), self.W.data_ptr(),
            self.H.data_ptr(), self.eps, kind, self.obj.data_ptr(), self.ws.data_ptr(), self.ws_bytes,
            self.stream())
        capi.check(err, "gpumlib_nmf_objective")
        return float(self.obj.cpu()[0])

    def get_w(self):
        return self.W.cpu().numpy().copy()

    def get_h(self):
        return self.H.cpu().numpy().copy()
