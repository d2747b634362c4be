"""Minimal ctypes binding of the HIP runtime that libgpumlib_nmf.so links
(/opt/rocm/lib/libamdhip64.so.7): device buffers, streams, events.

The harness must use the same runtime instance as the library: device
allocations, streams and events are runtime objects, and a second HIP runtime
in the process (PyTorch wheels bundle their own) would hand the library handles
it does not know.  Loading by the library's own soname returns the instance
the dynamic linker already mapped for libgpumlib_nmf.so.
"""
import ctypes
import os

import numpy as np

_c = ctypes
_lib = None

H2D, D2H, D2D = 1, 2, 3


def _load():
    global _lib
    if _lib is not None:
        return _lib
    from . import capi

    capi.load()  # maps libgpumlib_nmf.so and, through it, its HIP runtime
    cands = ["libamdhip64.so.7", os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so.7")]
    err = None
    for c in cands:
        try:
            L = ctypes.CDLL(c)
            break
        except OSError as e:  # pragma: no cover - environment dependent
            err = e
    else:
        raise OSError(f"HIP runtime not found: {err}")
    vp, sz, i = _c.c_void_p, _c.c_size_t, _c.c_int
    sigs = {
        "hipMalloc": [_c.POINTER(vp), sz],
        "hipFree": [vp],
        "hipMemcpy": [vp, vp, sz, i],
        "hipMemcpyAsync": [vp, vp, sz, i, vp],
        "hipMemset": [vp, i, sz],
        "hipStreamCreate": [_c.POINTER(vp)],
        "hipStreamDestroy": [vp],
        "hipStreamSynchronize": [vp],
        "hipDeviceSynchronize": [],
        "hipEventCreate": [_c.POINTER(vp)],
        "hipEventDestroy": [vp],
        "hipEventRecord": [vp, vp],
        "hipEventSynchronize": [vp],
        "hipEventElapsedTime": [_c.POINTER(_c.c_float), vp, vp],
        "hipGetDeviceCount": [_c.POINTER(i)],
        "hipSetDevice": [i],
        "hipGetLastError": [],
        "hipMemGetInfo": [_c.POINTER(sz), _c.POINTER(sz)],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = i
    L.hipGetErrorString.argtypes = [i]
    L.hipGetErrorString.restype = _c.c_char_p
    _lib = L
    return L


def check(err, what="HIP call"):
    if err != 0:
        msg = _load().hipGetErrorString(err).decode()
        raise RuntimeError(f"{what} failed: {msg} ({err})")


def device_count():
    n = _c.c_int(0)
    err = _load().hipGetDeviceCount(_c.byref(n))
    return n.value if err == 0 else 0


def synchronize():
    check(_load().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = _c.c_size_t(), _c.c_size_t()
    check(_load().hipMemGetInfo(_c.byref(free), _c.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class Buffer:
    """A device allocation of nbytes (freed on close/garbage collection)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = _c.c_void_p()
        check(_load().hipMalloc(_c.byref(p), max(self.nbytes, 1)), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if stream is None:
            check(_load().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        else:
            check(_load().hipMemcpyAsync(self.ptr, a.ctypes.data, a.nbytes, H2D, stream.handle), "hipMemcpyAsync")
            stream.synchronize()

    def download(self, dtype, count=None):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype=dtype)
        check(_load().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, other, nbytes=None):
        nbytes = other.nbytes if nbytes is None else nbytes
        check(_load().hipMemcpy(self.ptr, other.ptr, nbytes, D2D), "hipMemcpy D2D")

    def fill_zero(self):
        check(_load().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def close(self):
        if getattr(self, "ptr", None):
            _load().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = _c.c_void_p()
        check(_load().hipStreamCreate(_c.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        check(_load().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if getattr(self, "handle", None):
            _load().hipStreamDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = _c.c_void_p()
        check(_load().hipEventCreate(_c.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        check(_load().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        check(_load().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later):
        ms = _c.c_float()
        check(_load().hipEventElapsedTime(_c.byref(ms), self.handle, later.handle), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _load().hipEventDestroy(self.handle)
        except Exception:
            pass
