"""The reduction framework on the GPU (gpumlib_reduce_*, the drop-in for GPUMLib's
Reduction class, reference src/cuda/reduction/reduction.h:41-157) against the oracle,
the reference's self-check harness semantics (reductionapp.cpp:30-109, through
gpumlib_amd/bin/reduction_app), and the host + device containers
(tests/native/test_containers.cpp).

Bar: min / max values and indices bit-exact (lowest index among equal extremes);
sums within one float rounding of the FP64 sum (the library accumulates in FP64);
results reproducible bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from gpumlib_amd import capi

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gpumlib_amd", "bin")
FLT_MAX = float(np.finfo(np.float32).max)


class Reducer:
    def __init__(self, native, hip, x, offset=0):
        """x copied to the device at element offset `offset` (offset 1 misaligns it)."""
        self.lib, self.hip = native, hip
        self.n = x.size
        buf = np.zeros(x.size + offset + 4, np.float32)
        buf[offset:offset + x.size] = x
        self.buf = hip.Buffer.from_array(buf)
        self.ptr = self.buf.ptr + 4 * offset
        self.ws_bytes = int(native.gpumlib_reduce_workspace_bytes(self.n))
        self.ws = hip.Buffer(self.ws_bytes)
        self.out = hip.Buffer(8)
        self.stream = hip.Stream()

    def sum(self, factor=1.0):
        capi.check(self.lib.gpumlib_reduce_sum(self.ptr, self.n, factor, self.out.ptr, self.ws.ptr, self.ws_bytes,
                                               self.stream.handle))
        self.stream.synchronize()
        return float(self.out.download(np.float32, 1)[0])

    def extreme(self, is_max):
        fn = self.lib.gpumlib_reduce_max if is_max else self.lib.gpumlib_reduce_min
        capi.check(fn(self.ptr, self.n, self.out.ptr, self.out.ptr + 4, self.ws.ptr, self.ws_bytes,
                      self.stream.handle))
        self.stream.synchronize()
        raw = self.out.download(np.uint8, 8)
        return float(raw[:4].view(np.float32)[0]), int(raw[4:].view(np.int32)[0])


def _check(orc, red, x):
    s = red.sum()
    ref = orc.reduce_sum(x)
    if not np.isfinite(ref):  # a NaN or an infinity in x propagates into the sum
        assert s == ref or (np.isnan(s) and np.isnan(ref))
    else:
        assert abs(s - ref) <= 2.0 ** -23 * abs(ref) + 1e-12 * float(np.abs(x.astype(np.float64)).sum()) + 1e-30
    for is_max in (False, True):
        v, i = red.extreme(is_max)
        rv, ri = orc.reduce_extreme(x, is_max)
        assert i == ri, (is_max, i, ri)
        assert v == rv or (np.isnan(v) and np.isnan(rv))


SIZES = [0, 1, 3, 4, 5, 255, 4096, 4097, 65539, 1 << 20, 3_000_017]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("offset", [0, 1])
def test_random_vs_oracle(native, orc, cuda, n, offset):
    print(f"case n={n} offset={offset}", flush=True)
    rng = np.random.default_rng(n + 7 * offset)
    x = (rng.random(n, dtype=np.float32) * 2 - 1).astype(np.float32)
    _check(orc, Reducer(native, cuda, x, offset), x)


def test_ties_pick_lowest_index(native, orc, cuda):
    rng = np.random.default_rng(60)
    x = rng.integers(-3, 4, size=2_000_003).astype(np.float32)  # each extreme occurs ~285k times
    red = Reducer(native, cuda, x)
    assert red.extreme(False) == (-3.0, int(np.argmin(x)))
    assert red.extreme(True) == (3.0, int(np.argmax(x)))
    y = np.zeros(1 << 18, np.float32)
    y[[70000, 5, 200000]] = 2.0  # the maximum at three places, the minimum (0) everywhere else
    red = Reducer(native, cuda, y)
    assert red.extreme(True) == (2.0, 5) and red.extreme(False) == (0.0, 0)


def test_special_values(native, orc, cuda):
    x = np.full(50_001, 0.5, np.float32)
    x[::7] = np.nan
    x[12345] = -np.inf
    x[40000] = np.inf
    _check(orc, Reducer(native, cuda, x), x)
    red = Reducer(native, cuda, x)
    assert red.extreme(False) == (-np.inf, 12345) and red.extreme(True) == (np.inf, 40000)
    e = Reducer(native, cuda, np.zeros(0, np.float32))
    assert e.sum() == 0.0 and e.extreme(False) == (FLT_MAX, -1) and e.extreme(True) == (-FLT_MAX, -1)
    f = Reducer(native, cuda, np.array([FLT_MAX, np.inf], np.float32))
    assert f.extreme(False) == (FLT_MAX, -1)
    g = Reducer(native, cuda, np.full(3, np.nan, np.float32))
    assert g.extreme(True) == (-FLT_MAX, -1)


def test_factor_and_reproducible(native, orc, cuda):
    rng = np.random.default_rng(61)
    x = rng.random(10_000_019, dtype=np.float32)
    red = Reducer(native, cuda, x)
    a = red.sum(1.0 / x.size)
    ref = orc.reduce_sum(x, 1.0 / x.size)
    assert abs(a - ref) <= 2.0 ** -23 * abs(ref) * 2
    assert all(red.sum(1.0 / x.size) == a for _ in range(3))
    assert red.extreme(True) == red.extreme(True)


def test_full_size_sum_of_v(native, orc, cuda):
    """A 16384 x 8192 FP32 matrix (512 MiB): the property that the sum of column sums
    equals the sum, and the FP64 oracle sum."""
    n, m = 16384, 8192
    rng = np.random.default_rng(62)
    x = rng.random(n * m, dtype=np.float32)
    red = Reducer(native, cuda, x)
    s = red.sum()
    ref = orc.reduce_sum(x)
    assert abs(s - ref) <= 2.0 ** -23 * ref
    cols = x.reshape(m, n).sum(axis=1, dtype=np.float64)
    assert abs(cols.sum() - ref) <= 1e-9 * ref
    v, i = red.extreme(True)
    assert (v, i) == orc.reduce_extreme(x, True)


def test_reduction_app_self_check(native, cuda):
    """The self-check harness: Sum, Average, MinIndex, MaxIndex against host loops."""
    p = subprocess.run([os.path.join(BIN, "reduction_app"), "1000003", "3", "12345"], capture_output=True,
                       text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = [l for l in p.stdout.splitlines() if ":" in l and not l.startswith("Random")]
    assert len(lines) == 6 and all(l.endswith("ok") for l in lines)
    usage = subprocess.run([os.path.join(BIN, "reduction_app")], capture_output=True, text=True, timeout=60)
    assert usage.returncode == 0 and "usage" in usage.stdout


def test_native_containers(native, cuda):
    """CudaArray / CudaMatrix / HostMatrix3D / DeviceMatrix3D / CudaMatrix3D checks."""
    p = subprocess.run([os.path.join(BIN, "test_containers")], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failed" in p.stdout
