import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the native HIP library on cuda:0)")


@pytest.fixture(scope="session")
def native():
    """The native library; built in-tree first if it is missing (CPU side only)."""
    from gpumlib_amd import capi

    if not os.path.exists(capi.LIB_PATH):
        from gpumlib_amd import build as b

        b.build()
    return capi.load()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def cuda(native):
    """The library's HIP runtime (gpumlib_amd.hip); fails loudly without a GPU."""
    from gpumlib_amd import hip

    assert hip.device_count() > 0, "GPU tests need a visible MI355X"
    return hip


def objective_allowance(ref, v_sq_sum, r, d_wh):
    """Absolute difference allowed between a GPU Frobenius objective and a reference one, ref.

    The two are |V - W H|^2 for factors whose products differ by d_wh = |W H - W' H'|_F
    (measured), with the GPU's residuals formed in FP32 from r products each, an error of at
    most 4 r 2^-24 |v| per entry (a loose bound on r-term FP32 dot-product rounding).  With
    D = d_wh + 4 r 2^-24 |V|_F the objectives differ by at most 2 sqrt(ref) D + D^2
    (Cauchy-Schwarz).  This matters only near an exact fit (ref ~ 0), where the relative
    difference of the objectives says nothing about the factors."""
    D = d_wh + 4.0 * r * 2.0 ** -24 * np.sqrt(v_sq_sum)
    return 2.0 * np.sqrt(max(ref, 0.0)) * D + D * D


def rel_fro(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
