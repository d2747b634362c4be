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


def fp32_objective_allowance(ref, v_sq_sum, r):
    """Absolute error the GPU's FP32 residuals may add to a Frobenius objective near ref.

    Each residual d = v - (W H)(i, u) is formed in FP32 from r products, so it carries an
    error |delta| <= 4 r 2^-24 |v| (a loose bound on r-term FP32 dot-product rounding); the
    objective sum d^2 then moves by at most 2 |d| |delta| <= 2 sqrt(ref) 4 r 2^-24 |V|_F
    (Cauchy-Schwarz).  It matters only when the fit is nearly exact (ref ~ 0)."""
    return 2.0 * np.sqrt(max(ref, 0.0)) * 4.0 * r * 2.0 ** -24 * np.sqrt(v_sq_sum)


def rel_fro(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
