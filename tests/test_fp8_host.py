"""FP8 storage conversion (OCP E4M3 + E8M0 block scale) against a brute-force
nearest-value search over all 256 codes (ties to the even code)."""
import ctypes

import numpy as np
import pytest


def code_values():
    vals = []
    for q in range(256):
        s, E, m = q >> 7, (q >> 3) & 0xF, q & 7
        if E == 15 and m == 7:
            vals.append(np.nan)
            continue
        v = m / 8 * 2.0 ** -6 if E == 0 else (1 + m / 8) * 2.0 ** (E - 7)
        vals.append(-v if s else v)
    return np.array(vals)


VALS = code_values()


def brute(x):
    """nearest non-negative E4M3 code to x >= 0 (saturating), ties to even code."""
    cand = np.arange(0x7F)  # 0x00 .. 0x7e (0x7f is NaN)
    d = np.abs(VALS[cand] - x)
    best = d.min()
    ties = cand[d == best]
    even = ties[ties % 2 == 0]
    return int(even[0] if len(even) else ties[0])


def quant(native, x, scale):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.size, np.uint8)
    native.gpumlib_nmf_fp8_quantize(x.ctypes.data_as(ctypes.c_void_p), x.size, scale,
                                    out.ctypes.data_as(ctypes.c_void_p))
    return out


def test_decode_table(native):
    for q in range(256):
        if q & 0x7F == 0x7F:
            continue  # NaN codes
        got = native.gpumlib_nmf_fp8_decode(q, 127)
        assert got == np.float32(VALS[q]) and np.signbit(got) == bool(q & 0x80)
    assert native.gpumlib_nmf_fp8_decode(0x7E, 127) == 448.0
    assert native.gpumlib_nmf_fp8_decode(0x38, 130) == 8.0  # 1.0 * 2^3


def test_encode_matches_brute_force(native):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.random(3000) * 500, rng.random(1000) * 0.02, rng.random(500) * 2e-3,
                         VALS[:0x7F],                                    # exact codes
                         (VALS[1:0x7F] + VALS[:0x7E]) / 2,               # exact midpoints: ties
                         [0.0, 448.0, 449.0, 464.0, 1e6, 2.0 ** -10, 2.0 ** -9, 3 * 2.0 ** -10]])
    xs = xs.astype(np.float32)
    got = quant(native, xs, 127)
    expect = np.array([brute(float(x)) for x in xs])
    bad = np.nonzero(got != expect)[0]
    assert bad.size == 0, [(float(xs[i]), int(got[i]), int(expect[i])) for i in bad[:5]]


@pytest.mark.parametrize("amax", [1e-3, 0.7, 1.0, 447.0, 448.0, 449.0, 3e4, 2.5e7])
def test_scale_choice(native, amax):
    s = native.gpumlib_nmf_fp8_scale_for(amax)
    unit = 2.0 ** (s - 127)
    assert amax / unit <= 448.0          # fits
    assert amax / (unit / 2) > 448.0     # the smallest power of two that fits
    q = quant(native, np.array([amax]), s)[0]
    assert abs(native.gpumlib_nmf_fp8_decode(int(q), s) - amax) <= amax * 2 ** -4
