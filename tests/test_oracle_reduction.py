"""Pins the oracle's reductions (Reduction::Sum / Average / Min / MinIndex / Max /
MaxIndex, reference src/cuda/reduction/reduction.h:41-157) without the reference:
sums against math.fsum (exact rounding of the exact sum); extremes and their indices
against numpy's argmin / argmax (first occurrence) on NaN-free data; the start
values FLT_MAX / -FLT_MAX, strict comparisons (first of equal extremes, NaN never
taken) and the empty array on hand-made cases.
"""
import math

import numpy as np

FLT_MAX = float(np.finfo(np.float32).max)


def test_sum_matches_fsum(orc):
    rng = np.random.default_rng(50)
    for n in [0, 1, 7, 1000, 100_003]:
        x = (rng.random(n, dtype=np.float32) * 2 - 1).astype(np.float32)
        exact = math.fsum(float(v) for v in x)
        assert abs(orc.reduce_sum(x) - exact) <= 1e-12 * max(1.0, float(np.abs(x).sum()))
        assert abs(orc.reduce_sum(x, 1.0 / max(n, 1)) - exact / max(n, 1)) <= 1e-12 * max(1.0, float(np.abs(x).sum()))


def test_extremes_match_numpy(orc):
    rng = np.random.default_rng(51)
    for n in [1, 2, 999, 65537]:
        x = rng.integers(-50, 50, size=n).astype(np.float32)  # many ties
        v, i = orc.reduce_extreme(x, False)
        assert i == int(np.argmin(x)) and v == x.min()
        v, i = orc.reduce_extreme(x, True)
        assert i == int(np.argmax(x)) and v == x.max()


def test_extreme_special_cases(orc):
    x = np.array([np.nan, 3.0, np.nan, -2.0, -2.0, np.inf, -np.inf], np.float32)
    assert orc.reduce_extreme(x, False) == (-math.inf, 6)
    assert orc.reduce_extreme(x, True) == (math.inf, 5)
    y = np.array([np.nan, 1.0, 1.0, np.nan], np.float32)
    assert orc.reduce_extreme(y, False) == (1.0, 1) and orc.reduce_extreme(y, True) == (1.0, 1)
    e = np.zeros(0, np.float32)
    assert orc.reduce_extreme(e, False) == (FLT_MAX, -1) and orc.reduce_extreme(e, True) == (-FLT_MAX, -1)
    z = np.array([FLT_MAX, np.inf], np.float32)  # neither is strictly below the start value
    assert orc.reduce_extreme(z, False) == (FLT_MAX, -1)
    assert orc.reduce_extreme(np.full(3, np.nan, np.float32), True) == (-FLT_MAX, -1)
