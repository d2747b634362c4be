"""The seeded input generators (C, in the product library) against a numpy
restatement of the same published recipe (SplitMix64; Exp via -log(1-U);
Poisson by inversion / PTRS).  Uniforms and Euclidean V are bit-exact."""
import numpy as np
import pytest

from gpumlib_amd import nmf as G

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix(x):
    with np.errstate(over="ignore"):
        x = (x + np.uint64(0x9E3779B97F4A7C15)) & M64
        x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
        x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
        return x ^ (x >> np.uint64(31))


def key(seed, stream):
    with np.errstate(over="ignore"):
        return mix(np.uint64(seed) ^ ((np.uint64(stream) * np.uint64(0xD1B54A32D192ED03)) & M64))


def draws(seed, stream, count):
    with np.errstate(over="ignore"):
        return mix(key(seed, stream) + np.arange(count, dtype=np.uint64))


def u01f(x):
    return (x >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / 16777216.0)


def u01d(x):
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def test_rng_known_values(native):
    # SplitMix64 finaliser of 0 after one golden-ratio step (published test value)
    assert int(mix(np.uint64(0))) == 0xE220A8397B1DCDAF
    for k in (0, 1, 12345):
        assert int(native.gpumlib_nmf_rng_u64(7, 3, k)) == int(draws(7, 3, k + 1)[k])


def test_uniform_bit_exact(native):
    n, m, r = 37, 23, 5
    W, H = G.gen_init(n, m, r, seed=99)
    assert np.array_equal(W, u01f(draws(99, 4, n * r)))
    assert np.array_equal(H, u01f(draws(99, 5, r * m)))
    assert W.min() >= 0.0 and W.max() < 1.0


def test_euclidean_v_bit_exact(native):
    n, m, r0, seed = 64, 48, 7, 0x5EED
    V = G.gen_euclidean(n, m, r0, seed)
    W0 = u01f(draws(seed, 1, n * r0)).astype(np.float64).reshape(r0, n).T  # column-major n x r0
    H0 = u01f(draws(seed, 2, r0 * m)).astype(np.float64).reshape(m, r0).T  # column-major r0 x m
    Z = u01f(draws(seed, 3, n * m)).astype(np.float64).reshape(m, n).T
    acc = np.zeros((n, m))
    for a in range(r0):  # same summation order as the C generator (a ascending, no FMA)
        acc = acc + W0[:, a:a + 1] * H0[a:a + 1, :]
    acc = acc + np.float64(np.float32(0.01)) * Z
    assert np.array_equal(V, G.colmajor(acc.astype(np.float32)))


def test_poisson_statistics(native):
    n, m, r0, seed = 400, 300, 6, 0x5EED + 3
    V = G.gen_poisson(n, m, r0, seed)
    W0 = (-np.log1p(-u01d(draws(seed, 1, n * r0)))).astype(np.float32).astype(np.float64).reshape(r0, n).T
    H0 = (-np.log1p(-u01d(draws(seed, 2, r0 * m)))).astype(np.float32).astype(np.float64).reshape(m, r0).T
    lam = G.colmajor(W0 @ H0).astype(np.float64)
    assert np.all(V >= 0) and np.all(V == np.round(V))
    z = (V - lam) / np.sqrt(lam)
    assert abs(z.mean()) < 0.02          # unbiased, 120k draws
    assert abs(z.var() - 1.0) < 0.03     # Poisson variance = mean
    small, big = lam < 10, lam >= 10     # both samplers exercised
    assert small.sum() > 1000 and big.sum() > 1000
    for mask in (small, big):
        assert abs(z[mask].mean()) < 0.05 and abs(z[mask].var() - 1.0) < 0.08


def test_generators_deterministic(native):
    a = G.gen_poisson(50, 40, 3, 5)
    b = G.gen_poisson(50, 40, 3, 5)
    c = G.gen_poisson(50, 40, 3, 6)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.array_equal(G.gen_euclidean(50, 40, 3, 5), G.gen_euclidean(50, 40, 3, 5))
