// Seeded host-side generators for NMF inputs.  Compiled without floating-point
// contraction so that every value is a fixed function of (seed, index): the
// same bits come out on any host and for any thread count.
//
// Counter-based SplitMix64: value k of stream s under seed is
//   mix(mix(seed ^ (s * 0xD1B54A32D192ED03)) + k)
// uniform float  = (x >> 40) * 2^-24  in [0, 1)
// uniform double = (x >> 11) * 2^-53  in [0, 1)
// Streams: 1 W0, 2 H0, 3 noise, 4 initial W, 5 initial H, 6 Poisson draws.
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/gpumlib_nmf.h"
#include "nmf_fp8.hpp"

// Runs body(lo, hi) over [0, count) split into contiguous blocks on up to 16
// host threads.  Every element is computed independently, so the split does
// not change any value.
template <class F>
static void parallel_for(long count, F body) {
    long nt = (long)std::thread::hardware_concurrency();
    nt = std::max(1L, std::min(nt, 16L));
    if (count < 4096 || nt == 1) {
        body(0L, count);
        return;
    }
    std::vector<std::thread> th;
    const long per = (count + nt - 1) / nt;
    for (long t = 0; t < nt; t++) {
        const long lo = t * per, hi = std::min(count, lo + per);
        if (lo >= hi) break;
        th.emplace_back([=] { body(lo, hi); });
    }
    for (auto& x : th) x.join();
}

static inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
}

static inline float u01f(uint64_t x) { return (float)(x >> 40) * (1.0f / 16777216.0f); }
static inline double u01d(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

// Poisson(lam) for one element from its own draw sequence d(j) = u01d(mix(key + j)).
// lam < 10: inversion by sequential search; otherwise the PTRS transformed
// rejection method of Hoermann (1993), "The transformed rejection method for
// generating Poisson random variables", Insurance: Mathematics and Economics 12.
static double poisson_draw(uint64_t key, double lam) {
    uint64_t j = 0;
    if (lam <= 0.0) return 0.0;
    if (lam < 10.0) {
        const double u = u01d(mix64(key + j++));
        double p = exp(-lam), F = p, x = 0.0;
        while (u > F && x < 1000.0) {
            x += 1.0;
            p *= lam / x;
            F += p;
        }
        return x;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double U = u01d(mix64(key + j++)) - 0.5;
        const double V = u01d(mix64(key + j++));
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
    }
}

extern "C" {

uint64_t gpumlib_nmf_rng_u64(uint64_t seed, uint64_t stream, uint64_t k) { return mix64(stream_key(seed, stream) + k); }

void gpumlib_nmf_gen_uniform_f32(uint64_t seed, uint64_t stream, size_t count, float* out) {
    const uint64_t key = stream_key(seed, stream);
    parallel_for((long)count, [&](long lo, long hi) {
        for (long k = lo; k < hi; k++) out[k] = u01f(mix64(key + (uint64_t)k));
    });
}

// V(i,u) = (float)( sum_a W0(i,a) H0(a,u)  [double, a ascending]  + noise * U(i,u) )
void gpumlib_nmf_gen_euclidean_v(uint64_t seed, int n, int m, int r0, float noise, float* V) {
    const uint64_t k1 = stream_key(seed, 1), k2 = stream_key(seed, 2), k3 = stream_key(seed, 3);
    double* w0 = (double*)malloc(sizeof(double) * (size_t)n * r0);  // row-major copy: w0[i*r0 + a]
    double* h0 = (double*)malloc(sizeof(double) * (size_t)m * r0);  // h0[u*r0 + a]
    parallel_for((long)n * r0, [&](long lo, long hi) {
        for (long e = lo; e < hi; e++) {
            const long a = e / n, i = e - a * n;
            w0[i * r0 + a] = (double)u01f(mix64(k1 + (uint64_t)e));
        }
    });
    parallel_for((long)m * r0, [&](long lo, long hi) {
        for (long e = lo; e < hi; e++) h0[e] = (double)u01f(mix64(k2 + (uint64_t)e));
    });
    parallel_for((long)m, [&](long lo, long hi) {
        for (long u = lo; u < hi; u++) {
            const double* h = h0 + (size_t)u * r0;
            for (int i = 0; i < n; i++) {
                const double* w = w0 + (size_t)i * r0;
                double acc = 0.0;
                for (int a = 0; a < r0; a++) acc += w[a] * h[a];
                const double z = (double)u01f(mix64(k3 + (uint64_t)u * n + i));
                acc += (double)noise * z;
                V[(size_t)u * n + i] = (float)acc;
            }
        }
    });
    free(w0);
    free(h0);
}

// W0(i,a), H0(a,u) = (float)(-log(1 - U)) with U a 53-bit uniform (Exp(1) = Gamma(1,1));
// lam(i,u) = sum_a W0 H0 in double; V(i,u) = Poisson(lam) from the element's own draws.
void gpumlib_nmf_gen_poisson_v(uint64_t seed, int n, int m, int r0, float* V) {
    const uint64_t k1 = stream_key(seed, 1), k2 = stream_key(seed, 2), k6 = stream_key(seed, 6);
    double* w0 = (double*)malloc(sizeof(double) * (size_t)n * r0);  // w0[i*r0 + a]
    double* h0 = (double*)malloc(sizeof(double) * (size_t)m * r0);  // h0[u*r0 + a]
    parallel_for((long)n * r0, [&](long lo, long hi) {
        for (long e = lo; e < hi; e++) {
            const long a = e / n, i = e - a * n;
            w0[i * r0 + a] = (double)(float)(-log1p(-u01d(mix64(k1 + (uint64_t)e))));
        }
    });
    parallel_for((long)m * r0, [&](long lo, long hi) {
        for (long e = lo; e < hi; e++) h0[e] = (double)(float)(-log1p(-u01d(mix64(k2 + (uint64_t)e))));
    });
    parallel_for((long)m, [&](long lo, long hi) {
        for (long u = lo; u < hi; u++) {
            const double* h = h0 + (size_t)u * r0;
            for (int i = 0; i < n; i++) {
                const double* w = w0 + (size_t)i * r0;
                double lam = 0.0;
                for (int a = 0; a < r0; a++) lam += w[a] * h[a];
                const uint64_t e = (uint64_t)u * n + i;
                V[e] = (float)poisson_draw(mix64(k6 + e), lam);
            }
        }
    });
    free(w0);
    free(h0);
}

void gpumlib_nmf_fp8_quantize(const float* x, size_t count, int scale, uint8_t* out) {
    for (size_t k = 0; k < count; k++) out[k] = gpumlib_nmf::f32_to_e4m3(gpumlib_nmf::e8m0_unscale(x[k], scale));
}

int gpumlib_nmf_fp8_scale_for(float amax) { return gpumlib_nmf::e8m0_for_amax(amax); }

float gpumlib_nmf_fp8_decode(uint8_t q, int scale) {
    return gpumlib_nmf::e4m3_to_f32(q) * gpumlib_nmf::e8m0_value(scale);
}

void gpumlib_nmf_gen_init(uint64_t seed, int n, int m, int r, float* W, float* H) {
    gpumlib_nmf_gen_uniform_f32(seed, 4, (size_t)n * r, W);
    gpumlib_nmf_gen_uniform_f32(seed, 5, (size_t)r * m, H);
}

}  // extern "C"
