// OCP FP8 E4M3 ("e4m3fn": bias 7, no infinities, max 448, S.1111.111 = NaN) and
// E8M0 block scales (value 2^(s-127)), as consumed by the gfx950 scaled MFMA.
// Host and device share these routines so that packing can be checked bit for
// bit on the CPU.  Inputs are finite and non-negative (NMF data and factors).
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define NMF_HD __host__ __device__
#else
#define NMF_HD
#endif

namespace gpumlib_nmf {

NMF_HD inline uint32_t f32_bits(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    return b;
}
NMF_HD inline float bits_f32(uint32_t b) {
    float x;
    memcpy(&x, &b, 4);
    return x;
}

// Round-to-nearest-even, saturating at 448.  x >= 0, finite.
NMF_HD inline uint8_t f32_to_e4m3(float x) {
    const uint32_t b = f32_bits(x) & 0x7fffffffu;
    if (b == 0) return 0;
    const int e = (int)(b >> 23) - 127;
    const uint32_t full = (b & 0x7fffffu) | ((b >> 23) ? 0x800000u : 0u);  // 24-bit significand
    if (e > 8) return 0x7e;
    if (e >= -6) {  // normal range of e4m3
        uint32_t q = (b & 0x7fffffu) >> 20;
        const uint32_t rem = b & 0xfffffu, half = 0x80000u;
        if (rem > half || (rem == half && (q & 1u))) q++;
        int E = e + 7;
        if (q == 8) {
            q = 0;
            E++;
        }
        if (E > 15 || (E == 15 && q == 7)) return 0x7e;
        return (uint8_t)((E << 3) | q);
    }
    // subnormal: value = q * 2^-9 with q in [0, 8]; x = full * 2^(e-23)
    const int shift = 14 - e;  // full * 2^(e-23+9) = full >> (14 - e)
    if (shift >= 32) return 0;
    uint32_t q = full >> shift;
    const uint32_t rem = full & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint8_t)q;  // q == 8 encodes 2^-6, the smallest normal (E=1, m=0)
}

// Value of an OCP E4M3 code (the NaN codes 0x7f / 0xff are not decoded; the NMF stores hold
// non-negative codes only, the GEMM entry point takes either sign).
NMF_HD inline float e4m3_to_f32(uint8_t v) {
    const int E = (v >> 3) & 0xf, q = v & 7;
    const float mag =
        (E == 0) ? (float)q * (1.0f / 512.0f) : bits_f32((uint32_t)(E - 7 + 127) << 23 | ((uint32_t)q << 20));
    return (v & 0x80) ? -mag : mag;
}

// Smallest power-of-two scale 2^(s-127) with amax / scale <= 448 (s clamped to [0, 254]).
NMF_HD inline int e8m0_for_amax(float amax) {
    const uint32_t b = f32_bits(amax) & 0x7fffffffu;
    if (b == 0) return 127;
    const int E = (int)(b >> 23) - 127;
    int s = E - 8 + (((b & 0x7fffffu) > 0x600000u) ? 1 : 0);  // 448 = 1.75 * 2^8
    s += 127;
    if (s < 0) s = 0;
    if (s > 254) s = 254;
    return s;
}

// x / 2^(s-127), exact for s in range (power-of-two scaling).
NMF_HD inline float e8m0_unscale(float x, int s) { return x * bits_f32((uint32_t)(254 - s) << 23); }
NMF_HD inline float e8m0_value(int s) { return bits_f32((uint32_t)s << 23); }

}  // namespace gpumlib_nmf
