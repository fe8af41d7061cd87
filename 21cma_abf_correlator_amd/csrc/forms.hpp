// forms.hpp -- spectrum representations between the F- and the X-engine.
//
//   FP32: complex64, the drop-in numerics.
//   FP16: y*g as two IEEE half values (round to nearest even) packed in 32 bits,
//         real part in the low half; g = gain of the (input, channel), a power of
//         two chosen so that the spectrum's RMS maps to about 2^8, far from the
//         half range limits (6e-5 .. 65504).
//   INT8: clamp(rint(y*g), -7, 7) for real and imaginary part, two int8 packed in
//         16 bits (real in the low byte); the symmetric range keeps negation exact.
//         g = 1/(0.35 sigma) of each component, the step of a near-optimal
//         uniform quantizer for a Gaussian (Max 1960).
// The X-engine multiplies these (exactly: half*half and int8*int8 products are
// exact in FP32 / int32) and the dump divides by g_j g_k in FP64.
//
// Written for both the device (hipcc) and the CPU build of the same code; the
// includer defines CORR_HD.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace corr {
namespace forms {

constexpr int kInt8Max = 7;
constexpr float kFp16TargetRms = 256.0f;
constexpr float kInt8StepSigma = 0.35f;

// float -> IEEE half bits, round to nearest even, subnormals kept, overflow to inf.
CORR_HD inline uint16_t f32_to_f16_bits(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bit_cast(uint16_t, (_Float16)f);
#else
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u)  // inf or nan
        return (uint16_t)(sign | 0x7C00u | (ax > 0x7F800000u ? 0x200u : 0u));
    if (ax >= 0x477FF000u)  // rounds to >= 65520: overflow to inf
        return (uint16_t)(sign | 0x7C00u);
    if (ax >= 0x38800000u) {  // normal half
        const uint32_t mant = ax & 0x7FFFFFu;
        const uint32_t exp = (ax >> 23) - 127 + 15;
        uint32_t h = (exp << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1FFFu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // carries into the exponent
        return (uint16_t)(sign | h);
    }
    // subnormal half (or zero): value = m * 2^-24
    const int e = (int)(ax >> 23);           // biased float exponent
    if (e < 102) return (uint16_t)sign;      // below half the smallest subnormal
    const uint32_t mant = (ax & 0x7FFFFFu) | 0x800000u;
    const int shift = 126 - e;               // 14 .. 24 for e in [102, 112]
    uint32_t h = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
#endif
}

// IEEE half bits -> float (exact).
CORR_HD inline float f16_bits_to_f32(uint16_t h) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (float)__builtin_bit_cast(_Float16, h);
#else
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1Fu, mant = h & 0x3FFu;
    uint32_t x;
    if (exp == 0) {
        if (mant == 0) {
            x = sign;
        } else {
            float v = (float)mant * (1.0f / 16777216.0f);
            std::memcpy(&x, &v, 4);
            x |= sign;
        }
    } else if (exp == 31) {
        x = sign | 0x7F800000u | (mant << 13);
    } else {
        x = sign | ((exp - 15 + 127) << 23) | (mant << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
#endif
}

CORR_HD inline uint32_t pack_fp16(float2 y, float g) {
#if defined(__HIP_DEVICE_COMPILE__)
    // as a two-half vector the compiler writes the high half in place
    // (v_fma_mixhi_f16) instead of shifting and or-ing it in
    typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
    const half2_t v = {(_Float16)(y.x * g), (_Float16)(y.y * g)};
    return __builtin_bit_cast(uint32_t, v);
#else
    return (uint32_t)f32_to_f16_bits(y.x * g) | ((uint32_t)f32_to_f16_bits(y.y * g) << 16);
#endif
}

// Two IEEE halves (one 32-bit word) with each infinity replaced by the largest
// finite half of its sign, +-65504; finite values are unchanged. The FP16 form's
// F-engine converts y*g with round to nearest even, so a bin beyond the half
// range (|y*g| >= 65520, i.e. 48 dB above the calibrated RMS of 256) is stored
// as +-inf; the X-engine reads every spectrum word through this function, which
// gives exactly the products of y*g clamped to +-65504 before the conversion
// (the values in [65504, 65520) round to 65504 either way). NaN cannot occur:
// y and g are finite.
CORR_HD inline uint32_t saturate_fp16x2(uint32_t w) {
    // On the bit patterns: the magnitude of each half (sign cleared) is capped
    // at 0x7BFF = 65504 and the sign put back; inf (0x7C00) is the only
    // magnitude above it that can occur. Integer ops: exact, no canonicalisation.
#if defined(__HIP_DEVICE_COMPILE__)
    typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));
    const u16x2_t cap = {(unsigned short)0x7BFFu, (unsigned short)0x7BFFu};
    const u16x2_t mag = __builtin_bit_cast(u16x2_t, w & 0x7FFF7FFFu);
    return (w & 0x80008000u) | __builtin_bit_cast(uint32_t, __builtin_elementwise_min(mag, cap));
#else
    uint32_t out = w & 0x80008000u;
    for (int h = 0; h < 2; ++h) {
        const uint32_t m = (w >> (16 * h)) & 0x7FFFu;
        out |= (m < 0x7BFFu ? m : 0x7BFFu) << (16 * h);
    }
    return out;
#endif
}

CORR_HD inline int quant_int8(float v) {
    const float r = rintf(v);
    return r > (float)kInt8Max ? kInt8Max : (r < -(float)kInt8Max ? -kInt8Max : (int)r);
}

CORR_HD inline uint16_t pack_int8(float2 y, float g) {
#if defined(__HIP_DEVICE_COMPILE__)
    // Clamping to +-7 before rounding gives the same integer (the bounds are
    // integers). Adding 1.5 * 2^23 then rounds to nearest even, as rintf does,
    // and leaves the integer's two's-complement low byte in the sum's low bits.
    constexpr float kMagic = 12582912.0f;
    const float a = fminf(fmaxf(y.x * g, -(float)kInt8Max), (float)kInt8Max) + kMagic;
    const float b = fminf(fmaxf(y.y * g, -(float)kInt8Max), (float)kInt8Max) + kMagic;
    return (uint16_t)((__builtin_bit_cast(uint32_t, a) & 0xFFu) | ((__builtin_bit_cast(uint32_t, b) & 0xFFu) << 8));
#else
    return (uint16_t)((uint32_t)(quant_int8(y.x * g) & 0xFF) | ((uint32_t)(quant_int8(y.y * g) & 0xFF) << 8));
#endif
}

// Spectrum buffer of the FP16 and int8 forms: channels in tiles of `tile` (16
// or 32, the X-engine's channels per workgroup), each tile's frames
// contiguous, [station][n_freq/tile][fcap frames][tile channels] (fcap = the
// buffer's frame capacity). The X-engine then reads one tile's frames as a
// contiguous run (the [station][frame][channel] order would have it read
// pieces of a row 64 KB apart, which halves the HBM rate), and the F-engine
// writes pieces of tile*element bytes (whole 128-byte lines for 32 FP16
// channels).
CORR_HD inline size_t spec_index(int st, int f, int c, int nf, int fcap, int tile) {
    return (((size_t)st * (size_t)(nf / tile) + (size_t)(c / tile)) * (size_t)fcap + (size_t)f) * (size_t)tile +
           (size_t)(c % tile);
}

}  // namespace forms
}  // namespace corr
