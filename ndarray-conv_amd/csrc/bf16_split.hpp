// BF16 operands for the matrix cores: round-to-nearest-even conversion, the
// hi/lo split of the "BF16x3" products (a*b ~ ah*bh + ah*bl + al*bh) and the
// three-term split of the "BF16x6" products (a*b ~ a0b0 + a0b1 + a1b0 + a0b2 +
// a1b1 + a2b0).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ndc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Round to nearest even bf16 (the bit pattern in the low 16 bits).
__device__ __forceinline__ uint32_t bf16_bits(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

// f = hi + lo (+ a residual of at most 2^-17 |f|), packed as hi | lo << 16.
__device__ __forceinline__ uint32_t split_pack(float f) {
    const uint32_t h = bf16_bits(f);
    const float hf = __builtin_bit_cast(float, h << 16);
    const uint32_t l = bf16_bits(f - hf);
    return h | (l << 16);
}

// f = t[0] + ... + t[P-1] (+ a residual of at most 2^-(8P) |f|): each term is
// the bf16 rounding of what the previous terms left, and every difference is
// exact in FP32.  P = 2 is the hi/lo split of BF16x3; P = 3 is the three-term
// split of BF16x6.
template <int P>
__device__ __forceinline__ void split_terms(float f, uint16_t* t) {
    float r = f;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const uint32_t h = bf16_bits(r);
        t[p] = uint16_t(h);
        if (p + 1 < P) r -= __builtin_bit_cast(float, h << 16);
    }
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// The P-term split of two values at once: t[p] holds term p of f0 in its low
// and of f1 in its high half.  The conversion rounds to nearest even, as
// bf16_bits does, and gfx950 converts both values in one instruction.
template <int P>
__device__ __forceinline__ void split_pair(float f0, float f1, uint32_t* t) {
    float r0 = f0, r1 = f1;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const bf16x2 h = {__bf16(r0), __bf16(r1)};
        t[p] = __builtin_bit_cast(uint32_t, h);
        if (p + 1 < P) {
            r0 -= __builtin_bit_cast(float, t[p] << 16);
            r1 -= __builtin_bit_cast(float, t[p] & 0xFFFF0000u);
        }
    }
}

// Eight packed (hi | lo << 16) values to the two 8-element bf16 operands:
// hi[i] holds elements 2i, 2i+1 of the hi parts, lo[i] of the lo parts.
__device__ __forceinline__ void unpack8(const uint32_t* p, u32x4& hi, u32x4& lo) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        hi[i] = __builtin_amdgcn_perm(p[2 * i + 1], p[2 * i], 0x05040100u);
        lo[i] = __builtin_amdgcn_perm(p[2 * i + 1], p[2 * i], 0x07060302u);
    }
}

}  // namespace ndc
