// Closed form of the rank-1 border map (padding/half_dim.rs:30-343 on one
// axis), computed per position instead of tabulated, for batches of many
// rank-1 signals whose maps would not fit in memory.  It agrees with the
// tabulated map of plan.cpp position by position (checked by the CPU tests).
#pragma once

#include <cstdint>

#include "plan.hpp"

#ifndef __HIPCC__
#define NDC_HD inline
#else
#define NDC_HD __host__ __device__ __forceinline__
#endif

namespace ndc {

// Border tags after resolution: Zeros is a constant fill.
enum Side1D : int32_t { kS1Const = 0, kS1Reflect = 1, kS1Replicate = 2, kS1Circular = 3 };

NDC_HD int32_t side1d(int32_t border_tag) {
    switch (border_tag) {
        case NDC_BORDER_REFLECT: return kS1Reflect;
        case NDC_BORDER_REPLICATE: return kS1Replicate;
        case NDC_BORDER_CIRCULAR: return kS1Circular;
        default: return kS1Const;
    }
}

// Front pad cell p (< p0) after the front fill.  I is the index type: int64_t
// on the host and for batches, int32_t inside the LDS-resident FFT passes.
template <typename I>
NDC_HD I resolve_front(I p, I n, I p0, int32_t front) {
    switch (front) {
        case kS1Const: return kConstFront;
        case kS1Replicate: return 0;
        case kS1Reflect: {
            // src = 2 p0 - p lies after the interior start; the back pad is
            // not written yet during the front fill
            const I src = 2 * p0 - p;
            return src < p0 + n ? src - p0 : I(kUnwritten);
        }
        default: {
            // circular: src = n + p; a front cell not written yet keeps the
            // initial value
            const I src = n + p;
            return src < p0 ? I(kUnwritten) : src - p0;
        }
    }
}

// Source of padded position p in [0, n + p0 + p1): an index into x (>= 0) or
// kConstFront / kConstBack / kUnwritten.
template <typename I>
NDC_HD I resolve1d(I p, I n, I p0, int32_t front, int32_t back) {
    if (p < p0) return resolve_front(p, n, p0, front);
    if (p < p0 + n) return p - p0;
    switch (back) {
        case kS1Const: return kConstBack;
        case kS1Replicate: return n - 1;
        case kS1Reflect: {
            // src = 2 (p0 + n - 1) - p; reaching into the front pad reads the
            // value the front fill left there
            const I src = 2 * (p0 + n - 1) - p;
            return src >= p0 ? src - p0 : resolve_front(src, n, p0, front);
        }
        default:
            // circular back: periodic in n (each cell copies the one n before)
            return (p - p0) % n;
    }
}

}  // namespace ndc
