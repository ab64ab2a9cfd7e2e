// Small in-register DFT butterflies shared by the FFT kernels.
#pragma once

#include "device_common.hpp"

namespace ndc {
namespace fftb {

// ---------------------------------------------------------------------------
// Small DFT butterflies.  Forward uses W = exp(-2*pi*i/P); inverse the
// conjugate, unnormalised (rustfft's convention).
// ---------------------------------------------------------------------------

template <typename C>
__device__ __forceinline__ void bfly2(C* v) {
    C a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
}

template <typename C>
__device__ __forceinline__ void bfly4(C* v, bool inv) {
    C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    C s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
    C r = inv ? cmul_i(d13) : cmul_negi(d13);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, r);
    v[3] = csub(d02, r);
}

template <typename C>
__device__ __forceinline__ void bfly8(C* v, bool inv) {
    using R = decltype(C{}.x);
    const R h = R(0.70710678118654757);
    C a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        a[r] = cadd(v[r], v[r + 4]);
        b[r] = csub(v[r], v[r + 4]);
    }
    // b[r] *= W_8^r
    if (!inv) {
        b[1] = cmake<C>(h * (b[1].x + b[1].y), h * (b[1].y - b[1].x));
        b[2] = cmul_negi(b[2]);
        b[3] = cmake<C>(h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y));
    } else {
        b[1] = cmake<C>(h * (b[1].x - b[1].y), h * (b[1].x + b[1].y));
        b[2] = cmul_i(b[2]);
        b[3] = cmake<C>(-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y));
    }
    bfly4(a, inv);
    bfly4(b, inv);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[2 * q] = a[q];
        v[2 * q + 1] = b[q];
    }
}

// cos/sin(2*pi*m/P), m = 0 .. (P-1)/2
template <int P> struct OddConst;
template <> struct OddConst<3> {
    static constexpr double c[2] = {1.0, -0.5};
    static constexpr double s[2] = {0.0, 0.86602540378443865};
};
template <> struct OddConst<5> {
    static constexpr double c[3] = {1.0, 0.30901699437494745, -0.80901699437494745};
    static constexpr double s[3] = {0.0, 0.95105651629515353, 0.58778525229247314};
};
template <> struct OddConst<7> {
    static constexpr double c[4] = {1.0, 0.62348980185873359, -0.22252093395631434, -0.90096886790241915};
    static constexpr double s[4] = {0.0, 0.78183148246802980, 0.97492791218182362, 0.43388373911755812};
};

// y[k] = v0 + sum_n a_n cos(th n k) -/+ i sum_n b_n sin(th n k), with
// a_n = v[n] + v[P-n], b_n = v[n] - v[P-n].
template <int P, typename C>
__device__ __forceinline__ void bfly_odd(C* v, bool inv) {
    using R = decltype(C{}.x);
    constexpr int H = (P - 1) / 2;
    C a[H + 1], b[H + 1];
#pragma unroll
    for (int n = 1; n <= H; ++n) {
        a[n] = cadd(v[n], v[P - n]);
        b[n] = csub(v[n], v[P - n]);
    }
    C y0 = v[0];
#pragma unroll
    for (int n = 1; n <= H; ++n) y0 = cadd(y0, a[n]);
    C out[P];
    out[0] = y0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        C A = v[0], B = cmake<C>(R(0), R(0));
#pragma unroll
        for (int n = 1; n <= H; ++n) {
            const int m = (n * k) % P;
            const R cs = R(OddConst<P>::c[m <= H ? m : P - m]);
            const R sn = m <= H ? R(OddConst<P>::s[m]) : R(-OddConst<P>::s[P - m]);
            A = cadd(A, cscale(a[n], cs));
            B = cadd(B, cscale(b[n], sn));
        }
        // forward: -i*B ; inverse: +i*B
        const C t = inv ? cmul_i(B) : cmul_negi(B);
        out[k] = cadd(A, t);
        out[P - k] = csub(A, t);
    }
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = out[k];
}

// cos and sin of 2*pi*e/16, e = 0..9
struct W16Const {
    static constexpr double c[10] = {1.0, 0.92387953251128674, 0.70710678118654757, 0.38268343236508977, 0.0,
                                     -0.38268343236508977, -0.70710678118654757, -0.92387953251128674, -1.0,
                                     -0.92387953251128674};
    static constexpr double s[10] = {0.0, 0.38268343236508977, 0.70710678118654757, 0.92387953251128674, 1.0,
                                     0.92387953251128674, 0.70710678118654757, 0.38268343236508977, 0.0,
                                     -0.38268343236508977};
};

// 16-point DFT as 4 x 4: n = n1 + 4 n2, k = 4 k1 + k2,
// X[k] = sum_n1 W4^(n1 k1) W16^(n1 k2) sum_n2 v[n1 + 4 n2] W4^(n2 k2).
template <typename C>
__device__ __forceinline__ void bfly16(C* v, bool inv) {
    using R = decltype(C{}.x);
    C y[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
        C t[4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) t[n2] = v[n1 + 4 * n2];
        bfly4(t, inv);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            const int e = n1 * k2;
            if (e == 0) y[n1][k2] = t[k2];
            else if (e == 4) y[n1][k2] = inv ? cmul_i(t[k2]) : cmul_negi(t[k2]);
            else y[n1][k2] = cmulc(t[k2], cmake<C>(R(W16Const::c[e]), R(-W16Const::s[e])), inv);
        }
    }
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
        C t[4];
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) t[n1] = y[n1][k2];
        bfly4(t, inv);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) v[4 * k1 + k2] = t[k1];
    }
}

template <int P, typename C>
__device__ __forceinline__ void bfly(C* v, bool inv) {
    if constexpr (P == 2) bfly2(v);
    else if constexpr (P == 4) bfly4(v, inv);
    else if constexpr (P == 8) bfly8(v, inv);
    else if constexpr (P == 16) bfly16(v, inv);
    else bfly_odd<P>(v, inv);
}

}  // namespace fftb
}  // namespace ndc
