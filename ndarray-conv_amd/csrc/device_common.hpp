// Device-side element arithmetic shared by the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "plan.hpp"

namespace ndc {

// Complex numbers are stored as float2 / double2 (re, im), the layout of
// num_complex::Complex<T>.
template <typename R> struct CxOf;
template <> struct CxOf<float> { using type = float2; };
template <> struct CxOf<double> { using type = double2; };
template <typename R> using Cx = typename CxOf<R>::type;

template <typename C> __host__ __device__ __forceinline__ C cmake(decltype(C{}.x) re, decltype(C{}.x) im) {
    C c;
    c.x = re;
    c.y = im;
    return c;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmake<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmake<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    return cmake<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b) when inv, a * b otherwise
template <typename C> __device__ __forceinline__ C cmulc(C a, C b, bool conj_b) {
    if (conj_b) return cmake<C>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
    return cmul(a, b);
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmake<C>(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C{}.x) s) { return cmake<C>(a.x * s, a.y * s); }
// -i * a
template <typename C> __device__ __forceinline__ C cmul_negi(C a) { return cmake<C>(a.y, -a.x); }
// i * a
template <typename C> __device__ __forceinline__ C cmul_i(C a) { return cmake<C>(-a.y, a.x); }

// Element arithmetic of the direct convolution, in the element type itself:
// floats round after every product and every sum (the reference's
// `tmp_res += a * b`, conv/mod.rs:191-193); integers wrap like the
// reference's release build, which makes them order independent.
template <typename T, typename Enable = void> struct Arith {
    __device__ static __forceinline__ T zero() { return T(0); }
    __device__ static __forceinline__ bool is_zero(T v) { return v == T(0); }
    __device__ static __forceinline__ T mac(T acc, T a, T b) { return acc + a * b; }
};

template <typename T>
struct Arith<T, typename std::enable_if<std::is_integral<T>::value>::type> {
    using U = typename std::make_unsigned<T>::type;
    using W = typename std::conditional<(sizeof(T) < 4), uint32_t, U>::type;
    __device__ static __forceinline__ T zero() { return T(0); }
    __device__ static __forceinline__ bool is_zero(T v) { return v == T(0); }
    __device__ static __forceinline__ T mac(T acc, T a, T b) {
        W p = W(U(a)) * W(U(b));
        W s = W(U(acc)) + p;
        return T(U(s));
    }
};

template <> struct Arith<float2> {
    __device__ static __forceinline__ float2 zero() { return cmake<float2>(0.f, 0.f); }
    __device__ static __forceinline__ bool is_zero(float2 v) { return v.x == 0.f && v.y == 0.f; }
    __device__ static __forceinline__ float2 mac(float2 acc, float2 a, float2 b) { return cadd(acc, cmul(a, b)); }
};
template <> struct Arith<double2> {
    __device__ static __forceinline__ double2 zero() { return cmake<double2>(0.0, 0.0); }
    __device__ static __forceinline__ bool is_zero(double2 v) { return v.x == 0.0 && v.y == 0.0; }
    __device__ static __forceinline__ double2 mac(double2 acc, double2 a, double2 b) { return cadd(acc, cmul(a, b)); }
};

}  // namespace ndc
