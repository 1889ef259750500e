// cx.h — complex arithmetic and the Philox4x32-10 counter-based generator,
// usable from host and device code.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SHHD_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define SHHD_HD inline
#endif

namespace shhd {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T re, im;
};

template <typename T> SHHD_HD cx<T> mk(T r, T i) { cx<T> z; z.re = r; z.im = i; return z; }
template <typename T> SHHD_HD cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T> SHHD_HD cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T> SHHD_HD cx<T> operator*(cx<T> a, cx<T> b) {
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T> SHHD_HD cx<T> operator*(T s, cx<T> a) { return mk(s * a.re, s * a.im); }
template <typename T> SHHD_HD cx<T> conj(cx<T> a) { return mk(a.re, -a.im); }
// multiply by i and by -i
template <typename T> SHHD_HD cx<T> mul_i(cx<T> a) { return mk(-a.im, a.re); }
template <typename T> SHHD_HD cx<T> mul_mi(cx<T> a) { return mk(a.im, -a.re); }
template <typename T> SHHD_HD cx<T> czero() { return mk(T(0), T(0)); }

#if defined(__HIPCC__)
// Loads and stores of device (global) memory through pointers the compiler
// cannot place in an address space (they come out of descriptor tables):
// global_load / global_store with a scalar base and a 32-bit lane offset
// instead of flat accesses with 64-bit lane addresses, which also count
// against the LDS wait counter.
template <typename T> struct vec2_of;
template <> struct vec2_of<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct vec2_of<float> { typedef float type __attribute__((ext_vector_type(2))); };

template <typename T>
__device__ __forceinline__ cx<T> gload(const cx<T>* p) {
    using V = typename vec2_of<T>::type;
    const V v = *(const __attribute__((address_space(1))) V*)(p);
    return mk(v.x, v.y);
}
template <typename T>
__device__ __forceinline__ void gstore(cx<T>* p, cx<T> z) {
    using V = typename vec2_of<T>::type;
    V v;
    v.x = z.re;
    v.y = z.im;
    *(__attribute__((address_space(1))) V*)(p) = v;
}
// Element i of an array whose base is uniform across the wave: a scalar base
// and a 32-bit lane byte offset (i * sizeof, which must stay below 4 GiB: every
// array addressed this way is one field or plane)
template <typename T>
__device__ __forceinline__ cx<T> gload(const cx<T>* base, uint32_t i) {
    return gload(reinterpret_cast<const cx<T>*>(reinterpret_cast<const char*>(base) + i * uint32_t(sizeof(cx<T>))));
}
template <typename T>
__device__ __forceinline__ void gstore(cx<T>* base, uint32_t i, cx<T> z) {
    gstore(reinterpret_cast<cx<T>*>(reinterpret_cast<char*>(base) + i * uint32_t(sizeof(cx<T>))), z);
}
// Elements i and i + 1 (i even, base aligned to two elements) in one store
template <typename T> struct vec4_of;
template <> struct vec4_of<double> { typedef double type __attribute__((ext_vector_type(4))); };
template <> struct vec4_of<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <typename T>
__device__ __forceinline__ void gstore2(cx<T>* base, uint32_t i, cx<T> a, cx<T> b) {
    using V = typename vec4_of<T>::type;
    V v;
    v.x = a.re;
    v.y = a.im;
    v.z = b.re;
    v.w = b.im;
    *(__attribute__((address_space(1))) V*)(reinterpret_cast<char*>(base) + i * uint32_t(sizeof(cx<T>))) = v;
}
#endif

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t v[4]; };

SHHD_HD void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    const uint64_t p = static_cast<uint64_t>(a) * static_cast<uint64_t>(b);
    hi = static_cast<uint32_t>(p >> 32);
    lo = static_cast<uint32_t>(p);
}

SHHD_HD u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(M0, ctr.v[0], hi0, lo0);
        mulhilo(M1, ctr.v[2], hi1, lo1);
        u32x4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += W0;
        k1 += W1;
    }
    return ctr;
}

}  // namespace shhd
