// Shared device helpers for the galario_amd HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace gl {

// Interleaved complex, same memory layout as std::complex<T> / numpy complex.
template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T re, im;
};

__device__ __forceinline__ float ffma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double ffma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ void fsincos(float x, float* s, float* c) { sincosf(x, s, c); }
__device__ __forceinline__ void fsincos(double x, double* s, double* c) { sincos(x, s, c); }

template <typename T>
__device__ __forceinline__ cpx<T> mk(T re, T im) { return cpx<T>{re, im}; }
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.re, -a.im}; }
template <typename T>
__device__ __forceinline__ cpx<T> cscale(T s, cpx<T> a) { return {s * a.re, s * a.im}; }

// Full complex product with fused multiply-adds (FFT-internal twiddles only;
// the pipeline stages that follow galario's arithmetic never use it).
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {ffma(a.re, b.re, -a.im * b.im), ffma(a.re, b.im, a.im * b.re)};
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Wave64 sum of a double via cross-lane shuffles.
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

}  // namespace gl
