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

// Cross-lane moves through DPP row patterns: a modifier of a VALU move, no LDS round trip
// (shuffles compile to ds_bpermute, which goes through the LDS crossbar).
enum : int {
    DPP_XOR1 = 0xB1,             // quad_perm [1,0,3,2]: lane l <- l ^ 1
    DPP_XOR2 = 0x4E,             // quad_perm [2,3,0,1]: lane l <- l ^ 2
    DPP_ROW_MIRROR = 0x140,      // lane l <- 15 - l within each row of 16
    DPP_ROW_HALF_MIRROR = 0x141, // lane l <- 7 - l within each half row of 8
    DPP_ROW_ROR4 = 0x124,        // row rotation by 4 lanes within each row of 16
    DPP_ROW_ROR8 = 0x128         // row rotation by 8 lanes within each row of 16
};

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const int lo = dpp_i32<CTRL>(__double2loint(x)), hi = dpp_i32<CTRL>(__double2hiint(x));
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_f64(double x, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                            __builtin_amdgcn_readlane(__double2loint(x), lane));
}

// Wave64 sum of a double, the same value in every lane; all 64 lanes must be active.  Within each
// row of 16 lanes by DPP (pairs, quads, half rows, rows: every lane of a row ends with the same
// row sum), then the four row sums r0..r3 (lanes 0, 16, 32, 48) as (r0 + r1) + (r2 + r3).  Every
// kernel reduces with this one tree, so sums of the same terms agree bit for bit across kernels.
__device__ __forceinline__ double wave_sum(double x) {
    x += dpp_f64<DPP_XOR1>(x);
    x += dpp_f64<DPP_XOR2>(x);
    x += dpp_f64<DPP_ROW_HALF_MIRROR>(x);
    x += dpp_f64<DPP_ROW_MIRROR>(x);
    return (lane_f64(x, 0) + lane_f64(x, 16)) + (lane_f64(x, 32) + lane_f64(x, 48));
}

// Three wave64 sums at once, in a fixed order of their own (not wave_sum's); all 64 lanes active,
// results in every lane.  The lanes trade halves of the vector: after the pair step even lanes carry
// the pair sums of (a, b) and odd lanes those of (c, 0), after the quad step lane l carries one
// quad sum (l mod 4 = 0: a, 1: c, 2: b, 3: 0); then rows of 16 by rotations by 4 and 8 (lanes 4
// apart carry the same sum) and the rows by two shuffles.  About half the instructions of three
// wave_sum calls.
__device__ __forceinline__ void wave_sum3(double& a, double& b, double& c, int lane) {
    const bool o1 = lane & 1, o2 = lane & 2;
    double k0 = o1 ? c : a, k1 = o1 ? 0.0 : b;
    k0 += dpp_f64<DPP_XOR1>(o1 ? a : c);
    k1 += dpp_f64<DPP_XOR1>(o1 ? b : 0.0);
    double k = o2 ? k1 : k0;
    k += dpp_f64<DPP_XOR2>(o2 ? k0 : k1);
    k += dpp_f64<DPP_ROW_ROR4>(k);
    k += dpp_f64<DPP_ROW_ROR8>(k);
    k += __shfl_xor(k, 16, 64);
    k += __shfl_xor(k, 32, 64);
    a = lane_f64(k, 0);
    c = lane_f64(k, 1);
    b = lane_f64(k, 2);
}

// Wave64 (min, max, min, max) of four ints at once by the same scheme (lane l mod 4 after the quad
// step: 0 -> lo0, 1 -> lo1, 2 -> hi0, 3 -> hi1; lanes with l & 2 take maxima); results uniform.
__device__ __forceinline__ void wave_minmax2(int& lo0, int& hi0, int& lo1, int& hi1, int lane) {
    const bool o1 = lane & 1, o2 = lane & 2;
    int k0 = o1 ? lo1 : lo0, k1 = o1 ? hi1 : hi0;
    k0 = min(k0, dpp_i32<DPP_XOR1>(o1 ? lo0 : lo1));
    k1 = max(k1, dpp_i32<DPP_XOR1>(o1 ? hi0 : hi1));
    int k = o2 ? k1 : k0;
    int r = dpp_i32<DPP_XOR2>(o2 ? k0 : k1);
    k = o2 ? max(k, r) : min(k, r);
    r = dpp_i32<DPP_ROW_ROR4>(k);
    k = o2 ? max(k, r) : min(k, r);
    r = dpp_i32<DPP_ROW_ROR8>(k);
    k = o2 ? max(k, r) : min(k, r);
    r = __shfl_xor(k, 16, 64);
    k = o2 ? max(k, r) : min(k, r);
    r = __shfl_xor(k, 32, 64);
    k = o2 ? max(k, r) : min(k, r);
    lo0 = __builtin_amdgcn_readlane(k, 0);
    lo1 = __builtin_amdgcn_readlane(k, 1);
    hi0 = __builtin_amdgcn_readlane(k, 2);
    hi1 = __builtin_amdgcn_readlane(k, 3);
}

// Wave64 minimum / maximum of an int, the same value in every lane (all 64 lanes active).
__device__ __forceinline__ int wave_min(int x) {
    x = min(x, dpp_i32<DPP_XOR1>(x));
    x = min(x, dpp_i32<DPP_XOR2>(x));
    x = min(x, dpp_i32<DPP_ROW_HALF_MIRROR>(x));
    x = min(x, dpp_i32<DPP_ROW_MIRROR>(x));
    return min(min(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16)),
               min(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48)));
}
__device__ __forceinline__ int wave_max(int x) {
    x = max(x, dpp_i32<DPP_XOR1>(x));
    x = max(x, dpp_i32<DPP_XOR2>(x));
    x = max(x, dpp_i32<DPP_ROW_HALF_MIRROR>(x));
    x = max(x, dpp_i32<DPP_ROW_MIRROR>(x));
    return max(max(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16)),
               max(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48)));
}

}  // namespace gl
