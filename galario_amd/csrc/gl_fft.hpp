// Workgroup-resident Stockham FFT for gfx950 (wave64), N = 2^LOG2N points.
//
// One workgroup computes one complex N-point transform.  Each of the NT = N/E
// threads keeps E = 16 elements (fewer for N < 1024) in registers; a pass of
// radix R is a register-resident R-point DFT, and consecutive passes exchange
// data through LDS (N reals + 1/16 padding, real parts then imaginary parts,
// so the stride-R stores of the early passes spread over all banks).  The first pass reads its inputs
// straight from a caller functor and the last pass hands its outputs straight
// to a caller functor, so both touch HBM coalesced (thread j owns elements
// j + r*N/R) and the transform costs 2 (E>=16: log16 N - 1) LDS round trips.
//
// Pruning: the image rows/columns of galario's pipeline are non-zero only on
// [0, Rs) U [N-Rs, N) (the swept box after the quadrant shift).  When
// Rs <= N/R the first radix-R pass has at most two non-zero inputs per
// butterfly (r = 0 and r = R-1) and the pass is evaluated as
// X[k] = x0 + x_{R-1} W_R^{-k}, which also skips the loads of the zeros.
//
// Twiddles: W_N^e = exp(-2 pi i e/N) from a host-built table (forward sign);
// per radix-16 butterfly four table reads (W^k, W^2k, W^4k, W^8k) and eleven
// products give the 15 outer twiddles.  Internal radix-16 constants are exact
// compile-time literals.
#pragma once

#include <type_traits>

#include "gl_common.hpp"

namespace gl {

// cos / sin of 2 pi m / 16, m = 0..15.
__device__ constexpr double kC16[16] = {
    1.0, 0.92387953251128673848, 0.70710678118654752440, 0.38268343236508978178,
    0.0, -0.38268343236508978178, -0.70710678118654752440, -0.92387953251128673848,
    -1.0, -0.92387953251128673848, -0.70710678118654752440, -0.38268343236508978178,
    0.0, 0.38268343236508978178, 0.70710678118654752440, 0.92387953251128673848};
__device__ constexpr double kS16[16] = {
    0.0, 0.38268343236508978178, 0.70710678118654752440, 0.92387953251128673848,
    1.0, 0.92387953251128673848, 0.70710678118654752440, 0.38268343236508978178,
    0.0, -0.38268343236508978178, -0.70710678118654752440, -0.92387953251128673848,
    -1.0, -0.92387953251128673848, -0.70710678118654752440, -0.38268343236508978178};

// x * exp(-+ 2 pi i M / 16): forward (INV=false) uses the minus sign.
template <int M, bool INV, typename T>
__device__ __forceinline__ cpx<T> rot16(cpx<T> x) {
    constexpr int m = ((M % 16) + 16) % 16;
    constexpr int mm = INV ? (16 - m) % 16 : m;   // exp(+i a) = exp(-i (2pi - a))
    if constexpr (mm == 0) {
        return x;
    } else if constexpr (mm == 4) {               // * (-i)
        return {x.im, -x.re};
    } else if constexpr (mm == 8) {               // * (-1)
        return {-x.re, -x.im};
    } else if constexpr (mm == 12) {              // * (+i)
        return {-x.im, x.re};
    } else if constexpr (mm == 2) {               // * h(1 - i)
        constexpr T h = T(0.70710678118654752440);
        return {h * (x.re + x.im), h * (x.im - x.re)};
    } else if constexpr (mm == 6) {               // * h(-1 - i)
        constexpr T h = T(0.70710678118654752440);
        return {h * (x.im - x.re), -h * (x.re + x.im)};
    } else if constexpr (mm == 10) {              // * h(-1 + i)
        constexpr T h = T(0.70710678118654752440);
        return {-h * (x.re + x.im), h * (x.re - x.im)};
    } else if constexpr (mm == 14) {              // * h(1 + i)
        constexpr T h = T(0.70710678118654752440);
        return {h * (x.re - x.im), h * (x.re + x.im)};
    } else {
        constexpr T c = T(kC16[mm]);
        constexpr T s = T(-kS16[mm]);
        return {ffma(x.re, c, -x.im * s), ffma(x.re, s, x.im * c)};
    }
}

template <bool INV, typename T>
__device__ __forceinline__ void dft2(cpx<T>& a, cpx<T>& b) {
    const cpx<T> t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <bool INV, typename T>
__device__ __forceinline__ void dft4(cpx<T>& x0, cpx<T>& x1, cpx<T>& x2, cpx<T>& x3) {
    const cpx<T> s02 = cadd(x0, x2), d02 = csub(x0, x2);
    const cpx<T> s13 = cadd(x1, x3), d13 = rot16<4, INV>(csub(x1, x3));
    x0 = cadd(s02, s13);
    x2 = csub(s02, s13);
    x1 = cadd(d02, d13);
    x3 = csub(d02, d13);
}

// In-register R-point DFT (R = 2, 4, 8, 16), natural order in and out.
// R = R1 * R2 with R1 = min(R, 4): n = R2*n1 + n2, k = k1 + R1*k2.
template <int R, bool INV, typename T>
__device__ __forceinline__ void dft(cpx<T>* v) {
    if constexpr (R == 2) {
        dft2<INV>(v[0], v[1]);
    } else if constexpr (R == 4) {
        dft4<INV>(v[0], v[1], v[2], v[3]);
    } else {
        constexpr int R1 = 4, R2 = R / 4;
        cpx<T> y[R];
        // inner DFT4 over n1 for each n2, then the W_R^{n2 k1} twiddles
        static_for<0, R2>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            cpx<T> a0 = v[n2], a1 = v[R2 + n2], a2 = v[2 * R2 + n2], a3 = v[3 * R2 + n2];
            dft4<INV>(a0, a1, a2, a3);
            y[n2 * R1 + 0] = a0;
            y[n2 * R1 + 1] = rot16<(16 / R) * n2 * 1, INV>(a1);
            y[n2 * R1 + 2] = rot16<(16 / R) * n2 * 2, INV>(a2);
            y[n2 * R1 + 3] = rot16<(16 / R) * n2 * 3, INV>(a3);
        });
        // outer DFT_{R2} over n2 for each k1 -> X[k1 + 4 k2]
        static_for<0, R1>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            if constexpr (R2 == 2) {
                cpx<T> a = y[0 * R1 + k1], b = y[1 * R1 + k1];
                dft2<INV>(a, b);
                v[k1] = a;
                v[k1 + 4] = b;
            } else {
                cpx<T> a0 = y[0 * R1 + k1], a1 = y[1 * R1 + k1], a2 = y[2 * R1 + k1], a3 = y[3 * R1 + k1];
                dft4<INV>(a0, a1, a2, a3);
                v[k1] = a0;
                v[k1 + 4] = a1;
                v[k1 + 8] = a2;
                v[k1 + 12] = a3;
            }
        });
    }
}

// X[k] = x0 + x_{R-1} W_R^{(R-1)k}: an R-point DFT whose only non-zero
// inputs are the first and the last.
template <int R, bool INV, typename T>
__device__ __forceinline__ void dft_first_last(cpx<T>* v) {
    const cpx<T> a = v[0], b = v[R - 1];
    static_for<0, R>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        v[k] = cadd(a, rot16<((16 / R) * (R - 1) * k) % 16, INV>(b));
    });
}

template <int LOG2N>
struct FftShape {
    static constexpr int N = 1 << LOG2N;
    static constexpr int LOG2E = LOG2N >= 10 ? 4 : LOG2N - 6;
    static constexpr int E = 1 << LOG2E;
    static constexpr int NT = N / E;
    static constexpr int NFULL = LOG2N / LOG2E;      // radix-E passes
    static constexpr int LOG2REM = LOG2N % LOG2E;    // one extra pass of radix 2^LOG2REM
    static constexpr int NPASS = NFULL + (LOG2REM ? 1 : 0);
    static constexpr int LDS_ELEMS = N + (N >> 4);
    static_assert(LOG2N >= 7 && LOG2N <= 14, "workgroup FFT covers 128 <= N <= 16384");
    // radix of pass p (0-based): radix E everywhere except pass 1 when a remainder exists
    static constexpr int log2radix(int p) { return (LOG2REM && p == 1) ? LOG2REM : LOG2E; }
    static constexpr int ns_before(int p) {
        int s = 0;
        for (int i = 0; i < p; ++i) s += log2radix(i);
        return 1 << s;
    }
};

__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }

// SPLIT: exchange real parts then imaginary parts through one buffer of
// LDS_ELEMS reals (4 barriers per exchange, half the LDS: more workgroups per
// CU); otherwise both parts at once through 2 * LDS_ELEMS reals (2 barriers:
// lower latency, for kernels with few workgroups).
template <typename T, int LOG2N, bool INV, bool SPLIT = true>
struct BlockFft {
    using S = FftShape<LOG2N>;
    static constexpr int N = S::N, E = S::E, NT = S::NT;
    static constexpr int LDS_REALS = SPLIT ? S::LDS_ELEMS : 2 * S::LDS_ELEMS;   // LDS the caller provides
    static constexpr int TWC = 8;           // twiddle bases per pass and thread

    static __device__ __forceinline__ cpx<T> twid(const cpx<T>* __restrict__ tw, int e) {
        const cpx<T> w = tw[e];
        return INV ? cconj(w) : w;
    }

    // Table reads of pass P: for each of the thread's E/R butterflies the bases
    // W^e1, W^2e1, W^4e1, W^8e1 (as many as log2 R).  They depend only on the
    // thread index; they are issued before the LDS exchange that precedes the
    // pass, so the exchange covers their latency.
    template <int P>
    static __device__ __forceinline__ void preload(cpx<T> (&tc)[TWC], int tid, const cpx<T>* __restrict__ tw) {
        constexpr int LR = S::log2radix(P);
        constexpr int R = 1 << LR;
        constexpr int NS = S::ns_before(P);
        static_assert((E / R) * LR <= TWC, "twiddle cache too small");
#pragma unroll
        for (int b = 0; b < E / R; ++b) {
            const int j = tid + b * NT;
            const int e1 = (j & (NS - 1)) * (N / (NS * R));
#pragma unroll
            for (int q = 0; q < LR; ++q) tc[b * LR + q] = twid(tw, e1 << q);
        }
    }

    // v[r] *= W_N^{r e1}, r = 1..R-1, from the bases W^e1, W^2e1, W^4e1, W^8e1
    template <int R>
    static __device__ __forceinline__ void apply_twiddles(cpx<T>* v, const cpx<T>* base) {
        const cpx<T> t1 = base[0];
        v[1] = cmul(v[1], t1);
        if constexpr (R >= 4) {
            const cpx<T> t2 = base[1];
            const cpx<T> t3 = cmul(t2, t1);
            v[2] = cmul(v[2], t2);
            v[3] = cmul(v[3], t3);
            if constexpr (R >= 8) {
                const cpx<T> t4 = base[2];
                const cpx<T> t5 = cmul(t4, t1), t6 = cmul(t4, t2), t7 = cmul(t4, t3);
                v[4] = cmul(v[4], t4);
                v[5] = cmul(v[5], t5);
                v[6] = cmul(v[6], t6);
                v[7] = cmul(v[7], t7);
                if constexpr (R >= 16) {
                    const cpx<T> t8 = base[3];
                    v[8] = cmul(v[8], t8);
                    v[9] = cmul(v[9], cmul(t8, t1));
                    v[10] = cmul(v[10], cmul(t8, t2));
                    v[11] = cmul(v[11], cmul(t8, t3));
                    v[12] = cmul(v[12], cmul(t8, t4));
                    v[13] = cmul(v[13], cmul(t8, t5));
                    v[14] = cmul(v[14], cmul(t8, t6));
                    v[15] = cmul(v[15], cmul(t8, t7));
                }
            }
        }
    }

    // twiddle + DFT of pass P on the registers (E/R butterflies per thread)
    template <int P>
    static __device__ __forceinline__ void compute(cpx<T> (&v)[E], const cpx<T> (&tc)[TWC]) {
        constexpr int LR = S::log2radix(P);
        constexpr int R = 1 << LR;
        constexpr int NS = S::ns_before(P);
#pragma unroll
        for (int b = 0; b < E / R; ++b) {
            if constexpr (NS > 1) apply_twiddles<R>(&v[b * R], &tc[b * LR]);
            dft<R, INV>(&v[b * R]);
        }
    }

    // one component (0: real, 1: imaginary) of the results of pass P -> LDS
    // at their Stockham output positions
    template <int P, int PART>
    static __device__ __forceinline__ void store_lds(T* lds, const cpx<T> (&v)[E], int tid) {
        constexpr int R = 1 << S::log2radix(P);
        constexpr int NS = S::ns_before(P);
#pragma unroll
        for (int b = 0; b < E / R; ++b) {
            const int j = tid + b * NT;
            const int k = j & (NS - 1);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[lpad(base + r * NS)] = PART ? v[b * R + r].im : v[b * R + r].re;
        }
    }

    // one component of the inputs of pass P <- LDS
    template <int P, int PART>
    static __device__ __forceinline__ void load_lds(cpx<T> (&v)[E], const T* lds, int tid) {
        constexpr int R = 1 << S::log2radix(P);
#pragma unroll
        for (int b = 0; b < E / R; ++b) {
            const int j = tid + b * NT;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const T x = lds[lpad(j + r * (N / R))];
                if (PART) v[b * R + r].im = x;
                else v[b * R + r].re = x;
            }
        }
    }

    // Output 0 of the last pass only: X[tid] = sum_r v[r] W^{r e1}, by Horner's
    // rule on the base twiddle W^{e1} (the last pass is radix E, one butterfly
    // per thread).
    template <int P>
    static __device__ __forceinline__ void last_output0(cpx<T> (&v)[E], const cpx<T> (&tc)[TWC]) {
        const cpx<T> w = tc[0];
        cpx<T> acc = v[E - 1];
#pragma unroll
        for (int r = E - 2; r >= 0; --r) acc = cadd(cmul(acc, w), v[r]);
        v[0] = acc;
    }

    // Passes 1..NPASS-1 after the first pass has been computed in registers:
    // exchange through LDS (real parts, then imaginary parts, so the buffer
    // holds N reals), twiddle, DFT.  On return thread tid holds X[tid + r*N/E]
    // in v[r] (the last pass is radix E with one butterfly per thread); with
    // only0, v[0] = X[tid] alone (outputs below N/E).
    template <int P = 1>
    static __device__ __forceinline__ void rest(cpx<T> (&v)[E], T* lds, int tid, const cpx<T>* __restrict__ tw,
                                                bool only0) {
        if constexpr (P < S::NPASS) {
            cpx<T> tc[TWC];
            preload<P>(tc, tid, tw);          // issued before the exchange, used after it
            if constexpr (SPLIT) {
                store_lds<P - 1, 0>(lds, v, tid);
                __syncthreads();
                load_lds<P, 0>(v, lds, tid);
                __syncthreads();
                store_lds<P - 1, 1>(lds, v, tid);
                __syncthreads();
                load_lds<P, 1>(v, lds, tid);
                __syncthreads();
            } else {
                store_lds<P - 1, 0>(lds, v, tid);
                store_lds<P - 1, 1>(lds + S::LDS_ELEMS, v, tid);
                __syncthreads();
                load_lds<P, 0>(v, lds, tid);
                load_lds<P, 1>(v, lds + S::LDS_ELEMS, tid);
                __syncthreads();
            }
            if constexpr (P == S::NPASS - 1) {
                if (only0) {
                    last_output0<P>(v, tc);
                    return;
                }
            }
            compute<P>(v, tc);
            rest<P + 1>(v, lds, tid, tw, only0);
        }
    }

    // Full transform.  load(n) returns input element n; it is called only for
    // n in [0, Rs) U [N-Rs, N) (zeros elsewhere).  lds: LDS_ELEMS reals.  On
    // return v[r] holds X[tid + r*N/E] (only0: v[0] = X[tid] only), and every
    // thread has finished reading LDS.
    template <typename Loader>
    static __device__ __forceinline__ void run(cpx<T> (&v)[E], T* lds, int tid, int Rs,
                                               const cpx<T>* __restrict__ tw, Loader&& load, bool only0 = false) {
        constexpr int R0 = E;            // first pass is radix E, one butterfly per thread
        constexpr int STRIDE = N / R0;   // == NT
        const int j = tid;
        if (Rs <= STRIDE) {
            // at most x[j] (r = 0) and x[j + (R0-1) N/R0] (r = R0-1) are non-zero
            const cpx<T> z = {T(0), T(0)};
            v[0] = (j < Rs) ? load(j) : z;
            v[R0 - 1] = (j >= STRIDE - Rs) ? load(j + (R0 - 1) * STRIDE) : z;
            dft_first_last<R0, INV>(v);
        } else {
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const int n = j + r * STRIDE;
                const bool in = (n < Rs) || (n >= N - Rs);
                v[r] = in ? load(n) : cpx<T>{T(0), T(0)};
            }
            dft<R0, INV>(v);
        }
        rest<1>(v, lds, tid, tw, only0);
    }
};

// One wave's 512-point transform at the fractional frequencies K + k0/P
// (K = 0..511): Z[K] = sum_m y[m] W_512^(m (K + k0/P)), 64 lanes x 8
// elements, BlockFft's three radix-8 Stockham passes at N = 512 with every
// twiddle exponent q of pass p (q = lane mod 8^p) moved to q + k0/P, i.e.
// W_N^((P q + k0) 512 / 8^(p+1)) from the N-point table (N = 512 P).  Pass 0's
// factors W_N^(64 k0 r) are the same for every lane: the caller applies them
// with the input.  The exchanges go through the wave's own LDS region, real
// parts then imaginary parts (slot layout against bank conflicts: slot()), with no
// workgroup barrier (LDS operations of one wave execute in issue order; the
// fences keep the compiler from moving them).  Pass twiddles W^e1 come from
// one table read each (pass 2's as W_512^lane W_N^k0, a contiguous read of
// the 512-point table), W^2e1 and W^4e1 by squaring.  On entry
// v[r] = y[lane + 64 r] W_N^(64 k0 r); on return v[r] = Z[lane + 64 r].
// PAD selects the LDS layout of the exchanges (below).
template <typename T, bool PAD = false>
struct ResidueFft {
    using S = FftShape<9>;
    using B = BlockFft<T, 9, false, true>;
    static constexpr int E = S::E, TWC = B::TWC;
    static constexpr int LDS_REALS = PAD ? S::N + S::N / 8 : S::N;   // per wave
    static_assert(S::NT == 64 && E == 8 && S::NPASS == 3, "one wave of 8 elements per lane");

    // LDS slot of position i.  PAD = false: i ^ ((i >> 3) & 15) keeps the 512 positions in 512 slots and
    // spreads each exchange instruction's 16-lane groups over 32 distinct 4-byte banks (positions 8
    // apart, as pass 0 writes them, land in different banks).  PAD = true: i + (i >> 3) in 576 slots, a
    // few bank conflicts more, but every exchange address is a per-lane base plus a constant (positions
    // base + r, base + 8 r and lane + 64 r go to slots base' + r, base' + 9 r and lane' + 72 r), so the
    // offsets fold into the LDS instructions: the column pass's code has 33 fewer VALU and 16 fewer LDS
    // instructions, and its workgroups fit on a CU as before (its LDS is sized by the output transpose).  The adjoint column pass
    // keeps the 512-slot form, where 576 slots would cost a workgroup per CU.
    static __device__ __forceinline__ int slot(int i) { return PAD ? i + (i >> 3) : i ^ ((i >> 3) & 15); }

    template <int P, int PART>
    static __device__ __forceinline__ void store_lds(T* lds, const cpx<T> (&v)[E], int lane) {
        constexpr int NS = S::ns_before(P);
        const int k = lane & (NS - 1);
        const int base = (lane - k) * E + k;
#pragma unroll
        for (int r = 0; r < E; ++r) lds[slot(base + r * NS)] = PART ? v[r].im : v[r].re;
    }

    template <int P, int PART>
    static __device__ __forceinline__ void load_lds(cpx<T> (&v)[E], const T* lds, int lane) {
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const T x = lds[slot(lane + r * (S::N / E))];
            if (PART) v[r].im = x;
            else v[r].re = x;
        }
    }

    static __device__ __forceinline__ void wave_sync() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    template <int P>
    static __device__ __forceinline__ void exchange(cpx<T> (&v)[E], T* lds, int lane) {
        store_lds<P - 1, 0>(lds, v, lane);
        wave_sync();
        load_lds<P, 0>(v, lds, lane);
        wave_sync();
        store_lds<P - 1, 1>(lds, v, lane);
        wave_sync();
        load_lds<P, 1>(v, lds, lane);
        wave_sync();
    }

    static __device__ __forceinline__ void bases(cpx<T> (&tc)[TWC], cpx<T> w) {
        tc[0] = w;
        tc[1] = cmul(w, w);
        tc[2] = cmul(tc[1], tc[1]);
    }

    static __device__ __forceinline__ void run(cpx<T> (&v)[E], T* lds, int lane, const cpx<T>* __restrict__ tw,
                                               const cpx<T>* __restrict__ tw512, int P, int k0) {
        dft<E, false>(v);
        cpx<T> tc[TWC];
        bases(tc, tw[(P * (lane & 7) + k0) * 8]);
        exchange<1>(v, lds, lane);
        B::template compute<1>(v, tc);
        bases(tc, cmul(tw512[lane], tw[k0]));
        exchange<2>(v, lds, lane);
        B::template compute<2>(v, tc);
    }
};

}  // namespace gl
