// Adjoint (reverse-mode) chain of chi2 for the profile path on the real, even
// spectrum.  galario has no gradient; this differentiates its arithmetic:
//
//   V  = sgn(Z) A e^{i theta},  A = sum_k b_k |y_k|,  Z = sum_k b_k y_k
//        (y_k the four real corners, b_k the bilinear weights), theta =
//        2 pi (u dRA + v dDec) (the rotated form urot dRArot + vrot dDecrot
//        is the same number in exact arithmetic, so d theta/dPA = 0)
//   chi2 = sum w |V - O|^2,  G = dchi2/dV = 2 w (V - O)
//   dchi2/dA     = sgn(Z) Re(conj(G) e^{i theta})
//   dchi2/dtheta = Im(G) Re(V) - Re(G) Im(V)
//   dchi2/dy_k   = dchi2/dA b_k sgn(y_k)          (the phase of a real Z is
//                                                  locally constant)
//   PA moves the sampling point: dq/durot = s/duv, dt/dvrot = v_origin s/duv,
//   durot/dPA = -vrot, dvrot/dPA = urot.
//
// The corner gradients are scattered with FP64 atomics into Hh, the adjoint
// of the half spectrum H (both mirror rows of a frequency land on the same
// element, which is the adjoint of reading H at |r - N/2|).  The two passes
// are then transposed: each is sum_k c[k] cos(2 pi k m / N) over an even
// sequence, so its adjoint is the same even-real transform of the weighted
// even extension (weight 1/2 on interior input frequencies) times the
// multiplicity of each output index (1 at 0 and N/2, 2 elsewhere).  The row
// adjoint ends in the sweep adjoint: every pixel feeds I[iR] with weight
// sr (1 - alpha) and I[iR+1] with sr alpha, and inc through dr/dinc =
// (x/cos inc)^2 tan(inc) / r; the central pixel is linear in I[0..iIN+1].
#pragma once

#include <climits>

#include "gl_kernels.hpp"

namespace gl {

__device__ __forceinline__ void atomic_add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }

// ---------------------------------------------------------------------------
// sample + its adjoint: chi2 and (PA, dRA, dDec) partials per workgroup,
// corner gradients scattered into Hh.
//
// Global FP64 atomics on gfx950 are resolved beyond the per-XCD L2, so many
// atomics on one address serialise: with log-uniform baselines ~10% of all
// points fall in the few cells around the origin.  Each workgroup therefore
// first accumulates its corner gradients in LDS:
//   * cells with column < GHOT and folded row < GHOT go to a GHOT x GHOT LDS
//     tile, flushed into one of GREP replicas of that tile (by workgroup
//     index); the adjoint column pass folds the replicas into Hh;
//   * other cells go to an LDS window over the bounding box of the
//     workgroup's cells when it holds at most GWIN cells (points are stored in
//     a locality order, so it usually does), flushed with one atomic per
//     non-zero cell; otherwise straight to Hh.
// ---------------------------------------------------------------------------
constexpr int GHOT = 16;          // hot tile: columns [0, GHOT) x folded rows [0, GHOT)
constexpr int GREP = 32;          // replicas of the hot tile in global memory
constexpr int GWIN = 3072;        // LDS window capacity (cells): 26 KB with the hot tile, 6 waves per SIMD
                                  // (the single call's; the batched form's autotuner also tries 2048 and 4096)
constexpr int GHOT_ELEMS = GREP * GHOT * GHOT;

template <typename T>
struct GradArgs {
    double* Hh;          // adjoint of H, same layout (h_index, ldH), zero on entry
    double* hot;         // GREP x GHOT x GHOT replicas [rep][col][row], zero on entry
    double* partials;    // 4 streams of per-workgroup partials: chi2, dPA, ddRA, ddDec
    int ldp;             // stride between the streams (>= number of workgroups)
};

__device__ __forceinline__ void lds_add_f64(double* p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Storage T (H and the dataset streams of `a`), arithmetic P with the scalars of `sc` (P = T, sc = a:
// galario's arithmetic in that precision; T = float, P = double: the FP32-bulk form with FP64
// per-point arithmetic).  The reverse mode runs in double either way.  WIN: LDS window capacity
// (cells); a larger window keeps more workgroups off the direct atomics, a smaller one leaves room
// for more workgroups per CU.
template <typename T, typename P, int WIN = GWIN>
__device__ __forceinline__ void sample_grad_sym_body(const SampleArgs<T>& a, const SampleArgs<P>& sc,
                                                     const GradArgs<T>& g, int blk) {
    __shared__ double win[WIN];
    __shared__ double hot[GHOT * GHOT];
    __shared__ double red[4][4];
    __shared__ int box[4][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int i = blk * 256 + tid;
    double c2 = 0.0, gpa = 0.0, gra = 0.0, gde = 0.0;
    double gk[4] = {0.0, 0.0, 0.0, 0.0};         // corners (fl_u, m0), (fl_u, m1), (fu1, m0), (fu1, m1)
    int cu0 = 0, cu1 = 0, cm0 = 0, cm1 = 0;
    bool live = false, oob = false;
    for (int j = tid; j < GHOT * GHOT; j += 256) hot[j] = 0.0;
    if (i < a.nd) {
        const P u0 = P(a.u[i]), v0 = P(a.v[i]);
        const GridPos<P> p = grid_pos<true>(sc, u0, v0);
        oob = p.oob;
        const P q = p.q, t = p.t;
        const int half = a.nrow / 2;
        cu0 = p.fl_u;
        cu1 = p.fu1;
        cm0 = abs(p.fl_v - half);
        cm1 = abs(p.fv1 - half);
        const P y0 = P(a.H[h_index(cu0, cm0, a.ldH)]), y1 = P(a.H[h_index(cu0, cm1, a.ldH)]);
        const P y3 = P(a.H[h_index(cu1, cm0, a.ldH)]), y2 = P(a.H[h_index(cu1, cm1, a.ldH)]);
        const P z = (t * (q * ((y0 + y2) - (y1 + y3))) + (t * (y1 - y0) + q * (y3 - y0))) + y0;
        const P a0 = fabs(y0), a1 = fabs(y1), a2 = fabs(y2), a3 = fabs(y3);
        P amp = a0;
        amp += (a3 - a0) * q;
        amp += (a1 - a0) * t;
        amp += (((a0 - a1) + a2) - a3) * t * q;
        const P sg = copysign(P(1), z);
        P cth = P(1), sth = P(0);
        if (sc.phase) phase_factor(sc, p.urot, p.vrot, &sth, &cth);
        const P V0 = copysign(amp, z);
        const P Vre = V0 * cth, Vim = V0 * sth;
        const P w = P(a.w[i]), ore = P(a.re[i]), oim = P(a.im[i]);
        const P sw = sqrt(w);
        const P cr = (Vre - ore) * sw, ci = (Vim - oim) * sw;
        if constexpr (std::is_same<P, double>::value) c2 = cr * cr + ci * ci;
        else c2 = double(cr) * cr + double(ci) * ci;
        // reverse mode in double
        const double Gre = 2.0 * double(w) * (double(Vre) - double(ore));
        const double Gim = 2.0 * double(w) * (double(Vim) - double(oim));
        const double gA = double(sg) * (Gre * double(cth) + Gim * double(sth));
        const double gth = Gim * double(Vre) - Gre * double(Vim);
        const double dq = double(1 - t) * double(a3 - a0) + double(t) * double(a2 - a1);
        const double dt = double(1 - q) * double(a1 - a0) + double(q) * double(a2 - a3);
        double s_duv;
        if constexpr (std::is_same<P, double>::value) s_duv = p.sign_u * sc.inv_duv;   // (= sign / duv exactly)
        else s_duv = double(p.sign_u) / double(sc.duv);
        gpa = gA * (dq * s_duv * (-double(p.vrot)) + dt * double(sc.v_origin) * s_duv * double(p.urot));
        const double two_pi = 6.283185307179586476925286766559;
        gra = gth * two_pi * double(u0);
        gde = gth * two_pi * double(v0);
        if (!oob) {
            const auto sgn = [](P y) { return double((y > P(0)) - (y < P(0))); };
            gk[0] = gA * (double(1 - t) * double(1 - q)) * sgn(y0);
            gk[1] = gA * (double(t) * double(1 - q)) * sgn(y1);
            gk[2] = gA * (double(1 - t) * double(q)) * sgn(y3);
            gk[3] = gA * (double(t) * double(q)) * sgn(y2);
            live = true;
        } else {
            c2 = gpa = gra = gde = double(NAN);
        }
    }
    if (oob && a.flag) atomicOr(a.flag, 1u);
    // bounding box of the workgroup's cells, columns in pairs (the window follows h_index)
    int b0 = live ? (cu0 >> 1) : INT_MAX, b1 = live ? (cu1 >> 1) : -1;
    int b2 = live ? min(cm0, cm1) : INT_MAX, b3 = live ? max(cm0, cm1) : -1;
    wave_minmax2(b0, b1, b2, b3, lane);
    if (lane == 0) {
        box[wv][0] = b0;
        box[wv][1] = b1;
        box[wv][2] = b2;
        box[wv][3] = b3;
    }
    __syncthreads();
    const int pmin = min(min(box[0][0], box[1][0]), min(box[2][0], box[3][0]));
    const int pmax = max(max(box[0][1], box[1][1]), max(box[2][1], box[3][1]));
    const int mmin = min(min(box[0][2], box[1][2]), min(box[2][2], box[3][2]));
    const int mmax = max(max(box[0][3], box[1][3]), max(box[2][3], box[3][3]));
    const int wp = pmax - pmin + 1, wm = mmax - mmin + 1, w2 = 2 * wm;
    const bool use_win = pmax >= 0 && wp * w2 <= WIN;      // (wp, wm <= 2^13, no overflow)
    const int nwin = use_win ? wp * w2 : 0;
    for (int j = tid; j < nwin; j += 256) win[j] = 0.0;
    __syncthreads();
    if (use_win) {
        // LDS additions, corner by corner.  Neighbouring lanes (points in locality order) often add to the
        // same cell: lanes l and l ^ 1, then l and l ^ 2, with the same cell merge their values into the
        // lower lane first (DPP moves), which halves to quarters the same-address LDS atomics where the
        // points are dense.  key: window offset, or 2^30 + hot-tile offset; -1: nothing to add.
        const int cu[4] = {cu0, cu0, cu1, cu1}, cm[4] = {cm0, cm1, cm0, cm1};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool hotc = cu[k] < GHOT && cm[k] < GHOT;
            int key = (live && gk[k] != 0.0)
                          ? (hotc ? (1 << 30) + cu[k] * GHOT + cm[k]
                                  : ((cu[k] >> 1) - pmin) * w2 + 2 * (cm[k] - mmin) + (cu[k] & 1))
                          : -1;
            double val = gk[k];
            int ko = dpp_i32<DPP_XOR1>(key);
            double vo = dpp_f64<DPP_XOR1>(val);
            if (key >= 0 && key == ko) {
                if (lane & 1) key = -1;
                else val += vo;
            }
            ko = dpp_i32<DPP_XOR2>(key);
            vo = dpp_f64<DPP_XOR2>(val);
            if (key >= 0 && key == ko) {
                if (lane & 2) key = -1;
                else val += vo;
            }
            if (key >= (1 << 30)) lds_add_f64(&hot[key - (1 << 30)], val);
            else if (key >= 0) lds_add_f64(&win[key], val);
        }
    } else {
        // Straight to Hh.  The FP64 atomics are served per 64-byte line and wave instruction, and
        // a point's four corners lie in one 32-byte run of Hh when its lower column is even
        // (h_index), so lane quads issue them together: in round r, lane l carries corner l & 3
        // (gk order) of point (l >> 2) + 16 r of its wave.
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int src = (lane >> 2) + 16 * r, corner = lane & 3;
            const int c_lo = __shfl(cu0, src, 64), c_hi = __shfl(cu1, src, 64);
            const int m_a = __shfl(cm0, src, 64), m_b = __shfl(cm1, src, 64);
            const double g0 = __shfl(gk[0], src, 64), g1 = __shfl(gk[1], src, 64);
            const double g2 = __shfl(gk[2], src, 64), g3 = __shfl(gk[3], src, 64);
            const int col = (corner >> 1) ? c_hi : c_lo;
            const int m = (corner & 1) ? m_b : m_a;
            const double x = corner == 0 ? g0 : corner == 1 ? g1 : corner == 2 ? g2 : g3;
            if (x != 0.0) {
                if (col < GHOT && m < GHOT) lds_add_f64(&hot[col * GHOT + m], x);
                else atomic_add_f64(g.Hh + h_index(col, m, a.ldH), x);
            }
        }
    }
    __syncthreads();
    {
        double* rep = g.hot + size_t(blk % GREP) * (GHOT * GHOT);
        for (int j = tid; j < GHOT * GHOT; j += 256) {
            const double x = hot[j];
            if (x != 0.0) atomic_add_f64(rep + j, x);
        }
        // window cell j = cp * w2 + 2 mm + (column & 1) is Hh offset ((pmin + cp) ldH + mmin) * 2 + j - cp * w2
        // (consecutive j, consecutive addresses); cp from a float reciprocal (exact after one
        // correction, j < 2^24)
        const float inv_w2 = 1.0f / float(w2);
        for (int j = tid; j < nwin; j += 256) {
            const double x = win[j];
            if (x != 0.0) {
                int c = int(float(j) * inv_w2);
                c += (c + 1) * w2 <= j;
                c -= c * w2 > j;
                atomic_add_f64(g.Hh + (size_t(pmin + c) * a.ldH + mmin) * 2 + (j - c * w2), x);
            }
        }
    }
    // per-workgroup partials of chi2, dPA, ddRA, ddDec (fixed order; chi2 in the forward kernels' order,
    // so the gradient call's chi2 equals the forward call's bit for bit)
    c2 = wave_sum(c2);
    wave_sum3(gpa, gra, gde, lane);
    if (lane == 0) {
        red[0][wv] = c2;
        red[1][wv] = gpa;
        red[2][wv] = gra;
        red[3][wv] = gde;
    }
    __syncthreads();
    if (tid < 4) g.partials[size_t(tid) * g.ldp + blk] = (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_sample_grad_sym(SampleArgs<T> a, GradArgs<T> g) {
    sample_grad_sym_body<T, T>(a, a, g, blockIdx.x);
}

// Destinations of the 4 partial streams (chi2, dPA, ddRA, ddDec).
struct Dst4 {
    double* d[4];
};

// ---------------------------------------------------------------------------
// adjoint of the column pass: Hh (two columns per workgroup) -> Th[m][l],
// m = 0..Rs'; Hh is set back to zero after it is read.
// ---------------------------------------------------------------------------
template <typename T>
struct AdjColArgs {
    double* Hh;
    double* hot;         // GREP replicas of the hot tile, added to Hh columns < GHOT, rows < GHOT
    int ldH;
    int N, Rs;           // Rs': outputs m = 0..Rs'
    const cpx<T>* tw;
    T* Th;               // row-major [Rs'+1][ldTh]
    int ldTh;
    const double* part;  // optional: 4 streams of np partials (stride np), each summed into dst by one of
    int np;              // 8 extra workgroups after the nfft FFT workgroups
    Dst4 dst;
    int nfft;
    double rc2;          // > 0: Hh is zero beyond the store radius sqrt(rc2) (stored_rows)
    const cpx<T>* tw512; // W_512^e table (k_adjcol_sub)
};

// Deterministic sum *dst = sum_i p[i] of n partials in the order of
// k_sum_partials (1024 virtual threads, thread t summing p[t], p[t + 1024],
// ...; the 16 virtual waves' sums added in order), so chi2 equals the forward
// call's bit for bit; run by one workgroup of NT threads, NT a multiple of 64:
// real wave w plays virtual waves w, w + NT/64, ... (below 16), lane for lane.
template <int NT>
__device__ __forceinline__ void sum_partials_in_block(const double* __restrict__ p, int n, double* dst,
                                                      double* red16) {
    static_assert(NT % 64 == 0 && NT <= 1024, "whole waves");
    constexpr int NW = NT / 64;
    constexpr int V = (16 + NW - 1) / NW;     // virtual waves per real wave (the last ones may be beyond 16)
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    double acc[V];
#pragma unroll
    for (int q = 0; q < V; ++q) {
        const int vw = wv + q * NW;
        acc[q] = vw < 16 ? sum_stride1024(p, vw * 64 + lane, n, 0.0) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < V; ++q) {
        const int vw = wv + q * NW;
        const double x = wave_sum(acc[q]);
        if (vw < 16 && lane == 0) red16[vw] = x;
    }
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        for (int w = 0; w < 16; ++w) s += red16[w];
        if (dst) *dst = s;
    }
}

__device__ __forceinline__ double mult(int m, int N) { return (m == 0 || m == N / 2) ? 1.0 : 2.0; }
__device__ __forceinline__ double ext_weight(int k, int N) { return (k == 0 || k == N / 2) ? 1.0 : 0.5; }

// PARTIALS: workgroups blk >= nfft sum the sample kernel's partial streams
// (the batched form sums them in its own reduction kernel).
template <typename T, int LOG2N, bool PARTIALS = true>
__device__ __forceinline__ void adjcol_sym_body(const AdjColArgs<T>& a, int blk, int nblk) {
    using S = FftShape<LOG2N>;
    using F = BlockFft<T, LOG2N, false>;
    constexpr int N = S::N, NT = S::NT, E = S::E;
    constexpr int C = N / 2 + 1;
    constexpr int NPAIR = (C + 1) / 2;
    __shared__ T lds[S::LDS_ELEMS];
    if (PARTIALS && blk >= a.nfft) {        // extra workgroups: the partial sums of the sample kernel
        const int j = blk - a.nfft;
        __shared__ double red16[16];
        if (a.part && j < 4) sum_partials_in_block<NT>(a.part + size_t(j) * a.np, a.np, a.dst.d[j], red16);
        return;
    }
    const int pair = xcd_column(blk, a.nfft);
    if (pair >= NPAIR) return;
    const int la = 2 * pair, lb = la + 1;
    const bool has_b = lb < C;
    const int tid = threadIdx.x;
    cpx<double>* hab = reinterpret_cast<cpx<double>*>(a.Hh + h_index(la, 0, a.ldH));   // columns la, lb (h_index)
    const bool hot_cols = la < GHOT;   // GHOT is even: lb < GHOT too
    // rows the sample kernel can have touched (column lb: no more); beyond them Hh, and the
    // replicas of the hot tile, hold zeros
    const int kc = stored_rows(a.rc2, la, N);
    if (kc < 0) {                      // no sample reaches these columns: their adjoint rows are zero
        for (int m = tid; m <= a.Rs; m += NT) {
            T* o = a.Th + size_t(m) * a.ldTh + la;
            o[0] = T(0);
            if (has_b) o[1] = T(0);
        }
        return;
    }
    const auto hot_sum = [&](int l, int p) {
        double s = 0.0;
        for (int r = 0; r < GREP; ++r) s += a.hot[(r * GHOT + l) * GHOT + p];
        return s;
    };
    cpx<T> v[E];
    F::run(v, lds, tid, N / 2, a.tw, [&](int k) -> cpx<T> {
        const int p = k <= N / 2 ? k : N - k;
        const double wk = ext_weight(k, N);
        const cpx<double> x = p <= kc ? hab[p] : cpx<double>{0.0, 0.0};
        double xa = x.re, xb = has_b ? x.im : 0.0;
        if (hot_cols && p < GHOT) {
            xa += hot_sum(la, p);
            if (has_b) xb += hot_sum(lb, p);
        }
        return cpx<T>{T(wk * xa), T(wk * xb)};
    }, a.Rs < NT);        // outputs m <= Rs' only
    // all reads of Hh are done (the FFT synchronised after its first pass): clear it for the next call
    for (int p = tid; p <= kc; p += NT) hab[p] = cpx<double>{0.0, 0.0};   // (slot lb of the last pair: padding)
    if (hot_cols)
        for (int j = tid; j < GREP * GHOT; j += NT) {
            const int r = j / GHOT, p = j % GHOT;
            a.hot[(r * GHOT + la) * GHOT + p] = 0.0;
            if (has_b) a.hot[(r * GHOT + lb) * GHOT + p] = 0.0;
        }
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int m = tid + r * NT;
        if (r * NT <= a.Rs && m <= a.Rs) {
            const T cm = T(mult(m, N));
            T* o = a.Th + size_t(m) * a.ldTh + la;
            o[0] = cm * v[r].re;
            if (has_b) o[1] = cm * v[r].im;
        }
    }
}

template <typename T, int LOG2N>
__global__ __launch_bounds__(FftShape<LOG2N>::NT) void k_adjcol_sym(AdjColArgs<T> a) {
    adjcol_sym_body<T, LOG2N>(a, blockIdx.x, gridDim.x);
}

// Residue form of the adjoint column pass (N >= 1024, profile within 255 pixels: Rs' <= 256; the single
// call's: a shorter path per workgroup and 84 instead of 166 VGPRs, so all of one call's workgroups are
// resident at once, +2-3% on the gradient call; the batched form, bound by throughput, keeps the full
// transform).  Its
// outputs are rows m = 0..Rs' of the N-point transform of the even extension z[k] = w_k x[fold(k)]
// of a column pair of Hh.  With Y_k0 the 512-point transform of z[k0 + P K] (P = N / 512),
//   Z[m] = sum over k0 = 0..P/2 of  W_N^(m k0) Y_k0[m]  +  W_N^(-m k0) Y_k0[512 - m],
// the second term for k0 other than 0 and P/2: it is residue P - k0, whose sequence is residue k0's
// reversed (z is even).  One wave per residue k0: its 512-point transform (ResidueFft at k0 = 0) leaves
// Y[lane + 64 s] in v[s], and Y[512 - m] of m = lane + 64 s sits in lane 64 - lane, slot 7 - s (lane 0:
// its own slot 8 - s).  The column pair is staged in LDS, real parts then imaginary parts, for the
// waves' strided gathers (row p at p + p/8: rows P apart fall in different banks) and cleared as soon
// as it is read; the waves' terms are summed in LDS in residue order.
template <typename T, int LOG2N>
__device__ __forceinline__ void adjcol_sub_body(const AdjColArgs<T>& a, int blk) {
    using CS = ResidueShape<LOG2N>;
    using F = ResidueFft<T>;
    constexpr int N = CS::N, P = CS::P, NT = CS::NT, NK = CS::NK;
    constexpr int C = N / 2 + 1, NPAIR = (C + 1) / 2;
    constexpr int PER = (C + NT - 1) / NT;              // staged rows per thread
    constexpr int MOUT = 260;                           // output rows 0..Rs' (<= 256) per residue
    constexpr size_t STG_BYTES = size_t(N / 2 + N / 16 + 1) * sizeof(double);
    constexpr size_t EXCH_BYTES = size_t(NK) * F::LDS_REALS * sizeof(T);
    constexpr size_t SUM_BYTES = size_t(NK) * MOUT * sizeof(cpx<T>);
    constexpr size_t BYTES = STG_BYTES + EXCH_BYTES > SUM_BYTES ? STG_BYTES + EXCH_BYTES : SUM_BYTES;
    __shared__ __align__(16) unsigned char smem[BYTES];
    double* stg = reinterpret_cast<double*>(smem);
    T* exch = reinterpret_cast<T*>(smem + STG_BYTES);
    cpx<T>* sums = reinterpret_cast<cpx<T>*>(smem);
    const int tid = threadIdx.x;
    if (blk >= a.nfft) {               // extra workgroups: the partial sums of the sample kernel
        const int j = blk - a.nfft;
        __shared__ double red16[16];
        if (a.part && j < 4) sum_partials_in_block<NT>(a.part + size_t(j) * a.np, a.np, a.dst.d[j], red16);
        return;
    }
    const int pair = xcd_column(blk, a.nfft);
    if (pair >= NPAIR) return;
    const int la = 2 * pair, lb = la + 1;
    const bool has_b = lb < C;
    const int lane = tid & 63;
    const int k0 = __builtin_amdgcn_readfirstlane(tid >> 6);
    cpx<double>* hab = reinterpret_cast<cpx<double>*>(a.Hh + h_index(la, 0, a.ldH));   // columns la, lb
    const bool hot_cols = la < GHOT;   // GHOT is even: lb < GHOT too
    const int kc = stored_rows(a.rc2, la, N);
    if (kc < 0) {                      // no sample reaches these columns: their adjoint rows are zero
        for (int m = tid; m <= a.Rs; m += NT) {
            T* o = a.Th + size_t(m) * a.ldTh + la;
            o[0] = T(0);
            if (has_b) o[1] = T(0);
        }
        return;
    }
    // rows 0..kc of the pair (+ the hot tile's replicas), each read by one thread and then cleared
    cpx<double> xs[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int p = tid + q * NT;
        xs[q] = cpx<double>{0.0, 0.0};
        if (p <= kc) {
            xs[q] = hab[p];
            if (hot_cols && p < GHOT) {
                double sa = 0.0, sb = 0.0;
                for (int r = 0; r < GREP; ++r) {
                    sa += a.hot[(r * GHOT + la) * GHOT + p];
                    sb += a.hot[(r * GHOT + lb) * GHOT + p];
                }
                xs[q].re += sa;
                if (has_b) xs[q].im += sb;
            }
            hab[p] = cpx<double>{0.0, 0.0};
        }
        if (p < C) stg[p + (p >> 3)] = xs[q].re;
    }
    __syncthreads();                   // (every read of the hot replicas is done)
    if (hot_cols)
        for (int j = tid; j < GREP * GHOT; j += NT) {
            const int r = j / GHOT, p = j % GHOT;
            a.hot[(r * GHOT + la) * GHOT + p] = 0.0;
            if (has_b) a.hot[(r * GHOT + lb) * GHOT + p] = 0.0;
        }
    cpx<T> v[8];
    int sp[8];
    double wk[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int k = k0 + P * (lane + 64 * r);
        const int p = k <= N / 2 ? k : N - k;
        sp[r] = p + (p >> 3);
        wk[r] = ext_weight(k, N);
        v[r].re = T(wk[r] * stg[sp[r]]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int p = tid + q * NT;
        if (p < C) stg[p + (p >> 3)] = has_b ? xs[q].im : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r].im = T(wk[r] * stg[sp[r]]);
    F::run(v, exch + k0 * F::LDS_REALS, lane, a.tw, a.tw512, P, 0);
    // this residue's terms of rows m = lane + 64 s
    const bool mirror = k0 != 0 && 2 * k0 != P;
    const int src = (64 - lane) & 63;
    cpx<T> c[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        const int m = lane + 64 * s;
        const cpx<T> w = a.tw[(m * k0) & (N - 1)];
        c[s] = cmul(v[s], w);
        if (mirror) {
            const T yr = __shfl(v[7 - s].re, src, 64), yi = __shfl(v[7 - s].im, src, 64);
            const cpx<T> ym = lane == 0 ? v[(8 - s) & 7] : cpx<T>{yr, yi};
            c[s] = cadd(c[s], cmul(ym, cconj(w)));
        }
    }
    __syncthreads();                   // every wave is done with the staged pair and its exchange region
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        const int m = lane + 64 * s;
        if (m <= a.Rs) sums[k0 * MOUT + m] = c[s];
    }
    __syncthreads();
    for (int m = tid; m <= a.Rs; m += NT) {
        cpx<T> acc = sums[m];
#pragma unroll
        for (int k = 1; k < NK; ++k) acc = cadd(acc, sums[k * MOUT + m]);
        const T cm = T(mult(m, N));
        T* o = a.Th + size_t(m) * a.ldTh + la;
        o[0] = cm * acc.re;
        if (has_b) o[1] = cm * acc.im;
    }
}

template <typename T, int LOG2N>
__global__ __launch_bounds__(ResidueShape<LOG2N>::NT) void k_adjcol_sub(AdjColArgs<T> a) {
    adjcol_sub_body<T, LOG2N>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// adjoint of the row pass (two rows per workgroup) fused with the sweep
// adjoint: Th rows -> pixel gradients -> dchi2/dI (atomics) and dchi2/dinc.
// ---------------------------------------------------------------------------
template <typename T>
struct AdjRowArgs {
    SweepPrm<T> sw;
    const T* Th;
    int ldTh;
    int N, Rs, nrows;    // pixels j = 0..Rs', rows m = 0..nrows-1
    const cpx<T>* tw;
    double tan_inc;
    double* gI;          // nr, zeroed by the forward row pass
    double* ginc;        // 1, zeroed by the forward row pass
    const cpx<T>* tw512; // W_512^e table (k_adjrow_sub)
};

// dchi2/d(pixel value) = gz: accumulate into dI (workgroup-local copy in LDS)
// and return the inc term.
template <typename T>
__device__ __forceinline__ double sweep_pixel_adjoint(int xi, int yi, double gz, const SweepPrm<T>& p, double tan_inc,
                                                      double* gI) {
    if (gz == 0.0) return 0.0;
    const T x = T(xi) * p.dxy;
    const T y = T(yi) * p.dxy;
    T r;
    double xs2;
    if constexpr (std::is_same<T, double>::value) {
        const double xs = x / p.cos_inc;
        xs2 = xs * xs;
        r = sqrt(xs * xs + y * y);
    } else {
        const double xs = double(x / p.cos_inc);
        xs2 = xs * xs;
        r = float(sqrt(xs * xs + double(y) * double(y)));
    }
    T fi = floor((r - p.Rmin) / p.dR);
    fi = fi > T(0) ? fi : T(0);
    if (!(fi <= T(p.nr - 2))) return 0.0;
    const int iR = int(fi);
    const double alpha = (double(r) - double(iR) * double(p.dR) - double(p.Rmin)) / double(p.dR);
    const double sr = double(p.sr_to_px);
    lds_add_f64(gI + iR, gz * sr * (1.0 - alpha));
    lds_add_f64(gI + iR + 1, gz * sr * alpha);
    if (r == T(0)) return 0.0;
    const double dvdr = sr * (double(p.I[iR + 1]) - double(p.I[iR])) / double(p.dR);
    return gz * dvdr * xs2 * tan_inc / double(r);
}

// Bins i = t, t + nt, ... of the central pixel's adjoint (all threads t < nt
// of a workgroup together cover bins 0..iIN+1).
template <typename T>
__device__ __forceinline__ void central_adjoint(double gz, const SweepPrm<T>& p, double* gI, int t, int nt) {
    const int iIN = p.iIN;
    const double Rmin = p.Rmin, dR = p.dR, h = double(p.dxy) / 2.0;
    const double area = h * h - Rmin * Rmin;
    const double sc = gz * double(p.sr_to_px) / area;
    const double RiIN = Rmin + iIN * dR;
    const double delta = h - RiIN, beta = delta / dR;
    for (int i = t; i <= iIN + 1; i += nt) {
        if (i == 0) lds_add_f64(gI + 0, sc * dR * Rmin);
        if (i >= 1 && i < iIN) lds_add_f64(gI + i, sc * 2.0 * dR * (Rmin + i * dR));
        if (i == iIN) lds_add_f64(gI + iIN, sc * (dR * RiIN + delta * (RiIN + h * (1.0 - beta))));
        if (i == iIN + 1) lds_add_f64(gI + iIN + 1, sc * delta * h * beta);
    }
}

constexpr int ADJROW_LDS_NR = 2048;   // profile bins accumulated in LDS by the adjoint row pass

// SPLIT: as in rowpass_sym_body (the batched form uses the half-LDS exchange).
template <typename T, int LOG2N, bool SPLIT = false>
__device__ __forceinline__ void adjrow_sym_body(const AdjRowArgs<T>& a, int blk) {
    using S = FftShape<LOG2N>;
    using F = BlockFft<T, LOG2N, false, SPLIT>;
    constexpr int N = S::N, NT = S::NT, E = S::E;
    __shared__ T lds[F::LDS_REALS];
    __shared__ double red[16];
    __shared__ double gIl[ADJROW_LDS_NR];    // this workgroup's share of dchi2/dI (nr <= ADJROW_LDS_NR)
    const int tid = threadIdx.x;
    for (int k = tid; k < a.sw.nr; k += NT) gIl[k] = 0.0;   // ordered before use by the FFT's barriers
    const int ma = 2 * blk, mb = ma + 1;
    const bool hb = mb < a.nrows;
    const T* ra = a.Th + size_t(ma) * a.ldTh;
    const T* rb = ra + a.ldTh;
    cpx<T> v[E];
    F::run(v, lds, tid, N / 2, a.tw, [&](int k) -> cpx<T> {
        const int l = k <= N / 2 ? k : N - k;
        const T wk = T(ext_weight(k, N));
        return cpx<T>{wk * ra[l], hb ? wk * rb[l] : T(0)};
    }, a.Rs < NT);        // pixels j <= Rs' only
    __shared__ double gcen;
    double ginc = 0.0;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int j = tid + r * NT;
        if (r * NT <= a.Rs && j <= a.Rs) {
            const double cj = mult(j, N);
            const double gza = cj * double(v[r].re), gzb = cj * double(v[r].im);
            if (ma == 0 && j == 0) gcen = gza;      // the central pixel: spread over the workgroup below
            else ginc += sweep_pixel_adjoint(j, ma, gza, a.sw, a.tan_inc, gIl);
            if (hb) ginc += sweep_pixel_adjoint(j, mb, gzb, a.sw, a.tan_inc, gIl);
        }
    }
    if (ma == 0) {                                 // (uniform per workgroup)
        __syncthreads();
        central_adjoint(gcen, a.sw, gIl, tid, NT);
    }
    const double b = block_sum_any(ginc, red);    // (its barrier also orders the LDS sums before the flush)
    if (tid == 0 && b != 0.0) atomic_add_f64(a.ginc, b);
    for (int k = tid; k < a.sw.nr; k += NT) {      // consecutive bins: coalesced atomics
        const double x = gIl[k];
        if (x != 0.0) atomic_add_f64(a.gI + k, x);
    }
}

template <typename T, int LOG2N>
__global__ __launch_bounds__(FftShape<LOG2N>::NT) void k_adjrow_sym(AdjRowArgs<T> a) {
    adjrow_sym_body<T, LOG2N>(a, blockIdx.x);
}

// Residue form of the adjoint row pass (N >= 1024, Rs' <= 256).  Its outputs are the pixels
// j = 0..Rs' of the N-point transform of the weighted even extension of two Th rows (real and
// imaginary parts), so it is the residue adjoint column transform over rows instead of column pairs:
// the rows are staged in LDS (real, then imaginary parts), one wave per residue class k0 = 0..P/2 of
// the input frequencies takes the 512-point transform of its decimated sequence, and the waves'
// terms W_N^(j k0) Y_k0[j] (with the mirror term of residue P - k0) are summed in LDS in residue
// order.  The pixel adjoints, the central pixel and the dI flush are those of adjrow_sym_body.  Its
// 5 waves per workgroup (N = 4096) run their transforms without workgroup barriers, which shortens
// the single workgroup path that the pass's ~100 workgroups take.
template <typename T, int LOG2N>
__device__ __forceinline__ void adjrow_sub_body(const AdjRowArgs<T>& a, int blk) {
    using CS = ResidueShape<LOG2N>;
    using F = ResidueFft<T>;
    constexpr int N = CS::N, P = CS::P, NT = CS::NT, NK = CS::NK;
    constexpr int C = N / 2 + 1;
    constexpr int PER = (C + NT - 1) / NT;              // staged frequencies per thread
    constexpr int MOUT = 260;                           // pixels 0..Rs' (<= 256) per residue
    constexpr size_t STG_BYTES = size_t(N / 2 + N / 16 + 1) * sizeof(T);
    constexpr size_t EXCH_BYTES = size_t(NK) * F::LDS_REALS * sizeof(T);
    constexpr size_t SUM_BYTES = size_t(NK) * MOUT * sizeof(cpx<T>);
    constexpr size_t BYTES = STG_BYTES + EXCH_BYTES > SUM_BYTES ? STG_BYTES + EXCH_BYTES : SUM_BYTES;
    __shared__ __align__(16) unsigned char smem[BYTES];
    __shared__ double gIl[ADJROW_LDS_NR];    // this workgroup's share of dchi2/dI (nr <= ADJROW_LDS_NR)
    __shared__ double red[16];
    __shared__ double gcen;
    T* stg = reinterpret_cast<T*>(smem);
    T* exch = reinterpret_cast<T*>(smem + STG_BYTES);
    cpx<T>* sums = reinterpret_cast<cpx<T>*>(smem);
    const int tid = threadIdx.x, lane = tid & 63;
    const int k0 = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int k = tid; k < a.sw.nr; k += NT) gIl[k] = 0.0;   // ordered before use by the barriers below
    const int ma = 2 * blk, mb = ma + 1;
    const bool hb = mb < a.nrows;
    const T* ra = a.Th + size_t(ma) * a.ldTh;
    const T* rb = ra + a.ldTh;
    // row ma staged (frequency p at p + p/8: the waves' stride-P gathers spread over the banks), row mb
    // held in registers until the first gather is done
    T xb[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int p = tid + q * NT;
        xb[q] = (hb && p < C) ? rb[p] : T(0);
        if (p < C) stg[p + (p >> 3)] = ra[p];
    }
    __syncthreads();
    cpx<T> v[8];
    int sp[8];
    T wk[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int k = k0 + P * (lane + 64 * r);
        const int p = k <= N / 2 ? k : N - k;
        sp[r] = p + (p >> 3);
        wk[r] = T(ext_weight(k, N));
        v[r].re = wk[r] * stg[sp[r]];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int p = tid + q * NT;
        if (p < C) stg[p + (p >> 3)] = xb[q];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r].im = wk[r] * stg[sp[r]];
    F::run(v, exch + k0 * F::LDS_REALS, lane, a.tw, a.tw512, P, 0);
    // this residue's terms of pixels j = lane + 64 s
    const bool mirror = k0 != 0 && 2 * k0 != P;
    const int src = (64 - lane) & 63;
    cpx<T> c[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        const int j = lane + 64 * s;
        const cpx<T> w = a.tw[(j * k0) & (N - 1)];
        c[s] = cmul(v[s], w);
        if (mirror) {
            const T yr = __shfl(v[7 - s].re, src, 64), yi = __shfl(v[7 - s].im, src, 64);
            const cpx<T> ym = lane == 0 ? v[(8 - s) & 7] : cpx<T>{yr, yi};
            c[s] = cadd(c[s], cmul(ym, cconj(w)));
        }
    }
    __syncthreads();                   // every wave is done with the staged rows and its exchange region
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        const int j = lane + 64 * s;
        if (j <= a.Rs) sums[k0 * MOUT + j] = c[s];
    }
    __syncthreads();
    double ginc = 0.0;
    for (int j = tid; j <= a.Rs; j += NT) {
        cpx<T> acc = sums[j];
#pragma unroll
        for (int k = 1; k < NK; ++k) acc = cadd(acc, sums[k * MOUT + j]);
        const double cj = mult(j, N);
        const double gza = cj * double(acc.re), gzb = cj * double(acc.im);
        if (ma == 0 && j == 0) gcen = gza;      // the central pixel: spread over the workgroup below
        else ginc += sweep_pixel_adjoint(j, ma, gza, a.sw, a.tan_inc, gIl);
        if (hb) ginc += sweep_pixel_adjoint(j, mb, gzb, a.sw, a.tan_inc, gIl);
    }
    if (ma == 0) {                                 // (uniform per workgroup)
        __syncthreads();
        central_adjoint(gcen, a.sw, gIl, tid, NT);
    }
    const double b = block_sum_any(ginc, red);    // (its barrier also orders the LDS sums before the flush)
    if (tid == 0 && b != 0.0) atomic_add_f64(a.ginc, b);
    for (int k = tid; k < a.sw.nr; k += NT) {      // consecutive bins: coalesced atomics
        const double x = gIl[k];
        if (x != 0.0) atomic_add_f64(a.gI + k, x);
    }
}

template <typename T, int LOG2N>
__global__ __launch_bounds__(ResidueShape<LOG2N>::NT) void k_adjrow_sub(AdjRowArgs<T> a) {
    adjrow_sub_body<T, LOG2N>(a, blockIdx.x);
}

}  // namespace gl
