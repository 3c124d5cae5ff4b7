// Batched chi2Profile: many walkers' parameter sets (profile, inc, PA, dRA,
// dDec) over several datasets.  galario has no batched form (its callers loop
// over chi2Profile calls, one per walker and dataset); these kernels run the
// same per-call arithmetic with the walker taken from blockIdx.y:
//
//   * a setup kernel computes each walker's scalars from its (inc, PA, dRA,
//     dDec) with the formulas the single-call path evaluates on the host
//     (create_image_h's cos(inc), uv_rotate_h's cos/sin(PA) and rotated
//     offsets, apply_phase_sampled_h's skip test and 2 pi scaling);
//   * the row pass, column pass, sample and adjoint kernels are the bodies of
//     the single-call kernels with per-walker pointers (a slot of the chunk
//     of walkers being processed holds that walker's rows and spectrum);
//   * per-workgroup partials land in [walker][stream][block] and one
//     workgroup per walker sums them in the fixed order of the single-call
//     reduction, dataset by dataset.
#pragma once

#include "gl_adjoint.hpp"

namespace gl {

template <typename T>
struct WalkerPrm {
    T cos_inc;
    int rotate;
    T cosPA, sinPA;
    int phase;
    T dRA2, dDec2;       // 2 pi dRArot, 2 pi dDecrot
    T dRAt, dDect;       // 2 dRArot, 2 dDecrot (the phase in half turns)
    double tan_inc;
    // the same rotation and phase scalars in double (the FP32-bulk form's FP64 per-point arithmetic)
    double cosPAd, sinPAd, dRA2d, dDec2d, dRAtd, dDectd;
    int phased;
};

// geom: [W][4] = inc, PA, dRA, dDec, of type G (T, or double for the FP32-bulk form with FP64
// geometry).  The T scalars follow the single call's host setup in T (setup_profile,
// setup_sample); the double ones are the FP64 single call's, from the G values.  Also clears the
// gradient accumulators gI[0..W*nr) and ggeom[0..4W) when given.
template <typename T, typename G>
__global__ __launch_bounds__(256) void k_batch_setup(int W, const G* __restrict__ geom, WalkerPrm<T>* wp, double* gI,
                                                     int nr, double* ggeom) {
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < size_t(W)) {
        const T PA = T(geom[4 * i + 1]), dRA = T(geom[4 * i + 2]), dDec = T(geom[4 * i + 3]);
        WalkerPrm<T> p;
        p.cos_inc = T(cos(double(geom[4 * i])));
        p.tan_inc = tan(double(geom[4 * i]));
        T dRArot, dDecrot;
        if (PA == T(0)) {
            p.rotate = 0;
            p.cosPA = T(1);
            p.sinPA = T(0);
            dRArot = dRA;
            dDecrot = dDec;
        } else {
            p.rotate = 1;
            p.cosPA = T(cos(double(PA)));
            p.sinPA = T(sin(double(PA)));
            dRArot = dRA * p.cosPA - dDec * p.sinPA;
            dDecrot = dRA * p.sinPA + dDec * p.cosPA;
        }
        p.phase = !(dRArot == T(0) && dDecrot == T(0));
        p.dRA2 = T(double(dRArot) * (2. * double(T(M_PI))));
        p.dDec2 = T(double(dDecrot) * (2. * double(T(M_PI))));
        p.dRAt = T(2) * dRArot;
        p.dDect = T(2) * dDecrot;
        const double PAd = double(geom[4 * i + 1]), dRAd = double(geom[4 * i + 2]), dDecd = double(geom[4 * i + 3]);
        double dRArotd = dRAd, dDecrotd = dDecd;
        p.cosPAd = 1.0;
        p.sinPAd = 0.0;
        if (PAd != 0.0) {
            p.cosPAd = cos(PAd);
            p.sinPAd = sin(PAd);
            dRArotd = dRAd * p.cosPAd - dDecd * p.sinPAd;
            dDecrotd = dRAd * p.sinPAd + dDecd * p.cosPAd;
        }
        p.phased = !(dRArotd == 0.0 && dDecrotd == 0.0);
        p.dRA2d = dRArotd * (2. * M_PI);
        p.dDec2d = dDecrotd * (2. * M_PI);
        p.dRAtd = 2.0 * dRArotd;
        p.dDectd = 2.0 * dDecrotd;
        wp[i] = p;
    }
    if (gI && i < size_t(W) * size_t(nr)) gI[i] = 0.0;
    if (ggeom && i < size_t(W) * 4) ggeom[i] = 0.0;
}

// Per-walker pointers: walker w = w0 + blockIdx.y uses slot blockIdx.y.
template <typename T>
struct BatchPtrs {
    const T* I;              // profiles [W][nr]
    int nr;
    const WalkerPrm<T>* wp;  // [W]
    int w0;
    T* Ts;                   // rows of each slot (stride sTs)
    size_t sTs;
    T* H;                    // half spectrum of each slot (stride sH)
    size_t sH;
    double* Hh;              // adjoint half spectrum of each slot (stride sHh), zero between calls
    size_t sHh;
    size_t hot_off;          // hot-tile replicas at Hh + slot * sHh + hot_off
    T* Th;                   // adjoint rows of each slot (stride sTh)
    size_t sTh;
    double* gI;              // [W][nr]
    double* ggeom;           // [W][4]; the inc entry is accumulated here
    double* part;            // [W][4][ldPart]
    int ldPart;
    int partOff;             // first block of the dataset being sampled
};

template <typename T, int LOG2N>
__global__ __launch_bounds__(FftShape<LOG2N>::NT) void k_rowpass_sym_b(SymRowArgs<T> a, BatchPtrs<T> b) {
    const int slot = blockIdx.y, w = b.w0 + slot;
    a.sw.I = b.I + size_t(w) * b.nr;
    a.sw.cos_inc = b.wp[w].cos_inc;
    a.Ts = b.Ts + slot * b.sTs;
    a.zero = nullptr;
    a.zero1 = nullptr;
    rowpass_sym_body<T, LOG2N, true>(a, blockIdx.x);   // many walkers: half-LDS exchange
}

template <typename T, int LOG2N>
__global__ __launch_bounds__(FftShape<LOG2N>::NT) void k_colpass_sym_b(SymColArgs<T> a, BatchPtrs<T> b) {
    const int slot = blockIdx.y;
    a.Ts = b.Ts + slot * b.sTs;
    a.H = b.H + slot * b.sH;
    colpass_sym_body<T, LOG2N>(a, blockIdx.x, gridDim.x);
}

// Many walkers fill the chip: the half-LDS transpose and, at N = 4096, at most 72 VGPRs (7 waves per
// SIMD) let five 5-wave workgroups share a CU (+5% on the batched forward against four).
template <typename T, int LOG2N>
__global__ __launch_bounds__(ResidueShape<LOG2N>::NT)
__attribute__((amdgpu_waves_per_eu(LOG2N == 12 ? 7 : 1))) void k_colpass_sub_b(SymColArgs<T> a, BatchPtrs<T> b) {
    const int slot = blockIdx.y;
    a.Ts = b.Ts + slot * b.sTs;
    a.H = b.H + slot * b.sH;
    colpass_sub_body<T, LOG2N, true>(a, xcd_column(blockIdx.x, gridDim.x));
}

template <typename T>
__device__ __forceinline__ void walker_sample_args(SampleArgs<T>& a, const BatchPtrs<T>& b, int slot, int w) {
    const WalkerPrm<T> p = b.wp[w];
    a.H = b.H + slot * b.sH;
    a.rotate = p.rotate;
    a.cosPA = p.cosPA;
    a.sinPA = p.sinPA;
    a.phase = p.phase;
    a.dRA2 = p.dRA2;
    a.dDec2 = p.dDec2;
    a.dRAt = p.dRAt;
    a.dDect = p.dDect;
}

// Scalars of the FP64 per-point arithmetic for walker w on the dataset of `a` (only the fields
// grid_pos and apply_phase_pt read).
template <typename T>
__device__ __forceinline__ SampleArgs<double> walker_scalars_f64(const SampleArgs<T>& a, const BatchPtrs<T>& b,
                                                                 int w) {
    const WalkerPrm<T> p = b.wp[w];
    SampleArgs<double> sc{};
    sc.nrow = a.nrow;
    sc.ncol = a.ncol;
    sc.v_origin = double(a.v_origin);
    sc.duv = a.duv_d;
    sc.inv_duv = a.inv_duv_d;
    sc.rho2 = a.rho2;
    sc.rotate = p.rotate;
    sc.cosPA = p.cosPAd;
    sc.sinPA = p.sinPAd;
    sc.phase = p.phased;
    sc.dRA2 = p.dRA2d;
    sc.dDec2 = p.dDec2d;
    sc.dRAt = p.dRAtd;
    sc.dDect = p.dDectd;
    return sc;
}

// KW walkers per workgroup on the same points: the dataset streams (u, v, Re,
// Im, w) are read once for KW walkers.  PPT blocks of 256 points per workgroup
// (more loads in flight per wave); each block's partial per walker is the same
// fixed-order sum as in the one-walker kernel, whatever KW and PPT.
//
// MIX (T = float only): FP32 storage with the position, interpolation, phase and residual in FP64
// (walker_scalars_f64).
template <typename T, int KW, bool MIX = false, int PPT = 1>
__global__ __launch_bounds__(256) void k_sample_multi_b(SampleArgs<T> a, BatchPtrs<T> b, int nw, int nblocks) {
    using P = std::conditional_t<MIX, double, T>;
    __shared__ double red[PPT][KW][4];
    const int s0 = blockIdx.y * KW;
    double acc[PPT][KW];
    bool oob_any = false;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
#pragma unroll
        for (int k = 0; k < KW; ++k) acc[q][k] = 0.0;
        const int i = (blockIdx.x * PPT + q) * 256 + threadIdx.x;
        if (i < a.nd) {
            const P u0 = P(a.u[i]), v0 = P(a.v[i]);
            const P re = P(a.re[i]), im = P(a.im[i]);
            const P sw = sqrt(P(a.w[i]));
#pragma unroll
            for (int k = 0; k < KW; ++k) {
                const int slot = s0 + k;
                if (slot < nw) {
                    SampleArgs<T> ak = a;
                    walker_sample_args(ak, b, slot, b.w0 + slot);
                    bool oob = false;
                    cpx<P> V;
                    if constexpr (MIX)
                        V = sample_point_sym2<T, double>(ak, walker_scalars_f64(a, b, b.w0 + slot), u0, v0, oob);
                    else
                        V = sample_point_sym(ak, u0, v0, oob);
                    oob_any |= oob;
                    const P cr = (V.re - re) * sw;
                    const P ci = (V.im - im) * sw;
                    if constexpr (std::is_same<P, double>::value) acc[q][k] = cr * cr + ci * ci;
                    else acc[q][k] = double(cr) * cr + double(ci) * ci;
                }
            }
        }
    }
    if (oob_any && a.flag) atomicOr(a.flag, 1u);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
#pragma unroll
        for (int k = 0; k < KW; ++k) {
            const double x = wave_sum(acc[q][k]);
            if (lane == 0) red[q][k][wv] = x;
        }
    }
    __syncthreads();
    if (threadIdx.x < PPT * KW) {
        const int q = threadIdx.x / KW, k = threadIdx.x % KW;
        const int blk = blockIdx.x * PPT + q;
        if (s0 + k < nw && blk < nblocks)
            b.part[size_t(b.w0 + s0 + k) * 4 * b.ldPart + b.partOff + blk] =
                (red[q][k][0] + red[q][k][1]) + (red[q][k][2] + red[q][k][3]);
    }
}

template <typename T, bool MIX = false, int WIN = GWIN>
__global__ __launch_bounds__(256) void k_sample_grad_sym_b(SampleArgs<T> a, BatchPtrs<T> b) {
    const int slot = blockIdx.y, w = b.w0 + slot;
    walker_sample_args(a, b, slot, w);
    double* hh = b.Hh + slot * b.sHh;
    const GradArgs<T> g{hh, hh + b.hot_off, b.part + size_t(w) * 4 * b.ldPart + b.partOff, b.ldPart};
    if constexpr (MIX) sample_grad_sym_body<T, double, WIN>(a, walker_scalars_f64(a, b, w), g, blockIdx.x);
    else sample_grad_sym_body<T, T, WIN>(a, a, g, blockIdx.x);
}

template <typename T, int LOG2N>
__global__ __launch_bounds__(FftShape<LOG2N>::NT) void k_adjcol_sym_b(AdjColArgs<T> a, BatchPtrs<T> b) {
    const int slot = blockIdx.y;
    a.Hh = b.Hh + slot * b.sHh;
    a.hot = a.Hh + b.hot_off;
    a.Th = b.Th + slot * b.sTh;
    adjcol_sym_body<T, LOG2N, false>(a, blockIdx.x, gridDim.x);   // partials: k_batch_reduce
}

template <typename T, int LOG2N>
__global__ __launch_bounds__(FftShape<LOG2N>::NT) void k_adjrow_sym_b(AdjRowArgs<T> a, BatchPtrs<T> b) {
    const int slot = blockIdx.y, w = b.w0 + slot;
    a.sw.I = b.I + size_t(w) * b.nr;
    a.sw.cos_inc = b.wp[w].cos_inc;
    a.Th = b.Th + slot * b.sTh;
    a.tan_inc = b.wp[w].tan_inc;
    a.gI = b.gI + size_t(w) * b.nr;
    a.ginc = b.ggeom + size_t(w) * 4;
    adjrow_sym_body<T, LOG2N, true>(a, blockIdx.x);    // many walkers: half-LDS exchange
}

constexpr int BATCH_MAX_SETS = 32;

struct BatchSets {
    int n;
    int off[BATCH_MAX_SETS];   // first block of each dataset in a partial stream
    int nb[BATCH_MAX_SETS];    // blocks of each dataset
};

// One workgroup per walker: chi2 of each dataset (the single-call
// k_sum_partials order), their sum over datasets, and with nstream = 4 the
// PA, dRA, dDec gradients (the same order per dataset, summed over
// datasets).
__global__ __launch_bounds__(1024) void k_batch_reduce(const double* __restrict__ part, int ldPart, BatchSets sets,
                                                       int nstream, double* chi2_sets, double* chi2_tot,
                                                       double* ggeom) {
    __shared__ double red[16];
    const int w = blockIdx.x;
    double tot[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < sets.n; ++s) {
        for (int k = 0; k < nstream; ++k) {
            const double* p = part + (size_t(w) * 4 + k) * ldPart + sets.off[s];
            double acc = sum_stride1024(p, threadIdx.x, sets.nb[s], 0.0);
            acc = wave_sum(acc);
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
            __syncthreads();
            if (threadIdx.x == 0) {
                double v = 0.0;
                for (int q = 0; q < 16; ++q) v += red[q];
                if (k == 0 && chi2_sets) chi2_sets[size_t(w) * sets.n + s] = v;
                tot[k] += v;
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        if (chi2_tot) chi2_tot[w] = tot[0];
        if (ggeom && nstream == 4) {
            ggeom[size_t(w) * 4 + 1] = tot[1];
            ggeom[size_t(w) * 4 + 2] = tot[2];
            ggeom[size_t(w) * 4 + 3] = tot[3];
        }
    }
}

}  // namespace gl
