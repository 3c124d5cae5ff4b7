// Pipeline kernels: profile/image -> shifted R2C spectrum -> sampled visibilities -> chi2.
//
// Data layout in HBM (one evaluation):
//   Tm  : row-transformed image, only the 2*Rs rows that can be non-zero,
//         row-major [2*Rs][ldT] complex (ldT >= N/2+1).  Compact row kk maps
//         to quadrant-shifted image row kk (kk < Rs) or N-2*Rs+kk (kk >= Rs).
//   S   : shifted spectrum, column-major [N/2+1][ldS] complex: element
//         (v-row r, u-column l) at l*ldS + r, r = (k + N/2) mod N for FFT row
//         k (the reference's fftshift_axis0 folded into the stores).
//         The four interpolation corners are then two 32-byte pairs.
//
// All stages that restate galario arithmetic (sweep, central pixel,
// interpolation, phase, chi2 terms) are compiled without FMA contraction so
// they round exactly as the reference's -std=c++11 x86-64 build.
#pragma once

#include <type_traits>

#include "gl_common.hpp"
#include "gl_fft.hpp"

namespace gl {

enum { SRC_SWEEP = 0, SRC_IMAGE_SHIFT = 1, SRC_IMAGE_PLAIN = 2 };
enum { OUT_COLMAJOR_SHIFT = 0, OUT_ROWMAJOR_PLAIN = 1 };
enum { MODE_VIS = 0, MODE_CHI2 = 1 };

// ---------------------------------------------------------------------------
// sweep (galario.cpp:1048-1071) and central pixel (galario.cpp:1214-1232)
// ---------------------------------------------------------------------------
template <typename T>
struct SweepPrm {
    const T* I;        // radial profile, nr values (device)
    int nr;
    T Rmin, dR, dxy, cos_inc, sr_to_px;
    int iIN;           // floor((dxy/2 - Rmin)/dR), validated on the host
    int rmax;          // half side of the swept box (0: only the central pixel)
};

// Pixel value at x = xi*dxy, y = yi*dxy (xi = N/2 - column, yi = N/2 - row).
template <typename T>
__device__ __forceinline__ T sweep_pixel(int xi, int yi, const SweepPrm<T>& p) {
    const T x = T(xi) * p.dxy;
    const T y = T(yi) * p.dxy;
    T r;
    if constexpr (std::is_same<T, double>::value) {
        const double xs = x / p.cos_inc;
        r = sqrt(xs * xs + y * y);
    } else {
        // pow(float, 2) and the sum are double in the reference; r is stored as float
        const double xs = double(x / p.cos_inc);
        r = float(sqrt(xs * xs + double(y) * double(y)));
    }
    T fi = floor((r - p.Rmin) / p.dR);
    fi = fi > T(0) ? fi : T(0);
    if (!(fi <= T(p.nr - 2))) return T(0);
    const int iR = int(fi);
    const T I0 = p.I[iR], I1 = p.I[iR + 1];
    return p.sr_to_px * (I0 + (r - T(iR) * p.dR - p.Rmin) * (I1 - I0) / p.dR);
}

// Central pixel (create_image_d's central value, galario.cpp:1135-1154): the
// profile's flux inside the pixel's inscribed disc over its area.  I(i)
// returns profile value i; it is read for i = 0..iIN+1.
template <typename T, class Prof>
__device__ __forceinline__ T central_value_of(const SweepPrm<T>& p, Prof I) {
    const int iIN = p.iIN;
    const T Rmin = p.Rmin, dR = p.dR;
    T flux = T(0);
    for (int i = 1; i < iIN; ++i) flux += (Rmin + dR * T(i)) * I(i);
    flux = T(double(flux) * 2.0);
    flux += Rmin * I(0) + (Rmin + T(iIN) * dR) * I(iIN);
    flux *= dR;
    const double half = double(p.dxy) / 2.0;
    if constexpr (std::is_same<T, double>::value) {
        const double I_interp = (I(iIN + 1) - I(iIN)) / dR * (half - (Rmin + dR * iIN)) + I(iIN);
        flux += ((Rmin + iIN * dR) * I(iIN) + half * I_interp) * (half - (Rmin + iIN * dR));
        const double area = half * half - Rmin * Rmin;
        return p.sr_to_px * flux / area;
    } else {
        const float edge = Rmin + dR * float(iIN);
        const float I_interp = float(double((I(iIN + 1) - I(iIN)) / dR) * (half - double(edge)) + double(I(iIN)));
        const double term = (double(edge * I(iIN)) + half * double(I_interp)) * (half - double(edge));
        flux = float(double(flux) + term);
        const float area = float(half * half - double(Rmin) * double(Rmin));
        return p.sr_to_px * flux / area;
    }
}

template <typename T>
__device__ __noinline__ T central_value(const SweepPrm<T>& p) {
    return central_value_of(p, [&](int i) { return p.I[i]; });
}

// Value of quadrant-shifted image pixel (row k, column n) built from a profile.
template <typename T>
__device__ __forceinline__ T sweep_shifted(int k, int n, int N, const SweepPrm<T>& p) {
    if (k == 0 && n == 0) return central_value(p);
    if (p.rmax == 0) return T(0);
    const int xi = n < N / 2 ? -n : N - n;
    const int yi = k < N / 2 ? -k : N - k;
    return sweep_pixel(xi, yi, p);
}

// ---------------------------------------------------------------------------
// row pass: 2 real rows per workgroup packed as one complex N-point FFT
// ---------------------------------------------------------------------------
template <typename T>
struct RowArgs {
    SweepPrm<T> sw;              // SRC_SWEEP
    const T* img;                // SRC_IMAGE_*: N rows of ldimg reals (image in the first N)
    int ldimg;
    int N, Rs, ldT;
    const cpx<T>* tw;            // W_N^e table
    cpx<T>* Tm;                  // out: [2*Rs][ldT]
};

__device__ __forceinline__ int compact_to_row(int kk, int Rs, int N) { return kk < Rs ? kk : kk + N - 2 * Rs; }

template <typename T, int SRC>
__device__ __forceinline__ T row_source(const RowArgs<T>& a, int k, int n) {
    const int N = a.N;
    if constexpr (SRC == SRC_SWEEP) {
        return sweep_shifted(k, n, N, a.sw);
    } else if constexpr (SRC == SRC_IMAGE_SHIFT) {
        const int i = k + N / 2 < N ? k + N / 2 : k - N / 2;
        const int j = n + N / 2 < N ? n + N / 2 : n - N / 2;
        return a.img[size_t(i) * a.ldimg + j];
    } else {
        return a.img[size_t(k) * a.ldimg + n];
    }
}

template <typename T, int LOG2N, int SRC>
__global__ __launch_bounds__(FftShape<LOG2N>::NT) void k_rowpass(RowArgs<T> a) {
    using S = FftShape<LOG2N>;
    using F = BlockFft<T, LOG2N, false>;
    constexpr int N = S::N, NT = S::NT, E = S::E;
    __shared__ T lds[S::LDS_ELEMS];
    const int tid = threadIdx.x;
    const int kk_a = 2 * blockIdx.x, kk_b = kk_a + 1;
    const int ka = compact_to_row(kk_a, a.Rs, N), kb = compact_to_row(kk_b, a.Rs, N);
    cpx<T> v[E];
    F::run(v, lds, tid, a.Rs, a.tw, [&](int n) -> cpx<T> {
        return cpx<T>{row_source<T, SRC>(a, ka, n), row_source<T, SRC>(a, kb, n)};
    });
    // X_a[l] = (Z[l] + conj Z[N-l]) / 2, X_b[l] = (Z[l] - conj Z[N-l]) / 2i, one component at a time
    const int C = N / 2 + 1;
    cpx<T>* outa = a.Tm + size_t(kk_a) * a.ldT;
    cpx<T>* outb = a.Tm + size_t(kk_b) * a.ldT;
#pragma unroll
    for (int r = 0; r < E; ++r) lds[lpad(tid + r * NT)] = v[r].re;
    __syncthreads();
    for (int l = tid; l < C; l += NT) {
        const T zr = lds[lpad(l)], wr = lds[lpad((N - l) & (N - 1))];
        outa[l].re = T(0.5) * (zr + wr);
        outb[l].im = T(0.5) * (wr - zr);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; ++r) lds[lpad(tid + r * NT)] = v[r].im;
    __syncthreads();
    for (int l = tid; l < C; l += NT) {
        const T zi = lds[lpad(l)], wi = lds[lpad((N - l) & (N - 1))];
        outa[l].im = T(0.5) * (zi - wi);
        outb[l].re = T(0.5) * (zi + wi);
    }
}

// ---------------------------------------------------------------------------
// column pass: one column of the row-transformed image per workgroup
// ---------------------------------------------------------------------------
template <typename T>
struct ColArgs {
    const cpx<T>* Tm;
    int N, Rs, ldT;
    const cpx<T>* tw;
    cpx<T>* S;
    size_t ldS;      // OUT_COLMAJOR_SHIFT: column stride; OUT_ROWMAJOR_PLAIN: unused (row stride C)
};

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of columns so that the strided reads of Tm share L2 lines.
__device__ __forceinline__ int xcd_column(int b, int nblocks) {
    const int per = nblocks >> 3;
    return (b & 7) * per + (b >> 3);
}

template <typename T, int LOG2N, int OUT>
__global__ __launch_bounds__(FftShape<LOG2N>::NT) void k_colpass(ColArgs<T> a) {
    using S = FftShape<LOG2N>;
    using F = BlockFft<T, LOG2N, false>;
    constexpr int N = S::N, NT = S::NT, E = S::E;
    constexpr int C = N / 2 + 1;
    __shared__ T lds[S::LDS_ELEMS];
    const int col = xcd_column(blockIdx.x, gridDim.x);
    if (col >= C) return;
    const int tid = threadIdx.x;
    const int Rs = a.Rs;
    const cpx<T>* src = a.Tm + col;
    const size_t ldT = a.ldT;
    cpx<T> v[E];
    F::run(v, lds, tid, Rs, a.tw, [&](int n) -> cpx<T> {
        const int kk = n < Rs ? n : n - N + 2 * Rs;
        return src[size_t(kk) * ldT];
    });
    if constexpr (OUT == OUT_COLMAJOR_SHIFT) {
        cpx<T>* dst = a.S + size_t(col) * a.ldS;
#pragma unroll
        for (int r = 0; r < E; ++r) dst[(tid + r * NT + N / 2) & (N - 1)] = v[r];
    } else {
#pragma unroll
        for (int r = 0; r < E; ++r) a.S[size_t(tid + r * NT) * C + col] = v[r];
    }
}

// ---------------------------------------------------------------------------
// generic even N (not a power of two, or too small/large for the LDS FFT):
// direct DFT with exact twiddle indices, O(N) per output.
// ---------------------------------------------------------------------------
template <typename T, int SRC>
__global__ __launch_bounds__(256) void k_rowdft_generic(RowArgs<T> a) {
    extern __shared__ unsigned char smem_raw[];
    T* x = reinterpret_cast<T*>(smem_raw);           // 2*Rs support values
    const int N = a.N, Rs = a.Rs, C = N / 2 + 1;
    const int kk = blockIdx.x;
    const int k = compact_to_row(kk, Rs, N);
    for (int s = threadIdx.x; s < 2 * Rs; s += blockDim.x) {
        const int n = s < Rs ? s : s + N - 2 * Rs;
        x[s] = row_source<T, SRC>(a, k, n);
    }
    __syncthreads();
    cpx<T>* out = a.Tm + size_t(kk) * a.ldT;
    for (int l = threadIdx.x; l < C; l += blockDim.x) {
        T re = 0, im = 0;
        int e = 0;                                     // (n*l) mod N for n = 0, 1, ...
        for (int s = 0; s < Rs; ++s) {
            const cpx<T> w = a.tw[e];
            re = ffma(x[s], w.re, re);
            im = ffma(x[s], w.im, im);
            e += l; if (e >= N) e -= N;
        }
        e = int((long long)(N - Rs) * l % N);
        for (int s = Rs; s < 2 * Rs; ++s) {
            const cpx<T> w = a.tw[e];
            re = ffma(x[s], w.re, re);
            im = ffma(x[s], w.im, im);
            e += l; if (e >= N) e -= N;
        }
        out[l] = cpx<T>{re, im};
    }
}

template <typename T, int OUT>
__global__ __launch_bounds__(256) void k_coldft_generic(ColArgs<T> a) {
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* x = reinterpret_cast<cpx<T>*>(smem_raw);  // 2*Rs support values
    const int N = a.N, Rs = a.Rs, C = N / 2 + 1;
    const int col = blockIdx.x;
    for (int s = threadIdx.x; s < 2 * Rs; s += blockDim.x) x[s] = a.Tm[size_t(s) * a.ldT + col];
    __syncthreads();
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        T re = 0, im = 0;
        int e = 0;
        for (int s = 0; s < Rs; ++s) {
            const cpx<T> w = a.tw[e];
            re = ffma(x[s].re, w.re, ffma(-x[s].im, w.im, re));
            im = ffma(x[s].re, w.im, ffma(x[s].im, w.re, im));
            e += k; if (e >= N) e -= N;
        }
        e = int((long long)(N - Rs) * k % N);
        for (int s = Rs; s < 2 * Rs; ++s) {
            const cpx<T> w = a.tw[e];
            re = ffma(x[s].re, w.re, ffma(-x[s].im, w.im, re));
            im = ffma(x[s].re, w.im, ffma(x[s].im, w.re, im));
            e += k; if (e >= N) e -= N;
        }
        if constexpr (OUT == OUT_COLMAJOR_SHIFT) {
            int r = k + N / 2; if (r >= N) r -= N;
            a.S[size_t(col) * a.ldS + r] = cpx<T>{re, im};
        } else {
            a.S[size_t(k) * C + col] = cpx<T>{re, im};
        }
    }
}

// ---------------------------------------------------------------------------
// profile path with the image symmetry: a swept radial profile depends on
// x^2 and y^2 only, and the quadrant-shifted image z[k][n] is exactly even in
// both indices (z[N-k][n] = z[k][n] = z[k][N-n]; the box rows/columns that
// have no mirror inside the box are zero, since there r >= rmax*dxy puts the
// radial index at nr-1 or beyond).  Its DFT is therefore real and even in both
// frequencies: the row pass packs rows m_a, m_b as one complex FFT and reads
// the two real results off the real and imaginary parts; the column pass does
// the same for two columns and stores frequencies 0..N/2 only.
//   Ts : real, row transform of image row m = |y| pixel index (m = 0..Rs') at
//        u-frequency l (l < C), in 64-byte blocks of 2 rows x 4 frequencies
//        (ts_index): a row pass workgroup (two rows) writes whole lines, and a
//        column pair reads 16 bytes per row, two rows per line.
//   H  : real, S at u-column l and |v-frequency| m (S at shifted row r is
//        m = |r - N/2|), columns interleaved in pairs (h_index): column pair
//        {2p, 2p+1} is a run of ldH (m, 2p), (m, 2p+1) pairs, so the column pass
//        stores one 16-byte pair per row and a sample's four corners (columns
//        c, c+1, rows m0, m1) share one 32-byte run when c is even.
// ---------------------------------------------------------------------------
// Offset of H element (column l, row m); ldH (>= N/2 + 1) rows per column.
__device__ __forceinline__ size_t h_index(int l, int m, int ldH) {
    return (size_t(l >> 1) * size_t(ldH) + size_t(m)) * 2 + size_t(l & 1);
}

// Offset of Ts element (row m, frequency l); ldT (>= C, a multiple of 8) reals per row.
__device__ __forceinline__ size_t ts_index(int m, int l, int ldT) {
    return (size_t(m >> 1) * size_t(ldT) + size_t(l >> 2) * 4) * 2 + size_t(((m & 1) << 2) + (l & 3));
}

template <typename T>
struct SymRowArgs {
    SweepPrm<T> sw;
    int N, Rs;            // symmetric support half-width Rs' (columns [0,Rs') U [N-Rs',N))
    int nrows;            // Rs' + 1 rows m = 0..Rs'
    int ldT;              // reals per row of Ts (>= C, a multiple of 8; see ts_index)
    const cpx<T>* tw;
    T* Ts;
    double* zero;         // optional: workgroup 0 clears zero[0..nzero) and *zero1 (gradient accumulators)
    int nzero;
    double* zero1;
};

// SPLIT: the half-LDS exchange (more workgroups per CU: the batched form's many
// walkers); otherwise both parts per round trip (lower latency: one call's
// ~Rs'/2 workgroups).
template <typename T, int LOG2N, bool SPLIT = false>
__device__ __forceinline__ void rowpass_sym_body(const SymRowArgs<T>& a, int blk) {
    using S = FftShape<LOG2N>;
    using F = BlockFft<T, LOG2N, false, SPLIT>;
    constexpr int N = S::N, NT = S::NT, E = S::E;
    constexpr int C = N / 2 + 1;
    __shared__ T lds[F::LDS_REALS];
    __shared__ T Ic[NT];
    const int tid = threadIdx.x;
    const int ma = 2 * blk, mb = ma + 1;
    const bool hb = mb < a.nrows;
    const bool cen = ma == 0;     // this workgroup's row ma = 0 holds the central pixel
    if (blk == 0) {
        if (a.zero)
            for (int k = tid; k < a.nzero; k += NT) a.zero[k] = 0.0;
        if (a.zero1 && tid == 0) *a.zero1 = 0.0;
    }
    // The central pixel is row 0's input n = 0: a constant that adds its value
    // to every output frequency.  The transform runs without it and the value
    // is added at the store, so its profile reads (issued here, beside the
    // sweep's) and its serial arithmetic stay off the transform's path.
    T ipre = T(0);
    if (cen && tid < a.sw.nr) ipre = a.sw.I[tid];
    cpx<T> v[E];
    F::run(v, lds, tid, a.Rs, a.tw, [&](int n) -> cpx<T> {
        const int xi = n < N / 2 ? -n : N - n;
        const T za = (ma == 0 && n == 0) ? T(0) : sweep_pixel(xi, ma, a.sw);
        const T zb = hb ? sweep_pixel(xi, mb, a.sw) : T(0);
        return cpx<T>{za, zb};
    });
    T c = T(0);
    if (cen) {
        Ic[tid] = ipre;
        __syncthreads();
        c = central_value_of(a.sw, [&](int i) { return i < NT ? Ic[i] : a.sw.I[i]; });
    }
    // rows ma, mb of Ts (mb = nrows is a padding row): four consecutive lanes fill one line
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int k = tid + r * NT;
        if (r * NT < C && k < C) {
            T* o = a.Ts + ts_index(ma, k, a.ldT);
            o[0] = cen ? v[r].re + c : v[r].re;
            o[4] = v[r].im;
        }
    }
}

template <typename T, int LOG2N>
__global__ __launch_bounds__(FftShape<LOG2N>::NT) void k_rowpass_sym(SymRowArgs<T> a) {
    rowpass_sym_body<T, LOG2N>(a, blockIdx.x);
}

template <typename T>
struct SymColArgs {
    const T* Ts;
    int N, Rs, ldT;
    const cpx<T>* tw;
    T* H;
    int ldH;
    double rc2;                  // > 0: store only cells within sqrt(rc2) (cells) of the origin (stored_rows)
    const cpx<T>* tw512;         // W_512^e table (k_colpass_sub)
};

// Highest row of column l of a half spectrum that a sample within the store
// radius sqrt(rc2) can read (-1: none); N/2 when there is no radius (rc2 <= 0).
__device__ __forceinline__ int stored_rows(double rc2, int l, int N) {
    if (!(rc2 > 0.0)) return N / 2;
    const double r = rc2 - double(l) * double(l);
    if (r < 0.0) return -1;
    const int k = int(sqrt(r));
    return k < N / 2 ? k : N / 2;
}

// Column N/2 of H (the one column without a partner among the N/2 + 1):
// H[N/2][k] = sum over the loaded rows n of Ts[|n|][N/2] cos(2 pi n k / N), i.e.
// Ts[0] + 2 sum_{0<m<Rs} Ts[m] cos(2 pi m k/N) + Ts[Rs] cos(2 pi Rs k/N), as a
// direct sum (FP64 accumulation) by one wave per output k = blk + q N/4.
template <typename T, int LOG2N, int NW = (FftShape<LOG2N>::NT / 64 > 0 ? FftShape<LOG2N>::NT / 64 : 1)>
__device__ __forceinline__ void colpass_nyquist(const SymColArgs<T>& a, int blk) {
    constexpr int N = 1 << LOG2N;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kmax = stored_rows(a.rc2, N / 2, N);
    for (int q = wv; q < 3; q += NW) {
        const int k = blk + q * (N / 4);
        if (k > kmax) break;
        double acc = 0.0;
        for (int m = lane; m <= a.Rs; m += 64) {
            const double wm = (m == 0 || m == a.Rs) ? 1.0 : 2.0;
            acc += wm * double(a.Ts[ts_index(m, N / 2, a.ldT)]) * double(a.tw[(m * k) & (N - 1)].re);
        }
        acc = wave_sum(acc);
        if (lane == 0) a.H[h_index(N / 2, k, a.ldH)] = T(acc);
    }
}

template <typename T, int LOG2N>
__device__ __forceinline__ void colpass_sym_body(const SymColArgs<T>& a, int blk, int nblk) {
    using S = FftShape<LOG2N>;
    using F = BlockFft<T, LOG2N, false>;
    constexpr int N = S::N, NT = S::NT, E = S::E;
    constexpr int NPAIR = N / 4;                  // pairs of columns 0..N/2-1; column N/2 separately
    __shared__ T lds[S::LDS_ELEMS];
    const int pair = xcd_column(blk, nblk);
    if (pair >= NPAIR) return;
    const int la = 2 * pair;
    const int tid = threadIdx.x;
    const int kc = stored_rows(a.rc2, la, N);   // rows a sample can read (column la + 1: no more)
    if (kc >= 0) {                              // (else no sample reaches these columns)
        // columns la, la + 1 of Ts: one aligned 16-byte pair per row m (la even)
        cpx<T> v[E];
        F::run(v, lds, tid, a.Rs, a.tw, [&](int n) -> cpx<T> {
            const int m = n < N / 2 ? n : N - n;
            return *reinterpret_cast<const cpx<T>*>(a.Ts + ts_index(m, la, a.ldT));
        });
        cpx<T>* hab = reinterpret_cast<cpx<T>*>(a.H + h_index(la, 0, a.ldH));   // columns la, la + 1
#pragma unroll
        for (int r = 0; r <= E / 2; ++r) {
            const int k = tid + r * NT;
            if (k <= kc) hab[k] = v[r];
        }
    }
    colpass_nyquist<T, LOG2N>(a, pair);
}

// Pruned column transforms.  Column pair (la, la + 1) of Ts is nonzero only
// for signed positions |n'| < Rs (row |n'| of Ts; rows rmax = Rs - 1 and Rs lie
// beyond the swept profile and hold zeros, unless Rs = N/2).  With Rs <= 256
// that support fits in the window n' in [-256, 256) of 512 positions, so
// output k = k0 + P K (P = N / 512) of the N-point transform is the 512-point
// transform, at K, of z[m] = x[m'] W_N^(m' k0), m' = m or m - 512 the signed
// position of window slot m.  The input is even, so X[N - k] = X[k]: residues
// k0 = 0..P/2 give every row 0..N/2 (residue P - k0 as the mirror of residue
// k0).  One wave per residue (NK = P/2 + 1 waves per workgroup, one workgroup
// per column pair): 40% fewer FP64 operations per column pair than the full
// transform and a shorter path per workgroup (no barrier inside the
// transform), for about the same LDS traffic (each residue wave reads the
// whole window, and the outputs, rows P apart per wave, are transposed
// through LDS into contiguous row stores).
template <int LOG2N>
struct ResidueShape {
    static constexpr int N = 1 << LOG2N, P = N / 512, NK = P / 2 + 1, NT = 64 * NK;
    static_assert(LOG2N >= 9 && LOG2N <= 13, "residue transforms for 512 <= N <= 8192");
};
constexpr int RESIDUE_MAX_RS = 256;

template <typename T, int LOG2N, bool SPLITROW = false>
__device__ __forceinline__ void colpass_sub_body(const SymColArgs<T>& a, int pair) {
    using CS = ResidueShape<LOG2N>;
    using F = ResidueFft<T, true>;
    constexpr int N = CS::N, P = CS::P;
    // LDS: the window (read once per workgroup) and the waves' exchange regions, then the output
    // rows 0..N/2 (row i at i + i/8: the residue waves write rows P apart, which the padding spreads
    // over the banks), read back in row order for contiguous stores.  SPLITROW (the batched form, bound
    // by throughput): the rows pass through LDS real parts first, then imaginary parts, in half the
    // space, so more workgroups fit on a CU
    constexpr size_t FFT_BYTES = 512 * sizeof(cpx<T>) + CS::NK * F::LDS_REALS * sizeof(T);
    constexpr size_t ROW_BYTES = (N / 2 + N / 16 + 1) * (SPLITROW ? sizeof(T) : sizeof(cpx<T>));
    __shared__ __align__(16) unsigned char smem[FFT_BYTES > ROW_BYTES ? FFT_BYTES : ROW_BYTES];
    cpx<T>* xin = reinterpret_cast<cpx<T>*>(smem);
    T* lds = reinterpret_cast<T*>(smem + 512 * sizeof(cpx<T>));
    cpx<T>* rows = reinterpret_cast<cpx<T>*>(smem);
    const int tid = threadIdx.x, lane = tid & 63;
    const int k0 = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int la = 2 * pair;
    const int kc = stored_rows(a.rc2, la, N);   // rows a sample can read (column la + 1: no more)
    if (kc >= 0) {
        // window slot m holds signed position m' = m (m < 256) or m - 512; P = 1: the whole
        // column, position -256 = N/2 included
#pragma unroll
        for (int q = 0; q < (512 + CS::NT - 1) / CS::NT; ++q) {
            const int m = tid + q * CS::NT;
            if (m < 512) {
                const int mp = m < 256 ? m : m - 512;
                const bool in = mp >= 0 ? mp < a.Rs : (P == 1 ? -mp <= a.Rs : -mp < a.Rs);
                const T* src = a.Ts + ts_index(in ? (mp >= 0 ? mp : -mp) : 0, la, a.ldT);
                const T re = src[0], im = src[1];
                xin[m] = in ? cpx<T>{re, im} : cpx<T>{T(0), T(0)};
            }
        }
        __syncthreads();
        // y[m] = x[m'] W_N^(-512 k0 [m' < 0]) with pass 0's W_N^(64 k0 r): one factor per r
        cpx<T> v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            v[r] = xin[lane + 64 * r];
            if (r != 0 && k0 != 0) v[r] = cmul(v[r], a.tw[(64 * k0 * (r < 4 ? r : r - 8)) & (N - 1)]);
        }
        F::run(v, lds + k0 * F::LDS_REALS, lane, a.tw, a.tw512, P, k0);
        __syncthreads();                             // every wave is done with the window and its region
        const bool mirror = k0 != 0 && 2 * k0 != P;
        cpx<T>* hab = reinterpret_cast<cpx<T>*>(a.H + h_index(la, 0, a.ldH));   // columns la, la + 1
        if constexpr (SPLITROW) {
            constexpr int PER = (N / 2 + 1 + CS::NT - 1) / CS::NT;
            T* rr = reinterpret_cast<T*>(smem);
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int k = k0 + P * (lane + 64 * r);
                const int row = k <= N / 2 ? k : (mirror ? N - k : -1);
                if (row >= 0 && row <= kc) rr[row + (row >> 3)] = v[r].re;
            }
            __syncthreads();
            T keep[PER];
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int row = tid + q * CS::NT;
                keep[q] = row <= kc ? rr[row + (row >> 3)] : T(0);
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int k = k0 + P * (lane + 64 * r);
                const int row = k <= N / 2 ? k : (mirror ? N - k : -1);
                if (row >= 0 && row <= kc) rr[row + (row >> 3)] = v[r].im;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int row = tid + q * CS::NT;
                if (row <= kc) hab[row] = cpx<T>{keep[q], rr[row + (row >> 3)]};
            }
        } else {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int k = k0 + P * (lane + 64 * r);
                const int row = k <= N / 2 ? k : (mirror ? N - k : -1);
                if (row >= 0 && row <= kc) rows[row + (row >> 3)] = v[r];
            }
            __syncthreads();
            for (int row = tid; row <= kc; row += CS::NT) hab[row] = rows[row + (row >> 3)];
        }
    }
    colpass_nyquist<T, LOG2N, CS::NK>(a, pair);
}

template <typename T, int LOG2N>
__global__ __launch_bounds__(ResidueShape<LOG2N>::NT) void k_colpass_sub(SymColArgs<T> a) {
    colpass_sub_body<T, LOG2N>(a, xcd_column(blockIdx.x, gridDim.x));
}

// At most 128 VGPRs (4 waves per SIMD): one call's N/4 workgroups then all fit
// on the chip at once (at 130 VGPRs a quarter of them waited for a second
// round); the cap costs a 20-byte spill at N=4096 FP64 and still saves ~6%.
// The batched kernel, whose many walkers fill the chip regardless, keeps the
// compiler's choice.
template <typename T, int LOG2N>
__global__ __launch_bounds__(FftShape<LOG2N>::NT) __attribute__((amdgpu_waves_per_eu(4))) void k_colpass_sym(
    SymColArgs<T> a) {
    colpass_sym_body<T, LOG2N>(a, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// rotate -> interpolate -> phase -> (visibilities | chi2)
// interpolate_core galario.cpp:749-815, apply_phase_sampled_core 877-884,
// uv_rotate_core 955-960, diff_weighted_core 1453-1461.
// ---------------------------------------------------------------------------
template <typename T>
struct SampleArgs {
    const cpx<T>* S;             // complex spectrum (general path)
    long long rs, cs;            // element (row, col) at row*rs + col*cs
    const T* H;                  // real half spectrum (symmetric path), h_index layout
    int ldH;
    int nrow, ncol;
    T v_origin, duv;
    int rotate;  T cosPA, sinPA;
    int phase;   T dRA2, dDec2;  // 2*pi*dRArot, 2*pi*dDecrot
    T dRAt, dDect;               // 2*dRArot, 2*dDecrot: the phase in half turns (double: sincospi)
    T inv_duv;                   // 1/duv (double: the grid position multiplies by it)
    int nd;
    const T* u; const T* v;
    cpx<T>* vis;                 // MODE_VIS
    const int* perm;             // MODE_VIS: point i is written to vis[perm[i]] (null: vis[i])
    const T* re; const T* im; const T* w;
    double* partials;            // MODE_CHI2: one per workgroup
    unsigned* flag;              // bit 0: a point fell outside the grid
    double rho2;                 // > 0: squared uv extent in cells; points beyond it count as outside
    double duv_d, inv_duv_d;     // duv and 1/duv in double (the FP32-bulk form's FP64 per-point arithmetic)
};

// Grid coordinates of a (rotated) point; oob when it is not inside the grid
// (or on its last row/column, where the weight beyond is zero).
template <typename T>
struct GridPos {
    int fl_u, fl_v, fu1, fv1;
    T q, t, sign_u, urot, vrot;
    bool oob;
};

// RECIP (the real, even spectrum, in double): the two divisions by duv become multiplications by
// 1/duv, within an ulp of the quotients (4e-13 cells at the grid's edge).  The interpolated
// phase is then 0 or pi, so the sample moves by that ulp only.  The general complex path keeps
// the divisions: galario's phase atan2(Z) amplifies a change of the position by 1/|Z|.
template <bool RECIP = false, typename T>
__device__ __forceinline__ GridPos<T> grid_pos(const SampleArgs<T>& a, T u0, T v0) {
    GridPos<T> g;
    g.urot = u0;
    g.vrot = v0;
    if (a.rotate) {
        g.urot = u0 * a.cosPA - v0 * a.sinPA;
        g.vrot = u0 * a.sinPA + v0 * a.cosPA;
    }
    g.sign_u = copysign(T(1), g.urot);
    T indu, indv;
    if constexpr (RECIP && std::is_same<T, double>::value) {
        indu = fabs(g.urot) * a.inv_duv;
        indv = T(a.nrow / 2) + a.v_origin * g.sign_u * g.vrot * a.inv_duv;
    } else {
        indu = fabs(g.urot) / a.duv;
        indv = T(a.nrow / 2) + a.v_origin * g.sign_u * g.vrot / a.duv;
    }
    g.oob = !(indu <= T(a.ncol - 1) && indv >= T(0) && indv <= T(a.nrow - 1));
    if (a.rho2 > 0.0) {   // beyond the declared extent: the spectrum there may not have been stored
        const double du = double(indu), dv = double(indv) - double(a.nrow / 2);
        if (du * du + dv * dv > a.rho2) g.oob = true;
    }
    g.fl_u = g.oob ? 0 : int(floor(indu));
    g.fl_v = g.oob ? 0 : int(floor(indv));
    g.q = indu - T(g.fl_u);
    g.t = indv - T(g.fl_v);
    g.fu1 = g.fl_u + 1 < a.ncol ? g.fl_u + 1 : g.fl_u;   // weight q == 0 there
    g.fv1 = g.fl_v + 1 < a.nrow ? g.fl_v + 1 : g.fl_v;   // weight t == 0 there
    return g;
}

// e^{i theta}, theta = urot 2 pi dRArot + vrot 2 pi dDecrot.  In double as sincospi of theta / pi
// (argument reduction by exact subtraction instead of the division by 2 pi of sincos); the angle
// agrees to an ulp.
template <typename T>
__device__ __forceinline__ void phase_factor(const SampleArgs<T>& a, T urot, T vrot, T* s, T* c) {
    if constexpr (std::is_same<T, double>::value) {
        sincospi(urot * a.dRAt + vrot * a.dDect, s, c);
    } else {
        fsincos(urot * a.dRA2 + vrot * a.dDec2, s, c);
    }
}

template <typename T>
__device__ __forceinline__ cpx<T> apply_phase_pt(const SampleArgs<T>& a, const GridPos<T>& g, cpx<T> V) {
    if (a.phase) {
        T s, c;
        phase_factor(a, g.urot, g.vrot, &s, &c);
        V = cpx<T>{V.re * c - V.im * s, V.re * s + V.im * c};
    }
    return V;
}

template <typename T>
__device__ __forceinline__ cpx<T> sample_point(const SampleArgs<T>& a, T u0, T v0, bool& oob) {
    const GridPos<T> g = grid_pos(a, u0, v0);
    oob = g.oob;
    const T q = g.q, t = g.t;
    const cpx<T> y0 = a.S[g.fl_v * a.rs + g.fl_u * a.cs];
    const cpx<T> y1 = a.S[g.fv1 * a.rs + g.fl_u * a.cs];
    const cpx<T> y2 = a.S[g.fv1 * a.rs + g.fu1 * a.cs];
    const cpx<T> y3 = a.S[g.fl_v * a.rs + g.fu1 * a.cs];
    // Z = t*(q*((y0+y2)-(y1+y3))) + (t*(y1-y0) + q*(y3-y0)) + y0
    const T df_re = (y0.re + y2.re) - (y1.re + y3.re), df_im = (y0.im + y2.im) - (y1.im + y3.im);
    const T t1_re = t * (q * df_re), t1_im = t * (q * df_im);
    const T t2_re = t * (y1.re - y0.re), t2_im = t * (y1.im - y0.im);
    const T t3_re = q * (y3.re - y0.re), t3_im = q * (y3.im - y0.im);
    const T z_re = (t1_re + (t2_re + t3_re)) + y0.re;
    const T z_im = (t1_im + (t2_im + t3_im)) + y0.im;
    const T ph = atan2(z_im, z_re) * g.sign_u;
    const T a0 = hypot(y0.re, y0.im), a1 = hypot(y1.re, y1.im);
    const T a2 = hypot(y2.re, y2.im), a3 = hypot(y3.re, y3.im);
    T amp = a0;
    amp += (a3 - a0) * q;
    amp += (a1 - a0) * t;
    amp += (((a0 - a1) + a2) - a3) * t * q;
    T sp, cp;
    fsincos(ph, &sp, &cp);
    cpx<T> V = apply_phase_pt(a, g, cpx<T>{amp * cp, amp * sp});
    if (oob) V = cpx<T>{T(NAN), T(NAN)};
    return V;
}

// Same on the real, even spectrum H: the corners are real, Z is real, the
// phase of galario's amplitude/phase interpolation is 0 or pi, so
// V = (copysign(A, Z), 0) before the phase shift.
//
// Storage and arithmetic may differ: the spectrum H is read from `a` (type T)
// while the position, interpolation and phase use the scalars of `sc` and
// arithmetic in P.  P = T (sc = a) is galario's arithmetic in that precision;
// T = float, P = double is the FP32-bulk form with FP64 per-point arithmetic.
template <typename T, typename P>
__device__ __forceinline__ cpx<P> sample_point_sym2(const SampleArgs<T>& a, const SampleArgs<P>& sc, P u0, P v0,
                                                    bool& oob) {
    const GridPos<P> g = grid_pos<true>(sc, u0, v0);
    oob = g.oob;
    const P q = g.q, t = g.t;
    const int half = a.nrow / 2;
    const int m0 = abs(g.fl_v - half), m1 = abs(g.fv1 - half);
    const P y0 = P(a.H[h_index(g.fl_u, m0, a.ldH)]), y1 = P(a.H[h_index(g.fl_u, m1, a.ldH)]);
    const P y3 = P(a.H[h_index(g.fu1, m0, a.ldH)]), y2 = P(a.H[h_index(g.fu1, m1, a.ldH)]);
    const P z = (t * (q * ((y0 + y2) - (y1 + y3))) + (t * (y1 - y0) + q * (y3 - y0))) + y0;
    const P a0 = fabs(y0), a1 = fabs(y1), a2 = fabs(y2), a3 = fabs(y3);
    P amp = a0;
    amp += (a3 - a0) * q;
    amp += (a1 - a0) * t;
    amp += (((a0 - a1) + a2) - a3) * t * q;
    cpx<P> V = apply_phase_pt(sc, g, cpx<P>{copysign(amp, z), P(0)});
    if (oob) V = cpx<P>{P(NAN), P(NAN)};
    return V;
}

template <typename T>
__device__ __forceinline__ cpx<T> sample_point_sym(const SampleArgs<T>& a, T u0, T v0, bool& oob) {
    return sample_point_sym2<T, T>(a, a, u0, v0, oob);
}

// Deterministic sum over a workgroup of any size <= 1024 (red: 16 doubles);
// every thread gets the result.
__device__ __forceinline__ double block_sum_any(double acc, double* red16) {
    acc = wave_sum(acc);
    const int nw = int((blockDim.x + 63) >> 6);
    if ((threadIdx.x & 63) == 0) red16[threadIdx.x >> 6] = acc;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < nw; ++k) s += red16[k];
    return s;
}

// Read-rate reference: the workgroups read n2 16-byte elements in a
// grid-stride loop and write one sum each.  A plain stream of the same bytes
// as a kernel bounds the rate that kernel's own reads can reach.
__global__ __launch_bounds__(256) void k_read_probe(const cpx<double>* __restrict__ p, size_t n2, double* out);

// Deterministic block sum of `acc` (blockDim 256) into *out on thread 0.
__device__ __forceinline__ double block_sum256(double acc, double* red4) {
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = acc;
    __syncthreads();
    return (red4[0] + red4[1]) + (red4[2] + red4[3]);
}

// acc + p[i0] + p[i0 + 1024] + ... (terms below n), added in that order; the
// loads of eight terms are issued together, one memory round trip per eight.
__device__ __forceinline__ double sum_stride1024(const double* __restrict__ p, int i0, int n, double acc) {
    for (; i0 < n; i0 += 8 * 1024) {
        double x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = i0 + q * 1024 < n ? p[i0 + q * 1024] : 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (i0 + q * 1024 < n) acc += x[q];
    }
    return acc;
}

__global__ __launch_bounds__(256) void k_read_probe(const cpx<double>* __restrict__ p, size_t n2, double* out) {
    __shared__ double red[16];
    double s = 0.0;
    const size_t stride = size_t(gridDim.x) * 256;
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n2; i += stride) {
        const cpx<double> x = p[i];
        s += x.re + x.im;
    }
    s = block_sum_any(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// Deterministic final sum of n per-workgroup partials (one workgroup of 1024
// threads, fixed order).  A separate launch: publishing partials across the
// 8 XCDs from inside one kernel would need a device-scope release per
// workgroup, which on gfx950 writes back the whole L2 each time.
template <typename T>
__global__ __launch_bounds__(1024) void k_sum_partials(const double* __restrict__ p, int n, double* out_d, T* out_t) {
    __shared__ double red[16];
    double acc = sum_stride1024(p, threadIdx.x, n, 0.0);
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < 16; ++k) s += red[k];
        if (out_d) out_d[0] = s;
        if (out_t) out_t[0] = T(s);
    }
}

template <typename T, int MODE, bool SYM>
__device__ __forceinline__ void sample_body(const SampleArgs<T>& a, int blk) {
    __shared__ double red[4];
    const int i = blk * 256 + threadIdx.x;
    double acc = 0.0;
    bool oob = false;
    if (i < a.nd) {
        const cpx<T> V = SYM ? sample_point_sym(a, a.u[i], a.v[i], oob) : sample_point(a, a.u[i], a.v[i], oob);
        if constexpr (MODE == MODE_VIS) {
            a.vis[a.perm ? a.perm[i] : i] = V;
        } else {
            const T sw = sqrt(a.w[i]);
            const T cr = (V.re - a.re[i]) * sw;
            const T ci = (V.im - a.im[i]) * sw;
            if constexpr (std::is_same<T, double>::value) acc = cr * cr + ci * ci;
            else acc = double(cr) * cr + double(ci) * ci;
        }
    }
    if (oob && a.flag) atomicOr(a.flag, 1u);
    if constexpr (MODE == MODE_CHI2) {
        const double b = block_sum256(acc, red);
        if (threadIdx.x == 0) a.partials[blk] = b;
    }
}

template <typename T, int MODE, bool SYM>
__global__ __launch_bounds__(256) void k_sample(SampleArgs<T> a) {
    sample_body<T, MODE, SYM>(a, blockIdx.x);
}

// Several sources with the same profile, inclination and PA at different
// offsets (galario's multi-source recipe: the visibilities of each source,
// phase-shifted by its own offset, summed; docs/tech-specs.rst:45-57).  They
// share the spectrum and the interpolated value V0 at each point, so
// V = sum_k flux_k * apply_phase(V0, dRA_k, dDec_k), k in order.
// src: nsrc x 3 (flux, dRA, dDec) in the caller's (unrotated) frame.
template <typename T>
__device__ __forceinline__ cpx<T> multisrc_point(const SampleArgs<T>& a, int i, int nsrc, const T* __restrict__ src,
                                                 bool& oob) {
    SampleArgs<T> b = a;
    b.phase = 0;
    const cpx<T> V0 = a.H ? sample_point_sym(b, a.u[i], a.v[i], oob) : sample_point(b, a.u[i], a.v[i], oob);
    T ur = a.u[i], vr = a.v[i];
    if (a.rotate) {
        ur = a.u[i] * a.cosPA - a.v[i] * a.sinPA;
        vr = a.u[i] * a.sinPA + a.v[i] * a.cosPA;
    }
    cpx<T> V{T(0), T(0)};
    for (int k = 0; k < nsrc; ++k) {
        const T f = src[3 * k], dRA = src[3 * k + 1], dDec = src[3 * k + 2];
        T dRAr = dRA, dDecr = dDec;
        if (a.rotate) {
            dRAr = dRA * a.cosPA - dDec * a.sinPA;
            dDecr = dRA * a.sinPA + dDec * a.cosPA;
        }
        cpx<T> Vk = V0;
        if (!(dRAr == T(0) && dDecr == T(0))) {
            SampleArgs<T> ak;          // this source's phase scalars, as setup_sample forms them
            ak.dRA2 = T(double(dRAr) * (2. * double(T(M_PI))));
            ak.dDec2 = T(double(dDecr) * (2. * double(T(M_PI))));
            ak.dRAt = T(2) * dRAr;
            ak.dDect = T(2) * dDecr;
            T sn, cs;
            phase_factor(ak, ur, vr, &sn, &cs);
            Vk = cpx<T>{V0.re * cs - V0.im * sn, V0.re * sn + V0.im * cs};
        }
        V.re += f * Vk.re;
        V.im += f * Vk.im;
    }
    if (oob) V = cpx<T>{T(NAN), T(NAN)};
    return V;
}

// PPT blocks of 256 points per workgroup (more loads in flight per wave);
// block q's chi2 partial is the same fixed-order sum as in k_sample.
template <typename T, int MODE, int PPT>
__global__ __launch_bounds__(256) void k_sample_multisrc(SampleArgs<T> a, int nsrc, const T* __restrict__ src,
                                                         int nblocks) {
    __shared__ double red[PPT][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double acc[PPT];
    bool oob_any = false;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        acc[q] = 0.0;
        const int i = (blockIdx.x * PPT + q) * 256 + tid;
        if (i < a.nd) {
            bool oob = false;
            const cpx<T> V = multisrc_point(a, i, nsrc, src, oob);
            oob_any |= oob;
            if constexpr (MODE == MODE_VIS) {
                a.vis[a.perm ? a.perm[i] : i] = V;
            } else {
                const T sw = sqrt(a.w[i]);
                const T cr = (V.re - a.re[i]) * sw;
                const T ci = (V.im - a.im[i]) * sw;
                if constexpr (std::is_same<T, double>::value) acc[q] = cr * cr + ci * ci;
                else acc[q] = double(cr) * cr + double(ci) * ci;
            }
        }
    }
    if (oob_any && a.flag) atomicOr(a.flag, 1u);
    if constexpr (MODE == MODE_CHI2) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const double x = wave_sum(acc[q]);
            if (lane == 0) red[q][wv] = x;
        }
        __syncthreads();
        if (tid < PPT) {
            const int b = blockIdx.x * PPT + tid;
            if (b < nblocks) a.partials[b] = (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]);
        }
    }
}

// chi2 with PPT blocks of 256 points per workgroup (more loads in flight per
// wave); block q's partial is the same fixed-order sum as in k_sample.
#ifndef GL_SAMPLE_PPT
#define GL_SAMPLE_PPT 2
#endif
constexpr int SAMPLE_PPT = GL_SAMPLE_PPT;

template <typename T, bool SYM, int PPT>
__global__ __launch_bounds__(256) void k_sample_chi2_n(SampleArgs<T> a, int nblocks) {
    __shared__ double red[PPT][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    double acc[PPT];
    bool oob_any = false;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        acc[q] = 0.0;
        const int i = (blockIdx.x * PPT + q) * 256 + tid;
        if (i < a.nd) {
            bool oob = false;
            const cpx<T> V = SYM ? sample_point_sym(a, a.u[i], a.v[i], oob) : sample_point(a, a.u[i], a.v[i], oob);
            oob_any |= oob;
            const T sw = sqrt(a.w[i]);
            const T cr = (V.re - a.re[i]) * sw;
            const T ci = (V.im - a.im[i]) * sw;
            if constexpr (std::is_same<T, double>::value) acc[q] = cr * cr + ci * ci;
            else acc[q] = double(cr) * cr + double(ci) * ci;
        }
    }
    if (oob_any && a.flag) atomicOr(a.flag, 1u);
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const double x = wave_sum(acc[q]);
        if (lane == 0) red[q][wv] = x;
    }
    __syncthreads();
    if (tid < PPT) {
        const int b = blockIdx.x * PPT + tid;
        if (b < nblocks) a.partials[b] = (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]);
    }
}

// ---------------------------------------------------------------------------
// expert stages
// ---------------------------------------------------------------------------
// Full nxy x nxy image of a profile (create_image_h without the padding columns).
template <typename T>
__global__ __launch_bounds__(256) void k_sweep_image(SweepPrm<T> p, int N, int ld, T* img) {
    const size_t idx = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= size_t(N) * ld) return;
    const int i = int(idx / ld), j = int(idx % ld);
    T val = T(0);
    const int lo = N / 2 - p.rmax, hi = N / 2 + p.rmax;
    if (i == N / 2 && j == N / 2) {
        val = central_value(p);
    } else if (i >= lo && i < hi && j >= lo && j < hi) {
        val = sweep_pixel(N / 2 - j, N / 2 - i, p);
    }
    img[idx] = val;
}

// uv_rotate_core (galario.cpp:955-960) over arrays.
template <typename T>
__global__ __launch_bounds__(256) void k_rotate(int nd, T c, T s, const T* u, const T* v, T* ur, T* vr) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nd) return;
    const T u0 = u[i], v0 = v[i];
    ur[i] = u0 * c - v0 * s;
    vr[i] = u0 * s + v0 * c;
}

// apply_phase_sampled_core (galario.cpp:877-884), in place; dRA2/dDec2 include 2*pi.
template <typename T>
__global__ __launch_bounds__(256) void k_apply_phase(int nd, T dRA2, T dDec2, const T* u, const T* v, cpx<T>* vis) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nd) return;
    const T ang = u[i] * dRA2 + v[i] * dDec2;
    T s, c;
    fsincos(ang, &s, &c);
    const cpx<T> V = vis[i];
    vis[i] = cpx<T>{V.re * c - V.im * s, V.re * s + V.im * c};
}

// chi2 terms of given visibilities (reduce_chi2, galario.cpp:1506-1538);
// `a` carries only the reduction field (partials).
template <typename T>
__global__ __launch_bounds__(256) void k_chi2_terms(int nd, const T* re, const T* im, const cpx<T>* vis, const T* w,
                                                    SampleArgs<T> a) {
    __shared__ double red[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (i < nd) {
        const T sw = sqrt(w[i]);
        const cpx<T> V = vis[i];
        const T cr = (V.re - re[i]) * sw;
        const T ci = (V.im - im[i]) * sw;
        if constexpr (std::is_same<T, double>::value) acc = cr * cr + ci * ci;
        else acc = double(cr) * cr + double(ci) * ci;
    }
    const double b = block_sum256(acc, red);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = b;
}

// Clears n doubles (a kernel rather than a memset, so that it may sit in a captured graph).
__global__ __launch_bounds__(256) void k_fill_zero(double* p, size_t n) {
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Quadrant swap of the N x N real image held in rows of ld reals
// (shift_core, galario.cpp:565-589); the padding columns are copied through.
template <typename T>
__global__ __launch_bounds__(256) void k_fftshift_padded(int N, int ld, const T* in, T* out) {
    const size_t idx = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= size_t(N) * ld) return;
    const int i = int(idx / ld), j = int(idx % ld);
    out[idx] = j < N ? in[size_t((i + N / 2) % N) * ld + (j + N / 2) % N] : in[idx];
}

// Swap of the upper and lower row halves of an nrow x ncol complex matrix
// (shift_axis0_core, galario.cpp:657-668).
template <typename T>
__global__ __launch_bounds__(256) void k_fftshift_axis0(int nrow, int ncol, cpx<T>* m) {
    const size_t idx = size_t(blockIdx.x) * 256 + threadIdx.x;
    const size_t half = size_t(nrow / 2) * ncol;
    if (idx >= half) return;
    const cpx<T> a = m[idx];
    m[idx] = m[idx + half];
    m[idx + half] = a;
}

}  // namespace gl
