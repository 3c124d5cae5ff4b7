// galario_amd: host side of the MI355X HIP implementation of galario's
// model-to-visibility path, exported through the C ABI in include/galario_amd.h.
//
// Per host thread and device there is one Workspace (stream, grow-only device
// buffers, pinned scalars); per device and size one twiddle table.  A call
// enqueues: [uploads] -> row pass -> column pass -> sample -> [reduce] ->
// [downloads + sync].  Argument checks mirror galario's (src/galario.cpp:89-121)
// and run before anything touches the GPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "galario_amd.h"
#include "gl_adjoint.hpp"
#include "gl_batch.hpp"
#include "gl_kernels.hpp"

using namespace gl;

// Scratch of the allocation-free chi2 entry point (galario_chi2_profile_noalloc_dev_*): the row
// spectra, the half spectrum and the per-workgroup chi2 partials of one evaluation.  One array per
// device, held until the process ends.  Host code never passes this name to a kernel: its
// per-device address comes from hipGetSymbolAddress at set-up (static_setup_current).
constexpr size_t SCRATCH_BYTES = size_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

// ------------------------------------------------------------------ errors
thread_local std::string t_err;

struct Err {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw Err{code, msg}; }

[[noreturn]] void invalid(const char* where, const std::string& msg) {
    fail(GALARIO_ERR_INVALID, std::string("Error in ") + where + ": " + msg);
}

void hipchk(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory) throw std::bad_alloc();
    fail(GALARIO_ERR_RUNTIME, std::string("Error in hip: ") + what + ": " + hipGetErrorString(e));
}
#define HC(x) hipchk((x), #x)

template <class F>
int guard(F&& f) {
    try {
        f();
        return GALARIO_OK;
    } catch (const Err& e) {
        t_err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        t_err = "Error in hip: out of device or host memory";
        return GALARIO_ERR_ALLOC;
    } catch (const std::exception& e) {
        t_err = e.what();
        return GALARIO_ERR_RUNTIME;
    } catch (...) {
        t_err = "unknown error";
        return GALARIO_ERR_RUNTIME;
    }
}

// CHECK_INPUT / CHECK_INPUTXY, src/galario.cpp:101-111
void check_input(int nx) {
    if (nx < 2) invalid("check input image", "x dimension = " + std::to_string(nx) + " is less than 2");
    if (nx % 2 != 0) invalid("check input image", "x dimension = " + std::to_string(nx) + " is odd");
}
void check_input_xy(int nx, int ny) {
    if (nx != ny)
        invalid("check input image",
                "Expect a square image but got shape (" + std::to_string(nx) + ", " + std::to_string(ny) + ")");
    check_input(nx);
}
void check_nd(int nd) {
    if (nd < 0) invalid("check input", "number of (u, v) points = " + std::to_string(nd) + " is negative");
}
// Device arrays of a _dev call: the profile always, u and v (and the observation when `obs`)
// whenever there are points.  Runs after the size and geometry checks.
template <typename T>
void check_dev_inputs(const T* I, int nd, const T* u, const T* v, bool obs, const T* re, const T* im, const T* w) {
    if (!I) invalid("check input", "null intensity pointer");
    if (nd > 0 && (!u || !v)) invalid("check input", "null u or v pointer");
    if (nd > 0 && obs && (!re || !im || !w)) invalid("check input", "null observed-visibility or weight pointer");
}

// ------------------------------------------------------------------ memory
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    template <class X>
    X* get(size_t n) {
        size_t b = n * sizeof(X);
        if (b < 256) b = 256;
        if (b > cap) {
            release();
            const hipError_t e = hipMalloc(&p, b);
            if (e != hipSuccess) {
                p = nullptr;
                (void)hipGetLastError();   // a failed allocation must not surface in a later launch check
                HC(e);                     // out of memory: std::bad_alloc (MemoryError in Python)
            }
            cap = b;
        }
        return static_cast<X*>(p);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct Workspace {
    int device = -1;
    hipStream_t own = nullptr;    // stream of the host-pointer entry points
    hipStream_t last = nullptr;   // stream of the last call that used the buffers
    bool has_last = false;        // (the null stream is a valid `last`)
    hipEvent_t ev = nullptr;
    DevBuf Tm, S, part, scal, flag_host, flag_dev;
    DevBuf Hh, Th;                // gradient: adjoint half spectrum (kept zero between calls), adjoint rows
    bool Hh_zero = false;         // Hh holds zeros (cleared by the adjoint column pass)
    DevBuf in[8];
    DevBuf out;
    double* h_scal = nullptr;     // pinned: [0] result, [1] flag
    unsigned dev_flag_dirty = 0;

    void release() {
        Tm.release(); S.release(); part.release(); scal.release(); Hh.release(); Th.release();
        Hh_zero = false;
        flag_host.release(); flag_dev.release(); out.release();
        for (auto& b : in) b.release();
        if (h_scal) (void)hipHostFree(h_scal);
        if (ev) (void)hipEventDestroy(ev);
        if (own) (void)hipStreamDestroy(own);
        h_scal = nullptr; ev = nullptr; own = nullptr; last = nullptr; has_last = false;
    }
};

thread_local std::map<int, Workspace*> t_ws;

// ------------------------------------------------------------------ stage timer
// When enabled, the _dev entry points record a HIP event before each stage and
// after the last one, on the stream the kernels run on; galario_stage_times
// synchronises and returns the summed durations per stage.
enum { ST_ROW = 0, ST_COL = 1, ST_SAMPLE = 2, ST_ADJCOL = 3, ST_ADJROW = 4, ST_REDUCE = 5, NSTAGE = 6 };
struct StageTimer {
    bool on = false;
    int rep_stage = -1, rep_n = 1;           // while on: launch that stage's kernel rep_n times
    int reps(int stage) const { return (on && stage == rep_stage) ? rep_n : 1; }
    std::vector<hipEvent_t> pool;            // free events
    std::vector<std::vector<std::pair<int, hipEvent_t>>> calls;   // per call: (stage, event) in order
    std::vector<std::pair<int, hipEvent_t>> cur;
    hipEvent_t get() {
        hipEvent_t e;
        if (!pool.empty()) { e = pool.back(); pool.pop_back(); }
        else HC(hipEventCreate(&e));
        return e;
    }
    void mark(int stage, hipStream_t s) {
        if (!on) return;
        hipEvent_t e = get();
        HC(hipEventRecord(e, s));
        cur.emplace_back(stage, e);
    }
    void end_call(hipStream_t s) {
        if (!on || cur.empty()) return;
        mark(NSTAGE, s);
        calls.push_back(std::move(cur));
        cur.clear();
    }
};
thread_local StageTimer t_timer;


// Workspace of the current device; orders `s` after the previous user stream.
Workspace& workspace(hipStream_t s, bool use_own) {
    int dev = 0;
    HC(hipGetDevice(&dev));
    Workspace*& w = t_ws[dev];
    if (!w) {
        // built aside and published only when complete: a failure leaves no half-made workspace
        struct Del {
            void operator()(Workspace* p) const { p->release(); delete p; }
        };
        std::unique_ptr<Workspace, Del> n(new Workspace());
        n->device = dev;
        HC(hipStreamCreateWithFlags(&n->own, hipStreamNonBlocking));
        HC(hipEventCreateWithFlags(&n->ev, hipEventDisableTiming));
        HC(hipHostMalloc(reinterpret_cast<void**>(&n->h_scal), 64, hipHostMallocDefault));
        unsigned* fh = n->flag_host.get<unsigned>(1);
        unsigned* fd = n->flag_dev.get<unsigned>(1);
        HC(hipMemsetAsync(fh, 0, sizeof(unsigned), n->own));
        HC(hipMemsetAsync(fd, 0, sizeof(unsigned), n->own));
        HC(hipStreamSynchronize(n->own));
        w = n.release();
    }
    hipStream_t use = use_own ? w->own : s;
    if (w->has_last && w->last != use) {
        HC(hipEventRecord(w->ev, w->last));
        HC(hipStreamWaitEvent(use, w->ev, 0));
    }
    w->last = use;
    w->has_last = true;
    return *w;
}

// ------------------------------------------------------------------ twiddles
std::mutex g_tw_mu;
std::map<std::tuple<int, int, int>, void*> g_tw;

// W_N^e = exp(-2 pi i e / N), e in [0, N), in long double, exact on the axes.
template <typename T>
const cpx<T>* twiddles(int N) {
    int dev = 0;
    HC(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_tw_mu);
    auto key = std::make_tuple(dev, N, int(sizeof(T)));
    auto it = g_tw.find(key);
    if (it != g_tw.end()) return static_cast<const cpx<T>*>(it->second);
    const long double pi = 3.141592653589793238462643383279502884L;
    cpx<T>* h = static_cast<cpx<T>*>(std::malloc(sizeof(cpx<T>) * size_t(N)));
    if (!h) throw std::bad_alloc();
    for (int e = 0; e < N; ++e) {
        long double c, s;
        const long long e4 = 4LL * e;
        if (e4 % N == 0) {
            const int q = int(e4 / N);           // quarter turns
            c = (q == 0) ? 1.0L : (q == 2 ? -1.0L : 0.0L);
            s = (q == 1) ? 1.0L : (q == 3 ? -1.0L : 0.0L);
        } else if (2LL * e > N) {                // conj symmetry: W^e = conj(W^{N-e})
            const long double a = 2.0L * pi * (long double)(N - e) / (long double)N;
            c = cosl(a);
            s = -sinl(a);
        } else {
            const long double a = 2.0L * pi * (long double)e / (long double)N;
            c = cosl(a);
            s = sinl(a);
        }
        h[e] = cpx<T>{T(c), T(-s)};
    }
    void* d = nullptr;
    hipError_t err = hipMalloc(&d, sizeof(cpx<T>) * size_t(N));
    if (err == hipSuccess) err = hipMemcpy(d, h, sizeof(cpx<T>) * size_t(N), hipMemcpyHostToDevice);
    std::free(h);
    HC(err);
    g_tw[key] = d;
    return static_cast<const cpx<T>*>(d);
}

// ------------------------------------------------------------------ launches
template <typename T>
constexpr int max_fast_log2() { return sizeof(T) == 8 ? 13 : 14; }

// log2(N) when the workgroup-resident FFT covers N, else -1 (generic path).
template <typename T>
int fast_log2(int N) {
    if (N < 128 || (N & (N - 1))) return -1;
    int L = 0;
    while ((1 << L) < N) ++L;
    return L <= max_fast_log2<T>() ? L : -1;
}

template <typename T, int L, int SRC>
void launch_rowpass_L(int pairs, hipStream_t s, const RowArgs<T>& a) {
    hipLaunchKernelGGL((k_rowpass<T, L, SRC>), dim3(pairs), dim3(FftShape<L>::NT), 0, s, a);
}
template <typename T, int L, int OUT>
void launch_colpass_L(int blocks, hipStream_t s, const ColArgs<T>& a) {
    hipLaunchKernelGGL((k_colpass<T, L, OUT>), dim3(blocks), dim3(FftShape<L>::NT), 0, s, a);
}

#define GL_L_SWITCH(T, CALL)                                                   \
    switch (L) {                                                                \
        case 7: CALL(7); break;                                                 \
        case 8: CALL(8); break;                                                 \
        case 9: CALL(9); break;                                                 \
        case 10: CALL(10); break;                                               \
        case 11: CALL(11); break;                                               \
        case 12: CALL(12); break;                                               \
        case 13: CALL(13); break;                                               \
        case 14:                                                                \
            if constexpr (sizeof(T) == 4) { CALL(14); break; }                  \
            [[fallthrough]];                                                    \
        default: fail(GALARIO_ERR_RUNTIME, "internal: no FFT kernel for this size"); \
    }


// The residue transforms (k_colpass_sub, and for the single gradient call
// k_adjcol_sub and k_adjrow_sub) when the profile's support fits their
// 512-row window and N >= 1024 (at N = 512 there is one residue, the whole
// column, and the workgroup transform is as fast); the full workgroup
// transform otherwise.
inline bool residue_fits(int L, int Rs) { return L >= 10 && L <= 13 && Rs <= RESIDUE_MAX_RS; }

#define GL_RESIDUE_SWITCH(CALL)                                                  \
    switch (L) {                                                                \
        case 10: CALL(10); break;                                               \
        case 11: CALL(11); break;                                               \
        case 12: CALL(12); break;                                               \
        case 13: CALL(13); break;                                               \
        default: fail(GALARIO_ERR_RUNTIME, "internal: no residue kernel for this size"); \
    }

template <typename T>
void launch_rowpass_sym(int L, int blocks, hipStream_t s, const SymRowArgs<T>& a) {
#define GL_CALL(LL) hipLaunchKernelGGL((k_rowpass_sym<T, LL>), dim3(blocks), dim3(FftShape<LL>::NT), 0, s, a)
    GL_L_SWITCH(T, GL_CALL)
#undef GL_CALL
}

template <typename T>
void launch_colpass_sym(int L, int blocks, hipStream_t s, const SymColArgs<T>& a) {
    if (residue_fits(L, a.Rs)) {
#define GL_CALL(LL) hipLaunchKernelGGL((k_colpass_sub<T, LL>), dim3(blocks), dim3(ResidueShape<LL>::NT), 0, s, a)
        GL_RESIDUE_SWITCH(GL_CALL)
#undef GL_CALL
        return;
    }
#define GL_CALL(LL) hipLaunchKernelGGL((k_colpass_sym<T, LL>), dim3(blocks), dim3(FftShape<LL>::NT), 0, s, a)
    GL_L_SWITCH(T, GL_CALL)
#undef GL_CALL
}

template <typename T>
void launch_adjcol_sym(int L, int blocks, hipStream_t s, const AdjColArgs<T>& a) {
    if (residue_fits(L, a.Rs)) {
#define GL_CALL(LL) hipLaunchKernelGGL((k_adjcol_sub<T, LL>), dim3(blocks), dim3(ResidueShape<LL>::NT), 0, s, a)
        GL_RESIDUE_SWITCH(GL_CALL)
#undef GL_CALL
        return;
    }
#define GL_CALL(LL) hipLaunchKernelGGL((k_adjcol_sym<T, LL>), dim3(blocks), dim3(FftShape<LL>::NT), 0, s, a)
    GL_L_SWITCH(T, GL_CALL)
#undef GL_CALL
}

template <typename T>
void launch_adjrow_sym(int L, int blocks, hipStream_t s, const AdjRowArgs<T>& a) {
    if (a.tw512 && residue_fits(L, a.Rs)) {
#define GL_CALL(LL) hipLaunchKernelGGL((k_adjrow_sub<T, LL>), dim3(blocks), dim3(ResidueShape<LL>::NT), 0, s, a)
        GL_RESIDUE_SWITCH(GL_CALL)
#undef GL_CALL
        return;
    }
#define GL_CALL(LL) hipLaunchKernelGGL((k_adjrow_sym<T, LL>), dim3(blocks), dim3(FftShape<LL>::NT), 0, s, a)
    GL_L_SWITCH(T, GL_CALL)
#undef GL_CALL
}

template <typename T, int SRC>
void launch_rowpass(int L, int pairs, hipStream_t s, const RowArgs<T>& a) {
    switch (L) {
        case 7: launch_rowpass_L<T, 7, SRC>(pairs, s, a); break;
        case 8: launch_rowpass_L<T, 8, SRC>(pairs, s, a); break;
        case 9: launch_rowpass_L<T, 9, SRC>(pairs, s, a); break;
        case 10: launch_rowpass_L<T, 10, SRC>(pairs, s, a); break;
        case 11: launch_rowpass_L<T, 11, SRC>(pairs, s, a); break;
        case 12: launch_rowpass_L<T, 12, SRC>(pairs, s, a); break;
        case 13: launch_rowpass_L<T, 13, SRC>(pairs, s, a); break;
        case 14:
            if constexpr (sizeof(T) == 4) { launch_rowpass_L<T, 14, SRC>(pairs, s, a); break; }
            [[fallthrough]];
        default: fail(GALARIO_ERR_RUNTIME, "internal: no row-pass kernel for this size");
    }
}

template <typename T, int OUT>
void launch_colpass(int L, int blocks, hipStream_t s, const ColArgs<T>& a) {
    switch (L) {
        case 7: launch_colpass_L<T, 7, OUT>(blocks, s, a); break;
        case 8: launch_colpass_L<T, 8, OUT>(blocks, s, a); break;
        case 9: launch_colpass_L<T, 9, OUT>(blocks, s, a); break;
        case 10: launch_colpass_L<T, 10, OUT>(blocks, s, a); break;
        case 11: launch_colpass_L<T, 11, OUT>(blocks, s, a); break;
        case 12: launch_colpass_L<T, 12, OUT>(blocks, s, a); break;
        case 13: launch_colpass_L<T, 13, OUT>(blocks, s, a); break;
        case 14:
            if constexpr (sizeof(T) == 4) { launch_colpass_L<T, 14, OUT>(blocks, s, a); break; }
            [[fallthrough]];
        default: fail(GALARIO_ERR_RUNTIME, "internal: no column-pass kernel for this size");
    }
}

inline unsigned blocks_for(size_t n, unsigned per) { return unsigned((n + per - 1) / per); }

template <typename K>
void allow_lds(K kernel, size_t bytes) {
    if (bytes > 48 * 1024)
        HC(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(bytes)));
}

// Shifted (or, for OUT_ROWMAJOR_PLAIN, plain) R2C spectrum of the rows given
// by `ra` (source fields filled by the caller).  Rs: support half-width.
template <typename T, int SRC, int OUT>
cpx<T>* build_spectrum(Workspace& w, hipStream_t s, RowArgs<T> ra, int N, int Rs, cpx<T>* S_out, size_t ldS) {
    const int C = N / 2 + 1;
    const int ldT = (C + 3) & ~3;
    cpx<T>* Tm = w.Tm.get<cpx<T>>(size_t(2 * Rs) * ldT);
    const cpx<T>* tw = twiddles<T>(N);
    ra.N = N;
    ra.Rs = Rs;
    ra.ldT = ldT;
    ra.tw = tw;
    ra.Tm = Tm;
    ColArgs<T> ca{Tm, N, Rs, ldT, tw, S_out, ldS};
    const int L = fast_log2<T>(N);
    if (L > 0) {
        t_timer.mark(ST_ROW, s);
        launch_rowpass<T, SRC>(L, Rs, s, ra);
        t_timer.mark(ST_COL, s);
        launch_colpass<T, OUT>(L, 8 * ((C + 7) / 8), s, ca);
    } else {
        const size_t lds_row = sizeof(T) * size_t(2 * Rs);
        if (lds_row > 160 * 1024) fail(GALARIO_ERR_INVALID, "Error in fft: image size not supported");
        allow_lds(k_rowdft_generic<T, SRC>, lds_row);
        hipLaunchKernelGGL((k_rowdft_generic<T, SRC>), dim3(2 * Rs), dim3(256), lds_row, s, ra);
        const size_t lds_col = sizeof(cpx<T>) * size_t(2 * Rs);
        if (lds_col > 160 * 1024) fail(GALARIO_ERR_INVALID, "Error in fft: image size not supported");
        allow_lds(k_coldft_generic<T, OUT>, lds_col);
        hipLaunchKernelGGL((k_coldft_generic<T, OUT>), dim3(C), dim3(256), lds_col, s, ca);
    }
    HC(hipGetLastError());
    return S_out;
}

// uv extent of a dataset whose points lie within uvmax of the origin (uv
// units; <= 0: not given).  Rotation keeps |(u, v)|, so at any PA a point's
// grid position lies within rho = uvmax / duv cells of the origin (rho2: with
// slack for the rotation's rounding) and its interpolation corners within
// rho + sqrt(2): the half spectrum is stored, and its adjoint read, only
// within rc = rho + 3 cells.
struct Extent {
    double rho2 = 0.0, rc2 = 0.0;
};

template <typename T>
Extent uv_extent(double uvmax, double duv) {
    Extent e;
    if (!(uvmax > 0.0) || !std::isfinite(uvmax) || !(duv > 0.0)) return e;
    const double rel = sizeof(T) == 8 ? 1e-12 : 1e-5;
    const double rho = uvmax / duv * (1.0 + rel) + 1e-6;
    e.rho2 = rho * rho;
    e.rc2 = (rho + 3.0) * (rho + 3.0);
    return e;
}

// Real, even half spectrum H of a swept profile (see k_rowpass_sym); returns
// false when the workgroup FFT does not cover N (the caller then uses the
// general complex path).
struct SymGeom {
    int L, C, Rs, nrows, pairs_r, pairs_c, ldT, ldH;
};

SymGeom sym_geom(int L, int N, int rmax) {
    SymGeom g;
    g.L = L;
    g.C = N / 2 + 1;
    g.Rs = rmax + 1 < N / 2 ? rmax + 1 : N / 2;      // symmetric support half-width
    g.nrows = g.Rs + 1;
    g.pairs_r = (g.nrows + 1) / 2;
    g.pairs_c = (g.C + 1) / 2;
    g.ldT = (g.C + 7) & ~7;                           // reals per row of Ts (2 pairs_r rows, ts_index)
    g.ldH = (N / 2 + 1 + 7) & ~7;
    return g;
}

// Real, even half spectrum H of a swept profile (see k_rowpass_sym); returns
// false when the workgroup FFT does not cover N (the caller then uses the
// general complex path).  zero/nzero: doubles the row pass clears (gradient
// accumulators of the same call).
template <typename T>
bool build_spectrum_sym(Workspace& w, hipStream_t s, const SweepPrm<T>& sw, int N, int rmax, const T** H_out,
                        int* ldH_out, double* zero = nullptr, int nzero = 0, double* zero1 = nullptr,
                        double rc2 = 0.0) {
    const int L = fast_log2<T>(N);
    if (L < 0) return false;
    const SymGeom g = sym_geom(L, N, rmax);
    T* Ts = w.Tm.get<T>(size_t(2 * g.pairs_r) * g.ldT);      // rows 0..nrows-1 and a padding row
    T* H = w.S.get<T>(size_t(g.C + 1) * g.ldH);
    const cpx<T>* tw = twiddles<T>(N);
    SymRowArgs<T> ra{sw, N, g.Rs, g.nrows, g.ldT, tw, Ts, zero, nzero, zero1};
    SymColArgs<T> ca{Ts, N, g.Rs, g.ldT, tw, H, g.ldH, rc2, residue_fits(L, g.Rs) ? twiddles<T>(512) : nullptr};
    t_timer.mark(ST_ROW, s);
    for (int r = t_timer.reps(ST_ROW); r > 0; --r) launch_rowpass_sym<T>(L, g.pairs_r, s, ra);
    t_timer.mark(ST_COL, s);
    for (int r = t_timer.reps(ST_COL); r > 0; --r) launch_colpass_sym<T>(L, N / 4, s, ca);   // N/4 pairs + column N/2
    HC(hipGetLastError());
    *H_out = H;
    *ldH_out = g.ldH;
    return true;
}

// ------------------------------------------------------------------ setup of a profile evaluation
template <typename T>
struct ProfileSetup {
    SweepPrm<T> sw;
    int rmax;
};

// Validation and derived scalars of create_image_h (src/galario.cpp:1186-1235),
// in dreal arithmetic as the reference computes them.
template <typename T>
ProfileSetup<T> setup_profile(int nr, T Rmin, T dR, T dxy, int nxy, T inc, int check_central) {
    check_input(nxy);
    if (!(dR > T(0)) || !(dxy > T(0)) || !std::isfinite(double(dR)) || !std::isfinite(double(dxy)))
        invalid("create image", "expect dR > 0 and dxy > 0");
    if (check_central) {
        const T ratio = T((double(dxy) / 2. - double(Rmin)) / double(dR));
        if (ratio < 5)
            invalid("create image",
                    "Expect (dxy/2-Rmin)/dR > 5, i.e. dR small enough to ensure reliable interpolation the inside "
                    "central pixel but got (dxy/2-Rmin)/dR = " +
                        std::to_string(ratio) + ". Try reducing dR.");
    }
    const double fin = std::floor(double(T(T(dxy / T(2)) - Rmin) / dR));
    if (!(fin >= 0.0) || !(fin + 1.0 < double(nr)))
        invalid("create image", "the central pixel needs intensity[floor((dxy/2-Rmin)/dR) + 1] but got " +
                                    std::to_string(nr) + " intensity values and floor((dxy/2-Rmin)/dR) = " +
                                    std::to_string(fin));
    ProfileSetup<T> ps{};
    ps.sw.nr = nr;
    ps.sw.Rmin = Rmin;
    ps.sw.dR = dR;
    ps.sw.dxy = dxy;
    ps.sw.cos_inc = T(std::cos(double(inc)));
    ps.sw.sr_to_px = dxy * dxy;
    ps.sw.iIN = int(fin);
    const double rm = std::ceil(double((Rmin + T(nr) * dR) / dxy));
    int rmax = rm > double(nxy / 2) ? nxy / 2 : (rm < 0.0 ? 0 : int(rm));
    ps.sw.rmax = rmax;
    ps.rmax = rmax;
    return ps;
}

template <typename T>
struct SampleSetup {
    int rotate;
    T cosPA, sinPA;
    int phase;
    T dRA2, dDec2;
    T dRAt, dDect;   // 2 dRArot, 2 dDecrot: the phase argument in half turns (the double kernels' sincospi)
    T dRArot, dDecrot;
};

// uv_rotate_h (src/galario.cpp:976-999) scalars and apply_phase_sampled_h's
// (908-921) skip test and 2*pi scaling, in dreal arithmetic.
template <typename T>
SampleSetup<T> setup_sample(T PA, T dRA, T dDec) {
    SampleSetup<T> s{};
    if (PA == T(0)) {
        s.rotate = 0;
        s.cosPA = T(1);
        s.sinPA = T(0);
        s.dRArot = dRA;
        s.dDecrot = dDec;
    } else {
        s.rotate = 1;
        s.cosPA = T(std::cos(double(PA)));
        s.sinPA = T(std::sin(double(PA)));
        s.dRArot = dRA * s.cosPA - dDec * s.sinPA;
        s.dDecrot = dRA * s.sinPA + dDec * s.cosPA;
    }
    s.phase = !(s.dRArot == T(0) && s.dDecrot == T(0));
    s.dRA2 = T(double(s.dRArot) * (2. * double(T(M_PI))));
    s.dDec2 = T(double(s.dDecrot) * (2. * double(T(M_PI))));
    s.dRAt = T(2) * s.dRArot;
    s.dDect = T(2) * s.dDecrot;
    return s;
}

template <typename T>
SampleArgs<T> sample_args(const cpx<T>* S, long long rs, long long cs, int nrow, int ncol, T v_origin, T duv,
                          const SampleSetup<T>& ss, int nd, const T* u, const T* v) {
    SampleArgs<T> a{};
    a.S = S;
    a.H = nullptr;
    a.ldH = 0;
    a.rs = rs;
    a.cs = cs;
    a.nrow = nrow;
    a.ncol = ncol;
    a.v_origin = v_origin;
    a.duv = duv;
    a.rotate = ss.rotate;
    a.cosPA = ss.cosPA;
    a.sinPA = ss.sinPA;
    a.phase = ss.phase;
    a.dRA2 = ss.dRA2;
    a.dDec2 = ss.dDec2;
    a.dRAt = ss.dRAt;
    a.dDect = ss.dDect;
    a.inv_duv = T(1) / duv;
    a.nd = nd;
    a.u = u;
    a.v = v;
    return a;
}

// chi2 of sample args `a` into chi2_d[0] (double) and, if given, chi2_t[0].
template <typename T>
void launch_chi2(Workspace& w, hipStream_t s, SampleArgs<T> a, const T* re, const T* im, const T* wt,
                 double* chi2_d, T* chi2_t) {
    const unsigned nb = blocks_for(size_t(a.nd), 256);
    a.re = re;
    a.im = im;
    a.w = wt;
    a.partials = w.part.get<double>(nb ? nb : 1);
    t_timer.mark(ST_SAMPLE, s);
    const unsigned ng = (nb + SAMPLE_PPT - 1) / SAMPLE_PPT;
    for (int r = t_timer.reps(ST_SAMPLE); r > 0; --r) {
        if (nb && a.H) hipLaunchKernelGGL((k_sample_chi2_n<T, true, SAMPLE_PPT>), dim3(ng), dim3(256), 0, s, a, int(nb));
        else if (nb) hipLaunchKernelGGL((k_sample_chi2_n<T, false, SAMPLE_PPT>), dim3(ng), dim3(256), 0, s, a, int(nb));
    }
    t_timer.mark(ST_REDUCE, s);
    hipLaunchKernelGGL((k_sum_partials<T>), dim3(1), dim3(1024), 0, s, a.partials, int(nb), chi2_d, chi2_t);
    HC(hipGetLastError());
}

template <typename T>
void launch_vis(hipStream_t s, SampleArgs<T> a, cpx<T>* vis) {
    a.vis = vis;
    const unsigned nb = blocks_for(size_t(a.nd), 256);
    t_timer.mark(ST_SAMPLE, s);
    if (nb && a.H) hipLaunchKernelGGL((k_sample<T, MODE_VIS, true>), dim3(nb), dim3(256), 0, s, a);
    else if (nb) hipLaunchKernelGGL((k_sample<T, MODE_VIS, false>), dim3(nb), dim3(256), 0, s, a);
    HC(hipGetLastError());
}

// Spectrum of a swept profile and the sample arguments that read it: the
// real, even half spectrum when the workgroup FFT covers nxy, else the
// general complex spectrum.
template <typename T>
SampleArgs<T> profile_spectrum(Workspace& w, hipStream_t s, ProfileSetup<T>& ps, const SampleSetup<T>& ss, int N,
                               T duv, int nd, const T* u, const T* v, double uvmax = 0.0) {
    const int C = N / 2 + 1;
    const T* H = nullptr;
    int ldH = 0;
    const Extent ex = uv_extent<T>(uvmax, double(duv));
    if (build_spectrum_sym<T>(w, s, ps.sw, N, ps.rmax, &H, &ldH, nullptr, 0, nullptr, ex.rc2)) {
        SampleArgs<T> sa = sample_args<T>(nullptr, 0, 0, N, C, T(1), duv, ss, nd, u, v);
        sa.H = H;
        sa.ldH = ldH;
        sa.rho2 = ex.rho2;
        return sa;
    }
    RowArgs<T> ra{};
    ra.sw = ps.sw;
    cpx<T>* S = w.S.get<cpx<T>>(size_t(C) * N);
    build_spectrum<T, SRC_SWEEP, OUT_COLMAJOR_SHIFT>(w, s, ra, N, ps.rmax > 0 ? ps.rmax : 1, S, size_t(N));
    return sample_args<T>(S, 1, N, N, C, T(1), duv, ss, nd, u, v);
}

// ------------------------------------------------------------------ host transfer helpers
template <typename X>
X* upload(DevBuf& b, const X* h, size_t n, hipStream_t s) {
    X* d = b.get<X>(n);
    if (n) HC(hipMemcpyAsync(d, h, n * sizeof(X), hipMemcpyHostToDevice, s));
    return d;
}

template <typename X>
void download(X* h, const X* d, size_t n, hipStream_t s) {
    if (n) HC(hipMemcpyAsync(h, d, n * sizeof(X), hipMemcpyDeviceToHost, s));
}

// Synchronise the own stream and raise if a point fell outside the grid.
void finish_host_call(Workspace& w, hipStream_t s, const unsigned* flag_d) {
    t_timer.end_call(s);
    unsigned* hf = reinterpret_cast<unsigned*>(w.h_scal + 1);
    HC(hipMemcpyAsync(hf, flag_d, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    if (*hf)
        invalid("interpolate",
                "at least one (u, v) point falls outside the Fourier grid (|u|/duv must not exceed ncol-1 and "
                "nrow/2 + v/duv must lie in [0, nrow-1]); increase nxy or reduce dxy");
}

unsigned* reset_host_flag(Workspace& w, hipStream_t s) {
    unsigned* f = w.flag_host.get<unsigned>(1);
    HC(hipMemsetAsync(f, 0, sizeof(unsigned), s));
    return f;
}

// ------------------------------------------------------------------ main functions (host pointers)
template <typename T>
int sample_profile_host(int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc, T dRA, T dDec, T duv, T PA,
                        int nd, const T* u, const T* v, T* vis, int check_central) {
    return guard([&] {
        check_nd(nd);
        ProfileSetup<T> ps = setup_profile<T>(nr, Rmin, dR, dxy, nxy, inc, check_central);
        SampleSetup<T> ss = setup_sample<T>(PA, dRA, dDec);
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        unsigned* flag = reset_host_flag(w, s);
        ps.sw.I = upload(w.in[0], I, size_t(nr), s);
        const T* du = upload(w.in[1], u, size_t(nd), s);
        const T* dv = upload(w.in[2], v, size_t(nd), s);
        SampleArgs<T> sa = profile_spectrum<T>(w, s, ps, ss, nxy, duv, nd, du, dv);
        sa.flag = flag;
        cpx<T>* dvis = w.out.get<cpx<T>>(size_t(nd));
        launch_vis(s, sa, dvis);
        download(reinterpret_cast<cpx<T>*>(vis), dvis, size_t(nd), s);
        finish_host_call(w, s, flag);
    });
}

template <typename T>
int chi2_profile_host(int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc, T dRA, T dDec, T duv, T PA, int nd,
                      const T* u, const T* v, const T* re, const T* im, const T* wt, int check_central, T* chi2) {
    return guard([&] {
        check_nd(nd);
        ProfileSetup<T> ps = setup_profile<T>(nr, Rmin, dR, dxy, nxy, inc, check_central);
        SampleSetup<T> ss = setup_sample<T>(PA, dRA, dDec);
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        unsigned* flag = reset_host_flag(w, s);
        ps.sw.I = upload(w.in[0], I, size_t(nr), s);
        const T* du = upload(w.in[1], u, size_t(nd), s);
        const T* dv = upload(w.in[2], v, size_t(nd), s);
        const T* dre = upload(w.in[3], re, size_t(nd), s);
        const T* dim = upload(w.in[4], im, size_t(nd), s);
        const T* dw = upload(w.in[5], wt, size_t(nd), s);
        SampleArgs<T> sa = profile_spectrum<T>(w, s, ps, ss, nxy, duv, nd, du, dv);
        sa.flag = flag;
        double* res = w.scal.get<double>(4);
        launch_chi2<T>(w, s, sa, dre, dim, dw, res, nullptr);
        download(w.h_scal, res, 1, s);
        finish_host_call(w, s, flag);
        *chi2 = T(w.h_scal[0]);
    });
}

template <typename T>
int sample_image_host(int nx, int ny, const T* image, T v_origin, T dRA, T dDec, T duv, T PA, int nd, const T* u,
                      const T* v, T* vis) {
    return guard([&] {
        check_nd(nd);
        check_input_xy(nx, ny);
        SampleSetup<T> ss = setup_sample<T>(PA, dRA, dDec);
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        unsigned* flag = reset_host_flag(w, s);
        const int N = nx, C = N / 2 + 1;
        RowArgs<T> ra{};
        ra.img = upload(w.in[0], image, size_t(N) * N, s);
        ra.ldimg = N;
        const T* du = upload(w.in[1], u, size_t(nd), s);
        const T* dv = upload(w.in[2], v, size_t(nd), s);
        cpx<T>* S = w.S.get<cpx<T>>(size_t(C) * N);
        build_spectrum<T, SRC_IMAGE_SHIFT, OUT_COLMAJOR_SHIFT>(w, s, ra, N, N / 2, S, size_t(N));
        SampleArgs<T> sa = sample_args<T>(S, 1, N, N, C, v_origin, duv, ss, nd, du, dv);
        sa.flag = flag;
        cpx<T>* dvis = w.out.get<cpx<T>>(size_t(nd));
        launch_vis(s, sa, dvis);
        download(reinterpret_cast<cpx<T>*>(vis), dvis, size_t(nd), s);
        finish_host_call(w, s, flag);
    });
}

template <typename T>
int chi2_image_host(int nx, int ny, const T* image, T v_origin, T dRA, T dDec, T duv, T PA, int nd, const T* u,
                    const T* v, const T* re, const T* im, const T* wt, T* chi2) {
    return guard([&] {
        check_nd(nd);
        check_input_xy(nx, ny);
        SampleSetup<T> ss = setup_sample<T>(PA, dRA, dDec);
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        unsigned* flag = reset_host_flag(w, s);
        const int N = nx, C = N / 2 + 1;
        RowArgs<T> ra{};
        ra.img = upload(w.in[0], image, size_t(N) * N, s);
        ra.ldimg = N;
        const T* du = upload(w.in[1], u, size_t(nd), s);
        const T* dv = upload(w.in[2], v, size_t(nd), s);
        const T* dre = upload(w.in[3], re, size_t(nd), s);
        const T* dim = upload(w.in[4], im, size_t(nd), s);
        const T* dw = upload(w.in[5], wt, size_t(nd), s);
        cpx<T>* S = w.S.get<cpx<T>>(size_t(C) * N);
        build_spectrum<T, SRC_IMAGE_SHIFT, OUT_COLMAJOR_SHIFT>(w, s, ra, N, N / 2, S, size_t(N));
        SampleArgs<T> sa = sample_args<T>(S, 1, N, N, C, v_origin, duv, ss, nd, du, dv);
        sa.flag = flag;
        double* res = w.scal.get<double>(4);
        launch_chi2<T>(w, s, sa, dre, dim, dw, res, nullptr);
        download(w.h_scal, res, 1, s);
        finish_host_call(w, s, flag);
        *chi2 = T(w.h_scal[0]);
    });
}

template <typename T>
int sweep_host(int nr, const T* I, T Rmin, T dR, int nxy, T dxy, T inc, T* image, int check_central) {
    return guard([&] {
        ProfileSetup<T> ps = setup_profile<T>(nr, Rmin, dR, dxy, nxy, inc, check_central);
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        ps.sw.I = upload(w.in[0], I, size_t(nr), s);
        const int ld = 2 * (nxy / 2 + 1);
        T* dimg = w.out.get<T>(size_t(nxy) * ld);
        hipLaunchKernelGGL((k_sweep_image<T>), dim3(blocks_for(size_t(nxy) * ld, 256)), dim3(256), 0, s, ps.sw, nxy,
                           ld, dimg);
        HC(hipGetLastError());
        download(image, dimg, size_t(nxy) * ld, s);
        HC(hipStreamSynchronize(s));
    });
}

template <typename T>
int uv_rotate_host(T PA, T dRA, T dDec, T* dRArot, T* dDecrot, int nd, const T* u, const T* v, T* urot, T* vrot) {
    return guard([&] {
        check_nd(nd);
        SampleSetup<T> ss = setup_sample<T>(PA, dRA, dDec);
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        const T* du = upload(w.in[1], u, size_t(nd), s);
        const T* dv = upload(w.in[2], v, size_t(nd), s);
        T* dur = w.in[3].get<T>(size_t(nd));
        T* dvr = w.in[4].get<T>(size_t(nd));
        if (ss.rotate) {
            if (nd)
                hipLaunchKernelGGL((k_rotate<T>), dim3(blocks_for(size_t(nd), 256)), dim3(256), 0, s, nd, ss.cosPA,
                                   ss.sinPA, du, dv, dur, dvr);
            HC(hipGetLastError());
        } else if (nd) {
            HC(hipMemcpyAsync(dur, du, size_t(nd) * sizeof(T), hipMemcpyDeviceToDevice, s));
            HC(hipMemcpyAsync(dvr, dv, size_t(nd) * sizeof(T), hipMemcpyDeviceToDevice, s));
        }
        download(urot, dur, size_t(nd), s);
        download(vrot, dvr, size_t(nd), s);
        HC(hipStreamSynchronize(s));
        *dRArot = ss.dRArot;
        *dDecrot = ss.dDecrot;
    });
}

template <typename T>
int copy_input_host(int nx, int ny, const T* realdata, void** out) {
    return guard([&] {
        check_input_xy(nx, ny);
        const size_t ld = size_t(2 * (ny / 2 + 1));
        T* buf = static_cast<T*>(std::calloc(size_t(nx) * ld, sizeof(T)));
        if (!buf) throw std::bad_alloc();
        for (int i = 0; i < nx; ++i) std::memcpy(buf + size_t(i) * ld, realdata + size_t(i) * ny, sizeof(T) * ny);
        *out = buf;
    });
}

template <typename T>
int fft2d_host(int nx, int ny, void* data) {
    return guard([&] {
        check_input_xy(nx, ny);
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        const int N = nx, C = N / 2 + 1, ld = 2 * C;
        RowArgs<T> ra{};
        ra.img = upload(w.in[0], static_cast<const T*>(data), size_t(N) * ld, s);
        ra.ldimg = ld;
        cpx<T>* S = w.S.get<cpx<T>>(size_t(C) * N);
        build_spectrum<T, SRC_IMAGE_PLAIN, OUT_ROWMAJOR_PLAIN>(w, s, ra, N, N / 2, S, size_t(C));
        download(static_cast<cpx<T>*>(data), S, size_t(C) * N, s);
        HC(hipStreamSynchronize(s));
    });
}

template <typename T>
int fftshift_host(int nx, int ny, void* data) {
    return guard([&] {
        check_input_xy(nx, ny);
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        const int N = nx, ld = 2 * (N / 2 + 1);
        const size_t n = size_t(N) * ld;
        const T* din = upload(w.in[0], static_cast<const T*>(data), n, s);
        T* dout = w.out.get<T>(n);
        hipLaunchKernelGGL((k_fftshift_padded<T>), dim3(blocks_for(n, 256)), dim3(256), 0, s, N, ld, din, dout);
        HC(hipGetLastError());
        download(static_cast<T*>(data), dout, n, s);
        HC(hipStreamSynchronize(s));
    });
}

template <typename T>
int fftshift_axis0_host(int nrow, int ncol, void* data) {
    return guard([&] {
        check_input(nrow);
        if (ncol < 0) invalid("check input", "ncol = " + std::to_string(ncol) + " is negative");
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        const size_t n = size_t(nrow) * ncol;
        cpx<T>* d = upload(w.in[0], static_cast<const cpx<T>*>(data), n, s);
        const size_t half = size_t(nrow / 2) * ncol;
        if (half)
            hipLaunchKernelGGL((k_fftshift_axis0<T>), dim3(blocks_for(half, 256)), dim3(256), 0, s, nrow, ncol, d);
        HC(hipGetLastError());
        download(static_cast<cpx<T>*>(data), d, n, s);
        HC(hipStreamSynchronize(s));
    });
}

template <typename T>
int interpolate_host(int nrow, int ncol, const void* data, T v_origin, int nd, const T* u, const T* v, T duv,
                     void* vis) {
    return guard([&] {
        check_nd(nd);
        if (nrow < 2 || ncol < 2)
            invalid("interpolate", "expect a matrix of at least 2 x 2 but got (" + std::to_string(nrow) + ", " +
                                       std::to_string(ncol) + ")");
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        unsigned* flag = reset_host_flag(w, s);
        const cpx<T>* dS = upload(w.in[0], static_cast<const cpx<T>*>(data), size_t(nrow) * ncol, s);
        const T* du = upload(w.in[1], u, size_t(nd), s);
        const T* dv = upload(w.in[2], v, size_t(nd), s);
        SampleSetup<T> ss{};   // no rotation, no phase
        SampleArgs<T> sa = sample_args<T>(dS, ncol, 1, nrow, ncol, v_origin, duv, ss, nd, du, dv);
        sa.flag = flag;
        cpx<T>* dvis = w.out.get<cpx<T>>(size_t(nd));
        launch_vis(s, sa, dvis);
        download(static_cast<cpx<T>*>(vis), dvis, size_t(nd), s);
        finish_host_call(w, s, flag);
    });
}

template <typename T>
int apply_phase_host(T dRA, T dDec, int nd, const T* u, const T* v, void* vis) {
    return guard([&] {
        check_nd(nd);
        if (dRA == T(0) && dDec == T(0)) return;   // apply_phase_sampled_h returns early
        const T dRA2 = T(double(dRA) * (2. * double(T(M_PI))));
        const T dDec2 = T(double(dDec) * (2. * double(T(M_PI))));
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        const T* du = upload(w.in[1], u, size_t(nd), s);
        const T* dv = upload(w.in[2], v, size_t(nd), s);
        cpx<T>* dvis = upload(w.out, static_cast<const cpx<T>*>(vis), size_t(nd), s);
        if (nd)
            hipLaunchKernelGGL((k_apply_phase<T>), dim3(blocks_for(size_t(nd), 256)), dim3(256), 0, s, nd, dRA2,
                               dDec2, du, dv, dvis);
        HC(hipGetLastError());
        download(static_cast<cpx<T>*>(vis), dvis, size_t(nd), s);
        HC(hipStreamSynchronize(s));
    });
}

template <typename T>
int reduce_chi2_host(int nd, const T* re, const T* im, const void* vis, const T* wt, T* chi2) {
    return guard([&] {
        check_nd(nd);
        Workspace& w = workspace(nullptr, true);
        hipStream_t s = w.own;
        const T* dre = upload(w.in[3], re, size_t(nd), s);
        const T* dim = upload(w.in[4], im, size_t(nd), s);
        const T* dw = upload(w.in[5], wt, size_t(nd), s);
        const cpx<T>* dvis = upload(w.out, static_cast<const cpx<T>*>(vis), size_t(nd), s);
        const unsigned nb = blocks_for(size_t(nd), 256);
        double* res = w.scal.get<double>(4);
        SampleArgs<T> a{};
        a.partials = w.part.get<double>(nb ? nb : 1);
        if (nb) hipLaunchKernelGGL((k_chi2_terms<T>), dim3(nb), dim3(256), 0, s, nd, dre, dim, dvis, dw, a);
        hipLaunchKernelGGL((k_sum_partials<T>), dim3(1), dim3(1024), 0, s, a.partials, int(nb), res, (T*)nullptr);
        HC(hipGetLastError());
        download(w.h_scal, res, 1, s);
        HC(hipStreamSynchronize(s));
        *chi2 = T(w.h_scal[0]);
    });
}

// ------------------------------------------------------------------ device-resident forms
template <typename T>
int chi2_profile_dev(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc, T dRA, T dDec, T duv,
                     T PA, int nd, const T* u, const T* v, const T* re, const T* im, const T* wt, int check_central,
                     double* chi2_d, double uvmax) {
    return guard([&] {
        check_nd(nd);
        ProfileSetup<T> ps = setup_profile<T>(nr, Rmin, dR, dxy, nxy, inc, check_central);
        SampleSetup<T> ss = setup_sample<T>(PA, dRA, dDec);
        check_dev_inputs(I, nd, u, v, true, re, im, wt);
        if (!chi2_d) invalid("check input", "null chi2 pointer");
        hipStream_t s = static_cast<hipStream_t>(stream);
        Workspace& w = workspace(s, false);
        ps.sw.I = I;
        SampleArgs<T> sa = profile_spectrum<T>(w, s, ps, ss, nxy, duv, nd, u, v, uvmax);
        sa.flag = w.flag_dev.get<unsigned>(1);
        launch_chi2<T>(w, s, sa, re, im, wt, chi2_d, nullptr);
        t_timer.end_call(s);
    });
}

template <typename T>
int sample_profile_dev(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc, T dRA, T dDec, T duv,
                       T PA, int nd, const T* u, const T* v, int check_central, const int* perm, void* vis_d,
                       double uvmax) {
    return guard([&] {
        check_nd(nd);
        ProfileSetup<T> ps = setup_profile<T>(nr, Rmin, dR, dxy, nxy, inc, check_central);
        SampleSetup<T> ss = setup_sample<T>(PA, dRA, dDec);
        check_dev_inputs<T>(I, nd, u, v, false, nullptr, nullptr, nullptr);
        if (nd > 0 && !vis_d) invalid("check input", "null visibility pointer");
        hipStream_t s = static_cast<hipStream_t>(stream);
        Workspace& w = workspace(s, false);
        ps.sw.I = I;
        SampleArgs<T> sa = profile_spectrum<T>(w, s, ps, ss, nxy, duv, nd, u, v, uvmax);
        sa.flag = w.flag_dev.get<unsigned>(1);
        sa.perm = perm;
        launch_vis(s, sa, static_cast<cpx<T>*>(vis_d));
        t_timer.end_call(s);
    });
}

// Several sources sharing the profile, inc and PA at their own offsets and
// fluxes (src: nsrc x 3 = flux, dRA, dDec, device): visibilities (vis_d,
// written through perm when given) or chi2 (chi2_d).
template <typename T>
int profile_multisrc_dev(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc, T duv, T PA, int nsrc,
                         const T* src, int nd, const T* u, const T* v, const T* re, const T* im, const T* wt,
                         int check_central, const int* perm, void* vis_d, double* chi2_d, double uvmax) {
    return guard([&] {
        check_nd(nd);
        if (nsrc < 0) invalid("multiple sources", "nsrc = " + std::to_string(nsrc) + " is negative");
        if (nsrc > 0 && !src) invalid("multiple sources", "null source array");
        ProfileSetup<T> ps = setup_profile<T>(nr, Rmin, dR, dxy, nxy, inc, check_central);
        SampleSetup<T> ss = setup_sample<T>(PA, T(0), T(0));
        check_dev_inputs(I, nd, u, v, chi2_d != nullptr, re, im, wt);
        if (!chi2_d && nd > 0 && !vis_d) invalid("check input", "null visibility pointer");
        hipStream_t s = static_cast<hipStream_t>(stream);
        Workspace& w = workspace(s, false);
        ps.sw.I = I;
        SampleArgs<T> sa = profile_spectrum<T>(w, s, ps, ss, nxy, duv, nd, u, v, uvmax);
        sa.flag = w.flag_dev.get<unsigned>(1);
        const unsigned nb = blocks_for(size_t(nd), 256);
        if (chi2_d) {
            sa.re = re;
            sa.im = im;
            sa.w = wt;
            sa.partials = w.part.get<double>(nb ? nb : 1);
            t_timer.mark(ST_SAMPLE, s);
            if (nb)
                hipLaunchKernelGGL((k_sample_multisrc<T, MODE_CHI2, SAMPLE_PPT>), dim3((nb + SAMPLE_PPT - 1) / SAMPLE_PPT),
                                   dim3(256), 0, s, sa, nsrc, src, int(nb));
            t_timer.mark(ST_REDUCE, s);
            hipLaunchKernelGGL((k_sum_partials<T>), dim3(1), dim3(1024), 0, s, sa.partials, int(nb), chi2_d,
                               static_cast<T*>(nullptr));
        } else {
            sa.vis = static_cast<cpx<T>*>(vis_d);
            sa.perm = perm;
            t_timer.mark(ST_SAMPLE, s);
            if (nb)
                hipLaunchKernelGGL((k_sample_multisrc<T, MODE_VIS, SAMPLE_PPT>), dim3((nb + SAMPLE_PPT - 1) / SAMPLE_PPT),
                                   dim3(256), 0, s, sa, nsrc, src, int(nb));
        }
        HC(hipGetLastError());
        t_timer.end_call(s);
    });
}

// chi2 and its gradient with respect to the profile (grad_I, nr doubles) and
// to inc, PA, dRA, dDec (grad_geom[0..3]), all on the device.
template <typename T>
int chi2_profile_grad_dev(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc, T dRA, T dDec,
                          T duv, T PA, int nd, const T* u, const T* v, const T* re, const T* im, const T* wt,
                          int check_central, double* chi2_d, double* grad_I, double* grad_geom, double uvmax) {
    return guard([&] {
        check_nd(nd);
        ProfileSetup<T> ps = setup_profile<T>(nr, Rmin, dR, dxy, nxy, inc, check_central);
        SampleSetup<T> ss = setup_sample<T>(PA, dRA, dDec);
        const int L = fast_log2<T>(nxy);
        if (L < 0)
            invalid("chi2 gradient", "nxy = " + std::to_string(nxy) +
                                         " is not supported by the gradient (power of two from 128 to " +
                                         std::to_string(1 << max_fast_log2<T>()) + ")");
        if (nr > ADJROW_LDS_NR)
            invalid("chi2 gradient", "nr = " + std::to_string(nr) + " profile bins (the gradient supports up to " +
                                         std::to_string(ADJROW_LDS_NR) + ")");
        if (!grad_I || !grad_geom || !chi2_d) invalid("chi2 gradient", "null output pointer");
        check_dev_inputs(I, nd, u, v, true, re, im, wt);
        hipStream_t s = static_cast<hipStream_t>(stream);
        Workspace& w = workspace(s, false);
        ps.sw.I = I;
        const SymGeom g = sym_geom(L, nxy, ps.rmax);
        // Hh is kept zero between calls (the adjoint column pass clears what it read); it is
        // cleared here only when new or after a call that did not reach that pass
        // (followed by the replicas of the hot tile, which that pass folds in and clears)
        const size_t nHh = size_t(g.C + 1) * g.ldH, nAll = nHh + GHOT_ELEMS;
        if (w.Hh.cap < nAll * sizeof(double)) w.Hh_zero = false;
        double* Hh = w.Hh.get<double>(nAll);
        double* hot = Hh + nHh;
        if (!w.Hh_zero) {
            hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(nAll, 256)), dim3(256), 0, s, Hh, nAll);
            HC(hipGetLastError());
            w.Hh_zero = true;
        }
        const T* H = nullptr;
        int ldH = 0;
        // the forward row pass clears the accumulators grad_I[0..nr) and grad_geom[0] (inc)
        const Extent ex = uv_extent<T>(uvmax, double(duv));
        build_spectrum_sym<T>(w, s, ps.sw, nxy, ps.rmax, &H, &ldH, grad_I, nr, grad_geom, ex.rc2);
        SampleArgs<T> sa = sample_args<T>(nullptr, 0, 0, nxy, g.C, T(1), duv, ss, nd, u, v);
        sa.H = H;
        sa.ldH = ldH;
        sa.rho2 = ex.rho2;
        sa.re = re;
        sa.im = im;
        sa.w = wt;
        sa.flag = w.flag_dev.get<unsigned>(1);
        const unsigned nb = blocks_for(size_t(nd), 256);
        double* part = w.part.get<double>(size_t(4) * (nb ? nb : 1));
        GradArgs<T> ga{Hh, hot, part, int(nb)};
        t_timer.mark(ST_SAMPLE, s);
        w.Hh_zero = false;
        if (nb) hipLaunchKernelGGL((k_sample_grad_sym<T>), dim3(nb), dim3(256), 0, s, sa, ga);
        HC(hipGetLastError());
        // adjoint passes; workgroup 0 of the adjoint column pass also sums the partials into chi2 and
        // the PA, dRA, dDec gradients (the order of k_sum_partials)
        const int ldTh = (g.C + 1 + 7) & ~7;
        T* Th = w.Th.get<T>(size_t(g.nrows) * ldTh);
        const cpx<T>* tw = twiddles<T>(nxy);
        const Dst4 dst{{chi2_d, grad_geom + 1, grad_geom + 2, grad_geom + 3}};
        const int nfft = 8 * ((g.pairs_c + 7) / 8);
        AdjColArgs<T> aca{Hh,   hot,  g.ldH, nxy,  g.Rs,    tw,
                          Th,   ldTh, part,  int(nb), dst, nfft,
                          ex.rc2, residue_fits(L, g.Rs) ? twiddles<T>(512) : nullptr};
        t_timer.mark(ST_ADJCOL, s);
        launch_adjcol_sym<T>(L, nfft + 8, s, aca);   // + 8 workgroups for the partial sums
        HC(hipGetLastError());
        w.Hh_zero = true;
        AdjRowArgs<T> ara{ps.sw,  Th,        ldTh,     nxy, g.Rs, g.nrows, tw, std::tan(double(inc)),
                          grad_I, grad_geom, aca.tw512};   // tw512 set: the residue form where it fits
        t_timer.mark(ST_ADJROW, s);
        launch_adjrow_sym<T>(L, g.pairs_r, s, ara);
        HC(hipGetLastError());
        t_timer.end_call(s);
    });
}

// ====================================================================== batched form
// A plan holds the datasets (device pointers owned by the caller), their
// grouping by grid (nxy, dxy): one spectrum per walker and group, sampled by
// every dataset of the group), slot buffers for a chunk of walkers, and per
// (number of walkers, gradient) the chunk size picked by timing and the
// captured graph.  Calls copy the caller's parameters into plan-owned
// buffers, launch the graph and copy the results out, all on the caller's
// stream.
constexpr uint32_t BATCH_MAGIC = 0x67616c62u;
// Launch-sequence modes of a plan: bit 0 gradient, bit 1 FP32 bulk with FP64 geometry, phase and
// per-point arithmetic (float plans only).
constexpr int MODE_GRAD = 1, MODE_F64_PHASE = 2;
constexpr size_t BATCH_SLOT_BUDGET = size_t(16) << 30;   // bytes of slot buffers at most

struct BatchPlanBase {
    uint32_t magic = BATCH_MAGIC;
    int prec = 0;          // 8 double, 4 float
    virtual ~BatchPlanBase() = default;
};

template <typename T>
struct BatchSet {
    int nxy;
    T dxy, duv;
    int nd;
    const T *u, *v, *re, *im, *w;
    double uvmax;          // <= 0: not given
    double duv_d;          // duv in double (the FP32-bulk form's FP64 per-point arithmetic)
};

template <typename T>
struct BatchGroup {
    int N;
    T dxy;
    SymGeom g;
    SweepPrm<T> sw;        // shared scalars (I and cos_inc are per walker)
    std::vector<int> sets;
    Extent ex;             // the widest extent of its datasets (none if one has none)
    bool ex_all = true;    // every dataset so far gives its extent
};

template <typename T>
struct BatchPlan : BatchPlanBase {
    int device = 0;
    int nr = 0, max_w = 0;
    std::vector<BatchSet<T>> sets;
    std::vector<int> nb, off;
    int ldPart = 1;
    std::vector<BatchGroup<T>> groups;
    size_t sTs = 0, sH = 0, sHh = 0, hot_off = 0, sTh = 0;
    int slots_ts = 0, slots_grad = 0;   // slots allocated (forward buffers, gradient buffers)
    bool Hh_zero = false;
    DevBuf wp, I, geom, part, Ts, H, Hh, Th, out_sets, out_tot, gI, ggeom, flag;
    hipStream_t cap = nullptr;
    hipStream_t last = nullptr;   // stream of the previous call (its buffers are reused)
    bool has_last = false;
    hipEvent_t ev = nullptr;
    struct Exec {
        int chunk = 0;
        int kw = 4;                    // walkers per sample workgroup (forward)
        int ppt = 1;                   // 256-point blocks per sample workgroup (forward)
        int win = GWIN;                // LDS window cells of the gradient sample kernel
        bool tile_fixed = false;       // ppt / win set by the caller: the autotuner keeps them
        hipGraphExec_t exec = nullptr;
    };
    std::map<std::pair<int, int>, Exec> runs;   // (nwalkers, mode)
    std::mutex mu;

    ~BatchPlan() override {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(device);
        (void)hipDeviceSynchronize();   // work of the last call may still read the buffers
        for (auto& kv : runs)
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        for (DevBuf* b : {&wp, &I, &geom, &part, &Ts, &H, &Hh, &Th, &out_sets, &out_tot, &gI, &ggeom, &flag})
            b->release();
        if (cap) (void)hipStreamDestroy(cap);
        if (ev) (void)hipEventDestroy(ev);
        (void)hipSetDevice(cur);
        magic = 0;
    }

    size_t slot_bytes(bool grad) const {
        return sizeof(T) * (sTs + sH) + (grad ? sizeof(double) * sHh + sizeof(T) * sTh : 0);
    }
    int max_chunk(bool grad, int W) const {
        const size_t c = BATCH_SLOT_BUDGET / slot_bytes(grad);
        return int(std::max<size_t>(1, std::min<size_t>(c, size_t(W))));
    }
    // drop every captured graph (they hold the addresses of the slot buffers)
    void drop_graphs() {
        for (auto& kv : runs)
            if (kv.second.exec) {
                (void)hipGraphExecDestroy(kv.second.exec);
                kv.second.exec = nullptr;
            }
    }
    // grow slot buffers to `chunk` slots (outside any capture)
    void ensure_slots(int chunk, bool grad, hipStream_t s) {
        if (chunk > slots_ts || (grad && chunk > slots_grad)) {
            HC(hipStreamSynchronize(s));   // in-flight work may still use the buffers being replaced
            drop_graphs();
        }
        if (chunk > slots_ts) {
            Ts.get<T>(sTs * chunk);
            H.get<T>(sH * chunk);
            slots_ts = chunk;
        }
        if (grad && chunk > slots_grad) {
            Hh.get<double>(sHh * chunk);
            Th.get<T>(sTh * chunk);
            slots_grad = chunk;
            Hh_zero = false;
        }
        if (grad && !Hh_zero) {
            const size_t n = sHh * size_t(slots_grad);
            hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(n, 256)), dim3(256), 0, s, Hh.get<double>(n), n);
            HC(hipGetLastError());
            Hh_zero = true;
        }
    }
};

template <typename T>
BatchPtrs<T> batch_ptrs(BatchPlan<T>& P, bool grad) {
    BatchPtrs<T> b{};
    b.I = static_cast<const T*>(P.I.p);
    b.nr = P.nr;
    b.wp = static_cast<const WalkerPrm<T>*>(P.wp.p);
    b.Ts = static_cast<T*>(P.Ts.p);
    b.sTs = P.sTs;
    b.H = static_cast<T*>(P.H.p);
    b.sH = P.sH;
    if (grad) {
        b.Hh = static_cast<double*>(P.Hh.p);
        b.sHh = P.sHh;
        b.hot_off = P.hot_off;
        b.Th = static_cast<T*>(P.Th.p);
        b.sTh = P.sTh;
        b.gI = static_cast<double*>(P.gI.p);
        b.ggeom = static_cast<double*>(P.ggeom.p);
    }
    b.part = static_cast<double*>(P.part.p);
    b.ldPart = P.ldPart;
    return b;
}

// Enqueue one batched evaluation of W walkers (parameters already in P.I /
// P.geom) in chunks of `chunk` walkers; only kernel launches (capturable).
template <typename T, bool MIX, int KW>
void launch_sample_multi_kw(int ppt, int nb, int nw, hipStream_t s, const SampleArgs<T>& a, const BatchPtrs<T>& b) {
    const dim3 grid((nb + ppt - 1) / ppt, (nw + KW - 1) / KW);
    switch (ppt) {
        case 1: hipLaunchKernelGGL((k_sample_multi_b<T, KW, MIX, 1>), grid, dim3(256), 0, s, a, b, nw, nb); break;
        case 2: hipLaunchKernelGGL((k_sample_multi_b<T, KW, MIX, 2>), grid, dim3(256), 0, s, a, b, nw, nb); break;
        default: fail(GALARIO_ERR_RUNTIME, "internal: point blocks per sample workgroup must be 1 or 2");
    }
}

template <typename T, bool MIX>
void launch_sample_multi(int kw, int ppt, int nb, int nw, hipStream_t s, const SampleArgs<T>& a,
                         const BatchPtrs<T>& b) {
    switch (kw) {
        case 1: launch_sample_multi_kw<T, MIX, 1>(ppt, nb, nw, s, a, b); break;
        case 2: launch_sample_multi_kw<T, MIX, 2>(ppt, nb, nw, s, a, b); break;
        case 4: launch_sample_multi_kw<T, MIX, 4>(ppt, nb, nw, s, a, b); break;
        case 8: launch_sample_multi_kw<T, MIX, 8>(ppt, nb, nw, s, a, b); break;
        default: fail(GALARIO_ERR_RUNTIME, "internal: walkers per sample workgroup must be 1, 2, 4 or 8");
    }
}

template <typename T, bool MIX>
void launch_sample_grad(int win, int nb, int nw, hipStream_t s, const SampleArgs<T>& a, const BatchPtrs<T>& b) {
    const dim3 grid(nb, nw);
    switch (win) {
        case 2048: hipLaunchKernelGGL((k_sample_grad_sym_b<T, MIX, 2048>), grid, dim3(256), 0, s, a, b); break;
        case GWIN: hipLaunchKernelGGL((k_sample_grad_sym_b<T, MIX, GWIN>), grid, dim3(256), 0, s, a, b); break;
        case 4096: hipLaunchKernelGGL((k_sample_grad_sym_b<T, MIX, 4096>), grid, dim3(256), 0, s, a, b); break;
        default: fail(GALARIO_ERR_RUNTIME, "internal: gradient window must be 2048, 3072 or 4096 cells");
    }
}

// Launch shape of one batched evaluation.
struct BatchShape {
    int chunk, kw, ppt, win;
};

template <typename T>
void batch_enqueue(BatchPlan<T>& P, hipStream_t s, int W, const BatchShape& sh, int mode) {
    const int nr = P.nr, chunk = sh.chunk;
    constexpr bool is_f32 = std::is_same<T, float>::value;   // (the mixed form exists for float plans only)
    const bool grad = mode & MODE_GRAD, mix = is_f32 && (mode & MODE_F64_PHASE);
    const size_t nz = std::max(size_t(W), grad ? size_t(W) * size_t(std::max(nr, 4)) : size_t(W));
    if constexpr (is_f32) {
        if (mix)
            hipLaunchKernelGGL((k_batch_setup<T, double>), dim3(blocks_for(nz, 256)), dim3(256), 0, s, W,
                               static_cast<const double*>(P.geom.p), static_cast<WalkerPrm<T>*>(P.wp.p),
                               grad ? static_cast<double*>(P.gI.p) : nullptr, nr,
                               grad ? static_cast<double*>(P.ggeom.p) : nullptr);
    }
    if (!mix)
        hipLaunchKernelGGL((k_batch_setup<T, T>), dim3(blocks_for(nz, 256)), dim3(256), 0, s, W,
                           static_cast<const T*>(P.geom.p), static_cast<WalkerPrm<T>*>(P.wp.p),
                           grad ? static_cast<double*>(P.gI.p) : nullptr, nr,
                           grad ? static_cast<double*>(P.ggeom.p) : nullptr);
    BatchPtrs<T> b = batch_ptrs(P, grad);
    unsigned* flag = static_cast<unsigned*>(P.flag.p);
    for (int w0 = 0; w0 < W; w0 += chunk) {
        const int nw = std::min(chunk, W - w0);
        b.w0 = w0;
        for (const BatchGroup<T>& G : P.groups) {
            const SymGeom& g = G.g;
            const int L = g.L;
            const cpx<T>* tw = twiddles<T>(G.N);
            SymRowArgs<T> ra{G.sw, G.N, g.Rs, g.nrows, g.ldT, tw, nullptr, nullptr, 0, nullptr};
            SymColArgs<T> ca{nullptr, G.N, g.Rs,     g.ldT,
                             tw,      nullptr, g.ldH, G.ex.rc2,
                             residue_fits(L, g.Rs) ? twiddles<T>(512) : nullptr};
            const int cblocks = G.N / 4;                      // column pairs 0..N/2-1, column N/2 inside
            const int ablocks = 8 * ((g.pairs_c + 7) / 8);    // adjoint: (C+1)/2 pairs
#define GL_CALL(LL) \
    hipLaunchKernelGGL((k_rowpass_sym_b<T, LL>), dim3(g.pairs_r, nw), dim3(FftShape<LL>::NT), 0, s, ra, b)
            GL_L_SWITCH(T, GL_CALL)
#undef GL_CALL
            if (residue_fits(L, g.Rs)) {
#define GL_CALL(LL) hipLaunchKernelGGL((k_colpass_sub_b<T, LL>), dim3(cblocks, nw), dim3(ResidueShape<LL>::NT), 0, s, ca, b)
                GL_RESIDUE_SWITCH(GL_CALL)
#undef GL_CALL
            } else {
#define GL_CALL(LL) \
    hipLaunchKernelGGL((k_colpass_sym_b<T, LL>), dim3(cblocks, nw), dim3(FftShape<LL>::NT), 0, s, ca, b)
                GL_L_SWITCH(T, GL_CALL)
#undef GL_CALL
            }
            for (int si : G.sets) {
                const BatchSet<T>& d = P.sets[si];
                if (P.nb[si] == 0) continue;
                SampleArgs<T> a{};
                a.ldH = g.ldH;
                a.nrow = G.N;
                a.ncol = g.C;
                a.v_origin = T(1);
                a.duv = d.duv;
                a.inv_duv = T(1) / d.duv;
                a.nd = d.nd;
                a.u = d.u;
                a.v = d.v;
                a.re = d.re;
                a.im = d.im;
                a.w = d.w;
                a.flag = flag;
                a.rho2 = G.ex.rho2;
                a.duv_d = d.duv_d;
                a.inv_duv_d = 1.0 / d.duv_d;
                b.partOff = P.off[si];
                const int nb = P.nb[si];
                if constexpr (is_f32) {
                    if (grad && mix) launch_sample_grad<T, true>(sh.win, nb, nw, s, a, b);
                    else if (mix) launch_sample_multi<T, true>(sh.kw, sh.ppt, nb, nw, s, a, b);
                }
                if (grad && !mix) launch_sample_grad<T, false>(sh.win, nb, nw, s, a, b);
                else if (!mix) launch_sample_multi<T, false>(sh.kw, sh.ppt, nb, nw, s, a, b);
            }
            if (grad) {
                // the full adjoint column transform: with many walkers in flight the batched passes are
                // bound by throughput, where the residue form's extra LDS traffic costs ~1-2%
                AdjColArgs<T> aca{nullptr, nullptr, g.ldH, G.N, g.Rs, tw, nullptr, 0, nullptr, 0, Dst4{}, ablocks,
                                  G.ex.rc2};
                aca.ldTh = int(P.sTh / size_t(g.nrows));
                AdjRowArgs<T> ara{G.sw, nullptr, aca.ldTh, G.N, g.Rs, g.nrows, tw, 0.0, nullptr, nullptr};
#define GL_CALL(LL)                                                                                              \
    hipLaunchKernelGGL((k_adjcol_sym_b<T, LL>), dim3(ablocks, nw), dim3(FftShape<LL>::NT), 0, s, aca, b);        \
    hipLaunchKernelGGL((k_adjrow_sym_b<T, LL>), dim3(g.pairs_r, nw), dim3(FftShape<LL>::NT), 0, s, ara, b)
                GL_L_SWITCH(T, GL_CALL)
#undef GL_CALL
            }
        }
    }
    BatchSets bs{};
    bs.n = int(P.sets.size());
    for (int i = 0; i < bs.n; ++i) {
        bs.off[i] = P.off[size_t(i)];
        bs.nb[i] = P.nb[size_t(i)];
    }
    hipLaunchKernelGGL(k_batch_reduce, dim3(W), dim3(1024), 0, s, static_cast<const double*>(P.part.p), P.ldPart, bs,
                       grad ? 4 : 1, static_cast<double*>(P.out_sets.p), static_cast<double*>(P.out_tot.p),
                       grad ? static_cast<double*>(P.ggeom.p) : nullptr);
    HC(hipGetLastError());
}

template <typename T>
int batch_create(int nsets, const BatchSet<T>* sets, int nr, T Rmin, T dR, int max_w, int check_central,
                 void** out) {
    return guard([&] {
        if (!out) invalid("batch", "null plan pointer");
        *out = nullptr;
        if (nsets < 1 || nsets > BATCH_MAX_SETS)
            invalid("batch", "number of datasets = " + std::to_string(nsets) + " (expect 1 to " +
                                 std::to_string(BATCH_MAX_SETS) + ")");
        if (!sets) invalid("batch", "null dataset array");
        if (max_w < 1) invalid("batch", "max_walkers must be >= 1");
        auto P = std::make_unique<BatchPlan<T>>();
        P->prec = int(sizeof(T));
        P->nr = nr;
        P->max_w = max_w;
        size_t blocks = 0;
        for (int i = 0; i < nsets; ++i) {
            const BatchSet<T>& d = sets[i];
            check_input(d.nxy);
            check_nd(d.nd);
            if (d.nd > 0 && (!d.u || !d.v || !d.re || !d.im || !d.w))
                invalid("batch", "dataset " + std::to_string(i) + " has a null array");
            if (!(d.duv > T(0))) invalid("batch", "dataset " + std::to_string(i) + ": expect duv > 0");
            const int L = fast_log2<T>(d.nxy);
            if (L < 0)
                invalid("batch", "dataset " + std::to_string(i) + ": nxy = " + std::to_string(d.nxy) +
                                     " is not supported by the batched form (power of two from 128 to " +
                                     std::to_string(1 << max_fast_log2<T>()) + ")");
            // same grid (nxy, dxy): share the group's spectrum
            int gi = -1;
            for (size_t k = 0; k < P->groups.size(); ++k)
                if (P->groups[k].N == d.nxy && P->groups[k].dxy == d.dxy) gi = int(k);
            if (gi < 0) {
                BatchGroup<T> G{};
                ProfileSetup<T> ps = setup_profile<T>(nr, Rmin, dR, d.dxy, d.nxy, T(0), check_central);
                G.N = d.nxy;
                G.dxy = d.dxy;
                G.sw = ps.sw;
                G.sw.I = nullptr;
                G.g = sym_geom(L, d.nxy, ps.rmax);
                P->groups.push_back(G);
                gi = int(P->groups.size()) - 1;
            }
            BatchGroup<T>& Gi = P->groups[size_t(gi)];
            Gi.sets.push_back(i);
            const Extent e = uv_extent<T>(d.uvmax, double(d.duv));
            Gi.ex_all = Gi.ex_all && e.rc2 > 0.0;
            Gi.ex.rho2 = Gi.ex_all ? std::max(Gi.ex.rho2, e.rho2) : 0.0;
            Gi.ex.rc2 = Gi.ex_all ? std::max(Gi.ex.rc2, e.rc2) : 0.0;
            P->sets.push_back(d);
            const int nb = int(blocks_for(size_t(d.nd), 256));
            P->off.push_back(int(blocks));
            P->nb.push_back(nb);
            blocks += size_t(nb);
        }
        P->ldPart = int(std::max<size_t>(blocks, 1));
        HC(hipGetDevice(&P->device));   // (all argument checks above run without a GPU)
        size_t maxHh = 0, maxTh = 0;
        for (const BatchGroup<T>& G : P->groups) {
            const SymGeom& g = G.g;
            P->sTs = std::max(P->sTs, size_t(2 * g.pairs_r) * g.ldT);
            P->sH = std::max(P->sH, size_t(g.C + 1) * g.ldH);
            maxHh = std::max(maxHh, size_t(g.C + 1) * g.ldH);
            const size_t ldTh = size_t((g.C + 1 + 7) & ~7);
            maxTh = std::max(maxTh, ldTh);
        }
        P->hot_off = (maxHh + 31) & ~size_t(31);
        P->sHh = P->hot_off + size_t(GHOT_ELEMS);
        // adjoint rows: nrows x ldTh per group, one stride for all
        size_t sTh = 0;
        for (const BatchGroup<T>& G : P->groups) sTh = std::max(sTh, size_t(G.g.nrows) * maxTh);
        P->sTh = sTh;
        // parameters, walker scalars, partials, outputs
        P->I.template get<T>(size_t(max_w) * nr);
        P->geom.template get<double>(size_t(max_w) * 4);   // T, or double geometry (MODE_F64_PHASE)
        P->wp.template get<WalkerPrm<T>>(size_t(max_w));
        P->part.template get<double>(size_t(max_w) * 4 * P->ldPart);
        P->out_sets.template get<double>(size_t(max_w) * nsets);
        P->out_tot.template get<double>(size_t(max_w));
        P->gI.template get<double>(size_t(max_w) * nr);
        P->ggeom.template get<double>(size_t(max_w) * 4);
        unsigned* f = P->flag.template get<unsigned>(1);
        HC(hipStreamCreateWithFlags(&P->cap, hipStreamNonBlocking));
        HC(hipEventCreateWithFlags(&P->ev, hipEventDisableTiming));
        HC(hipMemsetAsync(f, 0, sizeof(unsigned), P->cap));
        HC(hipStreamSynchronize(P->cap));
        *out = static_cast<BatchPlanBase*>(P.release());
    });
}

template <typename T>
BatchPlan<T>& plan_of(void* h) {
    auto* b = static_cast<BatchPlanBase*>(h);
    if (!b || b->magic != BATCH_MAGIC) invalid("batch", "not a batch plan");
    if (b->prec != int(sizeof(T))) invalid("batch", "plan precision does not match the entry point");
    auto& P = *static_cast<BatchPlan<T>*>(b);
    int dev = 0;
    HC(hipGetDevice(&dev));
    if (dev != P.device) invalid("batch", "the plan belongs to device " + std::to_string(P.device));
    return P;
}

// Average time of one eager evaluation of shape `sh` (after one warm-up).
template <typename T>
float batch_time(BatchPlan<T>& P, hipStream_t s, int W, const BatchShape& sh, int mode) {
    P.ensure_slots(sh.chunk, mode & MODE_GRAD, s);
    batch_enqueue(P, s, W, sh, mode);
    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    HC(hipEventRecord(e0, s));
    const int reps = 2;
    for (int r = 0; r < reps; ++r) batch_enqueue(P, s, W, sh, mode);
    HC(hipEventRecord(e1, s));
    HC(hipEventSynchronize(e1));
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

// Launch shape for W walkers, by timing eager evaluations on the caller's
// stream, one dimension at a time from (4 walkers per sample workgroup, one
// 256-point block per workgroup, 3072-cell gradient window): the chunk size
// (powers of two and W itself, within the slot budget); forward: the walkers
// per sample workgroup (1, 2, 4, 8), then the point blocks per workgroup
// (1, 2); gradient: the LDS window (2048, 3072, 4096 cells).  Point blocks and
// window set by the caller (tile_fixed) are kept.
template <typename T>
void batch_tune(BatchPlan<T>& P, hipStream_t s, int W, int mode, typename BatchPlan<T>::Exec& run) {
    const bool grad = mode & MODE_GRAD;
    const int cmax = P.max_chunk(grad, W);
    std::vector<int> cand;
    for (int c = 1; c < cmax; c *= 2) cand.push_back(c);
    cand.push_back(cmax);
    BatchShape best{cand.back(), 4, run.tile_fixed ? run.ppt : 1, run.tile_fixed ? run.win : GWIN};
    float tbest = 0.f;
    for (int c : cand) {
        BatchShape t = best;
        t.chunk = c;
        const float ms = batch_time(P, s, W, t, mode);
        if (c == cand.front() || ms < tbest) {
            tbest = ms;
            best.chunk = c;
        }
    }
    auto try_shape = [&](BatchShape t) {
        const float ms = batch_time(P, s, W, t, mode);
        if (ms < tbest) {
            tbest = ms;
            best = t;
        }
    };
    if (!grad) {
        for (int k : {1, 2, 8}) {
            BatchShape t = best;
            t.kw = k;
            if (k != best.kw) try_shape(t);
        }
        if (!run.tile_fixed) {
            BatchShape t = best;
            t.ppt = 2;
            try_shape(t);
        }
    } else if (!run.tile_fixed) {
        for (int wn : {2048, 4096}) {
            BatchShape t = best;
            t.win = wn;
            try_shape(t);
        }
    }
    run.chunk = best.chunk;
    run.kw = best.kw;
    run.ppt = best.ppt;
    run.win = best.win;
}

template <typename T>
BatchShape shape_of(const typename BatchPlan<T>::Exec& run) {
    return BatchShape{run.chunk, run.kw, run.ppt, run.win};
}

// geom of type G: T, or double with MODE_F64_PHASE (float plans only).
template <typename T, typename G>
int batch_run(void* h, void* stream, int W, const T* I, const G* geom, double* chi2_sets, double* chi2,
              double* gI, double* ggeom, int flags, bool grad) {
    return guard([&] {
        BatchPlan<T>& P = plan_of<T>(h);
        std::lock_guard<std::mutex> lk(P.mu);
        constexpr bool mix = !std::is_same<G, T>::value;
        static_assert(!mix || (std::is_same<T, float>::value && std::is_same<G, double>::value),
                      "FP64 geometry is the float plans' mixed form");
        const int mode = (grad ? MODE_GRAD : 0) | (mix ? MODE_F64_PHASE : 0);
        if (W < 1 || W > P.max_w)
            invalid("batch", "nwalkers = " + std::to_string(W) + " (plan holds 1 to " + std::to_string(P.max_w) + ")");
        if (!I || !geom) invalid("batch", "null parameter array");
        if (grad && (!gI || !ggeom)) invalid("batch", "null gradient output");
        if (grad && P.nr > ADJROW_LDS_NR)
            invalid("batch", "nr = " + std::to_string(P.nr) + " profile bins (the gradient supports up to " +
                                 std::to_string(ADJROW_LDS_NR) + ")");
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (P.has_last && P.last != s) {
            HC(hipEventRecord(P.ev, P.last));
            HC(hipStreamWaitEvent(s, P.ev, 0));
        }
        P.last = s;
        P.has_last = true;
        if (I != P.I.p)
            HC(hipMemcpyAsync(P.I.p, I, sizeof(T) * size_t(W) * P.nr, hipMemcpyDeviceToDevice, s));
        if (geom != P.geom.p)
            HC(hipMemcpyAsync(P.geom.p, geom, sizeof(G) * size_t(W) * 4, hipMemcpyDeviceToDevice, s));
        const auto key = std::make_pair(W, mode);
        auto& run = P.runs[key];
        if (run.chunk == 0) batch_tune(P, s, W, mode, run);
        P.ensure_slots(run.chunk, grad, s);
        const BatchShape sh = shape_of<T>(run);
        if (flags & GALARIO_BATCH_EAGER) {
            batch_enqueue(P, s, W, sh, mode);
        } else {
            if (!run.exec) {
                // capture on the plan's stream (nothing runs during capture)
                HC(hipStreamSynchronize(s));
                hipGraph_t graph = nullptr;
                HC(hipStreamBeginCapture(P.cap, hipStreamCaptureModeThreadLocal));
                try {
                    batch_enqueue(P, P.cap, W, sh, mode);
                } catch (...) {
                    (void)hipStreamEndCapture(P.cap, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    throw;
                }
                HC(hipStreamEndCapture(P.cap, &graph));
                const hipError_t e = hipGraphInstantiate(&run.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HC(e);
            }
            HC(hipGraphLaunch(run.exec, s));
        }
        const size_t S = P.sets.size();
        if (chi2_sets)
            HC(hipMemcpyAsync(chi2_sets, P.out_sets.p, sizeof(double) * size_t(W) * S, hipMemcpyDeviceToDevice, s));
        if (chi2) HC(hipMemcpyAsync(chi2, P.out_tot.p, sizeof(double) * size_t(W), hipMemcpyDeviceToDevice, s));
        if (grad) {
            HC(hipMemcpyAsync(gI, P.gI.p, sizeof(double) * size_t(W) * P.nr, hipMemcpyDeviceToDevice, s));
            HC(hipMemcpyAsync(ggeom, P.ggeom.p, sizeof(double) * size_t(W) * 4, hipMemcpyDeviceToDevice, s));
        }
    });
}

int g_threads = 16;

// ====================================================================== allocation-free chi2
// Set-up per device (galario_init, galario_use_gpu): the device address of `scratch` and every
// twiddle table the allocation-free form can need, built (allocated) here so that its calls only
// look them up.
constexpr int MAX_DEVICES = 64;
std::mutex g_static_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

void static_setup_current() {
    int dev = 0;
    HC(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) fail(GALARIO_ERR_RUNTIME, "device index beyond the library's table");
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (g_scratch_ptr[dev]) return;
    void* p = nullptr;
    HC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(GALARIO_ERR_RUNTIME, "Error in hip: no device address for the scratch array");
    for (int L = 7; L <= max_fast_log2<double>(); ++L) (void)twiddles<double>(1 << L);
    for (int L = 7; L <= max_fast_log2<float>(); ++L) (void)twiddles<float>(1 << L);
    g_scratch_ptr[dev] = p;
}

// Twiddle table built at set-up (never allocates).
template <typename T>
const cpx<T>* twiddles_built(int dev, int N) {
    std::lock_guard<std::mutex> lock(g_tw_mu);
    auto it = g_tw.find(std::make_tuple(dev, N, int(sizeof(T))));
    if (it == g_tw.end()) fail(GALARIO_ERR_RUNTIME, "internal: twiddle table missing after set-up");
    return static_cast<const cpx<T>*>(it->second);
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// chi2_profile_dev with its scratch in `scratch`: the same checks and kernels (chi2 is bit-identical),
// no allocation and no library stream or event; points outside the grid make chi2 NaN.
template <typename T>
int chi2_profile_noalloc_dev(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc, T dRA, T dDec,
                             T duv, T PA, int nd, const T* u, const T* v, const T* re, const T* im, const T* wt,
                             int check_central, double* chi2_d, double uvmax) {
    return guard([&] {
        check_nd(nd);
        ProfileSetup<T> ps = setup_profile<T>(nr, Rmin, dR, dxy, nxy, inc, check_central);
        SampleSetup<T> ss = setup_sample<T>(PA, dRA, dDec);
        check_dev_inputs(I, nd, u, v, true, re, im, wt);
        if (!chi2_d) invalid("check input", "null chi2 pointer");
        const int L = fast_log2<T>(nxy);
        if (L < 0)
            invalid("chi2 without allocation", "nxy = " + std::to_string(nxy) + " (expect a power of two from 128 to " +
                                                   std::to_string(1 << max_fast_log2<T>()) + ")");
        const SymGeom g = sym_geom(L, nxy, ps.rmax);
        const size_t nb = blocks_for(size_t(nd), 256);
        const size_t o_Ts = align256(sizeof(double) * std::max<size_t>(nb, 1));
        const size_t o_H = align256(o_Ts + sizeof(T) * size_t(2 * g.pairs_r) * size_t(g.ldT));
        const size_t need = o_H + sizeof(T) * size_t(g.C + 1) * size_t(g.ldH);
        if (need > SCRATCH_BYTES)
            invalid("chi2 without allocation",
                    "this evaluation needs " + std::to_string(need) + " bytes of scratch (set by nxy, the swept radius "
                    "and the number of points); the library holds " + std::to_string(SCRATCH_BYTES));
        int dev = 0;
        HC(hipGetDevice(&dev));
        char* base = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_static_mu);
            if (dev >= 0 && dev < MAX_DEVICES) base = static_cast<char*>(g_scratch_ptr[dev]);
        }
        if (!base)
            invalid("chi2 without allocation", "device " + std::to_string(dev) +
                                                   " is not set up: call galario_init() (or galario_use_gpu()) "
                                                   "with it current first");
        const cpx<T>* tw = twiddles_built<T>(dev, nxy);
        double* part = reinterpret_cast<double*>(base);
        T* Ts = reinterpret_cast<T*>(base + o_Ts);
        T* H = reinterpret_cast<T*>(base + o_H);
        const Extent ex = uv_extent<T>(uvmax, double(duv));
        ps.sw.I = I;
        hipStream_t s = static_cast<hipStream_t>(stream);
        const SymRowArgs<T> ra{ps.sw, nxy, g.Rs, g.nrows, g.ldT, tw, Ts, nullptr, 0, nullptr};
        const SymColArgs<T> ca{Ts, nxy, g.Rs, g.ldT, tw, H, g.ldH, ex.rc2,
                               residue_fits(L, g.Rs) ? twiddles_built<T>(dev, 512) : nullptr};
        launch_rowpass_sym<T>(L, g.pairs_r, s, ra);
        launch_colpass_sym<T>(L, nxy / 4, s, ca);
        SampleArgs<T> sa = sample_args<T>(nullptr, 0, 0, nxy, g.C, T(1), duv, ss, nd, u, v);
        sa.H = H;
        sa.ldH = g.ldH;
        sa.rho2 = ex.rho2;
        sa.re = re;
        sa.im = im;
        sa.w = wt;
        sa.partials = part;
        sa.flag = nullptr;
        const unsigned ng = unsigned((nb + SAMPLE_PPT - 1) / SAMPLE_PPT);
        if (nb) hipLaunchKernelGGL((k_sample_chi2_n<T, true, SAMPLE_PPT>), dim3(ng), dim3(256), 0, s, sa, int(nb));
        hipLaunchKernelGGL((k_sum_partials<T>), dim3(1), dim3(1024), 0, s, part, int(nb), chi2_d,
                           static_cast<T*>(nullptr));
        HC(hipGetLastError());
    });
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* galario_last_error(void) { return t_err.c_str(); }

int galario_init(void) {
    return guard([] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {   // no GPU visible: nothing to set up
            (void)hipGetLastError();
            return;
        }
        static_setup_current();
    });
}

int galario_cleanup(void) {
    return guard([] {
        for (auto& kv : t_ws) {
            if (kv.second) {
                int cur = 0;
                HC(hipGetDevice(&cur));
                HC(hipSetDevice(kv.first));
                (void)hipDeviceSynchronize();
                kv.second->release();
                delete kv.second;
                HC(hipSetDevice(cur));
            }
        }
        t_ws.clear();
        // The twiddle tables stay: they are shared by every thread and baked into the
        // kernel arguments of captured batch graphs, so they live until the process ends
        // (a few hundred KB per device).
    });
}

int galario_threads(int num) {
    if (num > 0) g_threads = num;
    return g_threads;
}

int galario_ngpus(int* out) {
    return guard([&] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e == hipErrorNoDevice) n = 0;
        else HC(e);
        *out = n;
    });
}

int galario_use_gpu(int device_id) {
    return guard([&] {
        HC(hipSetDevice(device_id));
        static_setup_current();
    });
}

void galario_free(void* ptr) { std::free(ptr); }

#define GL_BOTH(MACRO) MACRO(f64, double) MACRO(f32, float)

#define GL_MAIN(SUF, T)                                                                                           \
    int galario_sample_profile_##SUF(int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc, T dRA, T dDec,     \
                                     T duv, T PA, int nd, const T* u, const T* v, T* vis, int cc) {              \
        return sample_profile_host<T>(nr, I, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, nd, u, v, vis, cc);    \
    }                                                                                                             \
    int galario_sample_image_##SUF(int nx, int ny, const T* image, T vo, T dRA, T dDec, T duv, T PA, int nd,     \
                                   const T* u, const T* v, T* vis) {                                             \
        return sample_image_host<T>(nx, ny, image, vo, dRA, dDec, duv, PA, nd, u, v, vis);                       \
    }                                                                                                             \
    int galario_chi2_profile_##SUF(int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc, T dRA, T dDec,       \
                                   T duv, T PA, int nd, const T* u, const T* v, const T* re, const T* im,        \
                                   const T* w, int cc, T* chi2) {                                                \
        return chi2_profile_host<T>(nr, I, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, nd, u, v, re, im, w, cc, \
                                    chi2);                                                                        \
    }                                                                                                             \
    int galario_chi2_image_##SUF(int nx, int ny, const T* image, T vo, T dRA, T dDec, T duv, T PA, int nd,       \
                                 const T* u, const T* v, const T* re, const T* im, const T* w, T* chi2) {        \
        return chi2_image_host<T>(nx, ny, image, vo, dRA, dDec, duv, PA, nd, u, v, re, im, w, chi2);             \
    }                                                                                                             \
    int galario_sweep_##SUF(int nr, const T* I, T Rmin, T dR, int nxy, T dxy, T inc, T* image, int cc) {         \
        return sweep_host<T>(nr, I, Rmin, dR, nxy, dxy, inc, image, cc);                                         \
    }                                                                                                             \
    int galario_uv_rotate_##SUF(T PA, T dRA, T dDec, T* dRArot, T* dDecrot, int nd, const T* u, const T* v,      \
                                T* urot, T* vrot) {                                                              \
        return uv_rotate_host<T>(PA, dRA, dDec, dRArot, dDecrot, nd, u, v, urot, vrot);                          \
    }                                                                                                             \
    int galario_copy_input_##SUF(int nx, int ny, const T* realdata, void** out) {                                \
        return copy_input_host<T>(nx, ny, realdata, out);                                                        \
    }                                                                                                             \
    int galario_fft2d_##SUF(int nx, int ny, void* data) { return fft2d_host<T>(nx, ny, data); }                  \
    int galario_fftshift_##SUF(int nx, int ny, void* data) { return fftshift_host<T>(nx, ny, data); }            \
    int galario_fftshift_axis0_##SUF(int nrow, int ncol, void* data) {                                           \
        return fftshift_axis0_host<T>(nrow, ncol, data);                                                         \
    }                                                                                                             \
    int galario_interpolate_##SUF(int nrow, int ncol, const void* data, T vo, int nd, const T* u, const T* v,    \
                                  T duv, void* vis) {                                                            \
        return interpolate_host<T>(nrow, ncol, data, vo, nd, u, v, duv, vis);                                    \
    }                                                                                                             \
    int galario_apply_phase_sampled_##SUF(T dRA, T dDec, int nd, const T* u, const T* v, void* vis) {            \
        return apply_phase_host<T>(dRA, dDec, nd, u, v, vis);                                                    \
    }                                                                                                             \
    int galario_reduce_chi2_##SUF(int nd, const T* re, const T* im, const void* vis, const T* w, T* chi2) {      \
        return reduce_chi2_host<T>(nd, re, im, vis, w, chi2);                                                    \
    }                                                                                                             \
    int galario_chi2_profile_dev_##SUF(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc,    \
                                       T dRA, T dDec, T duv, T PA, int nd, const T* u, const T* v, const T* re,  \
                                       const T* im, const T* w, int cc, double* chi2_d, double uvmax) {          \
        return chi2_profile_dev<T>(stream, nr, I, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, nd, u, v, re, im, \
                                   w, cc, chi2_d, uvmax);                                                        \
    }                                                                                                             \
    int galario_sample_profile_dev_##SUF(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy, T inc,  \
                                         T dRA, T dDec, T duv, T PA, int nd, const T* u, const T* v, int cc,     \
                                         const int* perm, void* vis_d, double uvmax) {                           \
        return sample_profile_dev<T>(stream, nr, I, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, nd, u, v, cc,   \
                                     perm, vis_d, uvmax);                                                        \
    }

GL_BOTH(GL_MAIN)

#define GL_GRAD(SUF, T)                                                                                           \
    int galario_chi2_profile_grad_dev_##SUF(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy,       \
                                            T inc, T dRA, T dDec, T duv, T PA, int nd, const T* u, const T* v,    \
                                            const T* re, const T* im, const T* w, int cc, double* chi2_d,         \
                                            double* grad_I, double* grad_geom, double uvmax) {                    \
        return chi2_profile_grad_dev<T>(stream, nr, I, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, nd, u, v, re, \
                                        im, w, cc, chi2_d, grad_I, grad_geom, uvmax);                             \
    }

GL_BOTH(GL_GRAD)

#define GL_NOALLOC(SUF, T)                                                                                        \
    int galario_chi2_profile_noalloc_dev_##SUF(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy,    \
                                               T inc, T dRA, T dDec, T duv, T PA, int nd, const T* u, const T* v, \
                                               const T* re, const T* im, const T* w, int cc, double* chi2_d,      \
                                               double uvmax) {                                                    \
        return chi2_profile_noalloc_dev<T>(stream, nr, I, Rmin, dR, dxy, nxy, inc, dRA, dDec, duv, PA, nd, u, v,  \
                                           re, im, w, cc, chi2_d, uvmax);                                         \
    }

GL_BOTH(GL_NOALLOC)

#define GL_MULTI(SUF, T)                                                                                          \
    int galario_sample_profile_multi_dev_##SUF(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy,    \
                                               T inc, T duv, T PA, int nsrc, const T* src, int nd, const T* u,    \
                                               const T* v, int cc, const int* perm, void* vis_d, double uvmax) {  \
        return profile_multisrc_dev<T>(stream, nr, I, Rmin, dR, dxy, nxy, inc, duv, PA, nsrc, src, nd, u, v,      \
                                       nullptr, nullptr, nullptr, cc, perm, vis_d, nullptr, uvmax);               \
    }                                                                                                             \
    int galario_chi2_profile_multi_dev_##SUF(void* stream, int nr, const T* I, T Rmin, T dR, T dxy, int nxy,      \
                                             T inc, T duv, T PA, int nsrc, const T* src, int nd, const T* u,      \
                                             const T* v, const T* re, const T* im, const T* w, int cc,            \
                                             double* chi2_d, double uvmax) {                                      \
        if (!chi2_d) {                                                                                            \
            t_err = "Error in multiple sources: null chi2 pointer";                                               \
            return GALARIO_ERR_INVALID;                                                                           \
        }                                                                                                         \
        return profile_multisrc_dev<T>(stream, nr, I, Rmin, dR, dxy, nxy, inc, duv, PA, nsrc, src, nd, u, v, re,  \
                                       im, w, cc, nullptr, nullptr, chi2_d, uvmax);                               \
    }

GL_BOTH(GL_MULTI)

#define GL_BATCH(SUF, T)                                                                                          \
    int galario_batch_create_##SUF(int nsets, const galario_dataset_##SUF* sets, int nr, T Rmin, T dR,           \
                                   int max_walkers, int check_central, void** plan) {                            \
        std::vector<BatchSet<T>> v;                                                                               \
        if (sets && nsets > 0 && nsets <= BATCH_MAX_SETS)                                                         \
            for (int i = 0; i < nsets; ++i)                                                                       \
                v.push_back(BatchSet<T>{sets[i].nxy, sets[i].dxy, T(sets[i].duv), sets[i].nd, sets[i].u, sets[i].v, \
                                        sets[i].vis_obs_re, sets[i].vis_obs_im, sets[i].weights,                 \
                                        sets[i].uvmax, double(sets[i].duv)});                                     \
        return batch_create<T>(nsets, v.empty() ? nullptr : v.data(), nr, Rmin, dR, max_walkers, check_central,  \
                               plan);                                                                             \
    }                                                                                                             \
    int galario_batch_chi2_profile_##SUF(void* plan, void* stream, int nwalkers, const T* I, const T* geom,      \
                                         double* chi2_sets, double* chi2, int flags) {                            \
        return batch_run<T, T>(plan, stream, nwalkers, I, geom, chi2_sets, chi2, nullptr, nullptr, flags, false); \
    }                                                                                                             \
    int galario_batch_chi2_profile_grad_##SUF(void* plan, void* stream, int nwalkers, const T* I, const T* geom, \
                                              double* chi2_sets, double* chi2, double* grad_I, double* grad_geom, \
                                              int flags) {                                                        \
        return batch_run<T, T>(plan, stream, nwalkers, I, geom, chi2_sets, chi2, grad_I, grad_geom, flags, true); \
    }

GL_BOTH(GL_BATCH)

int galario_batch_chi2_profile_f32x(void* plan, void* stream, int nwalkers, const float* I, const double* geom,
                                    double* chi2_sets, double* chi2, int flags) {
    return batch_run<float, double>(plan, stream, nwalkers, I, geom, chi2_sets, chi2, nullptr, nullptr, flags, false);
}

int galario_batch_chi2_profile_grad_f32x(void* plan, void* stream, int nwalkers, const float* I, const double* geom,
                                         double* chi2_sets, double* chi2, double* grad_I, double* grad_geom,
                                         int flags) {
    return batch_run<float, double>(plan, stream, nwalkers, I, geom, chi2_sets, chi2, grad_I, grad_geom, flags, true);
}

int galario_batch_destroy(void* plan) {
    return guard([&] {
        if (!plan) return;
        auto* b = static_cast<BatchPlanBase*>(plan);
        if (b->magic != BATCH_MAGIC) invalid("batch", "not a batch plan");
        delete b;
    });
}

int galario_batch_info(void* plan, int nwalkers, int mode, int* chunk, int* walkers_per_sample, int* graph_ready) {
    return guard([&] {
        auto* b = static_cast<BatchPlanBase*>(plan);
        if (!b || b->magic != BATCH_MAGIC) invalid("batch", "not a batch plan");
        int c = 0, k = 0, g = 0;
        auto look = [&](auto& P) {
            std::lock_guard<std::mutex> lk(P.mu);
            auto it = P.runs.find(std::make_pair(nwalkers, mode));
            if (it != P.runs.end()) {
                c = it->second.chunk;
                k = (mode & MODE_GRAD) ? 1 : it->second.kw;
                g = it->second.exec != nullptr;
            }
        };
        if (b->prec == 8) look(*static_cast<BatchPlan<double>*>(b));
        else look(*static_cast<BatchPlan<float>*>(b));
        if (chunk) *chunk = c;
        if (walkers_per_sample) *walkers_per_sample = k;
        if (graph_ready) *graph_ready = g;
    });
}

int galario_batch_set_chunk(void* plan, int nwalkers, int mode, int chunk, int walkers_per_sample) {
    return guard([&] {
        auto* b = static_cast<BatchPlanBase*>(plan);
        if (!b || b->magic != BATCH_MAGIC) invalid("batch", "not a batch plan");
        auto set = [&](auto& P) {
            std::lock_guard<std::mutex> lk(P.mu);
            if (nwalkers < 1 || nwalkers > P.max_w) invalid("batch", "nwalkers out of range");
            if (mode < 0 || mode > (MODE_GRAD | MODE_F64_PHASE) || ((mode & MODE_F64_PHASE) && P.prec != 4))
                invalid("batch", "mode = " + std::to_string(mode) + " (gradient 1, FP64 phase 2 on float plans)");
            const bool grad = mode & MODE_GRAD;
            if (chunk < 1 || chunk > P.max_chunk(grad, nwalkers))
                invalid("batch", "chunk = " + std::to_string(chunk) + " (expect 1 to " +
                                     std::to_string(P.max_chunk(grad, nwalkers)) + ")");
            if (walkers_per_sample != 1 && walkers_per_sample != 2 && walkers_per_sample != 4 &&
                walkers_per_sample != 8)
                invalid("batch", "walkers per sample workgroup must be 1, 2, 4 or 8");
            auto& run = P.runs[std::make_pair(nwalkers, mode)];
            if (run.exec) (void)hipGraphExecDestroy(run.exec);
            run.exec = nullptr;
            run.chunk = chunk;
            run.kw = walkers_per_sample;
        };
        if (b->prec == 8) set(*static_cast<BatchPlan<double>*>(b));
        else set(*static_cast<BatchPlan<float>*>(b));
    });
}

int galario_batch_tile(void* plan, int nwalkers, int mode, int* point_blocks, int* grad_window) {
    return guard([&] {
        auto* b = static_cast<BatchPlanBase*>(plan);
        if (!b || b->magic != BATCH_MAGIC) invalid("batch", "not a batch plan");
        int pb = 0, wn = 0;
        auto look = [&](auto& P) {
            std::lock_guard<std::mutex> lk(P.mu);
            auto it = P.runs.find(std::make_pair(nwalkers, mode));
            if (it != P.runs.end() && (it->second.chunk > 0 || it->second.tile_fixed)) {
                pb = (mode & MODE_GRAD) ? 1 : it->second.ppt;
                wn = (mode & MODE_GRAD) ? it->second.win : 0;
            }
        };
        if (b->prec == 8) look(*static_cast<BatchPlan<double>*>(b));
        else look(*static_cast<BatchPlan<float>*>(b));
        if (point_blocks) *point_blocks = pb;
        if (grad_window) *grad_window = wn;
    });
}

int galario_batch_set_tile(void* plan, int nwalkers, int mode, int point_blocks, int grad_window) {
    return guard([&] {
        auto* b = static_cast<BatchPlanBase*>(plan);
        if (!b || b->magic != BATCH_MAGIC) invalid("batch", "not a batch plan");
        auto set = [&](auto& P) {
            std::lock_guard<std::mutex> lk(P.mu);
            if (nwalkers < 1 || nwalkers > P.max_w) invalid("batch", "nwalkers out of range");
            if (mode < 0 || mode > (MODE_GRAD | MODE_F64_PHASE) || ((mode & MODE_F64_PHASE) && P.prec != 4))
                invalid("batch", "mode = " + std::to_string(mode) + " (gradient 1, FP64 phase 2 on float plans)");
            if (point_blocks != 1 && point_blocks != 2)
                invalid("batch", "point blocks per sample workgroup must be 1 or 2");
            if (grad_window != 2048 && grad_window != GWIN && grad_window != 4096)
                invalid("batch", "gradient window must be 2048, 3072 or 4096 cells");
            auto& run = P.runs[std::make_pair(nwalkers, mode)];
            if (run.exec) (void)hipGraphExecDestroy(run.exec);
            run.exec = nullptr;
            run.ppt = point_blocks;
            run.win = grad_window;
            run.tile_fixed = true;
        };
        if (b->prec == 8) set(*static_cast<BatchPlan<double>*>(b));
        else set(*static_cast<BatchPlan<float>*>(b));
    });
}

int galario_batch_take_oob_flag(void* plan, int* out) {
    return guard([&] {
        auto* b = static_cast<BatchPlanBase*>(plan);
        if (!b || b->magic != BATCH_MAGIC) invalid("batch", "not a batch plan");
        auto take = [&](auto& P) {
            std::lock_guard<std::mutex> lk(P.mu);
            unsigned h = 0;
            if (P.has_last) HC(hipStreamSynchronize(P.last));
            HC(hipMemcpy(&h, P.flag.p, sizeof(unsigned), hipMemcpyDeviceToHost));
            if (h) HC(hipMemset(P.flag.p, 0, sizeof(unsigned)));
            if (out) *out = h ? 1 : 0;
        };
        if (b->prec == 8) take(*static_cast<BatchPlan<double>*>(b));
        else take(*static_cast<BatchPlan<float>*>(b));
    });
}

int galario_read_probe(void* stream, const void* buf, long long bytes, int grid, double* out) {
    return guard([&] {
        if (bytes < 0 || bytes % 16 != 0) invalid("read_probe", "bytes must be a non-negative multiple of 16");
        if (grid < 1 || grid > 65536) invalid("read_probe", "grid must be in [1, 65536]");
        if (!buf || !out) invalid("read_probe", "null buffer");
        hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_read_probe, dim3(grid), dim3(256), 0, s, static_cast<const cpx<double>*>(buf),
                           size_t(bytes / 16), out);
        HC(hipGetLastError());
    });
}

int galario_stage_timer(int enable) {
    return guard([&] {
        t_timer.on = enable != 0;
        for (auto& c : t_timer.calls)
            for (auto& pe : c) t_timer.pool.push_back(pe.second);
        t_timer.calls.clear();
        for (auto& pe : t_timer.cur) t_timer.pool.push_back(pe.second);
        t_timer.cur.clear();
    });
}

int galario_stage_repeat(int stage, int reps) {
    return guard([&] {
        if (reps < 1) invalid("stage_repeat", "reps must be >= 1");
        if (stage != ST_ROW && stage != ST_COL && stage != ST_SAMPLE && stage != -1)
            invalid("stage_repeat", "only the row pass (0), column pass (1) and forward sample (2) repeat");
        t_timer.rep_stage = stage;
        t_timer.rep_n = reps;
    });
}

int galario_stage_times(double* ms, int nstage, int* ncalls) {
    return guard([&] {
        std::vector<double> acc(NSTAGE, 0.0);
        for (auto& c : t_timer.calls) {
            for (size_t k = 0; k + 1 < c.size(); ++k) {
                HC(hipEventSynchronize(c[k + 1].second));
                float t = 0.f;
                HC(hipEventElapsedTime(&t, c[k].second, c[k + 1].second));
                acc[size_t(c[k].first)] += double(t);
            }
        }
        for (int k = 0; k < nstage && k < NSTAGE; ++k) ms[k] = acc[size_t(k)];
        *ncalls = int(t_timer.calls.size());
    });
}

int galario_take_oob_flag(int* out) {
    return guard([&] {
        int dev = 0;
        HC(hipGetDevice(&dev));
        auto it = t_ws.find(dev);
        if (it == t_ws.end() || !it->second) {
            *out = 0;
            return;
        }
        Workspace& w = *it->second;
        unsigned* fd = w.flag_dev.get<unsigned>(1);
        hipStream_t s = w.has_last ? w.last : w.own;
        unsigned* hf = reinterpret_cast<unsigned*>(w.h_scal + 1);
        HC(hipMemcpyAsync(hf, fd, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        HC(hipMemsetAsync(fd, 0, sizeof(unsigned), s));
        HC(hipStreamSynchronize(s));
        *out = *hf ? 1 : 0;
    });
}

}  // extern "C"
