/*
 * galario_amd -- C ABI of the MI355X (gfx950) HIP implementation of galario's
 * model-to-visibility path.
 *
 * Conventions
 *  - Every entry point returns a status: GALARIO_OK (0) or one of the error
 *    codes below; galario_last_error() gives the calling thread's message.
 *    The codes map onto the C++ exceptions galario's core throws
 *    (src/galario.cpp:89-121,158-198; docs/py-api.rst:72-82):
 *        GALARIO_ERR_INVALID  <- std::invalid_argument  (Python ValueError)
 *        GALARIO_ERR_ALLOC    <- std::bad_alloc         (Python MemoryError)
 *        GALARIO_ERR_RUNTIME  <- std::runtime_error     (Python RuntimeError)
 *  - Suffix _f64 computes in double (galario "double"), _f32 in float
 *    (galario "single").  Complex arrays are interleaved (re, im) pairs, the
 *    layout of std::complex / numpy complex.
 *  - Host-pointer entry points (no _dev suffix) copy their inputs to the
 *    current GPU, compute there and copy the results back before returning;
 *    no host pointer is retained.  They are thread-safe: each host thread
 *    uses its own stream and device workspace.
 *  - _dev entry points take device pointers and a hipStream_t (as void*);
 *    they enqueue work and return without synchronising.  Calls from one
 *    host thread share that thread's workspace and are ordered across
 *    streams by the library.
 *  - check_central: 1 keeps galario's central-pixel guard
 *    (CHECK_CENTRAL_PIXEL, src/galario.cpp:113-121: (dxy/2-Rmin)/dR >= 5);
 *    0 skips it (the arithmetic below the guard stays defined whenever
 *    0 <= floor((dxy/2-Rmin)/dR) and floor(...)+1 < nr, which is checked).
 */
#ifndef GALARIO_AMD_H
#define GALARIO_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define GALARIO_OK 0
#define GALARIO_ERR_INVALID 1
#define GALARIO_ERR_ALLOC 2
#define GALARIO_ERR_RUNTIME 3

/* Message of the last failed call on this thread (empty if none). */
const char *galario_last_error(void);

/* ---- management: src/galario.h:51-58 ---------------------------------- */

/* replaces galario::init() (src/galario.h:52; python/__init_module__.py.in:22):
 * with a GPU visible, sets up the current device for the allocation-free
 * entry point (galario_chi2_profile_noalloc_dev_*); without one, does nothing */
int galario_init(void);
/* replaces galario::cleanup() (src/galario.h:53): frees the calling thread's
 * workspaces.  The FFT twiddle tables (shared by all threads and by captured
 * batch graphs) are kept until the process ends, so batch plans and other
 * threads' calls stay valid across a cleanup. */
int galario_cleanup(void);
/* replaces galario::threads(int) (src/galario.h:54): returns the setting
 * (default 16, kept for API compatibility; the kernels' launch shapes are fixed
 * per kernel) */
int galario_threads(int num);
/* replaces galario::ngpus() (src/galario.h:57) */
int galario_ngpus(int *out);
/* replaces galario::use_gpu(int) (src/galario.h:58): selects the device of
 * the calling thread and sets it up as galario_init does */
int galario_use_gpu(int device_id);
/* replaces galario::galario_free(void*) (src/galario.h:42) for buffers from
 * galario_copy_input_* */
void galario_free(void *ptr);

/* ---- main functions: src/galario_py.h:30-37 -------------------------- */

/* replaces galario::_sample_profile (src/galario_py.h:31) */
int galario_sample_profile_f64(int nr, const double *intensity, double Rmin, double dR, double dxy, int nxy,
                               double inc, double dRA, double dDec, double duv, double PA, int nd,
                               const double *u, const double *v, double *vis, int check_central);
int galario_sample_profile_f32(int nr, const float *intensity, float Rmin, float dR, float dxy, int nxy,
                               float inc, float dRA, float dDec, float duv, float PA, int nd,
                               const float *u, const float *v, float *vis, int check_central);

/* replaces galario::_sample_image (src/galario_py.h:33); image is nx x ny row-major */
int galario_sample_image_f64(int nx, int ny, const double *image, double v_origin, double dRA, double dDec,
                             double duv, double PA, int nd, const double *u, const double *v, double *vis);
int galario_sample_image_f32(int nx, int ny, const float *image, float v_origin, float dRA, float dDec,
                             float duv, float PA, int nd, const float *u, const float *v, float *vis);

/* replaces galario::_chi2_profile (src/galario_py.h:34) */
int galario_chi2_profile_f64(int nr, const double *intensity, double Rmin, double dR, double dxy, int nxy,
                             double inc, double dRA, double dDec, double duv, double PA, int nd,
                             const double *u, const double *v, const double *vis_obs_re,
                             const double *vis_obs_im, const double *weights, int check_central,
                             double *chi2);
int galario_chi2_profile_f32(int nr, const float *intensity, float Rmin, float dR, float dxy, int nxy,
                             float inc, float dRA, float dDec, float duv, float PA, int nd,
                             const float *u, const float *v, const float *vis_obs_re,
                             const float *vis_obs_im, const float *weights, int check_central,
                             float *chi2);

/* replaces galario::_chi2_image (src/galario_py.h:36) */
int galario_chi2_image_f64(int nx, int ny, const double *image, double v_origin, double dRA, double dDec,
                           double duv, double PA, int nd, const double *u, const double *v,
                           const double *vis_obs_re, const double *vis_obs_im, const double *weights,
                           double *chi2);
int galario_chi2_image_f32(int nx, int ny, const float *image, float v_origin, float dRA, float dDec,
                           float duv, float PA, int nd, const float *u, const float *v,
                           const float *vis_obs_re, const float *vis_obs_im, const float *weights,
                           float *chi2);

/* replaces galario::_sweep (src/galario_py.h:37); image: nxy x (nxy/2+1)
 * complex = nxy rows of 2*(nxy/2+1) reals, the image in the first nxy reals
 * of each row (padding set to zero) */
int galario_sweep_f64(int nr, const double *intensity, double Rmin, double dR, int nxy, double dxy, double inc,
                      double *image, int check_central);
int galario_sweep_f32(int nr, const float *intensity, float Rmin, float dR, int nxy, float dxy, float inc,
                      float *image, int check_central);

/* replaces galario::_uv_rotate (src/galario_py.h:38) */
int galario_uv_rotate_f64(double PA, double dRA, double dDec, double *dRArot, double *dDecrot, int nd,
                          const double *u, const double *v, double *urot, double *vrot);
int galario_uv_rotate_f32(float PA, float dRA, float dDec, float *dRArot, float *dDecrot, int nd,
                          const float *u, const float *v, float *urot, float *vrot);

/* ---- expert functions: src/galario_py.h:40-46 ------------------------- */

/* replaces galario::_copy_input (src/galario_py.h:40): returns in *out a
 * host buffer of nx x (ny/2+1) complex holding the rows of realdata padded to
 * 2*(ny/2+1) reals; free it with galario_free */
int galario_copy_input_f64(int nx, int ny, const double *realdata, void **out);
int galario_copy_input_f32(int nx, int ny, const float *realdata, void **out);

/* replaces galario::_fft2d (src/galario_py.h:41): in-place forward R2C 2-D
 * FFT of the padded buffer (unnormalised, e^{-2 pi i kn/N}) */
int galario_fft2d_f64(int nx, int ny, void *data);
int galario_fft2d_f32(int nx, int ny, void *data);

/* replaces galario::_fftshift (src/galario_py.h:42): in-place quadrant swap
 * of the real image held in the padded buffer */
int galario_fftshift_f64(int nx, int ny, void *data);
int galario_fftshift_f32(int nx, int ny, void *data);

/* replaces galario::_fftshift_axis0 (src/galario_py.h:43): in-place swap of
 * the row halves of an nrow x ncol complex matrix */
int galario_fftshift_axis0_f64(int nrow, int ncol, void *data);
int galario_fftshift_axis0_f32(int nrow, int ncol, void *data);

/* replaces galario::_interpolate (src/galario_py.h:44); data is an
 * nrow x ncol row-major complex matrix (the shifted R2C spectrum).  Points
 * outside the grid give GALARIO_ERR_INVALID (galario does not check them). */
int galario_interpolate_f64(int nrow, int ncol, const void *data, double v_origin, int nd, const double *u,
                            const double *v, double duv, void *vis);
int galario_interpolate_f32(int nrow, int ncol, const void *data, float v_origin, int nd, const float *u,
                            const float *v, float duv, void *vis);

/* replaces galario::_apply_phase_sampled (src/galario_py.h:45): vis in place */
int galario_apply_phase_sampled_f64(double dRA, double dDec, int nd, const double *u, const double *v,
                                    void *vis);
int galario_apply_phase_sampled_f32(float dRA, float dDec, int nd, const float *u, const float *v,
                                    void *vis);

/* replaces galario::_reduce_chi2 (src/galario_py.h:46) */
int galario_reduce_chi2_f64(int nd, const double *vis_obs_re, const double *vis_obs_im, const void *vis,
                            const double *weights, double *chi2);
int galario_reduce_chi2_f32(int nd, const float *vis_obs_re, const float *vis_obs_im, const void *vis,
                            const float *weights, float *chi2);

/* ---- device-resident forms (no counterpart in galario: the hot path with
 *      the data already in HBM, for torch callers and the benchmark) ------
 *
 * uvmax (last argument, wavelengths): a bound on sqrt(u^2 + v^2) over the
 * points, or <= 0 for none.  Rotation keeps that radius, so with a bound the
 * spectrum is stored (and its adjoint read) only within uvmax/duv + 3 cells
 * of the origin; a point beyond the bound is treated as outside the grid
 * (NaN result, flag set), never read from unstored cells. */

/* galario::chi2_profile (src/galario.cpp:1615) on device arrays; *chi2_d
 * (device, double) receives chi2.  Enqueued on `stream`, not synchronised.
 * Points outside the grid make chi2 NaN and set the flag read by
 * galario_take_oob_flag. */
int galario_chi2_profile_dev_f64(void *stream, int nr, const double *intensity, double Rmin, double dR,
                                 double dxy, int nxy, double inc, double dRA, double dDec, double duv,
                                 double PA, int nd, const double *u, const double *v, const double *vis_obs_re,
                                 const double *vis_obs_im, const double *weights, int check_central,
                                 double *chi2_d, double uvmax);
int galario_chi2_profile_dev_f32(void *stream, int nr, const float *intensity, float Rmin, float dR,
                                 float dxy, int nxy, float inc, float dRA, float dDec, float duv,
                                 float PA, int nd, const float *u, const float *v, const float *vis_obs_re,
                                 const float *vis_obs_im, const float *weights, int check_central,
                                 double *chi2_d, double uvmax);

/* galario::sample_profile (src/galario.cpp:1404) on device arrays; vis_d: nd
 * complex.  perm (device, may be null): point i is written to vis_d[perm[i]],
 * so a dataset stored in a locality order returns visibilities in the
 * caller's order. */
int galario_sample_profile_dev_f64(void *stream, int nr, const double *intensity, double Rmin, double dR,
                                   double dxy, int nxy, double inc, double dRA, double dDec, double duv,
                                   double PA, int nd, const double *u, const double *v, int check_central,
                                   const int *perm, void *vis_d, double uvmax);
int galario_sample_profile_dev_f32(void *stream, int nr, const float *intensity, float Rmin, float dR,
                                   float dxy, int nxy, float inc, float dRA, float dDec, float duv,
                                   float PA, int nd, const float *u, const float *v, int check_central,
                                   const int *perm, void *vis_d, double uvmax);

/* chi2 of a profile and its gradient (no counterpart in galario): *chi2_d,
 * grad_I[0..nr) = dchi2/dintensity, grad_geom[0..3] = dchi2/d(inc, PA, dRA,
 * dDec), all device doubles, enqueued on `stream`.  Requires nxy a power of
 * two in [128, 8192] (double) or [128, 16384] (float) and nr <= 2048.  dchi2/dintensity is
 * accumulated with atomics (summation order, hence the last bits, may vary
 * between calls). */
int galario_chi2_profile_grad_dev_f64(void *stream, int nr, const double *intensity, double Rmin, double dR,
                                      double dxy, int nxy, double inc, double dRA, double dDec, double duv,
                                      double PA, int nd, const double *u, const double *v,
                                      const double *vis_obs_re, const double *vis_obs_im, const double *weights,
                                      int check_central, double *chi2_d, double *grad_I, double *grad_geom,
                                      double uvmax);
int galario_chi2_profile_grad_dev_f32(void *stream, int nr, const float *intensity, float Rmin, float dR,
                                      float dxy, int nxy, float inc, float dRA, float dDec, float duv,
                                      float PA, int nd, const float *u, const float *v,
                                      const float *vis_obs_re, const float *vis_obs_im, const float *weights,
                                      int check_central, double *chi2_d, double *grad_I, double *grad_geom,
                                      double uvmax);

/* chi2 of a profile on device arrays without allocating device memory.  No
 * counterpart in galario: for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, checks and kernels as galario_chi2_profile_dev_* (chi2 is
 * bit-identical); the kernels' scratch -- the row spectra, the half spectrum
 * and the per-workgroup partials -- lives in one 48 MiB device array per GPU,
 * which the library holds from the first use of that GPU until the process
 * ends.  Limits (else GALARIO_ERR_INVALID, nothing enqueued): nxy a power of
 * two from 128 to 8192 (double) or 16384 (float), and
 *   8 ceil(nd/256) + s (2 ceil((Rs'+1)/2) ldT + (nxy/2+2) ldH) <= 48 MiB
 * (each term rounded up to 256 bytes; s = 8 or 4 bytes, Rs' = min(rmax+1,
 * nxy/2), ldT = ldH = nxy/2+1 rounded up to 8): nxy = 4096 in double with a
 * swept radius of 201 pixels and 1e6 points needs 37 MB; nxy = 8192 in double
 * does not fit.  Set-up: galario_init() or galario_use_gpu() with the device
 * current (they resolve the array's device address and build the FFT
 * twiddle tables); a device not set up gives GALARIO_ERR_INVALID.  Points
 * outside the grid make chi2 NaN (no flag is set).  Every call on one GPU
 * uses the same array: calls that overlap on one GPU (on any streams) race,
 * so order them. */
int galario_chi2_profile_noalloc_dev_f64(void *stream, int nr, const double *intensity, double Rmin, double dR,
                                         double dxy, int nxy, double inc, double dRA, double dDec, double duv,
                                         double PA, int nd, const double *u, const double *v,
                                         const double *vis_obs_re, const double *vis_obs_im,
                                         const double *weights, int check_central, double *chi2_d, double uvmax);
int galario_chi2_profile_noalloc_dev_f32(void *stream, int nr, const float *intensity, float Rmin, float dR,
                                         float dxy, int nxy, float inc, float dRA, float dDec, float duv,
                                         float PA, int nd, const float *u, const float *v, const float *vis_obs_re,
                                         const float *vis_obs_im, const float *weights, int check_central,
                                         double *chi2_d, double uvmax);

/* Several sources with the same profile, inclination and PA at their own
 * offsets (galario's multi-source recipe, docs/tech-specs.rst:45-57: sum the
 * visibilities of each source phase-shifted to its position; no counterpart
 * function in galario).  src: nsrc x 3 device array (flux, dRA, dDec); the
 * result is sum_k flux_k * sampleProfile(..., dRA_k, dDec_k), computed from
 * one spectrum and one interpolation per point (sources in order).  nsrc = 1,
 * flux 1 reproduces the single-source call bit for bit. */
int galario_sample_profile_multi_dev_f64(void *stream, int nr, const double *intensity, double Rmin, double dR,
                                         double dxy, int nxy, double inc, double duv, double PA, int nsrc,
                                         const double *src, int nd, const double *u, const double *v,
                                         int check_central, const int *perm, void *vis_d, double uvmax);
int galario_sample_profile_multi_dev_f32(void *stream, int nr, const float *intensity, float Rmin, float dR,
                                         float dxy, int nxy, float inc, float duv, float PA, int nsrc,
                                         const float *src, int nd, const float *u, const float *v,
                                         int check_central, const int *perm, void *vis_d, double uvmax);
int galario_chi2_profile_multi_dev_f64(void *stream, int nr, const double *intensity, double Rmin, double dR,
                                       double dxy, int nxy, double inc, double duv, double PA, int nsrc,
                                       const double *src, int nd, const double *u, const double *v,
                                       const double *vis_obs_re, const double *vis_obs_im, const double *weights,
                                       int check_central, double *chi2_d, double uvmax);
int galario_chi2_profile_multi_dev_f32(void *stream, int nr, const float *intensity, float Rmin, float dR,
                                       float dxy, int nxy, float inc, float duv, float PA, int nsrc,
                                       const float *src, int nd, const float *u, const float *v,
                                       const float *vis_obs_re, const float *vis_obs_im, const float *weights,
                                       int check_central, double *chi2_d, double uvmax);

/* ---- batched form (no counterpart in galario, whose callers loop over
 *      chi2Profile per walker and dataset, docs/quickstart.rst:147-174) ---- */

/* A dataset of the batched form: device arrays of nd points (the caller keeps
 * them alive and unchanged while a plan uses them), its grid nxy (a power of
 * two from 128 to 8192 in double, 16384 in float) and cell size dxy, duv
 * (galario passes 1/(nxy*dxy), python/libcommon.pyx:723), and uvmax as for
 * the device-resident forms (<= 0: none; a grid's spectrum is cut only when
 * every dataset on it gives one).  The float form takes duv in double: its
 * FP32 kernels round it to float, the FP32-bulk form with FP64 geometry
 * (galario_batch_chi2_profile_f32x) uses it as given. */
typedef struct {
    int nxy;
    double dxy, duv;
    int nd;
    const double *u, *v, *vis_obs_re, *vis_obs_im, *weights;
    double uvmax;
} galario_dataset_f64;
typedef struct {
    int nxy;
    float dxy;
    double duv;
    int nd;
    const float *u, *v, *vis_obs_re, *vis_obs_im, *weights;
    double uvmax;
} galario_dataset_f32;

/* Plan for up to max_walkers parameter sets, each a profile of nr values on
 * (Rmin, dR) and (inc, PA, dRA, dDec), evaluated against nsets (<= 32)
 * datasets on the current device.  Datasets on the same (nxy, dxy) share one
 * spectrum per walker.  check_central as above (checked here, per grid). */
int galario_batch_create_f64(int nsets, const galario_dataset_f64 *sets, int nr, double Rmin, double dR,
                             int max_walkers, int check_central, void **plan);
int galario_batch_create_f32(int nsets, const galario_dataset_f32 *sets, int nr, float Rmin, float dR,
                             int max_walkers, int check_central, void **plan);
int galario_batch_destroy(void *plan);

#define GALARIO_BATCH_EAGER 1   /* flag: launch kernel by kernel instead of the captured graph */

/* chi2 of nwalkers parameter sets: intensity [nwalkers][nr], geom
 * [nwalkers][4] = (inc, PA, dRA, dDec), device arrays; outputs (device
 * doubles, either may be null) chi2_sets [nwalkers][nsets] and chi2
 * [nwalkers] = sum over the datasets.  Enqueued on `stream`.  The first call
 * for a given nwalkers times chunk sizes (walkers whose spectra are built and
 * sampled together) on `stream` and keeps the fastest (synchronising), then
 * captures the launch sequence in a graph; later calls copy the parameters
 * in, launch the graph and copy the results out.  chi2_sets[w][s] equals the
 * single-call chi2 of walker w on dataset s (same kernels and summation
 * order; the per-walker cos/sin are computed on the device).  A point outside
 * a grid makes that chi2 NaN and sets the flag read by
 * galario_batch_take_oob_flag.  Calls on one plan are ordered across streams
 * by the library. */
int galario_batch_chi2_profile_f64(void *plan, void *stream, int nwalkers, const double *intensity,
                                   const double *geom, double *chi2_sets, double *chi2, int flags);
int galario_batch_chi2_profile_f32(void *plan, void *stream, int nwalkers, const float *intensity,
                                   const float *geom, double *chi2_sets, double *chi2, int flags);
/* Same with the gradient: grad_I [nwalkers][nr] = dchi2/dintensity and
 * grad_geom [nwalkers][4] = dchi2/d(inc, PA, dRA, dDec) of the sum over the
 * datasets (device doubles, required). */
int galario_batch_chi2_profile_grad_f64(void *plan, void *stream, int nwalkers, const double *intensity,
                                        const double *geom, double *chi2_sets, double *chi2, double *grad_I,
                                        double *grad_geom, int flags);
int galario_batch_chi2_profile_grad_f32(void *plan, void *stream, int nwalkers, const float *intensity,
                                        const float *geom, double *chi2_sets, double *chi2, double *grad_I,
                                        double *grad_geom, int flags);

/* FP32 bulk with FP64 geometry and phase, on a float plan: the profiles,
 * row and half spectra and the dataset streams stay float (half the bytes of
 * the double form), while geom [nwalkers][4] is double and every per-point
 * step after the spectrum -- rotation, grid position, interpolation, the
 * phase argument urot*2pi*dRArot + vrot*2pi*dDecrot and its sincos (the step
 * galario performs in dreal, src/galario.cpp:877-884,893-894), the residual
 * and chi2 -- runs in double on the float values; chi2 and the gradients
 * accumulate in double.  Same outputs as the calls above. */
int galario_batch_chi2_profile_f32x(void *plan, void *stream, int nwalkers, const float *intensity,
                                    const double *geom, double *chi2_sets, double *chi2, int flags);
int galario_batch_chi2_profile_grad_f32x(void *plan, void *stream, int nwalkers, const float *intensity,
                                         const double *geom, double *chi2_sets, double *chi2, double *grad_I,
                                         double *grad_geom, int flags);

/* Launch modes of a plan (the `mode` of galario_batch_info, _set_chunk, _tile, _set_tile):
 * each mode has its own timed launch shape and captured graph. */
#define GALARIO_BATCH_MODE_GRAD 1       /* the _grad calls */
#define GALARIO_BATCH_MODE_F64_PHASE 2  /* the _f32x calls (float plans) */
/* The launch shape picked for (nwalkers, mode): chunk size (0 before the
 * first call) and walkers per sample workgroup (the forward sample reads each
 * dataset point once for that many walkers; 1 with the gradient), and
 * whether its graph is captured.  Any pointer may be null. */
int galario_batch_info(void *plan, int nwalkers, int mode, int *chunk, int *walkers_per_sample,
                       int *graph_ready);
/* Overrides the timed choice for (nwalkers, mode): chunk size and walkers
 * per sample workgroup (1, 2, 4 or 8). */
int galario_batch_set_chunk(void *plan, int nwalkers, int mode, int chunk, int walkers_per_sample);
/* The tile of (nwalkers, mode)'s sample kernel, chosen with the chunk
 * (0 before the first call): forward -- point_blocks, the 256-point blocks
 * per sample workgroup (1 or 2; each block's partial is the same fixed-order
 * sum either way, so the outputs do not depend on it), grad_window 0;
 * gradient -- point_blocks 1 and grad_window, the LDS window of the
 * corner-gradient scatter in cells (2048, 3072 or 4096; a workgroup whose
 * corners fit it accumulates there, others add directly to the half
 * spectrum).  Any pointer may be null. */
int galario_batch_tile(void *plan, int nwalkers, int mode, int *point_blocks, int *grad_window);
/* Fixes the tile of (nwalkers, mode) (point_blocks 1 or 2, grad_window 2048,
 * 3072 or 4096): the timed choice of the first call keeps it and tunes the
 * rest of the launch shape. */
int galario_batch_set_tile(void *plan, int nwalkers, int mode, int point_blocks, int grad_window);
/* Synchronises the plan's last stream; returns (and clears) its out-of-grid flag. */
int galario_batch_take_oob_flag(void *plan, int *out);

/* Per-stage timing of the calling thread's _dev calls (no counterpart in
 * galario, whose GALARIO_TIMING build prints cudaEvent timings per stage,
 * src/galario.cpp:275-315): enable=1 starts recording a HIP event before each
 * stage (0 row pass, 1 column pass, 2 sample (forward, or forward and
 * corner-gradient scatter), 3 adjoint column pass, 4 adjoint row pass +
 * profile adjoint, 5 final reduction of the per-workgroup partials) and after
 * the last, on the call's stream; enable=0 stops.  Either clears what was
 * recorded. */
int galario_stage_timer(int enable);
/* While the stage timer is on, launch the kernel of `stage` (0 row pass,
 * 1 column pass, 2 forward sample; -1 none) `reps` times back to back inside
 * its timed interval, so that the interval / reps is that kernel's average
 * launch duration.  The three kernels are idempotent; results are unchanged. */
int galario_stage_repeat(int stage, int reps);
/* Waits for the recorded events and writes the summed milliseconds of each
 * stage to ms[0..nstage-1] and the number of timed calls to *ncalls. */
int galario_stage_times(double *ms, int nstage, int *ncalls);

/* Read-rate reference for the measurement tools (no counterpart in galario):
 * `grid` workgroups of 256 threads read `bytes` (a multiple of 16) of device
 * memory `buf` in 16-byte loads, grid-stride, and write one sum per
 * workgroup to the device array out[0..grid-1], on `stream`.  Timing it over
 * the bytes a kernel reads gives the rate a plain stream of that size
 * reaches on this GPU. */
int galario_read_probe(void *stream, const void *buf, long long bytes, int grid, double *out);

/* Synchronises the calling thread's work on the current device and returns
 * (and clears) its out-of-grid flag: 1 if any point of a _dev call since the
 * last take fell outside the Fourier grid. */
int galario_take_oob_flag(int *out);

#ifdef __cplusplus
}
#endif

#endif /* GALARIO_AMD_H */
