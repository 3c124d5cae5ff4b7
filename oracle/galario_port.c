/*
 * C/OpenMP restatement of galario's CPU loops (double precision).
 *
 * TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into
 * oracle/_build/libgalario_port.so and driven by oracle/port.py, which adds
 * the FFT stage with scipy.fft.  It is the timed CPU baseline and a second,
 * independent restatement that the numpy oracle is checked against.  The
 * shipped library never links or calls it.
 *
 * Every loop follows src/galario.cpp of galario v1.2.2 (cited per function):
 * same data layouts (FFTW's padded in-place R2C layout, interleaved complex),
 * same evaluation order, IEEE double with no FMA contraction
 * (-ffp-contract=off), OpenMP parallel loops where the reference has them.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* threads(): galario.cpp:367-390 (CPU branch). */
int port_threads(int num)
{
#ifdef _OPENMP
    if (num > 0) {
        omp_set_dynamic(0);
        omp_set_num_threads(num);
    }
    return omp_get_max_threads();
#else
    (void)num;
    return 1;
#endif
}

/* Image allocation as sample_profile does it per call (galario.cpp:1427):
 * nxy rows of 2*(nxy/2+1) reals. */
double *port_alloc_image(int nxy)
{
    size_t n = (size_t)nxy * (size_t)(2 * (nxy / 2 + 1));
    return (double *)aligned_alloc(64, ((n * sizeof(double) + 63) / 64) * 64);
}

void port_free(void *p) { free(p); }

/* create_image_h, galario.cpp:1186-1235, with sweep_core 1048-1071.
 * The central-pixel guard is the caller's (Python) business. */
void port_create_image(int nr, const double *intensity, double Rmin, double dR,
                       int nxy, double dxy, double inc, double *image)
{
    const int ncol = nxy / 2 + 1;
    const int rowsize = 2 * ncol;
    memset(image, 0, sizeof(double) * (size_t)nxy * (size_t)rowsize);

    const double cos_inc = cos(inc);
    int rmax = (int)ceil((Rmin + nr * dR) / dxy);
    if (rmax > nxy / 2) rmax = nxy / 2;
    const double sr_to_px = dxy * dxy;
    const int off = nxy / 2 - rmax;

#pragma omp parallel for
    for (int i = 0; i < 2 * rmax; ++i) {
        for (int j = 0; j < 2 * rmax; ++j) {
            const double x = (rmax - j) * dxy;
            const double y = (rmax - i) * dxy;
            const double xs = x / cos_inc;
            const double r = sqrt(xs * xs + y * y);
            int iR = (int)floor((r - Rmin) / dR);
            if (iR < 0) iR = 0;
            const size_t base = (size_t)(i + off) * rowsize + (size_t)(j + off);
            if (iR > nr - 2)
                image[base] = 0.0;
            else
                image[base] = sr_to_px * (intensity[iR] + (r - iR * dR - Rmin) *
                                          (intensity[iR + 1] - intensity[iR]) / dR);
        }
    }

    /* central pixel: trapezoid rule of 2 I(R) R dR over [Rmin, dxy/2] */
    const int iIN = (int)floor((dxy / 2 - Rmin) / dR);
    double flux = 0.;
    for (int i = 1; i < iIN; ++i)
        flux += (Rmin + dR * i) * intensity[i];
    flux *= 2.;
    flux += Rmin * intensity[0] + (Rmin + iIN * dR) * intensity[iIN];
    flux *= dR;
    const double I_interp = (intensity[iIN + 1] - intensity[iIN]) / dR *
                            (dxy / 2. - (Rmin + dR * iIN)) + intensity[iIN];
    flux += ((Rmin + iIN * dR) * intensity[iIN] + dxy / 2. * I_interp) *
            (dxy / 2. - (Rmin + iIN * dR));
    const double area = (dxy / 2.) * (dxy / 2.) - Rmin * Rmin;
    image[(size_t)(nxy / 2) * rowsize + nxy / 2] = sr_to_px * flux / area;
}

/* shift_h / shift_core, galario.cpp:565-624: quadrant swap on the padded real view. */
void port_shift(int nx, int ny, double *a)
{
    const int rowsize = 2 * (ny / 2 + 1);
#pragma omp parallel for
    for (int x = 0; x < nx / 2; ++x) {
        for (int y = 0; y < ny / 2; ++y) {
            const size_t src_ul = (size_t)x * rowsize + y;
            const size_t src_ur = src_ul + ny / 2;
            const size_t tgt_ul = src_ur + (size_t)(nx / 2) * rowsize;
            const size_t tgt_ur = tgt_ul - ny / 2;
            double t = a[src_ul]; a[src_ul] = a[tgt_ul]; a[tgt_ul] = t;
            t = a[src_ur]; a[src_ur] = a[tgt_ur]; a[tgt_ur] = t;
        }
    }
}

/* shift_axis0_h / core, galario.cpp:657-698: swap rows r <-> r + nrow/2 of a complex matrix. */
void port_shift_axis0(int nrow, int ncol, double *m)
{
#pragma omp parallel for
    for (int x = 0; x < nrow / 2; ++x) {
        for (int y = 0; y < ncol; ++y) {
            const size_t s = 2 * ((size_t)x * ncol + y);
            const size_t t = s + 2 * (size_t)(nrow / 2) * ncol;
            double a = m[s], b = m[s + 1];
            m[s] = m[t]; m[s + 1] = m[t + 1];
            m[t] = a; m[t + 1] = b;
        }
    }
}

/* uv_rotate_h, galario.cpp:976-999. */
void port_uv_rotate(double PA, double dRA, double dDec, double *dRArot, double *dDecrot,
                    int nd, const double *u, const double *v, double *urot, double *vrot)
{
    if (PA == 0.) {
        *dRArot = dRA;
        *dDecrot = dDec;
        memcpy(urot, u, sizeof(double) * (size_t)nd);
        memcpy(vrot, v, sizeof(double) * (size_t)nd);
        return;
    }
    const double c = cos(PA), s = sin(PA);
#pragma omp parallel for
    for (int i = 0; i < nd; ++i) {
        urot[i] = u[i] * c - v[i] * s;
        vrot[i] = u[i] * s + v[i] * c;
    }
    *dRArot = dRA * c - dDec * s;
    *dDecrot = dRA * s + dDec * c;
}

/* interpolate_h / interpolate_core, galario.cpp:749-838 (interleaved complex, row-major nrow x ncol). */
void port_interpolate(int nrow, int ncol, const double *data, double v_origin, int nd,
                      const double *u, const double *v, double duv, double *vis)
{
    const int half_nrow = nrow / 2;
#pragma omp parallel for
    for (int idx = 0; idx < nd; ++idx) {
        const double indu = fabs(u[idx]) / duv;
        const double sign_u = copysign(1., u[idx]);
        const double indv = half_nrow + v_origin * sign_u * v[idx] / duv;
        const int fl_u = (int)floor(indu);
        const int fl_v = (int)floor(indv);
        const double t = indv - fl_v;
        const double q = indu - fl_u;
        const long base = (long)fl_u + (long)fl_v * ncol;
        const double *y0 = data + 2 * base;
        const double *y1 = data + 2 * (base + ncol);
        const double *y2 = data + 2 * (base + ncol + 1);
        const double *y3 = data + 2 * (base + 1);
        /* Z = t*(q*((y0+y2)-(y1+y3))) + (t*(y1-y0) + q*(y3-y0)) + y0 */
        const double df_r = (y0[0] + y2[0]) - (y1[0] + y3[0]);
        const double df_i = (y0[1] + y2[1]) - (y1[1] + y3[1]);
        const double term1_r = t * (q * df_r), term1_i = t * (q * df_i);
        const double term2_r = t * (y1[0] - y0[0]), term2_i = t * (y1[1] - y0[1]);
        const double term3_r = q * (y3[0] - y0[0]), term3_i = q * (y3[1] - y0[1]);
        const double z_r = (term1_r + (term2_r + term3_r)) + y0[0];
        const double z_i = (term1_i + (term2_i + term3_i)) + y0[1];
        const double phase = atan2(z_i, z_r) * sign_u;
        const double a0 = hypot(y0[0], y0[1]), a1 = hypot(y1[0], y1[1]);
        const double a2 = hypot(y2[0], y2[1]), a3 = hypot(y3[0], y3[1]);
        double amp = a0;
        amp += (a3 - a0) * q;
        amp += (a1 - a0) * t;
        amp += (a0 - a1 + a2 - a3) * t * q;
        vis[2 * idx] = amp * cos(phase);
        vis[2 * idx + 1] = amp * sin(phase);
    }
}

/* apply_phase_sampled_h, galario.cpp:877-921. */
void port_apply_phase(double dRA, double dDec, int nd, const double *u, const double *v, double *vis)
{
    if (dRA == 0. && dDec == 0.) return;
    dRA *= 2. * M_PI;
    dDec *= 2. * M_PI;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < nd; ++x) {
        const double ang = u[x] * dRA + v[x] * dDec;
        const double c = cos(ang), s = sin(ang);
        const double re = vis[2 * x], im = vis[2 * x + 1];
        vis[2 * x] = re * c - im * s;
        vis[2 * x + 1] = re * s + im * c;
    }
}

/* reduce_chi2 (CPU branch), galario.cpp:1506-1538 with diff_weighted_core 1453-1461. */
double port_reduce_chi2(int nd, const double *re, const double *im, const double *vis, const double *w)
{
    double chi2 = 0.;
#pragma omp parallel for reduction(+ : chi2)
    for (int i = 0; i < nd; ++i) {
        const double sw = sqrt(w[i]);
        const double a = (vis[2 * i] - re[i]) * sw;
        const double b = (vis[2 * i + 1] - im[i]) * sw;
        chi2 += a * a + b * b;
    }
    return chi2;
}
