/* Runs every function of oracle/galario_port.c once on a small problem (nxy = 64) with consistent
 * buffers; built with AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitizers.py.
 * The spectrum step is a naive DFT of the shifted image into the padded R2C layout the port's
 * functions use (nxy rows of nxy/2 + 1 complex values). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

double *port_alloc_image(int nxy);
void port_free(void *p);
int port_threads(int num);
void port_create_image(int nr, const double *intensity, double Rmin, double dR, int nxy, double dxy, double inc,
                       double *image);
void port_shift(int nx, int ny, double *a);
void port_shift_axis0(int nrow, int ncol, double *m);
void port_uv_rotate(double PA, double dRA, double dDec, double *dRArot, double *dDecrot, int nd, const double *u,
                    const double *v, double *urot, double *vrot);
void port_interpolate(int nrow, int ncol, const double *data, double v_origin, int nd, const double *u,
                      const double *v, double duv, double *vis);
void port_apply_phase(double dRA, double dDec, int nd, const double *u, const double *v, double *vis);
double port_reduce_chi2(int nd, const double *re, const double *im, const double *vis, const double *w);

int main(void) {
    const int nxy = 64, ncol = nxy / 2 + 1, nr = 200, nd = 500;
    const double arcsec = 4.84813681109536e-06, dxy = 0.05 * arcsec, Rmin = 1e-4 * arcsec, dR = 0.002 * arcsec;
    const double duv = 1.0 / (nxy * dxy);
    port_threads(2);
    double *I = malloc(sizeof(double) * nr);
    for (int i = 0; i < nr; ++i) I[i] = exp(-0.5 * pow((Rmin + dR * i) / (0.1 * arcsec), 2));
    double *img = port_alloc_image(nxy);
    port_create_image(nr, I, Rmin, dR, nxy, dxy, 0.5, img);
    port_shift(nxy, nxy, img);
    /* naive R2C DFT of the real view (row stride 2 ncol) into a separate complex array */
    double *spec = port_alloc_image(nxy);
    for (int k = 0; k < nxy; ++k)
        for (int l = 0; l < ncol; ++l) {
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < nxy; ++i)
                for (int j = 0; j < nxy; ++j) {
                    const double a = -2.0 * M_PI * ((double)i * k + (double)j * l) / nxy;
                    const double x = img[(size_t)i * 2 * ncol + j];
                    sr += x * cos(a);
                    si += x * sin(a);
                }
            spec[((size_t)k * ncol + l) * 2] = sr;
            spec[((size_t)k * ncol + l) * 2 + 1] = si;
        }
    port_shift_axis0(nxy, ncol, spec);
    double *u = malloc(sizeof(double) * nd), *v = malloc(sizeof(double) * nd), *ur = malloc(sizeof(double) * nd),
           *vr = malloc(sizeof(double) * nd), *re = malloc(sizeof(double) * nd), *im = malloc(sizeof(double) * nd),
           *w = malloc(sizeof(double) * nd), *vis = malloc(sizeof(double) * 2 * nd);
    srand(7);
    for (int i = 0; i < nd; ++i) {
        const double b = duv * (1.0 + (nxy / 2 - 3) * (double)rand() / RAND_MAX), t = 2 * M_PI * rand() / RAND_MAX;
        u[i] = b * cos(t);
        v[i] = b * sin(t);
        re[i] = 1e-3 * rand() / RAND_MAX;
        im[i] = 1e-3 * rand() / RAND_MAX;
        w[i] = 1.0 + (double)rand() / RAND_MAX;
    }
    double dRAr = 0.0, dDecr = 0.0;
    port_uv_rotate(0.4, 0.01 * arcsec, -0.02 * arcsec, &dRAr, &dDecr, nd, u, v, ur, vr);
    port_interpolate(nxy, ncol, spec, 1.0, nd, ur, vr, duv, vis);
    port_apply_phase(dRAr, dDecr, nd, ur, vr, vis);
    const double chi2 = port_reduce_chi2(nd, re, im, vis, w);
    printf("chi2 %.17g\n", chi2);
    port_free(img);
    port_free(spec);
    free(I); free(u); free(v); free(ur); free(vr); free(re); free(im); free(w); free(vis);
    printf("%s\n", isfinite(chi2) && chi2 > 0 ? "ALL OK" : "FAILED");
    return isfinite(chi2) && chi2 > 0 ? 0 : 1;
}
