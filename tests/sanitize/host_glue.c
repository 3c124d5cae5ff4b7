/* Exercises the host side of libgalario_amd (argument checks, error strings, plan handling) in a
 * process without a GPU; built with AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_sanitizers.py against a host-only build of the library.  Every call below must fail
 * cleanly with a status and a message, before any GPU work. */
#include <stdio.h>
#include <string.h>

#include "galario_amd.h"

static int failures = 0;

static void expect_error(int status, const char *what) {
    const char *msg = galario_last_error();
    if (status == GALARIO_OK || !msg || !msg[0]) {
        printf("FAIL %s: status %d, message '%s'\n", what, status, msg ? msg : "(null)");
        ++failures;
    } else {
        printf("ok   %s: %d %s\n", what, status, msg);
    }
}

int main(void) {
    double I[100], u[10], v[10], re[10], im[10], w[10], vis[20], out = 0.0;
    float If[100], uf[10];
    for (int i = 0; i < 100; ++i) I[i] = If[i] = 1.0f;
    for (int i = 0; i < 10; ++i) u[i] = v[i] = re[i] = im[i] = w[i] = uf[i] = 1e3f * (float)(i + 1);
    const double arcsec = 4.84813681109536e-06;

    /* host-pointer entry points: size and profile checks come first */
    expect_error(galario_sample_profile_f64(100, I, 1e-6 * arcsec, 0.001 * arcsec, 0.2 * arcsec, 1023, 0.0, 0.0,
                                            0.0, 1.0, 0.0, 10, u, v, vis, 1),
                 "sample_profile odd nxy");
    expect_error(galario_chi2_profile_f64(100, I, 1e-4 * arcsec, 0.002 * arcsec, 0.01 * arcsec, 4096, 0.0, 0.0, 0.0,
                                          1.0, 0.0, 10, u, v, re, im, w, 1, &out),
                 "chi2_profile central-pixel guard");
    expect_error(galario_sweep_f64(3, I, 1e-6 * arcsec, 0.001 * arcsec, 64, 0.2 * arcsec, 0.0, vis, 1),
                 "sweep central pixel beyond the profile");
    expect_error(galario_sample_profile_f32(100, If, 1e-6f, 1e-9f, 1e-7f, 1023, 0.f, 0.f, 0.f, 1.f, 0.f, 10, uf, uf,
                                            (float *)vis, 1),
                 "sample_profile_f32 odd nxy");
    /* device entry points: null arrays and unsupported sizes are rejected on the host */
    expect_error(galario_chi2_profile_dev_f64(NULL, 100, NULL, 1e-6 * arcsec, 0.01 * arcsec, 0.2 * arcsec, 1024, 0.0,
                                              0.0, 0.0, 1.0, 0.0, 10, NULL, NULL, NULL, NULL, NULL, 0, NULL, 0.0),
                 "chi2_profile_dev null arrays");
    expect_error(galario_chi2_profile_grad_dev_f64(NULL, 3000, NULL, 1e-6 * arcsec, 0.0005 * arcsec, 0.01 * arcsec,
                                                   1024, 0.5, 0.0, 0.0, 1.0, 0.3, 10, NULL, NULL, NULL, NULL, NULL, 0,
                                                   NULL, NULL, NULL, 0.0),
                 "chi2_profile_grad_dev too many bins");
    expect_error(galario_chi2_profile_noalloc_dev_f64(NULL, 100, I, 1e-6 * arcsec, 0.01 * arcsec, 0.2 * arcsec,
                                                      1000, 0.0, 0.0, 0.0, 1.0, 0.0, 10, u, v, re, im, w,
                                                      0, NULL, 0.0),
                 "chi2_profile_noalloc_dev nxy not a power of two");
    /* batch plans */
    void *plan = NULL;
    galario_dataset_f64 ds;
    memset(&ds, 0, sizeof ds);
    ds.nxy = 512;
    ds.dxy = 0.01 * arcsec;
    ds.duv = 1.0 / (512 * ds.dxy);
    expect_error(galario_batch_create_f64(0, &ds, 100, 1e-6 * arcsec, 0.001 * arcsec, 4, 0, &plan), "batch no sets");
    expect_error(galario_batch_create_f64(1, NULL, 100, 1e-6 * arcsec, 0.001 * arcsec, 4, 0, &plan), "batch null sets");
    ds.nd = 10;
    expect_error(galario_batch_create_f64(1, &ds, 100, 1e-6 * arcsec, 0.001 * arcsec, 4, 0, &plan),
                 "batch dataset with null arrays");
    ds.nd = 0;
    ds.nxy = 1000;
    expect_error(galario_batch_create_f64(1, &ds, 100, 1e-6 * arcsec, 0.001 * arcsec, 4, 0, &plan),
                 "batch unsupported nxy");
    ds.nxy = 512;
    expect_error(galario_batch_create_f64(1, &ds, 100, 1e-6 * arcsec, 0.001 * arcsec, 0, 0, &plan),
                 "batch max_walkers 0");
    int a = 0, b = 0, c = 0;
    expect_error(galario_batch_info(NULL, 1, 0, &a, &b, &c), "batch_info null plan");
    expect_error(galario_batch_tile(NULL, 1, 0, &a, &b), "batch_tile null plan");
    expect_error(galario_batch_set_tile(NULL, 1, 0, 1, 3072), "batch_set_tile null plan");
    expect_error(galario_batch_set_chunk(NULL, 1, 0, 1, 4), "batch_set_chunk null plan");
    expect_error(galario_batch_chi2_profile_f64(NULL, NULL, 1, NULL, NULL, NULL, NULL, 0), "batch call null plan");
    expect_error(galario_batch_take_oob_flag(NULL, &a), "batch_take_oob_flag null plan");
    if (galario_batch_destroy(NULL) != GALARIO_OK) {
        printf("FAIL batch_destroy(NULL) should be a no-op\n");
        ++failures;
    }
    /* measurement helpers */
    expect_error(galario_read_probe(NULL, NULL, 17, 64, &out), "read_probe odd size");
    /* settings that need no GPU */
    const int t = galario_threads(0);
    printf("threads %d\n", t);
    printf("%s\n", failures ? "FAILED" : "ALL OK");
    return failures ? 1 : 0;
}
