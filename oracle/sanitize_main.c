/* sanitize_main.c -- drives every entry point of the CPU oracle (oracle.c,
 * oracle64.c) once under AddressSanitizer and UndefinedBehaviorSanitizer
 * (`make sanitize`).  TEST INFRASTRUCTURE ONLY.
 *
 * Cases: ragged and tiny sizes (1 x 1 up to odd sizes that do not halve
 * evenly), the three solvers at small sizes, and frames holding NaN, +Inf
 * and -Inf at the first pixel and inside.  The program checks status codes
 * only; the numbers are compared elsewhere.  Exit 0 means every call
 * returned and no sanitizer report was raised (the sanitizers abort on the
 * first report). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static uint32_t rng = 12345u;
static float frand(void)
{
    rng = rng * 1664525u + 1013904223u;
    return (float)(rng >> 8) / 16777216.0f;
}

static float *frame(int w, int h, float shift)
{
    float *f = malloc(sizeof(float) * (size_t)w * h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            f[(size_t)y * w + x] =
                (float)(128.0 + 60.0 * sin(0.21 * (x - shift)) * cos(0.17 * y) + 8.0 * frand());
    return f;
}

static uint8_t *quant(const float *f, size_t n)
{
    uint8_t *q = malloc(n);
    for (size_t i = 0; i < n; i++) {
        const float v = f[i] < 0 ? 0 : (f[i] > 255 ? 255 : f[i]);
        q[i] = (uint8_t)(v + 0.5f);
    }
    return q;
}

static int fails = 0;
#define EXPECT(cond, what)                                        \
    do {                                                          \
        if (!(cond)) {                                            \
            fprintf(stderr, "FAILED: %s (%s:%d)\n", what, __FILE__, __LINE__); \
            fails++;                                              \
        }                                                         \
    } while (0)

static void tvl1_case(int w, int h, int nscales, int nonfinite)
{
    orc_tvl1_params p;
    orc_tvl1_default_params(&p);
    p.nscales = nscales;
    p.max_iterations = 20;
    float *a = frame(w, h, 0.0f), *b = frame(w, h, 1.5f);
    const size_t n = (size_t)w * h;
    if (nonfinite == 1) { a[0] = NAN; b[n / 2] = NAN; }
    if (nonfinite == 2) { a[n / 3] = INFINITY; b[0] = -INFINITY; }
    const int ns = orc_tvl1_nscales(w, h, nscales, p.zfactor);
    float *u = calloc(2 * n, sizeof(float));
    int *it = calloc((size_t)ns * p.warps, sizeof(int));
    double *u64 = calloc(2 * n, sizeof(double));
    int nso = 0;
    int st = orc_tvl1_run_f32(a, b, w, h, &p, u, it, &nso);
    EXPECT(st == ORC_OK || st == ORC_ERR_SIGMA, "orc_tvl1_run_f32");
    st = o64_tvl1_run_f32(a, b, w, h, &p, u64, it, &nso);
    EXPECT(st == ORC_OK || st == ORC_ERR_SIGMA || (st == ORC_ERR_ARG && (w < 2 || h < 2)),
           "o64_tvl1_run_f32");
    if (!nonfinite) {
        uint8_t *qa = quant(a, n), *qb = quant(b, n);
        st = orc_tvl1_run_u8(qa, qb, w, h, &p, u, it, &nso);
        EXPECT(st == ORC_OK || st == ORC_ERR_SIGMA, "orc_tvl1_run_u8");
        if (st == ORC_OK && ns > 1) {
            int warps_s[32], max_s[32];
            for (int s = 0; s < 32; s++) { warps_s[s] = 2; max_s[s] = 7; }
            st = orc_tvl1_run_schedule_u8(qa, qb, w, h, &p, NULL, ns - 1, warps_s, max_s, u, it, &nso);
            EXPECT(st == ORC_OK, "orc_tvl1_run_schedule_u8");
        }
        double data = 0, tv = 0;
        (void)orc_tvl1_energy(a, b, u, u + n, w, h, p.lambda, &data, &tv);
        free(qa);
        free(qb);
    }
    free(a); free(b); free(u); free(it); free(u64);
}

static void pieces_case(int w, int h)
{
    const size_t n = (size_t)w * h;
    float *a = frame(w, h, 0.0f), *b = frame(w, h, 0.7f);
    float *dx = malloc(sizeof(float) * n), *dy = malloc(sizeof(float) * n);
    orc_centered_gradient(a, dx, dy, w, h);
    orc_forward_gradient(a, dx, dy, w, h);
    orc_divergence(dx, dy, b, w, h);
    float *c = frame(w, h, 0.3f);
    (void)orc_gaussian(c, w, h, 0.8);
    int nxx = 0, nyy = 0;
    orc_zoom_size(w, h, 0.5f, &nxx, &nyy);
    if (nxx > 0 && nyy > 0) {
        float *z = malloc(sizeof(float) * (size_t)nxx * nyy), *sc = malloc(sizeof(float) * n);
        (void)orc_zoom_out(a, z, w, h, 0.5f, sc);
        orc_zoom_in(z, c, nxx, nyy, w, h);
        free(z);
        free(sc);
    }
    const float coords[] = {-1e30f, -3.5f, 0.0f, 0.49f, (float)w - 1.5f, 1e30f, NAN, INFINITY, -INFINITY};
    for (int i = 0; i < 9; i++)
        for (int j = 0; j < 9; j++) {
            (void)orc_bicubic_at(a, coords[i], coords[j], w, h, 0);
            (void)orc_bicubic_at(a, coords[i], coords[j], w, h, 1);
            (void)orc_bilinear_at(a, w, h, w, coords[i], coords[j]);
            double *a64 = malloc(sizeof(double) * n);
            for (size_t k = 0; k < n; k++) a64[k] = a[k];
            (void)o64_bicubic_at(a64, coords[i], coords[j], w, h, 1);
            free(a64);
        }
    free(a); free(b); free(c); free(dx); free(dy);
}

static void hs_lk_case(int w, int h, int nonfinite)
{
    const size_t n = (size_t)w * h;
    float *a = frame(w, h, 0.0f), *b = frame(w, h, 1.0f);
    if (nonfinite) { a[0] = NAN; b[n / 2] = INFINITY; }
    float *u = calloc(n, sizeof(float)), *v = calloc(n, sizeof(float));
    double *u64 = calloc(n, sizeof(double)), *v64 = calloc(n, sizeof(double));
    int st = orc_hs_compute(a, b, w, h, w, 15.0f, 2, 2, 10, u, v);
    EXPECT(st == ORC_OK || st == ORC_ERR_ARG, "orc_hs_compute");
    st = o64_hs_compute(a, b, w, h, w, 15.0, 2, 2, 10, u64, v64);
    EXPECT(st == ORC_OK || st == ORC_ERR_ARG, "o64_hs_compute");
    st = orc_lk_pyr_compute(a, b, w, h, 3, 0.5f, 1.0f, 7, 3, 1e-6f, u, v);
    EXPECT(st == ORC_OK || st == ORC_ERR_ARG || st == ORC_ERR_SIGMA, "orc_lk_pyr_compute");
    st = o64_lk_pyr_compute(a, b, w, h, 3, 0.5, 1.0, 7, 3, 1e-6, u64, v64);
    EXPECT(st == ORC_OK || st == ORC_ERR_ARG || st == ORC_ERR_SIGMA, "o64_lk_pyr_compute");
    if (!nonfinite && w >= 8 && h >= 8) {
        const int T = 5;
        uint8_t *frames = malloc((size_t)T * n);
        for (int t = 0; t < T; t++) {
            float *f = frame(w, h, 0.5f * t);
            uint8_t *q = quant(f, n);
            memcpy(frames + (size_t)t * n, q, n);
            free(f);
            free(q);
        }
        st = orc_lk_single(frames, T, 2, w, h, 3, 5, u, v);
        EXPECT(st == ORC_OK || st == ORC_ERR_ARG, "orc_lk_single");
        free(frames);
    }
    free(a); free(b); free(u); free(v); free(u64); free(v64);
}

int main(void)
{
    orc_set_threads(1);
    static const int sizes[][2] = {{1, 1}, {2, 3}, {5, 4}, {17, 9}, {33, 21}, {64, 48}, {97, 61}};
    for (size_t k = 0; k < sizeof(sizes) / sizeof(sizes[0]); k++) {
        const int w = sizes[k][0], h = sizes[k][1];
        pieces_case(w, h);
        hs_lk_case(w, h, 0);
        tvl1_case(w, h, 3, 0);
    }
    tvl1_case(80, 60, 4, 1);
    tvl1_case(80, 60, 4, 2);
    hs_lk_case(72, 40, 1);
    orc_set_error_mode(ORC_ERROR_FLOAT_SEQUENTIAL);
    tvl1_case(48, 40, 2, 0);
    orc_set_error_mode(ORC_ERROR_DOUBLE_ROWS);
    if (fails) {
        fprintf(stderr, "%d call(s) returned an unexpected status\n", fails);
        return 1;
    }
    printf("oracle sanitizer run: all calls returned\n");
    return 0;
}
