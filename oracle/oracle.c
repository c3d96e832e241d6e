/*
 * oracle.c -- CPU restatement of the reference optical-flow solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; it exists so that tests/, the smoke check and the CPU
 * baseline leg of bench.py have a checker to compare the HIP path against.
 *
 * TV-L1 follows the reference's sequential C++ path operation by operation
 * (same float/double promotions, same evaluation order, no contraction), so
 * that the HIP kernels, which are written to the same arithmetic, can be
 * compared bit for bit.  Citations are file:line in the reference tree
 * (artecs-group/sycl-optic-flow).
 *
 *   src/tvl1/sequential/tvl1.cpp, tvl1.hpp:6-8
 *   src/tvl1/sequential/mask/mask.cpp, mask.hpp:7
 *   src/tvl1/sequential/zoom/zoom.cpp, zoom.hpp:4,14-26
 *   src/tvl1/sequential/bicubic_interpolation/bicubic_interpolation.cpp
 *
 * Deliberate differences from the reference (each documented where it occurs):
 *   - the per-iteration stopping error is accumulated in double in a fixed
 *     order (the reference uses an OpenMP float reduction whose order depends
 *     on the thread count), see orc_dual_tvl1;
 *   - errors are returned as status codes instead of `throw;`;
 *   - a Gaussian wider than the image height is rejected (the reference only
 *     checks the width and reads out of bounds in the column pass).
 *
 * MKL routines used by the reference (cblas_sscal/saxpy/isamax/isamin/dasum/
 * dscal, oneAPI MKL 2023.2.0 per README.md:14) are restated from their
 * published BLAS semantics: isamax/isamin return the first index of the
 * largest/smallest |x|; saxpy is y := a*x + y evaluated as a rounded product
 * followed by a rounded sum; dasum sums |x| left to right.  Whether MKL fuses
 * the saxpy multiply-add or reorders dasum on a given CPU is not observable
 * from the reference tree: that bit-level detail is "parity unpinned".
 *
 * Horn-Schunck and Lucas-Kanade have no CPU implementation in the reference;
 * their GPU code has out-of-bounds index bugs (src/hs/cuda/kernels/
 * warpingKernel.cuh:38-42, src/lk/cuda/kernels/kernels.cu:106).  The
 * restatements below implement the algorithms' stated intent with the
 * border rules written next to each function.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

/* tvl1.hpp:6-8, zoom.hpp:4, mask.hpp:7 */
#define ORC_PRESMOOTHING_SIGMA 0.8f
#define ORC_GRAD_IS_ZERO 1E-10f
#define ORC_ZOOM_SIGMA_ZERO 0.6f
#define ORC_GAUSSIAN_WINDOW 5

/* ------------------------------------------------------------------------- */
/* pyramid sizes                                                             */
/* ------------------------------------------------------------------------- */

/* tvl1.cpp:42-43.  N is computed as in the reference: hypot of the integer
 * sizes in double, log(1/zfactor) in float (std::log(float) overload), the
 * quotient in double, assigned to float, then min'ed with nscales in float and
 * truncated to int.  Clamped to >= 1 (the reference yields 0 for tiny images,
 * which is undefined behaviour there). */
int orc_tvl1_nscales(int width, int height, int nscales, float zfactor)
{
    const float logz = logf(1.0f / zfactor);
    const double q = log(hypot((double)width, (double)height) / 16.0) / (double)logz;
    const float N = (float)(1.0 + q);
    const float pick = (N < (float)nscales) ? N : (float)nscales;
    int ns = (int)pick;
    if (ns < 1) ns = 1;
    return ns;
}

/* zoom.hpp:14-26 */
void orc_zoom_size(int nx, int ny, float factor, int *nxx, int *nyy)
{
    *nxx = (int)((double)((float)nx * factor) + 0.5);
    *nyy = (int)((double)((float)ny * factor) + 0.5);
}

/* ------------------------------------------------------------------------- */
/* Gaussian (mask.cpp:204-274)                                               */
/* ------------------------------------------------------------------------- */

/* Tap count and normalised taps, mask.cpp:212-231.  Returns the tap count. */
int orc_gaussian_taps(double sigma, double *B, int max_taps)
{
    const double den = 2 * sigma * sigma;
    const double sPi = sigma * sqrt(M_PI * 2);
    const int size = (int)((double)ORC_GAUSSIAN_WINDOW * sigma) + 1;
    if (size > max_taps) return -1;
    for (int i = 0; i < size; i++)
        B[i] = 1 / sPi * exp(-i * i / den);
    /* cblas_dasum: sum of |B| left to right */
    double norm = 0.0;
    for (int i = 0; i < size; i++) norm += fabs(B[i]);
    norm = norm * 2 - B[0];
    /* cblas_dscal(size, 1/norm, B): B[i] = alpha * B[i] */
    const double alpha = 1 / norm;
    for (int i = 0; i < size; i++) B[i] = alpha * B[i];
    return size;
}

/* One filtered line.  `line` holds n samples; the window extends with a
 * whole-sample mirror on the low side (x[-k] = x[k], mask.cpp:240-241) and a
 * half-sample mirror on the high side (x[n-1+k] = x[n-k], mask.cpp:242).
 * Accumulation in double, pair sums in float (mask.cpp:247-250). */
static void gaussian_line(const float *line, float *out, int n, const double *B, int size)
{
    for (int x = 0; x < n; x++) {
        double sum = B[0] * (double)line[x];
        for (int j = 1; j < size; j++) {
            int xl = x - j, xr = x + j;
            if (xl < 0) xl = -xl;
            if (xr >= n) xr = 2 * n - 1 - xr;
            const float pair = line[xl] + line[xr];
            sum += B[j] * (double)pair;
        }
        out[x] = (float)sum;
    }
}

/* In-place separable Gaussian, rows then columns (mask.cpp:204-274).
 * Returns ORC_ERR_SIGMA if the window does not fit the image. */
int orc_gaussian(float *I, int xdim, int ydim, double sigma)
{
    double B[64];
    const int size = orc_gaussian_taps(sigma, B, 64);
    if (size < 0 || size > xdim || size > ydim) return ORC_ERR_SIGMA;
    const int nmax = xdim > ydim ? xdim : ydim;

    #pragma omp parallel
    {
        float *line = (float *)malloc(sizeof(float) * (size_t)nmax);
        float *res = (float *)malloc(sizeof(float) * (size_t)nmax);
        #pragma omp for
        for (int k = 0; k < ydim; k++) {
            memcpy(line, I + (size_t)k * xdim, sizeof(float) * (size_t)xdim);
            gaussian_line(line, res, xdim, B, size);
            memcpy(I + (size_t)k * xdim, res, sizeof(float) * (size_t)xdim);
        }
        #pragma omp for
        for (int k = 0; k < xdim; k++) {
            for (int i = 0; i < ydim; i++) line[i] = I[(size_t)i * xdim + k];
            gaussian_line(line, res, ydim, B, size);
            for (int i = 0; i < ydim; i++) I[(size_t)i * xdim + k] = res[i];
        }
        free(line);
        free(res);
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* bicubic interpolation (bicubic_interpolation.cpp:17-163)                  */
/* ------------------------------------------------------------------------- */

/* bicubic_interpolation.cpp:17-30 */
static inline int neumann_clamp(int x, int n, int *out)
{
    if (x < 0) { *out = 1; return 0; }
    if (x >= n) { *out = 1; return n - 1; }
    return x;
}

/* Keys cubic convolution (a = -1/2) on one 4-sample cell, written with the
 * reference's promotions: the 0.5/2.0/5.0/4.0/3.0 literals are double, the
 * two bare differences v2-v0 and v1-v2 are float, the result is rounded to
 * float (bicubic_interpolation.cpp:38-46). */
static inline float cubic_cell(const float v[4], float x)
{
    const double xd = (double)x;
    const float d20 = v[2] - v[0];
    const float d12 = v[1] - v[2];
    const double c3 = ((3.0 * (double)d12) + (double)v[3]) - (double)v[0];
    const double c2 = (((2.0 * (double)v[0]) - (5.0 * (double)v[1])) + (4.0 * (double)v[2])) - (double)v[3];
    const double inner = (double)d20 + xd * (c2 + xd * c3);
    return (float)((double)v[1] + (0.5 * xd) * inner);
}

/* float -> int truncation, defined for every input: NaN gives 0, values
 * beyond the int range saturate (C leaves both undefined; the GPU's
 * conversion instruction behaves this way). */
static inline int f2i_sat(double x)
{
    if (x != x) return 0;
    if (x >= 2147483647.0) return INT_MAX;
    if (x <= -2147483648.0) return INT_MIN;
    return (int)x;
}

/* A tap base index limited to [-4, n+4]: every tap within 2 of it then
 * clamps to the same pixel and sets `out` as without the limit, and the tap
 * offsets cannot overflow. */
static inline int tap_base(int i, int n) { return i < -4 ? -4 : (i > n + 4 ? n + 4 : i); }

/* bicubic_interpolation.cpp:74-133.  Tap columns (int)uu-sx .. (int)uu+2sx,
 * tap rows (int)vv-sx .. (int)vv+2sy: the row above uses the x sign, as in
 * the reference (:92); any clamped tap sets `out`.  Non-finite coordinates
 * take the conversion of f2i_sat. */
float orc_bicubic_at(const float *input, float uu, float vv, int nx, int ny, int border_out)
{
    const int sx = (uu < 0) ? -1 : 1;
    const int sy = (vv < 0) ? -1 : 1;
    int out = 0;
    const int iu = tap_base(f2i_sat(uu), nx), iv = tap_base(f2i_sat(vv), ny);
    const int x = neumann_clamp(iu, nx, &out);
    const int y = neumann_clamp(iv, ny, &out);
    const int mx = neumann_clamp(iu - sx, nx, &out);
    const int my = neumann_clamp(iv - sx, ny, &out);
    const int dx = neumann_clamp(iu + sx, nx, &out);
    const int dy = neumann_clamp(iv + sy, ny, &out);
    const int ddx = neumann_clamp(iu + 2 * sx, nx, &out);
    const int ddy = neumann_clamp(iv + 2 * sy, ny, &out);
    if (out && border_out) return 0.0f;

    const int cols[4] = {mx, x, dx, ddx};
    const int rows[4] = {my, y, dy, ddy};
    const float fy = vv - (float)y;
    const float fx = uu - (float)x;
    float v[4];
    for (int c = 0; c < 4; c++) {
        float col[4];
        for (int r = 0; r < 4; r++) col[r] = input[cols[c] + (size_t)nx * rows[r]];
        v[c] = cubic_cell(col, fy);
    }
    return cubic_cell(v, fx);
}

/* bicubic_interpolation.cpp:141-163 */
void orc_bicubic_warp(const float *input, const float *u, const float *v, float *output,
                      int nx, int ny, int border_out)
{
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const size_t p = (size_t)i * nx + j;
            const float uu = (float)j + u[p];
            const float vv = (float)i + v[p];
            output[p] = orc_bicubic_at(input, uu, vv, nx, ny, border_out);
        }
}

/* ------------------------------------------------------------------------- */
/* zoom (zoom.cpp:20-85)                                                     */
/* ------------------------------------------------------------------------- */

/* Blur sigma of zoom_out, zoom.cpp:38: float 0.6 times a double root,
 * rounded to float because the reference stores it in a float. */
float orc_zoom_sigma(float factor)
{
    return (float)((double)ORC_ZOOM_SIGMA_ZERO * sqrt(1.0 / (double)(factor * factor) - 1.0));
}

/* zoom.cpp:20-53.  `Is` is scratch of nx*ny floats. */
int orc_zoom_out(const float *I, float *Iout, int nx, int ny, float factor, float *Is)
{
    memcpy(Is, I, sizeof(float) * (size_t)nx * ny);
    int nxx, nyy;
    orc_zoom_size(nx, ny, factor, &nxx, &nyy);
    const float sigma = orc_zoom_sigma(factor);
    int st = orc_gaussian(Is, nx, ny, (double)sigma);
    if (st) return st;
    #pragma omp parallel for
    for (int i1 = 0; i1 < nyy; i1++)
        for (int j1 = 0; j1 < nxx; j1++) {
            const float i2 = (float)i1 / factor;
            const float j2 = (float)j1 / factor;
            Iout[(size_t)i1 * nxx + j1] = orc_bicubic_at(Is, j2, i2, nx, ny, 0);
        }
    return ORC_OK;
}

/* zoom.cpp:61-85 */
void orc_zoom_in(const float *I, float *Iout, int nx, int ny, int nxx, int nyy)
{
    const float factorx = ((float)nxx / (float)nx);
    const float factory = ((float)nyy / (float)ny);
    #pragma omp parallel for
    for (int i1 = 0; i1 < nyy; i1++)
        for (int j1 = 0; j1 < nxx; j1++) {
            const float i2 = (float)i1 / factory;
            const float j2 = (float)j1 / factorx;
            Iout[(size_t)i1 * nxx + j1] = orc_bicubic_at(I, j2, i2, nx, ny, 0);
        }
}

/* ------------------------------------------------------------------------- */
/* finite differences (mask.cpp:31-196)                                      */
/* ------------------------------------------------------------------------- */

/* mask.cpp:139-196: centred differences times 0.5 (double literal), with
 * one-sided differences (still times 0.5) on the border rows/columns. */
void orc_centered_gradient(const float *in, float *dx, float *dy, int nx, int ny)
{
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const size_t k = (size_t)i * nx + j;
            const int jl = j > 0 ? j - 1 : 0, jr = j < nx - 1 ? j + 1 : nx - 1;
            const int iu = i > 0 ? i - 1 : 0, id = i < ny - 1 ? i + 1 : ny - 1;
            const float ddx = in[(size_t)i * nx + jr] - in[(size_t)i * nx + jl];
            const float ddy = in[(size_t)id * nx + j] - in[(size_t)iu * nx + j];
            dx[k] = (float)(0.5 * (double)ddx);
            dy[k] = (float)(0.5 * (double)ddy);
        }
}

/* mask.cpp:86-131: forward differences, zero on the last column (x) and on
 * the last row (y). */
void orc_forward_gradient(const float *f, float *fx, float *fy, int nx, int ny)
{
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const size_t p = (size_t)i * nx + j;
            fx[p] = (j < nx - 1) ? f[p + 1] - f[p] : 0.0f;
            fy[p] = (i < ny - 1) ? f[p + nx] - f[p] : 0.0f;
        }
}

/* Divergence at one pixel with the Chambolle border rules of mask.cpp:31-77.
 * The first/last column cases associate left to right exactly as written in
 * the reference (:81-82), which is not the same rounding as dx + dy. */
static inline float div_at(const float *v1, const float *v2, int i, int j, int nx, int ny)
{
    const size_t p = (size_t)i * nx + j;
    const int first_row = (i == 0), last_row = (i == ny - 1);
    const int first_col = (j == 0), last_col = (j == nx - 1);
    if ((first_row || last_row) && (first_col || last_col)) {
        /* corners, mask.cpp:73-76 */
        const float a = first_col ? v1[p] : -v1[p - 1];
        if (first_row) return a + v2[p];
        return a - v2[p - nx];
    }
    if (first_row) return (v1[p] - v1[p - 1]) + v2[p];          /* :70 */
    if (last_row) return (v1[p] - v1[p - 1]) - v2[p - nx];      /* :71 */
    if (first_col) return (v1[p] + v2[p]) - v2[p - nx];         /* :81 */
    if (last_col) return (-v1[p - 1] + v2[p]) - v2[p - nx];     /* :82 */
    return (v1[p] - v1[p - 1]) + (v2[p] - v2[p - nx]);          /* :59 */
}

void orc_divergence(const float *v1, const float *v2, float *div, int nx, int ny)
{
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++)
            div[(size_t)i * nx + j] = div_at(v1, v2, i, j, nx, ny);
}

/* ------------------------------------------------------------------------- */
/* normalisation (tvl1.cpp:302-329)                                          */
/* ------------------------------------------------------------------------- */

static size_t isamax(const float *x, size_t n)
{
    size_t best = 0;
    float bv = fabsf(x[0]);
    for (size_t i = 1; i < n; i++)
        if (fabsf(x[i]) > bv) { bv = fabsf(x[i]); best = i; }
    return best;
}

static size_t isamin(const float *x, size_t n)
{
    size_t best = 0;
    float bv = fabsf(x[0]);
    for (size_t i = 1; i < n; i++)
        if (fabsf(x[i]) < bv) { bv = fabsf(x[i]); best = i; }
    return best;
}

/* Joint min/max of both images picked through isamax/isamin (largest and
 * smallest |x|), no-op when den <= 0, else 255*(I-min)/den in double. */
void orc_image_normalization(const float *I0, const float *I1, float *I0n, float *I1n, size_t size)
{
    const float a0 = I0[isamax(I0, size)], a1 = I1[isamax(I1, size)];
    const float b0 = I0[isamin(I0, size)], b1 = I1[isamin(I1, size)];
    const float mx = (a0 < a1) ? a1 : a0;   /* std::max */
    const float mn = (b1 < b0) ? b1 : b0;   /* std::min */
    const float den = mx - mn;
    if (den <= 0) {
        if (I0n != I0) memcpy(I0n, I0, sizeof(float) * size);
        if (I1n != I1) memcpy(I1n, I1, sizeof(float) * size);
        return;
    }
    #pragma omp parallel for
    for (size_t i = 0; i < size; i++) {
        I0n[i] = (float)(255.0 * (double)(I0[i] - mn) / (double)den);
        I1n[i] = (float)(255.0 * (double)(I1[i] - mn) / (double)den);
    }
}

/* ------------------------------------------------------------------------- */
/* TV-L1 (tvl1.cpp:182-294)                                                  */
/* ------------------------------------------------------------------------- */

/* Thresholding step TH for one pixel (tvl1.cpp:221-245). */
static inline void th_step(float rho_c, float wx, float wy, float u1, float u2, float l_t,
                           float *v1, float *v2)
{
    const float grad = (wx * wx) + (wy * wy);
    const float rho = rho_c + ((wx * u1) + (wy * u2));
    const float fi = -rho / grad;
    const int c1 = rho >= -l_t * grad;
    const int c2 = rho > l_t * grad;
    const int c3 = grad < ORC_GRAD_IS_ZERO;
    float d1 = l_t * wx, d2 = l_t * wy;
    if (c1) {
        d1 = fi * wx;
        d2 = fi * wy;
        if (c2) {
            d1 = -l_t * wx;
            d2 = -l_t * wy;
        } else if (c3) {
            d1 = d2 = 0.0f;
        }
    }
    *v1 = u1 + d1;
    *v2 = u2 + d2;
}

void orc_tvl1_th(const float *rho_c, const float *wx, const float *wy, const float *u1,
                 const float *u2, float l_t, float *v1, float *v2, size_t n)
{
    #pragma omp parallel for
    for (size_t i = 0; i < n; i++) th_step(rho_c[i], wx[i], wy[i], u1[i], u2[i], l_t, v1 + i, v2 + i);
}

/* Warp set-up of one warping (tvl1.cpp:198-212): warped I1 and its warped
 * derivatives; rho_c = I1w - I1wx*u1 - I1wy*u2 - I0 with u at warp start. */
void orc_tvl1_warp_setup(const float *I0, const float *I1, const float *I1x, const float *I1y,
                         const float *u1, const float *u2, float *I1w, float *I1wx,
                         float *I1wy, float *rho_c, int nx, int ny)
{
    orc_bicubic_warp(I1, u1, u2, I1w, nx, ny, 1);
    orc_bicubic_warp(I1x, u1, u2, I1wx, nx, ny, 1);
    orc_bicubic_warp(I1y, u1, u2, I1wy, nx, ny, 1);
    const size_t n = (size_t)nx * ny;
    #pragma omp parallel for
    for (size_t i = 0; i < n; i++)
        rho_c[i] = ((I1w[i] - I1wx[i] * u1[i]) - I1wy[i] * u2[i]) - I0[i];
}

/* How the stopping error of a warping is accumulated.
 *   ORC_ERROR_DOUBLE_ROWS (default): the float per-pixel terms summed in
 *     double, row by row, rows added in order -- independent of the thread
 *     count, and the order the GPU's fixed-order sum reproduces bit for bit.
 *   ORC_ERROR_FLOAT_SEQUENTIAL: the reference's `float error` accumulated
 *     pixel by pixel in index order and divided in float (tvl1.cpp:252-265
 *     as it runs with one OpenMP thread; with more threads the reference's
 *     order depends on the thread count and completion order).
 * The two can only disagree on a stopping decision when error/N lies within
 * the float sum's rounding of epsilon^2. */
static int g_error_mode = ORC_ERROR_DOUBLE_ROWS;

void orc_set_error_mode(int mode)
{
    g_error_mode = mode == ORC_ERROR_FLOAT_SEQUENTIAL ? mode : ORC_ERROR_DOUBLE_ROWS;
}

/* One primal-dual iteration from a given state (tvl1.cpp:218-291).  Updates
 * u1,u2,p11..p22 in place and returns sum over pixels of
 * (du1^2 + du2^2) accumulated in double (each pixel term in float). */
double orc_tvl1_iteration(const float *I1wx, const float *I1wy, const float *rho_c,
                          float *u1, float *u2, float *p11, float *p12, float *p21,
                          float *p22, int nx, int ny, float tau, float lambda, float theta,
                          float *scratch /* 6*nx*ny */)
{
    const size_t size = (size_t)nx * ny;
    const float l_t = lambda * theta;
    const float taut = tau / theta;
    float *v1 = scratch, *v2 = scratch + size;
    float *div1 = scratch + 2 * size, *div2 = scratch + 3 * size;
    double *rows = (double *)malloc(sizeof(double) * (size_t)ny);
    /* u before the update, kept only for the float replay of the error */
    float *u1k_save = NULL, *u2k_save = NULL;
    if (g_error_mode == ORC_ERROR_FLOAT_SEQUENTIAL) {
        u1k_save = (float *)malloc(sizeof(float) * 2 * size);
        u2k_save = u1k_save + size;
        memcpy(u1k_save, u1, sizeof(float) * size);
        memcpy(u2k_save, u2, sizeof(float) * size);
    }

    orc_tvl1_th(rho_c, I1wx, I1wy, u1, u2, l_t, v1, v2, size);
    orc_divergence(p11, p12, div1, nx, ny);
    orc_divergence(p21, p22, div2, nx, ny);

    /* u update and the stopping error.  Each pixel term is the reference's
     * float expression (tvl1.cpp:262-263); the sum runs in double, row by
     * row, rows added in order, so it does not depend on the thread count. */
    #pragma omp parallel for
    for (int i = 0; i < ny; i++) {
        double acc = 0.0;
        for (int j = 0; j < nx; j++) {
            const size_t p = (size_t)i * nx + j;
            const float u1k = u1[p], u2k = u2[p];
            u1[p] = v1[p] + theta * div1[p];
            u2[p] = v2[p] + theta * div2[p];
            const float e = (u1[p] - u1k) * (u1[p] - u1k) + (u2[p] - u2k) * (u2[p] - u2k);
            acc += (double)e;
        }
        rows[i] = acc;
    }
    double err = 0.0;
    for (int i = 0; i < ny; i++) err += rows[i];
    free(rows);
    if (g_error_mode == ORC_ERROR_FLOAT_SEQUENTIAL) {
        /* the reference's `float error` summed pixel by pixel in index order
         * (tvl1.cpp:252-264 with one OpenMP thread); the terms are the same
         * float expressions, recomputed from the new and the saved u */
        float ef = 0.0f;
        for (size_t p = 0; p < size; p++) {
            const float d1 = u1[p] - u1k_save[p], d2 = u2[p] - u2k_save[p];
            ef += d1 * d1 + d2 * d2;
        }
        err = (double)ef;
    }
    free(u1k_save);

    /* forward gradients go into the div/v scratch as in tvl1.cpp:268-269 */
    orc_forward_gradient(u1, div1, v1, nx, ny);
    orc_forward_gradient(u2, div2, v2, nx, ny);
    #pragma omp parallel for
    for (size_t i = 0; i < size; i++) {
        const float g1 = 1.0f + taut * hypotf(div1[i], v1[i]);
        const float g2 = 1.0f + taut * hypotf(div2[i], v2[i]);
        /* cblas_saxpy: y := a*x + y */
        const float q11 = p11[i] + taut * div1[i];
        const float q12 = p12[i] + taut * v1[i];
        const float q21 = p21[i] + taut * div2[i];
        const float q22 = p22[i] + taut * v2[i];
        p11[i] = q11 / g1;
        p12[i] = q12 / g1;
        p21[i] = q21 / g2;
        p22[i] = q22 / g2;
    }
    return err;
}

/* One scale of dual TV-L1 (tvl1.cpp:182-294).  `iters` receives the
 * iteration count of each warping. */
int orc_dual_tvl1(const float *I0, const float *I1, float *u1, float *u2, int nx, int ny,
                  const orc_tvl1_params *prm, int *iters)
{
    const size_t size = (size_t)nx * ny;
    float *buf = (float *)calloc(16 * size, sizeof(float));
    if (!buf) return ORC_ERR_ALLOC;
    float *I1x = buf, *I1y = buf + size, *I1w = buf + 2 * size, *I1wx = buf + 3 * size;
    float *I1wy = buf + 4 * size, *rho_c = buf + 5 * size;
    float *p11 = buf + 6 * size, *p12 = buf + 7 * size, *p21 = buf + 8 * size, *p22 = buf + 9 * size;
    float *scratch = buf + 10 * size;

    orc_centered_gradient(I1, I1x, I1y, nx, ny);
    /* p = 0 (calloc), tvl1.cpp:190-193 */
    const float eps2 = prm->epsilon * prm->epsilon;
    for (int w = 0; w < prm->warps; w++) {
        orc_tvl1_warp_setup(I0, I1, I1x, I1y, u1, u2, I1w, I1wx, I1wy, rho_c, nx, ny);
        int n = 0;
        double error = INFINITY;
        float error_f = INFINITY;
        const int fmode = g_error_mode == ORC_ERROR_FLOAT_SEQUENTIAL;
        while ((fmode ? error_f > eps2 : error > (double)eps2) && n < prm->max_iterations) {
            n++;
            const double s = orc_tvl1_iteration(I1wx, I1wy, rho_c, u1, u2, p11, p12, p21, p22,
                                                nx, ny, prm->tau, prm->lambda, prm->theta, scratch);
            error = s / (double)size;
            /* tvl1.cpp:265: float error divided by the size_t size, in float */
            error_f = (float)s;
            error_f /= (float)size;
        }
        if (iters) iters[w] = n;
    }
    free(buf);
    return ORC_OK;
}

/* OpenMP threads of the oracle's parallel loops (the reference runs with all
 * cores, OMP_NUM_THREADS); the results do not depend on the count. */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_tvl1_default_params(orc_tvl1_params *p)
{
    /* tvl1.hpp:12-14, tvl1.hpp:6 */
    p->tau = 0.25f;
    p->lambda = 0.15f;
    p->theta = 0.3f;
    p->nscales = 100;
    p->zfactor = 0.5f;
    p->warps = 5;
    p->epsilon = 0.01f;
    p->max_iterations = 300;
}

/* Multiscale driver, tvl1.cpp:129-175, from float frames already converted
 * (tvl1.cpp:118-123 converts u8 with a plain cast).  `u` receives planar
 * [u1 | u2] (2*W*H floats).  `iters` receives nscales*warps counts, finest
 * scale first.  Returns the number of scales used via *nscales_out. */
/* The multiscale driver (tvl1.cpp:139-175).  The warm-start extension,
 * which has no reference counterpart: the solve starts at scale `first`
 * (-1: the coarsest, as the reference) from `uinit` (planar [u1|u2] at that
 * scale) instead of the zero flow of tvl1.cpp:157-158 when not NULL; the
 * coarser scales are not solved. */
static int tvl1_run(const float *I0in, const float *I1in, int width, int height,
                    const orc_tvl1_params *prm, const float *uinit, int first, float *u,
                    int *iters, int *nscales_out, const int *warps_s, const int *max_s)
{
    const int ns = orc_tvl1_nscales(width, height, prm->nscales, prm->zfactor);
    if (nscales_out) *nscales_out = ns;
    const size_t size = (size_t)width * height;
    int nx[64], ny[64];
    float *I0s[64], *I1s[64], *u1s[64], *u2s[64];
    if (ns > 64) return ORC_ERR_ARG;
    int st = ORC_OK;
    float *scratch = (float *)malloc(sizeof(float) * size);
    for (int s = 0; s < ns; s++) {
        I0s[s] = (float *)malloc(sizeof(float) * size);
        I1s[s] = (float *)malloc(sizeof(float) * size);
        u1s[s] = s ? (float *)malloc(sizeof(float) * size) : u;
        u2s[s] = s ? (float *)malloc(sizeof(float) * size) : u + size;
    }
    nx[0] = width;
    ny[0] = height;
    memcpy(I0s[0], I0in, sizeof(float) * size);
    memcpy(I1s[0], I1in, sizeof(float) * size);
    orc_image_normalization(I0s[0], I1s[0], I0s[0], I1s[0], size);
    st = orc_gaussian(I0s[0], nx[0], ny[0], (double)ORC_PRESMOOTHING_SIGMA);
    if (!st) st = orc_gaussian(I1s[0], nx[0], ny[0], (double)ORC_PRESMOOTHING_SIGMA);
    for (int s = 1; s < ns && !st; s++) {
        orc_zoom_size(nx[s - 1], ny[s - 1], prm->zfactor, &nx[s], &ny[s]);
        st = orc_zoom_out(I0s[s - 1], I0s[s], nx[s - 1], ny[s - 1], prm->zfactor, scratch);
        if (!st) st = orc_zoom_out(I1s[s - 1], I1s[s], nx[s - 1], ny[s - 1], prm->zfactor, scratch);
    }
    if (first < 0 || first >= ns) first = ns - 1;
    if (iters)
        for (int i = 0; i < ns * prm->warps; i++) iters[i] = 0;
    if (!st) {
        const size_t nc = (size_t)nx[first] * ny[first];
        for (size_t i = 0; i < nc; i++) {
            u1s[first][i] = uinit ? uinit[i] : 0.0f;
            u2s[first][i] = uinit ? uinit[nc + i] : 0.0f;
        }
        const float invZ = 1 / prm->zfactor;
        /* per-scale limits (warm-start schedules): a copy of the parameters
         * with that scale's warpings and iteration cap */
        orc_tvl1_params ps[64];
        for (int s = 0; s < ns; s++) {
            ps[s] = *prm;
            if (warps_s && warps_s[s] > 0) ps[s].warps = warps_s[s];
            if (max_s && max_s[s] > 0) ps[s].max_iterations = max_s[s];
        }
        for (int s = first; s > 0 && !st; s--) {
            st = orc_dual_tvl1(I0s[s], I1s[s], u1s[s], u2s[s], nx[s], ny[s], &ps[s],
                               iters ? iters + (size_t)s * prm->warps : NULL);
            orc_zoom_in(u1s[s], u1s[s - 1], nx[s], ny[s], nx[s - 1], ny[s - 1]);
            orc_zoom_in(u2s[s], u2s[s - 1], nx[s], ny[s], nx[s - 1], ny[s - 1]);
            const size_t nf = (size_t)nx[s - 1] * ny[s - 1];
            for (size_t i = 0; i < nf; i++) {
                u1s[s - 1][i] = invZ * u1s[s - 1][i];   /* cblas_sscal */
                u2s[s - 1][i] = invZ * u2s[s - 1][i];
            }
        }
        if (!st) st = orc_dual_tvl1(I0s[0], I1s[0], u1s[0], u2s[0], nx[0], ny[0], &ps[0], iters);
    }
    for (int s = 0; s < ns; s++) {
        free(I0s[s]);
        free(I1s[s]);
        if (s) { free(u1s[s]); free(u2s[s]); }
    }
    free(scratch);
    return st;
}

int orc_tvl1_run_f32(const float *I0in, const float *I1in, int width, int height,
                     const orc_tvl1_params *prm, float *u, int *iters, int *nscales_out)
{
    return tvl1_run(I0in, I1in, width, height, prm, NULL, -1, u, iters, nscales_out, NULL, NULL);
}

/* tvl1.cpp:118-123 then the multiscale driver, from `uinit` (or zero). */
int orc_tvl1_run_init_u8(const uint8_t *I0, const uint8_t *I1, int width, int height,
                         const orc_tvl1_params *prm, const float *uinit, int first, float *u,
                         int *iters, int *nscales_out)
{
    const size_t size = (size_t)width * height;
    float *f0 = (float *)malloc(sizeof(float) * size);
    float *f1 = (float *)malloc(sizeof(float) * size);
    if (!f0 || !f1) { free(f0); free(f1); return ORC_ERR_ALLOC; }
    for (size_t i = 0; i < size; i++) {
        f0[i] = (float)I0[i];
        f1[i] = (float)I1[i];
    }
    const int st = tvl1_run(f0, f1, width, height, prm, uinit, first, u, iters, nscales_out, NULL,
                            NULL);
    free(f0);
    free(f1);
    return st;
}

/* The same with per-scale limits (no reference counterpart): warps_s[s]
 * warpings and max_s[s] iterations per warping at scale s (entries <= 0
 * keep prm's).  Iteration counts keep prm->warps slots per scale. */
int orc_tvl1_run_schedule_u8(const uint8_t *I0, const uint8_t *I1, int width, int height,
                             const orc_tvl1_params *prm, const float *uinit, int first,
                             const int *warps_s, const int *max_s, float *u, int *iters,
                             int *nscales_out)
{
    const size_t size = (size_t)width * height;
    float *f0 = (float *)malloc(sizeof(float) * size);
    float *f1 = (float *)malloc(sizeof(float) * size);
    if (!f0 || !f1) { free(f0); free(f1); return ORC_ERR_ALLOC; }
    for (size_t i = 0; i < size; i++) {
        f0[i] = (float)I0[i];
        f1[i] = (float)I1[i];
    }
    const int st = tvl1_run(f0, f1, width, height, prm, uinit, first, u, iters, nscales_out,
                            warps_s, max_s);
    free(f0);
    free(f1);
    return st;
}

int orc_tvl1_run_u8(const uint8_t *I0, const uint8_t *I1, int width, int height,
                    const orc_tvl1_params *prm, float *u, int *iters, int *nscales_out)
{
    return orc_tvl1_run_init_u8(I0, I1, width, height, prm, NULL, -1, u, iters, nscales_out);
}

/* TV-L1 energy of a flow (no reference counterpart; the warm start's
 * fallback test): lambda * sum |I1(x + u) - I0| + sum (|grad u1| + |grad u2|),
 * I1 sampled by the bicubic of bicubic_interpolation.cpp:108-133 with the
 * clamped border, forward differences as mask.cpp:86-131, |.| = hypotf. */
double orc_tvl1_energy(const float *I0, const float *I1, const float *u1, const float *u2, int nx,
                       int ny, float lambda, double *data_out, double *tv_out)
{
    double data = 0.0, tv = 0.0;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const size_t i = (size_t)y * nx + x;
            const float w = orc_bicubic_at(I1, (float)x + u1[i], (float)y + u2[i], nx, ny, 0);
            data += (double)fabsf(w - I0[i]);
            const float ax = x < nx - 1 ? u1[i + 1] - u1[i] : 0.0f;
            const float ay = y < ny - 1 ? u1[i + nx] - u1[i] : 0.0f;
            const float bx = x < nx - 1 ? u2[i + 1] - u2[i] : 0.0f;
            const float by = y < ny - 1 ? u2[i + nx] - u2[i] : 0.0f;
            tv += (double)hypotf(ax, ay) + (double)hypotf(bx, by);
        }
    if (data_out) *data_out = data;
    if (tv_out) *tv_out = tv;
    return (double)lambda * data + tv;
}

/* ------------------------------------------------------------------------- */
/* Horn-Schunck (intent of src/hs/cuda/flow.cu:112-208 and kernels/)        */
/* ------------------------------------------------------------------------- */

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Bilinear sample at (x, y) in pixel-centre coordinates with clamp-to-edge
 * addressing: the texture fetch the HS sample intends
 * (warpingKernel.cuh:28-43, upscaleKernel.cuh:26-40). */
float orc_bilinear_at(const float *src, int w, int h, int stride, float x, float y)
{
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    /* limited to [-1, w] first so that +1 cannot overflow (clamps unchanged) */
    const int x0 = clampi(f2i_sat(fx0), -1, w), y0 = clampi(f2i_sat(fy0), -1, h);
    const int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    const int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    const float a = src[(size_t)ya * stride + xa], b = src[(size_t)ya * stride + xb];
    const float c = src[(size_t)yb * stride + xa], d = src[(size_t)yb * stride + xb];
    const float top = (1.0f - ax) * a + ax * b;
    const float bot = (1.0f - ax) * c + ax * d;
    return (1.0f - ay) * top + ay * bot;
}

/* 2x2 box downscale, defined for every output pixel (the reference skips
 * the last row/column, downscaleKernel.cuh:32-33). */
void orc_hs_downscale(const float *src, int w, int h, int s, float *out, int nw, int nh, int ns)
{
    (void)w; (void)h;
    for (int iy = 0; iy < nh; iy++)
        for (int ix = 0; ix < nw; ix++) {
            const size_t sx = (size_t)ix * 2, sy = (size_t)iy * 2;
            const float a = src[sx + sy * s], b = src[sx + (sy + 1) * s];
            const float c = src[(sx + 1) + sy * s], d = src[(sx + 1) + (sy + 1) * s];
            out[(size_t)ix + (size_t)iy * ns] = 0.25f * (((a + b) + c) + d);
        }
}

/* Flow upscale x2: bilinear at ((ix-0.5)/2, (iy-0.5)/2) of the coarse field
 * times `scale` (the texture sample upscaleKernel.cuh:33-39 intends). */
void orc_hs_upscale(const float *src, int w, int h, int s, float *out, int nw, int nh, int ns,
                    float scale)
{
    for (int iy = 0; iy < nh; iy++)
        for (int ix = 0; ix < nw; ix++) {
            const float x = ((float)ix - 0.5f) * 0.5f;
            const float y = ((float)iy - 0.5f) * 0.5f;
            out[(size_t)ix + (size_t)iy * ns] = orc_bilinear_at(src, w, h, s, x, y) * scale;
        }
}

void orc_hs_warp(const float *src, int w, int h, int s, const float *u, const float *v, float *out)
{
    #pragma omp parallel for
    for (int iy = 0; iy < h; iy++)
        for (int ix = 0; ix < w; ix++) {
            const size_t pos = (size_t)ix + (size_t)iy * s;
            out[pos] = orc_bilinear_at(src, w, h, s, (float)ix + u[pos], (float)iy + v[pos]);
        }
}

/* 5-tap derivative [1 -8 0 8 -1]/12 with clamp-to-edge indices, evaluated
 * in the order of derivativesKernel.cuh:41-72. */
static inline float d5(const float *p, int c, int n, ptrdiff_t step)
{
    float t = p[clampi(c - 2, 0, n - 1) * step];
    t -= p[clampi(c - 1, 0, n - 1) * step] * 8.0f;
    t += p[clampi(c + 1, 0, n - 1) * step] * 8.0f;
    t -= p[clampi(c + 2, 0, n - 1) * step];
    t /= 12.0f;
    return t;
}

void orc_hs_derivatives(const float *I0, const float *I1w, int w, int h, int s, float *Ix,
                        float *Iy, float *Iz)
{
    #pragma omp parallel for
    for (int iy = 0; iy < h; iy++)
        for (int ix = 0; ix < w; ix++) {
            const size_t pos = (size_t)ix + (size_t)iy * s;
            const float *r0 = I0 + (size_t)iy * s, *r1 = I1w + (size_t)iy * s;
            float t0 = d5(r0, ix, w, 1);
            float t1 = d5(r1, ix, w, 1);
            Ix[pos] = (t0 + t1) * 0.5f;
            Iz[pos] = I1w[pos] - I0[pos];
            t0 = d5(I0 + ix, iy, h, s);
            t1 = d5(I1w + ix, iy, h, s);
            Iy[pos] = (t0 + t1) * 0.5f;
        }
}

/* One Jacobi sweep, solverKernel.cuh:145-152: 4-neighbour mean with
 * clamp-to-edge, alpha added to |grad I|^2 as given. */
void orc_hs_jacobi(const float *du0, const float *dv0, const float *Ix, const float *Iy,
                   const float *Iz, int w, int h, int s, float alpha, float *du1, float *dv1)
{
    #pragma omp parallel for
    for (int iy = 0; iy < h; iy++)
        for (int ix = 0; ix < w; ix++) {
            const size_t pos = (size_t)ix + (size_t)iy * s;
            const size_t l = (size_t)clampi(ix - 1, 0, w - 1) + (size_t)iy * s;
            const size_t r = (size_t)clampi(ix + 1, 0, w - 1) + (size_t)iy * s;
            const size_t up = (size_t)ix + (size_t)clampi(iy + 1, 0, h - 1) * s;
            const size_t dn = (size_t)ix + (size_t)clampi(iy - 1, 0, h - 1) * s;
            const float sumU = (((du0[l] + du0[r]) + du0[up]) + du0[dn]) * 0.25f;
            const float sumV = (((dv0[l] + dv0[r]) + dv0[up]) + dv0[dn]) * 0.25f;
            const float frac = ((Ix[pos] * sumU + Iy[pos] * sumV) + Iz[pos]) /
                               ((Ix[pos] * Ix[pos] + Iy[pos] * Iy[pos]) + alpha);
            du1[pos] = sumU - Ix[pos] * frac;
            dv1[pos] = sumV - Iy[pos] * frac;
        }
}

/* ComputeFlow (flow.cu:112-208): pyramid by /2, coarse to fine, nWarpIters
 * x (warp, derivatives, nSolverIters Jacobi sweeps, u += du), upscale by the
 * width/height ratio between levels.  Images and flow use `stride`. */
int orc_hs_compute(const float *I0, const float *I1, int width, int height, int stride,
                   float alpha, int nLevels, int nWarpIters, int nSolverIters, float *u, float *v)
{
    if (nLevels < 1 || width < 1 || height < 1 || stride < width) return ORC_ERR_ARG;
    int pW[32], pH[32], pS[32];
    float *pI0[32], *pI1[32];
    if (nLevels > 32) return ORC_ERR_ARG;
    const size_t fsize = (size_t)stride * height;
    int L = nLevels - 1;
    pW[L] = width; pH[L] = height; pS[L] = stride;
    pI0[L] = (float *)malloc(sizeof(float) * fsize);
    pI1[L] = (float *)malloc(sizeof(float) * fsize);
    memcpy(pI0[L], I0, sizeof(float) * fsize);
    memcpy(pI1[L], I1, sizeof(float) * fsize);
    for (; L > 0; --L) {
        const int nw = pW[L] / 2, nh = pH[L] / 2;
        const int ns = ((nw + 31) / 32) * 32;
        if (nw < 1 || nh < 1) {   /* too many levels for the frame */
            for (int k = L; k < nLevels; k++) {
                free(pI0[k]);
                free(pI1[k]);
            }
            return ORC_ERR_ARG;
        }
        pI0[L - 1] = (float *)calloc((size_t)ns * nh, sizeof(float));
        pI1[L - 1] = (float *)calloc((size_t)ns * nh, sizeof(float));
        orc_hs_downscale(pI0[L], pW[L], pH[L], pS[L], pI0[L - 1], nw, nh, ns);
        orc_hs_downscale(pI1[L], pW[L], pH[L], pS[L], pI1[L - 1], nw, nh, ns);
        pW[L - 1] = nw; pH[L - 1] = nh; pS[L - 1] = ns;
    }
    float *du0 = (float *)calloc(fsize, sizeof(float)), *dv0 = (float *)calloc(fsize, sizeof(float));
    float *du1 = (float *)calloc(fsize, sizeof(float)), *dv1 = (float *)calloc(fsize, sizeof(float));
    float *Ix = (float *)calloc(fsize, sizeof(float)), *Iy = (float *)calloc(fsize, sizeof(float));
    float *Iz = (float *)calloc(fsize, sizeof(float)), *tmp = (float *)calloc(fsize, sizeof(float));
    float *cu = (float *)calloc(fsize, sizeof(float)), *cv = (float *)calloc(fsize, sizeof(float));
    float *nu = (float *)calloc(fsize, sizeof(float)), *nv = (float *)calloc(fsize, sizeof(float));
    for (L = 0; L < nLevels; ++L) {
        const int w = pW[L], h = pH[L], s = pS[L];
        for (int wi = 0; wi < nWarpIters; ++wi) {
            memset(du0, 0, sizeof(float) * fsize);
            memset(dv0, 0, sizeof(float) * fsize);
            memset(du1, 0, sizeof(float) * fsize);
            memset(dv1, 0, sizeof(float) * fsize);
            orc_hs_warp(pI1[L], w, h, s, cu, cv, tmp);
            orc_hs_derivatives(pI0[L], tmp, w, h, s, Ix, Iy, Iz);
            for (int it = 0; it < nSolverIters; ++it) {
                orc_hs_jacobi(du0, dv0, Ix, Iy, Iz, w, h, s, alpha, du1, dv1);
                float *t = du0; du0 = du1; du1 = t;
                t = dv0; dv0 = dv1; dv1 = t;
            }
            for (int iy = 0; iy < h; iy++)
                for (int ix = 0; ix < w; ix++) {
                    const size_t pos = (size_t)ix + (size_t)iy * s;
                    cu[pos] = cu[pos] + du0[pos];
                    cv[pos] = cv[pos] + dv0[pos];
                }
        }
        if (L != nLevels - 1) {
            const float scaleX = (float)pW[L + 1] / (float)pW[L];
            const float scaleY = (float)pH[L + 1] / (float)pH[L];
            orc_hs_upscale(cu, w, h, s, nu, pW[L + 1], pH[L + 1], pS[L + 1], scaleX);
            orc_hs_upscale(cv, w, h, s, nv, pW[L + 1], pH[L + 1], pS[L + 1], scaleY);
            float *t = cu; cu = nu; nu = t;
            t = cv; cv = nv; nv = t;
        }
    }
    for (int iy = 0; iy < height; iy++)
        for (int ix = 0; ix < width; ix++) {
            const size_t pos = (size_t)ix + (size_t)iy * stride;
            u[pos] = cu[pos];
            v[pos] = cv[pos];
        }
    for (L = 0; L < nLevels; ++L) { free(pI0[L]); free(pI1[L]); }
    free(du0); free(dv0); free(du1); free(dv1); free(Ix); free(Iy); free(Iz); free(tmp);
    free(cu); free(cv); free(nu); free(nv);
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* Lucas-Kanade                                                              */
/* ------------------------------------------------------------------------- */

/* Central-difference filters of lucaskanade.cu:79-161 (spatial) and the
 * temporal ones; returns 0 for an unsupported size. */
int orc_lk_filter(int size, int temporal, float *f)
{
    switch (size) {
    case 1: f[0] = 1.0f; return 1;
    case 2:
        if (!temporal) return 0;
        f[0] = -1.0f; f[1] = 1.0f; return 1;
    case 3: f[0] = -1.0f / 2.0f; f[1] = 0.0f; f[2] = 1.0f / 2.0f; return 1;
    case 5:
        f[0] = 1.0f / 12.0f; f[1] = -2.0f / 3.0f; f[2] = 0.0f; f[3] = 2.0f / 3.0f; f[4] = -1.0f / 12.0f;
        return 1;
    case 7:
        f[0] = -1.0f / 60.0f; f[1] = 3.0f / 20.0f; f[2] = -3.0f / 4.0f; f[3] = 0.0f;
        f[4] = 3.0f / 4.0f; f[5] = -3.0f / 20.0f; f[6] = 1.0f / 60.0f;
        return 1;
    case 9:
        f[0] = 1.0f / 280.0f; f[1] = -4.0f / 105.0f; f[2] = 1.0f / 5.0f; f[3] = -4.0f / 5.0f;
        f[4] = 0.0f; f[5] = 4.0f / 5.0f; f[6] = -1.0f / 5.0f; f[7] = 4.0f / 105.0f;
        f[8] = -1.0f / 280.0f;
        return 1;
    default: return 0;
    }
}

/* Single-scale LK over a T-frame circular buffer of u8 frames
 * (lucaskanade.cu:54-71, kernels.cu:14-181):
 *   It = sum_k filt_t[k] * frame[(k+iframe+1)%T]        (kernels.cu:14-30)
 *   Ix, Iy: zero-padded correlation of frame[(iframe+1)%T] with the spatial
 *           filter (kernels.cu:44-121; the y border uses ny, not nx);
 *   window sums over w x w, Cramer solve when det != 0 (kernels.cu:134-181).
 * The ring of width (w-1)/2 that the reference leaves unwritten is 0 here. */
int orc_lk_single(const uint8_t *frames, int T, int iframe, int nx, int ny, int spac, int win,
                  float *Vx, float *Vy)
{
    float ft[16], fs[16];
    if (!orc_lk_filter(T, 1, ft) || !orc_lk_filter(spac, 0, fs)) return ORC_ERR_ARG;
    if (win < 1 || (win & 1) == 0) return ORC_ERR_ARG;
    const size_t n = (size_t)nx * ny;
    float *It = (float *)malloc(sizeof(float) * n), *Ix = (float *)malloc(sizeof(float) * n);
    float *Iy = (float *)malloc(sizeof(float) * n);
    const uint8_t *fr = frames + (size_t)((iframe + 1) % T) * n;
    const int fc = (spac - 1) / 2;
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const size_t p = (size_t)i * nx + j;
            float t = 0.0f;
            for (int k = 0; k < T; k++) t += ft[k] * (float)frames[(size_t)((k + iframe + 1) % T) * n + p];
            It[p] = t;
            float gx = 0.0f, gy = 0.0f;
            for (int k = 0; k < spac; k++) {
                const int jj = j + k - fc, ii = i + k - fc;
                if (jj >= 0 && jj < nx) gx += (float)fr[(size_t)i * nx + jj] * fs[k];
                if (ii >= 0 && ii < ny) gy += (float)fr[(size_t)ii * nx + j] * fs[k];
            }
            Ix[p] = gx;
            Iy[p] = gy;
        }
    const int wc = (win - 1) / 2;
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const size_t p = (size_t)i * nx + j;
            Vx[p] = 0.0f;
            Vy[p] = 0.0f;
            if (i < wc || j < wc || i >= ny - wc || j >= nx - wc) continue;
            float sxx = 0.0f, sxy = 0.0f, syy = 0.0f, sxt = 0.0f, syt = 0.0f;
            for (int ii = -wc; ii <= wc; ii++)
                for (int jj = -wc; jj <= wc; jj++) {
                    const size_t q = (size_t)(i + ii) * nx + (j + jj);
                    sxx += Ix[q] * Ix[q];
                    sxy += Ix[q] * Iy[q];
                    syy += Iy[q] * Iy[q];
                    sxt += Ix[q] * It[q];
                    syt += Iy[q] * It[q];
                }
            const float det = (sxx * syy - sxy * sxy);
            if (det != 0.0f) {
                Vx[p] = 1.0f / det * (syy * (-sxt) + (-sxy) * (-syt));
                Vy[p] = 1.0f / det * ((-sxy) * (-sxt) + sxx * (-syt));
            }
        }
    free(It); free(Ix); free(Iy);
    return ORC_OK;
}

/* Pyramidal iterative LK (the configuration the single-scale reference is
 * extended to): Gaussian pyramid built like the TV-L1 one (blur with
 * `sigma`, bicubic decimation at i/zf), coarse to fine; per level the
 * template gradients (centred, one-sided x0.5 borders as in
 * orc_centered_gradient) and the structure tensor G are summed over the
 * w x w window clipped to the image; each iteration warps I1 bilinearly by
 * the current flow, forms It = I1w - I0, sums b = (Ix It, Iy It) over the
 * window and adds d = -G^{-1} b when det G > det_eps (else d = 0).  The flow
 * is carried to the next finer level with zoom_in and multiplied by 1/zf.
 * Window sums are row sums then column sums, in float, ascending order. */
static void box_sum(const float *in, float *out, float *tmp, int nx, int ny, int wc)
{
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            float s = 0.0f;
            for (int k = -wc; k <= wc; k++) {
                const int jj = j + k;
                if (jj >= 0 && jj < nx) s += in[(size_t)i * nx + jj];
            }
            tmp[(size_t)i * nx + j] = s;
        }
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            float s = 0.0f;
            for (int k = -wc; k <= wc; k++) {
                const int ii = i + k;
                if (ii >= 0 && ii < ny) s += tmp[(size_t)ii * nx + j];
            }
            out[(size_t)i * nx + j] = s;
        }
}

int orc_lk_pyr_compute(const float *I0, const float *I1, int width, int height, int levels,
                       float zf, float sigma, int win, int iters, float det_eps, float *u, float *v)
{
    if (levels < 1 || levels > 16 || win < 1 || (win & 1) == 0 || iters < 0) return ORC_ERR_ARG;
    int nx[16], ny[16];
    float *P0[16], *P1[16];
    const size_t n0 = (size_t)width * height;
    float *scratch = (float *)malloc(sizeof(float) * n0);
    nx[0] = width; ny[0] = height;
    P0[0] = (float *)malloc(sizeof(float) * n0);
    P1[0] = (float *)malloc(sizeof(float) * n0);
    memcpy(P0[0], I0, sizeof(float) * n0);
    memcpy(P1[0], I1, sizeof(float) * n0);
    int st = ORC_OK;
    for (int l = 1; l < levels && !st; l++) {
        orc_zoom_size(nx[l - 1], ny[l - 1], zf, &nx[l], &ny[l]);
        const size_t nl = (size_t)nx[l] * ny[l];
        P0[l] = (float *)malloc(sizeof(float) * nl);
        P1[l] = (float *)malloc(sizeof(float) * nl);
        /* blur a copy, then sample at (j/zf, i/zf) as zoom.cpp:28-53 does.
         * This pyramid has no reference counterpart (the reference LK is
         * single-scale); at zf = 1/2 its sample is defined as the blurred
         * pixel (2j, 2i) itself, which the bicubic at those integer positions
         * also returns except for IEEE effects of its zero-weight terms (a
         * NaN or infinite neighbour, the sign of a zero). */
        for (int k = 0; k < 2 && !st; k++) {
            const float *src = k ? P1[l - 1] : P0[l - 1];
            float *dst = k ? P1[l] : P0[l];
            memcpy(scratch, src, sizeof(float) * (size_t)nx[l - 1] * ny[l - 1]);
            st = orc_gaussian(scratch, nx[l - 1], ny[l - 1], (double)sigma);
            if (st) break;
            for (int i1 = 0; i1 < ny[l]; i1++)
                for (int j1 = 0; j1 < nx[l]; j1++)
                    dst[(size_t)i1 * nx[l] + j1] =
                        zf == 0.5f ? scratch[(size_t)(2 * i1) * nx[l - 1] + 2 * j1]
                                   : orc_bicubic_at(scratch, (float)j1 / zf, (float)i1 / zf,
                                                                  nx[l - 1], ny[l - 1], 0);
        }
        if (st) {   /* levels 0..l hold buffers */
            for (int k = 0; k <= l; k++) {
                free(P0[k]);
                free(P1[k]);
            }
            free(scratch);
            return st;
        }
    }
    const int wc = (win - 1) / 2;
    float *cu = (float *)calloc(n0, sizeof(float)), *cv = (float *)calloc(n0, sizeof(float));
    float *nu = (float *)calloc(n0, sizeof(float)), *nv = (float *)calloc(n0, sizeof(float));
    float *Ix = (float *)malloc(sizeof(float) * n0), *Iy = (float *)malloc(sizeof(float) * n0);
    float *Gxx = (float *)malloc(sizeof(float) * n0), *Gxy = (float *)malloc(sizeof(float) * n0);
    float *Gyy = (float *)malloc(sizeof(float) * n0), *t0 = (float *)malloc(sizeof(float) * n0);
    float *t1 = (float *)malloc(sizeof(float) * n0), *bx = (float *)malloc(sizeof(float) * n0);
    float *by = (float *)malloc(sizeof(float) * n0);
    const float invZ = 1 / zf;
    for (int l = levels - 1; l >= 0; l--) {
        const int w = nx[l], h = ny[l];
        const size_t nl = (size_t)w * h;
        orc_centered_gradient(P0[l], Ix, Iy, w, h);
        for (size_t i = 0; i < nl; i++) t0[i] = Ix[i] * Ix[i];
        box_sum(t0, Gxx, t1, w, h, wc);
        for (size_t i = 0; i < nl; i++) t0[i] = Ix[i] * Iy[i];
        box_sum(t0, Gxy, t1, w, h, wc);
        for (size_t i = 0; i < nl; i++) t0[i] = Iy[i] * Iy[i];
        box_sum(t0, Gyy, t1, w, h, wc);
        for (int it = 0; it < iters; it++) {
            /* It then the two products */
            #pragma omp parallel for
            for (int i = 0; i < h; i++)
                for (int j = 0; j < w; j++) {
                    const size_t p = (size_t)i * w + j;
                    const float wv = orc_bilinear_at(P1[l], w, h, w, (float)j + cu[p], (float)i + cv[p]);
                    const float itv = wv - P0[l][p];
                    nu[p] = Ix[p] * itv;
                    nv[p] = Iy[p] * itv;
                }
            box_sum(nu, bx, t1, w, h, wc);
            box_sum(nv, by, t1, w, h, wc);
            #pragma omp parallel for
            for (size_t p = 0; p < nl; p++) {
                const float det = Gxx[p] * Gyy[p] - Gxy[p] * Gxy[p];
                if (det > det_eps) {
                    const float du = (Gxy[p] * by[p] - Gyy[p] * bx[p]) / det;
                    const float dv = (Gxy[p] * bx[p] - Gxx[p] * by[p]) / det;
                    cu[p] = cu[p] + du;
                    cv[p] = cv[p] + dv;
                }
            }
        }
        if (l > 0) {
            orc_zoom_in(cu, nu, w, h, nx[l - 1], ny[l - 1]);
            orc_zoom_in(cv, nv, w, h, nx[l - 1], ny[l - 1]);
            const size_t nf = (size_t)nx[l - 1] * ny[l - 1];
            for (size_t i = 0; i < nf; i++) {
                cu[i] = invZ * nu[i];
                cv[i] = invZ * nv[i];
            }
        }
    }
    memcpy(u, cu, sizeof(float) * n0);
    memcpy(v, cv, sizeof(float) * n0);
    for (int l = 0; l < levels; l++) { free(P0[l]); free(P1[l]); }
    free(scratch); free(cu); free(cv); free(nu); free(nv); free(Ix); free(Iy);
    free(Gxx); free(Gxy); free(Gyy); free(t0); free(t1); free(bx); free(by);
    return ORC_OK;
}
