/*
 * oracle64.c -- double-precision restatement of the three flow solvers.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  It is the FP64 oracle: the same algorithms as oracle.c
 * (and so as the reference), every quantity held and computed in double, so
 * that the float solvers (the GPU, which equals oracle.c bit for bit, and
 * the reference's float code) can be measured against a solution with
 * negligible rounding.  Only integer geometry is shared with oracle.c
 * (pyramid sizes, scale count); every arithmetic step is written again here
 * from the reference's definitions:
 *
 *   TV-L1  src/tvl1/sequential/tvl1.cpp:129-175 (multiscale driver),
 *          :182-294 (one scale), :302-329 (normalisation),
 *          mask/mask.cpp:31-131 (divergence, forward gradient),
 *          :139-196 (centred gradient), :204-274 (Gaussian),
 *          zoom/zoom.cpp:20-85, bicubic_interpolation.cpp:17-163.
 *   HS     src/hs/cuda/flow.cu:112-208 and its kernels/ headers, with the border
 *          rules oracle.c states (clamp-to-edge; the intended bilinear
 *          samples where the reference's index arithmetic is broken).
 *   LK     the pyramidal iterative LK of oracle.c (orc_lk_pyr_compute).
 *
 * Parameters arrive as the float values a caller gives the float solvers and
 * are widened to double (epsilon^2 is formed in double).  Frames are float
 * inputs and are read exactly.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define O64_GRAD_IS_ZERO 1E-10
#define O64_PRESMOOTHING_SIGMA 0.8
#define O64_ZOOM_SIGMA_ZERO 0.6

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ------------------------------------------------------------ Gaussian -- */

/* (int)(5 sigma) + 1 taps exp(-i^2/2sigma^2)/(sigma sqrt(2pi)), normalised
 * so that the symmetric kernel sums to one (mask.cpp:212-231). */
static int taps64(double sigma, double *B)
{
    const int size = (int)(5.0 * sigma) + 1;
    if (size < 1 || size > 64) return -1;
    double norm = 0.0;
    for (int i = 0; i < size; i++) {
        B[i] = exp(-(double)(i * i) / (2.0 * sigma * sigma)) / (sigma * sqrt(2.0 * M_PI));
        norm += B[i];
    }
    norm = 2.0 * norm - B[0];
    for (int i = 0; i < size; i++) B[i] /= norm;
    return size;
}

/* Separable blur, rows then columns; low side whole-sample mirror
 * x[-k] = x[k], high side half-sample mirror x[n-1+k] = x[n-k]
 * (mask.cpp:240-242, 261-263). */
int o64_gaussian(double *I, int nx, int ny, double sigma)
{
    double B[64];
    const int size = taps64(sigma, B);
    if (size < 0 || size > nx || size > ny) return ORC_ERR_SIGMA;
    const int nmax = nx > ny ? nx : ny;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)nmax * (nx > ny ? ny : nx));
    if (!tmp) return ORC_ERR_ALLOC;
    #pragma omp parallel for
    for (int i = 0; i < ny; i++) {
        const double *row = I + (size_t)i * nx;
        double *out = tmp + (size_t)i * nx;
        for (int x = 0; x < nx; x++) {
            double s = B[0] * row[x];
            for (int j = 1; j < size; j++) {
                int xl = x - j, xr = x + j;
                if (xl < 0) xl = -xl;
                if (xr >= nx) xr = 2 * nx - 1 - xr;
                s += B[j] * (row[xl] + row[xr]);
            }
            out[x] = s;
        }
    }
    #pragma omp parallel for
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++) {
            double s = B[0] * tmp[(size_t)y * nx + x];
            for (int j = 1; j < size; j++) {
                int yl = y - j, yr = y + j;
                if (yl < 0) yl = -yl;
                if (yr >= ny) yr = 2 * ny - 1 - yr;
                s += B[j] * (tmp[(size_t)yl * nx + x] + tmp[(size_t)yr * nx + x]);
            }
            I[(size_t)y * nx + x] = s;
        }
    free(tmp);
    return ORC_OK;
}

/* ------------------------------------------------------------- bicubic -- */

/* Keys cubic convolution, a = -1/2, on samples v[0..3] at offset x in [0,1)
 * from v[1] (bicubic_interpolation.cpp:38-46). */
static inline double keys(const double v[4], double x)
{
    return v[1] + 0.5 * x * (v[2] - v[0] +
                             x * (2.0 * v[0] - 5.0 * v[1] + 4.0 * v[2] - v[3] +
                                  x * (3.0 * (v[1] - v[2]) + v[3] - v[0])));
}

static inline int neumann(int x, int n, int *out)
{
    if (x < 0) { *out = 1; return 0; }
    if (x >= n) { *out = 1; return n - 1; }
    return x;
}

/* float -> int truncation defined for every input (NaN 0, saturating) and
 * the tap base limited to [-4, n+4], as in oracle.c. */
static inline int f2i_sat(double x)
{
    if (x != x) return 0;
    if (x >= 2147483647.0) return INT_MAX;
    if (x <= -2147483648.0) return INT_MIN;
    return (int)x;
}

static inline int tap_base(int i, int n) { return i < -4 ? -4 : (i > n + 4 ? n + 4 : i); }

/* bicubic_interpolation.cpp:74-133: taps at (int)uu - s .. (int)uu + 2s in x
 * and (int)vv - s .. +2s in y (s the sign of the coordinate; the row above
 * takes the x sign), clamped; with border_out, 0 when any tap was clamped. */
double o64_bicubic_at(const double *I, double uu, double vv, int nx, int ny, int border_out)
{
    const int sx = uu < 0 ? -1 : 1, sy = vv < 0 ? -1 : 1;
    int out = 0;
    const int iu = tap_base(f2i_sat(uu), nx), iv = tap_base(f2i_sat(vv), ny);
    const int x = neumann(iu, nx, &out), y = neumann(iv, ny, &out);
    const int mx = neumann(iu - sx, nx, &out), my = neumann(iv - sx, ny, &out);
    const int dx = neumann(iu + sx, nx, &out), dy = neumann(iv + sy, ny, &out);
    const int ddx = neumann(iu + 2 * sx, nx, &out), ddy = neumann(iv + 2 * sy, ny, &out);
    if (out && border_out) return 0.0;
    const int cols[4] = {mx, x, dx, ddx}, rows[4] = {my, y, dy, ddy};
    double v[4];
    for (int c = 0; c < 4; c++) {
        double col[4];
        for (int r = 0; r < 4; r++) col[r] = I[(size_t)rows[r] * nx + cols[c]];
        v[c] = keys(col, vv - y);
    }
    return keys(v, uu - x);
}

/* -------------------------------------------------------------- zoom ---- */

static int zoom_out64(const double *I, double *out, int nx, int ny, double f, int nxx, int nyy)
{
    double *Is = (double *)malloc(sizeof(double) * (size_t)nx * ny);
    if (!Is) return ORC_ERR_ALLOC;
    memcpy(Is, I, sizeof(double) * (size_t)nx * ny);
    const int st = o64_gaussian(Is, nx, ny, O64_ZOOM_SIGMA_ZERO * sqrt(1.0 / (f * f) - 1.0));
    if (!st) {
        #pragma omp parallel for
        for (int i = 0; i < nyy; i++)
            for (int j = 0; j < nxx; j++)
                out[(size_t)i * nxx + j] = o64_bicubic_at(Is, j / f, i / f, nx, ny, 0);
    }
    free(Is);
    return st;
}

/* zoom.cpp:61-85: sample at (j nx/nxx, i ny/nyy) */
static void zoom_in64(const double *I, double *out, int nx, int ny, int nxx, int nyy)
{
    const double fx = (double)nxx / nx, fy = (double)nyy / ny;
    #pragma omp parallel for
    for (int i = 0; i < nyy; i++)
        for (int j = 0; j < nxx; j++)
            out[(size_t)i * nxx + j] = o64_bicubic_at(I, j / fx, i / fy, nx, ny, 0);
}

/* ----------------------------------------------------- differences ------ */

/* mask.cpp:139-196: central differences / 2; one-sided differences / 2 on
 * the first and last rows and columns. */
static void centered_gradient64(const double *in, double *dx, double *dy, int nx, int ny)
{
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const int jl = j > 0 ? j - 1 : 0, jr = j < nx - 1 ? j + 1 : nx - 1;
            const int iu = i > 0 ? i - 1 : 0, id = i < ny - 1 ? i + 1 : ny - 1;
            dx[(size_t)i * nx + j] = 0.5 * (in[(size_t)i * nx + jr] - in[(size_t)i * nx + jl]);
            dy[(size_t)i * nx + j] = 0.5 * (in[(size_t)id * nx + j] - in[(size_t)iu * nx + j]);
        }
}

/* mask.cpp:86-131: forward differences; fx = 0 on the last column, fy = 0 on
 * the last row. */
static void forward_gradient64(const double *f, double *fx, double *fy, int nx, int ny)
{
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const size_t p = (size_t)i * nx + j;
            fx[p] = j < nx - 1 ? f[p + 1] - f[p] : 0.0;
            fy[p] = i < ny - 1 ? f[p + nx] - f[p] : 0.0;
        }
}

/* mask.cpp:31-77: the negative adjoint of forward_gradient64, i.e.
 * div = D_x^- v1 + D_y^- v2 with v1 taken as 0 left of column 0 and at
 * column nx-1, v2 as 0 above row 0 and at row ny-1. */
static void divergence64(const double *v1, const double *v2, double *div, int nx, int ny)
{
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            const size_t p = (size_t)i * nx + j;
            const double a = j < nx - 1 ? v1[p] : 0.0;
            const double b = j > 0 ? v1[p - 1] : 0.0;
            const double c = i < ny - 1 ? v2[p] : 0.0;
            const double d = i > 0 ? v2[p - nx] : 0.0;
            div[p] = (a - b) + (c - d);
        }
}

/* -------------------------------------------------------------- TV-L1 --- */

/* One primal-dual iteration in double (tvl1.cpp:218-291): TH step, u update
 * u = v + theta div p, dual step p = (p + (tau/theta) grad u) / g with
 * g = 1 + (tau/theta)|grad u|.  Returns sum (du1^2 + du2^2). */
double o64_tvl1_iteration(const double *wx, const double *wy, const double *rho_c, double *u1,
                          double *u2, double *p11, double *p12, double *p21, double *p22, int nx,
                          int ny, double tau, double lambda, double theta, double *scratch)
{
    const size_t n = (size_t)nx * ny;
    const double lt = lambda * theta, taut = tau / theta;
    double *v1 = scratch, *v2 = scratch + n, *d1 = scratch + 2 * n, *d2 = scratch + 3 * n;
    #pragma omp parallel for
    for (size_t i = 0; i < n; i++) {
        const double g = wx[i] * wx[i] + wy[i] * wy[i];
        const double rho = rho_c[i] + wx[i] * u1[i] + wy[i] * u2[i];
        double s1, s2;
        if (rho < -lt * g) {
            s1 = lt * wx[i];
            s2 = lt * wy[i];
        } else if (rho > lt * g) {
            s1 = -lt * wx[i];
            s2 = -lt * wy[i];
        } else if (g < O64_GRAD_IS_ZERO) {
            s1 = s2 = 0.0;
        } else {
            s1 = -rho / g * wx[i];
            s2 = -rho / g * wy[i];
        }
        v1[i] = u1[i] + s1;
        v2[i] = u2[i] + s2;
    }
    divergence64(p11, p12, d1, nx, ny);
    divergence64(p21, p22, d2, nx, ny);
    double err = 0.0;
    #pragma omp parallel for reduction(+ : err)
    for (size_t i = 0; i < n; i++) {
        const double a = v1[i] + theta * d1[i], b = v2[i] + theta * d2[i];
        err += (a - u1[i]) * (a - u1[i]) + (b - u2[i]) * (b - u2[i]);
        u1[i] = a;
        u2[i] = b;
    }
    forward_gradient64(u1, d1, v1, nx, ny);
    forward_gradient64(u2, d2, v2, nx, ny);
    #pragma omp parallel for
    for (size_t i = 0; i < n; i++) {
        const double g1 = 1.0 + taut * hypot(d1[i], v1[i]);
        const double g2 = 1.0 + taut * hypot(d2[i], v2[i]);
        p11[i] = (p11[i] + taut * d1[i]) / g1;
        p12[i] = (p12[i] + taut * v1[i]) / g1;
        p21[i] = (p21[i] + taut * d2[i]) / g2;
        p22[i] = (p22[i] + taut * v2[i]) / g2;
    }
    return err;
}

/* One scale (tvl1.cpp:182-294): p = 0, then per warping the bicubic warps of
 * I1 and its centred gradient (border_out: 0 outside), rho_c, and iterations
 * while sum(du^2)/N > epsilon^2 and n < max_iterations. */
int o64_dual_tvl1(const double *I0, const double *I1, double *u1, double *u2, int nx, int ny,
                  const orc_tvl1_params *prm, int *iters)
{
    const size_t n = (size_t)nx * ny;
    double *buf = (double *)calloc(14 * n, sizeof(double));
    if (!buf) return ORC_ERR_ALLOC;
    double *I1x = buf, *I1y = buf + n, *I1w = buf + 2 * n, *wx = buf + 3 * n, *wy = buf + 4 * n;
    double *rho_c = buf + 5 * n, *p11 = buf + 6 * n, *p12 = buf + 7 * n, *p21 = buf + 8 * n;
    double *p22 = buf + 9 * n, *scratch = buf + 10 * n;
    centered_gradient64(I1, I1x, I1y, nx, ny);
    const double eps2 = (double)prm->epsilon * (double)prm->epsilon;
    for (int w = 0; w < prm->warps; w++) {
        #pragma omp parallel for
        for (int i = 0; i < ny; i++)
            for (int j = 0; j < nx; j++) {
                const size_t p = (size_t)i * nx + j;
                const double x = j + u1[p], y = i + u2[p];
                I1w[p] = o64_bicubic_at(I1, x, y, nx, ny, 1);
                wx[p] = o64_bicubic_at(I1x, x, y, nx, ny, 1);
                wy[p] = o64_bicubic_at(I1y, x, y, nx, ny, 1);
                rho_c[p] = I1w[p] - wx[p] * u1[p] - wy[p] * u2[p] - I0[p];
            }
        int it = 0;
        double error = INFINITY;
        while (error > eps2 && it < prm->max_iterations) {
            it++;
            error = o64_tvl1_iteration(wx, wy, rho_c, u1, u2, p11, p12, p21, p22, nx, ny,
                                       prm->tau, prm->lambda, prm->theta, scratch) /
                    (double)n;
        }
        if (iters) iters[w] = it;
    }
    free(buf);
    return ORC_OK;
}

/* Multiscale TV-L1 (tvl1.cpp:129-175) of float frames: joint normalisation
 * to [0, 255] (skipped when max = min), Gaussian pre-smoothing (sigma 0.8),
 * pyramid by zoom_out, zero flow at the coarsest scale, coarse to fine with
 * zoom_in and a factor 1/zfactor between scales.  u receives planar
 * [u1 | u2] at full size; iters nscales*warps counts, finest scale first. */
int o64_tvl1_run_f32(const float *I0in, const float *I1in, int width, int height,
                     const orc_tvl1_params *prm, double *u, int *iters, int *nscales_out)
{
    const int ns = orc_tvl1_nscales(width, height, prm->nscales, prm->zfactor);
    if (nscales_out) *nscales_out = ns;
    if (ns > 64 || width < 2 || height < 2) return ORC_ERR_ARG;
    const size_t N = (size_t)width * height;
    int nx[64], ny[64];
    double *I0s[64] = {0}, *I1s[64] = {0}, *u1s[64] = {0}, *u2s[64] = {0};
    int st = ORC_OK;
    nx[0] = width;
    ny[0] = height;
    for (int s = 1; s < ns; s++) orc_zoom_size(nx[s - 1], ny[s - 1], prm->zfactor, &nx[s], &ny[s]);
    for (int s = 0; s < ns && !st; s++) {
        const size_t n = (size_t)nx[s] * ny[s];
        I0s[s] = (double *)malloc(sizeof(double) * n);
        I1s[s] = (double *)malloc(sizeof(double) * n);
        u1s[s] = s ? (double *)malloc(sizeof(double) * n) : u;
        u2s[s] = s ? (double *)malloc(sizeof(double) * n) : u + N;
        if (!I0s[s] || !I1s[s] || !u1s[s] || !u2s[s]) st = ORC_ERR_ALLOC;
    }
    if (!st) {
        double mx = -INFINITY, mn = INFINITY;
        for (size_t i = 0; i < N; i++) {
            const double a = I0in[i], b = I1in[i];
            mx = fmax(mx, fmax(a, b));
            mn = fmin(mn, fmin(a, b));
        }
        const double den = mx - mn;
        for (size_t i = 0; i < N; i++) {
            I0s[0][i] = den > 0 ? 255.0 * (I0in[i] - mn) / den : I0in[i];
            I1s[0][i] = den > 0 ? 255.0 * (I1in[i] - mn) / den : I1in[i];
        }
        st = o64_gaussian(I0s[0], width, height, O64_PRESMOOTHING_SIGMA);
        if (!st) st = o64_gaussian(I1s[0], width, height, O64_PRESMOOTHING_SIGMA);
    }
    for (int s = 1; s < ns && !st; s++) {
        st = zoom_out64(I0s[s - 1], I0s[s], nx[s - 1], ny[s - 1], prm->zfactor, nx[s], ny[s]);
        if (!st) st = zoom_out64(I1s[s - 1], I1s[s], nx[s - 1], ny[s - 1], prm->zfactor, nx[s], ny[s]);
    }
    if (iters)
        for (int i = 0; i < ns * prm->warps; i++) iters[i] = 0;
    if (!st) {
        const size_t nc = (size_t)nx[ns - 1] * ny[ns - 1];
        memset(u1s[ns - 1], 0, sizeof(double) * nc);
        memset(u2s[ns - 1], 0, sizeof(double) * nc);
        const double invz = 1.0 / (double)prm->zfactor;
        for (int s = ns - 1; s >= 0 && !st; s--) {
            st = o64_dual_tvl1(I0s[s], I1s[s], u1s[s], u2s[s], nx[s], ny[s], prm,
                               iters ? iters + (size_t)s * prm->warps : NULL);
            if (st || s == 0) break;
            zoom_in64(u1s[s], u1s[s - 1], nx[s], ny[s], nx[s - 1], ny[s - 1]);
            zoom_in64(u2s[s], u2s[s - 1], nx[s], ny[s], nx[s - 1], ny[s - 1]);
            const size_t nf = (size_t)nx[s - 1] * ny[s - 1];
            for (size_t i = 0; i < nf; i++) {
                u1s[s - 1][i] *= invz;
                u2s[s - 1][i] *= invz;
            }
        }
    }
    for (int s = 0; s < ns; s++) {
        free(I0s[s]);
        free(I1s[s]);
        if (s) { free(u1s[s]); free(u2s[s]); }
    }
    return st;
}

/* ----------------------------------------------------------------- HS --- */

/* Bilinear sample at (x, y), clamp-to-edge (warpingKernel.cuh:28-43 and
 * upscaleKernel.cuh:26-40 as intended). */
static double bilinear64(const double *src, int w, int h, double x, double y)
{
    const double fx = floor(x), fy = floor(y), ax = x - fx, ay = y - fy;
    const int x0 = clampi(f2i_sat(fx), -1, w), y0 = clampi(f2i_sat(fy), -1, h);
    const int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    const int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    const double a = src[(size_t)ya * w + xa], b = src[(size_t)ya * w + xb];
    const double c = src[(size_t)yb * w + xa], d = src[(size_t)yb * w + xb];
    return (1 - ay) * ((1 - ax) * a + ax * b) + ay * ((1 - ax) * c + ax * d);
}

/* 2x2 box mean of every output pixel (downscaleKernel.cuh:27-40). */
void o64_hs_downscale(const double *src, int w, int h, double *out, int nw, int nh)
{
    (void)h;
    for (int iy = 0; iy < nh; iy++)
        for (int ix = 0; ix < nw; ix++) {
            const size_t a = (size_t)(2 * iy) * w + 2 * ix;
            out[(size_t)iy * nw + ix] = 0.25 * (src[a] + src[a + w] + src[a + 1] + src[a + w + 1]);
        }
}

/* x2 upscale: bilinear at ((ix - 0.5)/2, (iy - 0.5)/2) times `scale`
 * (upscaleKernel.cuh:26-40). */
void o64_hs_upscale(const double *src, int w, int h, double *out, int nw, int nh, double scale)
{
    for (int iy = 0; iy < nh; iy++)
        for (int ix = 0; ix < nw; ix++)
            out[(size_t)iy * nw + ix] = bilinear64(src, w, h, (ix - 0.5) * 0.5, (iy - 0.5) * 0.5) * scale;
}

/* [1 -8 0 8 -1]/12 with clamped indices (derivativesKernel.cuh:41-72) */
static inline double d5_64(const double *p, int c, int n, ptrdiff_t step)
{
    return (p[clampi(c - 2, 0, n - 1) * step] - 8.0 * p[clampi(c - 1, 0, n - 1) * step] +
            8.0 * p[clampi(c + 1, 0, n - 1) * step] - p[clampi(c + 2, 0, n - 1) * step]) / 12.0;
}

/* One Jacobi sweep (solverKernel.cuh:145-152), clamp-to-edge neighbours. */
void o64_hs_jacobi(const double *du0, const double *dv0, const double *Ix, const double *Iy,
                   const double *Iz, int w, int h, double alpha, double *du1, double *dv1)
{
    #pragma omp parallel for
    for (int iy = 0; iy < h; iy++)
        for (int ix = 0; ix < w; ix++) {
            const size_t p = (size_t)iy * w + ix;
            const size_t l = (size_t)iy * w + clampi(ix - 1, 0, w - 1);
            const size_t r = (size_t)iy * w + clampi(ix + 1, 0, w - 1);
            const size_t up = (size_t)clampi(iy + 1, 0, h - 1) * w + ix;
            const size_t dn = (size_t)clampi(iy - 1, 0, h - 1) * w + ix;
            const double su = 0.25 * (du0[l] + du0[r] + du0[up] + du0[dn]);
            const double sv = 0.25 * (dv0[l] + dv0[r] + dv0[up] + dv0[dn]);
            const double frac = (Ix[p] * su + Iy[p] * sv + Iz[p]) / (Ix[p] * Ix[p] + Iy[p] * Iy[p] + alpha);
            du1[p] = su - Ix[p] * frac;
            dv1[p] = sv - Iy[p] * frac;
        }
}

/* ComputeFlow (flow.cu:112-208) in double on dense frames (stride = width
 * for the pyramid levels; the caller's stride is honoured on input and
 * output). */
int o64_hs_compute(const float *I0, const float *I1, int width, int height, int stride,
                   double alpha, int nLevels, int nWarpIters, int nSolverIters, double *u, double *v)
{
    if (nLevels < 1 || nLevels > 32 || width < 1 || height < 1 || stride < width) return ORC_ERR_ARG;
    int pW[32], pH[32];
    double *pI0[32] = {0}, *pI1[32] = {0};
    int L = nLevels - 1;
    pW[L] = width;
    pH[L] = height;
    const size_t N = (size_t)width * height;
    pI0[L] = (double *)malloc(sizeof(double) * N);
    pI1[L] = (double *)malloc(sizeof(double) * N);
    if (!pI0[L] || !pI1[L]) return ORC_ERR_ALLOC;
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            pI0[L][(size_t)y * width + x] = I0[(size_t)y * stride + x];
            pI1[L][(size_t)y * width + x] = I1[(size_t)y * stride + x];
        }
    for (; L > 0; --L) {
        const int nw = pW[L] / 2, nh = pH[L] / 2;
        if (nw < 1 || nh < 1) {   /* too many levels for the frame */
            for (int k = L; k < nLevels; k++) {
                free(pI0[k]);
                free(pI1[k]);
            }
            return ORC_ERR_ARG;
        }
        pI0[L - 1] = (double *)malloc(sizeof(double) * (size_t)nw * nh);
        pI1[L - 1] = (double *)malloc(sizeof(double) * (size_t)nw * nh);
        o64_hs_downscale(pI0[L], pW[L], pH[L], pI0[L - 1], nw, nh);
        o64_hs_downscale(pI1[L], pW[L], pH[L], pI1[L - 1], nw, nh);
        pW[L - 1] = nw;
        pH[L - 1] = nh;
    }
    double *m = (double *)calloc(12 * N, sizeof(double));
    if (!m) return ORC_ERR_ALLOC;
    double *du0 = m, *dv0 = m + N, *du1 = m + 2 * N, *dv1 = m + 3 * N, *Ix = m + 4 * N;
    double *Iy = m + 5 * N, *Iz = m + 6 * N, *I1w = m + 7 * N, *cu = m + 8 * N, *cv = m + 9 * N;
    double *nu = m + 10 * N, *nv = m + 11 * N;
    for (L = 0; L < nLevels; ++L) {
        const int w = pW[L], h = pH[L];
        const size_t n = (size_t)w * h;
        for (int wi = 0; wi < nWarpIters; ++wi) {
            #pragma omp parallel for
            for (int iy = 0; iy < h; iy++)
                for (int ix = 0; ix < w; ix++) {
                    const size_t p = (size_t)iy * w + ix;
                    I1w[p] = bilinear64(pI1[L], w, h, ix + cu[p], iy + cv[p]);
                }
            #pragma omp parallel for
            for (int iy = 0; iy < h; iy++)
                for (int ix = 0; ix < w; ix++) {
                    const size_t p = (size_t)iy * w + ix;
                    const double *r0 = pI0[L] + (size_t)iy * w, *r1 = I1w + (size_t)iy * w;
                    Ix[p] = 0.5 * (d5_64(r0, ix, w, 1) + d5_64(r1, ix, w, 1));
                    Iy[p] = 0.5 * (d5_64(pI0[L] + ix, iy, h, w) + d5_64(I1w + ix, iy, h, w));
                    Iz[p] = I1w[p] - pI0[L][p];
                }
            memset(du0, 0, sizeof(double) * n);
            memset(dv0, 0, sizeof(double) * n);
            for (int it = 0; it < nSolverIters; ++it) {
                o64_hs_jacobi(du0, dv0, Ix, Iy, Iz, w, h, alpha, du1, dv1);
                double *t = du0; du0 = du1; du1 = t;
                t = dv0; dv0 = dv1; dv1 = t;
            }
            for (size_t p = 0; p < n; p++) {
                cu[p] += du0[p];
                cv[p] += dv0[p];
            }
        }
        if (L != nLevels - 1) {
            o64_hs_upscale(cu, w, h, nu, pW[L + 1], pH[L + 1], (double)pW[L + 1] / pW[L]);
            o64_hs_upscale(cv, w, h, nv, pW[L + 1], pH[L + 1], (double)pH[L + 1] / pH[L]);
            double *t = cu; cu = nu; nu = t;
            t = cv; cv = nv; nv = t;
        }
    }
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            u[(size_t)y * stride + x] = cu[(size_t)y * width + x];
            v[(size_t)y * stride + x] = cv[(size_t)y * width + x];
        }
    for (L = 0; L < nLevels; ++L) { free(pI0[L]); free(pI1[L]); }
    free(m);
    return ORC_OK;
}

/* ----------------------------------------------------------------- LK --- */

/* Window sums of a w x w box clipped to the image. */
static void box_sum64(const double *in, double *out, double *tmp, int nx, int ny, int wc)
{
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            double s = 0.0;
            for (int k = -wc; k <= wc; k++)
                if (j + k >= 0 && j + k < nx) s += in[(size_t)i * nx + j + k];
            tmp[(size_t)i * nx + j] = s;
        }
    #pragma omp parallel for
    for (int i = 0; i < ny; i++)
        for (int j = 0; j < nx; j++) {
            double s = 0.0;
            for (int k = -wc; k <= wc; k++)
                if (i + k >= 0 && i + k < ny) s += tmp[(size_t)(i + k) * nx + j];
            out[(size_t)i * nx + j] = s;
        }
}

/* Pyramidal iterative LK (the algorithm of orc_lk_pyr_compute) in double:
 * Gaussian pyramid (blur sigma, bicubic decimation at i/zf), per level the
 * structure tensor of I0's centred gradient over the window, `iters`
 * Gauss-Newton steps d = -G^{-1} b with b the window sums of (Ix It, Iy It),
 * It = I1(x + u) - I0 (bilinear), applied where det G > det_eps; zoom_in and
 * 1/zf between levels. */
int o64_lk_pyr_compute(const float *I0, const float *I1, int width, int height, int levels,
                       double zf, double sigma, int win, int iters, double det_eps, double *u,
                       double *v)
{
    if (levels < 1 || levels > 16 || win < 1 || (win & 1) == 0 || iters < 0) return ORC_ERR_ARG;
    int nx[16], ny[16];
    double *P0[16] = {0}, *P1[16] = {0};
    const size_t n0 = (size_t)width * height;
    nx[0] = width;
    ny[0] = height;
    P0[0] = (double *)malloc(sizeof(double) * n0);
    P1[0] = (double *)malloc(sizeof(double) * n0);
    if (!P0[0] || !P1[0]) return ORC_ERR_ALLOC;
    for (size_t i = 0; i < n0; i++) {
        P0[0][i] = I0[i];
        P1[0][i] = I1[i];
    }
    int st = ORC_OK;
    for (int l = 1; l < levels && !st; l++) {
        orc_zoom_size(nx[l - 1], ny[l - 1], (float)zf, &nx[l], &ny[l]);
        const size_t nl = (size_t)nx[l] * ny[l];
        P0[l] = (double *)malloc(sizeof(double) * nl);
        P1[l] = (double *)malloc(sizeof(double) * nl);
        double *blur = (double *)malloc(sizeof(double) * (size_t)nx[l - 1] * ny[l - 1]);
        for (int k = 0; k < 2 && !st; k++) {
            memcpy(blur, k ? P1[l - 1] : P0[l - 1], sizeof(double) * (size_t)nx[l - 1] * ny[l - 1]);
            st = o64_gaussian(blur, nx[l - 1], ny[l - 1], sigma);
            if (st) break;
            double *dst = k ? P1[l] : P0[l];
            for (int i = 0; i < ny[l]; i++)
                for (int j = 0; j < nx[l]; j++)
                    dst[(size_t)i * nx[l] + j] =   /* zf = 1/2: the blurred pixel, as oracle.c */
                        zf == 0.5 ? blur[(size_t)(2 * i) * nx[l - 1] + 2 * j]
                                  : o64_bicubic_at(blur, j / zf, i / zf, nx[l - 1], ny[l - 1], 0);
        }
        free(blur);
        if (st) {   /* levels 0..l hold buffers */
            for (int k = 0; k <= l; k++) {
                free(P0[k]);
                free(P1[k]);
            }
            return st;
        }
    }
    const int wc = (win - 1) / 2;
    double *m = (double *)calloc(13 * n0, sizeof(double));
    if (!m) return ORC_ERR_ALLOC;
    double *cu = m, *cv = m + n0, *nu = m + 2 * n0, *nv = m + 3 * n0, *Ix = m + 4 * n0;
    double *Iy = m + 5 * n0, *Gxx = m + 6 * n0, *Gxy = m + 7 * n0, *Gyy = m + 8 * n0;
    double *t0 = m + 9 * n0, *t1 = m + 10 * n0, *bx = m + 11 * n0, *by = m + 12 * n0;
    for (int l = levels - 1; l >= 0; l--) {
        const int w = nx[l], h = ny[l];
        const size_t nl = (size_t)w * h;
        centered_gradient64(P0[l], Ix, Iy, w, h);
        for (size_t i = 0; i < nl; i++) t0[i] = Ix[i] * Ix[i];
        box_sum64(t0, Gxx, t1, w, h, wc);
        for (size_t i = 0; i < nl; i++) t0[i] = Ix[i] * Iy[i];
        box_sum64(t0, Gxy, t1, w, h, wc);
        for (size_t i = 0; i < nl; i++) t0[i] = Iy[i] * Iy[i];
        box_sum64(t0, Gyy, t1, w, h, wc);
        for (int it = 0; it < iters; it++) {
            #pragma omp parallel for
            for (int i = 0; i < h; i++)
                for (int j = 0; j < w; j++) {
                    const size_t p = (size_t)i * w + j;
                    const double It = bilinear64(P1[l], w, h, j + cu[p], i + cv[p]) - P0[l][p];
                    nu[p] = Ix[p] * It;
                    nv[p] = Iy[p] * It;
                }
            box_sum64(nu, bx, t1, w, h, wc);
            box_sum64(nv, by, t1, w, h, wc);
            for (size_t p = 0; p < nl; p++) {
                const double det = Gxx[p] * Gyy[p] - Gxy[p] * Gxy[p];
                if (det > det_eps) {
                    cu[p] += (Gxy[p] * by[p] - Gyy[p] * bx[p]) / det;
                    cv[p] += (Gxy[p] * bx[p] - Gxx[p] * by[p]) / det;
                }
            }
        }
        if (l > 0) {
            zoom_in64(cu, nu, w, h, nx[l - 1], ny[l - 1]);
            zoom_in64(cv, nv, w, h, nx[l - 1], ny[l - 1]);
            const size_t nf = (size_t)nx[l - 1] * ny[l - 1];
            for (size_t i = 0; i < nf; i++) {
                cu[i] = nu[i] / zf;
                cv[i] = nv[i] / zf;
            }
        }
    }
    memcpy(u, cu, sizeof(double) * n0);
    memcpy(v, cv, sizeof(double) * n0);
    for (int l = 0; l < levels; l++) { free(P0[l]); free(P1[l]); }
    free(m);
    return ORC_OK;
}
