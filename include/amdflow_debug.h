/* amdflow_debug.h -- per-kernel entry points of libamdflow.so for parity
 * tests.  Each takes host arrays, runs one GPU kernel (or one fixed kernel
 * sequence) on them and copies the result back; they exist so that every
 * stage of the solver path can be compared with its CPU counterpart.  They
 * allocate device memory per call and are not meant for production use.
 * Sizes are width x height row-major float planes unless noted. */
#ifndef AMDFLOW_DEBUG_H
#define AMDFLOW_DEBUG_H

#include <stddef.h>
#include <stdint.h>

#include "amdflow.h"

#ifdef __cplusplus
extern "C" {
#endif

/* gaussian(I, xdim, ydim, sigma, buffer), mask.cpp:204-274 (out of place). */
int of_dbg_gaussian(const float *in, float *out, int width, int height, double sigma);
/* zoom_out(I, Iout, nx, ny, factor, ...), zoom.cpp:20-53. */
int of_dbg_zoom_out(const float *in, float *out, int width, int height, float factor);
/* zoom_in(I, Iout, nx, ny, nxx, nyy), zoom.cpp:61-85, times `scale`. */
int of_dbg_zoom_in(const float *in, float *out, int width, int height, int nxx, int nyy,
                   float scale);
/* centered_gradient, mask.cpp:139-196. */
int of_dbg_centered_gradient(const float *in, float *dx, float *dy, int width, int height);
/* image_normalization of two frames, tvl1.cpp:302-329 (in place). */
int of_dbg_normalize(float *I0, float *I1, int width, int height);
/* warp set-up of one warping, tvl1.cpp:198-212: I1wx, I1wy, rho_c. */
int of_dbg_warp_setup(const float *I0, const float *I1, const float *I1x, const float *I1y,
                      const float *u1, const float *u2, float *I1wx, float *I1wy, float *rho_c,
                      int width, int height);
/* `iters` primal-dual iterations (tvl1.cpp:218-291) from the given state,
 * with launch shape (iters_per_launch, rows_per_wave) and no stopping
 * test; state arrays are updated in place; err_sums receives the per
 * iteration sum of (du1^2+du2^2) (double). */
int of_dbg_tvl1_iterations(const float *I1wx, const float *I1wy, const float *rho_c, float *u1,
                           float *u2, float *p11, float *p12, float *p21, float *p22, int width,
                           int height, float tau, float lambda, float theta, int iters,
                           int iters_per_launch, int rows_per_wave, double *err_sums);
/* hypot as the solver evaluates it on the GPU, element-wise. */
int of_dbg_hypot(const float *a, const float *b, float *out, size_t n);

/* One scale of the solver, dualTVL1(I0, I1, u1, u2, nx, ny)
 * (tvl1.cpp:182-294), on float frames taken as they are (no normalisation,
 * pre-smoothing or pyramid): centred gradient of I1, p = 0, then
 * params->warps warpings of warp set-up and iterations under the stopping
 * rule, with the product's launch shapes.  u1, u2 hold the initial flow and
 * receive the result; iters receives the params->warps iteration counts.
 * params->nscales is ignored. */
int of_dbg_tvl1_dual(const float *I0, const float *I1, float *u1, float *u2, int width,
                     int height, const of_tvl1_params *params, int *iters);

/* Horn-Schunck pyramid steps on dense planes (stride = width): the 2x2 box
 * downscale of a w x h plane to nw x nh with 2*nw <= w, 2*nh <= h
 * (downscaleKernel.cuh:27-40), and the x2 flow upscale of a w x h field to
 * nw x nh times `scale` (upscaleKernel.cuh:26-40). */
int of_dbg_hs_downscale(const float *src, int w, int h, float *out, int nw, int nh);
int of_dbg_hs_upscale(const float *src, int w, int h, float *out, int nw, int nh, float scale);

/* Horn-Schunck: `iters` Jacobi sweeps (solverKernel.cuh:145-152) from
 * (du, dv) with launch shape (sweeps_per_launch K, rows_per_wave R). */
int of_dbg_hs_jacobi(const float *du, const float *dv, const float *Ix, const float *Iy,
                     const float *Iz, int width, int height, float alpha, int iters, int K, int R,
                     float *du_out, float *dv_out);
/* Horn-Schunck: bilinear warp of I1 by (u, v), then Ix, Iy, Iz
 * (warpingKernel.cuh, derivativesKernel.cuh intent). */
int of_dbg_hs_warp_derivatives(const float *I0, const float *I1, const float *u, const float *v,
                               int width, int height, float *I1w, float *Ix, float *Iy, float *Iz);
/* Warm-start network: D = A * B for 16x16 row-major fp32 matrices with
 * bf16-rounded operands, through the convolution kernel's MFMA path (a
 * 1x16 image of 16 channels, centre tap only). */
int of_dbg_cnn_mfma(const float *A, const float *B, float *D);
/* While `on` is nonzero, launches of the persistent warping kernel report at
 * their first grid barrier that the grid was not co-resident (every
 * workgroup arrives, then all end), so the solve takes its host-paced
 * repeat; a test of that path. */
int of_dbg_persist_force_not_resident(int on);
/* Allocation failure injection: after this call the first `n` device or
 * pinned host allocations of the library succeed and every later one fails
 * with OF_ERR_ALLOC, as a real failure would (n < 0: none fails).  The call
 * also restarts the attempt count of of_dbg_alloc_attempts. */
int of_dbg_fail_alloc_after(long long n);
/* Allocation attempts since the last of_dbg_fail_alloc_after. */
long long of_dbg_alloc_attempts(void);
/* Device and pinned host blocks the library holds now. */
long long of_dbg_live_allocs(void);

#ifdef __cplusplus
}
#endif
#endif
