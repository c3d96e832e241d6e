// tvl1_pixel.hpp -- per-pixel bodies of the TV-L1 scale and warping kernels,
// shared by the one-pair kernels (tvl1_kernels.hip) and the batched
// kernels that cover one scale of many pairs per launch (batch_lockstep.hip),
// so both produce the same bits.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"

namespace amdflow {

// (I1, I1x, I1y, 0) at pixel (i, j): the centred gradient of mask.cpp:152-209
// (one-sided differences times 0.5 on the borders, evaluated in double).
__device__ __forceinline__ void gradient_packed_px(const float* __restrict__ in,
                                                   float4* __restrict__ out, int nx, int ny, int i,
                                                   int j)
{
    const int jl = j > 0 ? j - 1 : 0, jr = j < nx - 1 ? j + 1 : nx - 1;
    const int iu = i > 0 ? i - 1 : 0, id = i < ny - 1 ? i + 1 : ny - 1;
    const float ddx = in[(size_t)i * nx + jr] - in[(size_t)i * nx + jl];
    const float ddy = in[(size_t)id * nx + j] - in[(size_t)iu * nx + j];
    out[(size_t)i * nx + j] = make_float4(in[(size_t)i * nx + j], (float)(0.5 * (double)ddx),
                                          (float)(0.5 * (double)ddy), 0.0f);
}

// Warping set-up at pixel (i, j) (tvl1.cpp:237-254): I1, I1x, I1y sampled
// at x + u with border_out = true, then I1wx, I1wy and
// rho_c = I1w - I1wx*u1 - I1wy*u2 - I0.  `ub` is the current flow buffer
// (u1 plane, u2 plane at +nx*ny).
__device__ __forceinline__ void warp_setup_px(const float* __restrict__ I0,
                                              const float4* __restrict__ src,
                                              const float* __restrict__ ub, float* __restrict__ wx,
                                              float* __restrict__ wy, float* __restrict__ rhoc,
                                              int nx, int ny, int i, int j)
{
    const size_t plane = (size_t)nx * ny;
    const size_t p = (size_t)i * nx + j;
    const float u1 = ub[p], u2 = ub[plane + p];
    const float uu = (float)j + u1;
    const float vv = (float)i + u2;
    const BicubicTaps t = bicubic_taps(uu, vv, nx, ny);
    float w = 0.0f, gx = 0.0f, gy = 0.0f;
    if (!t.out) bicubic_sample3(src, nx, t, w, gx, gy);
    wx[p] = gx;
    wy[p] = gy;
    rhoc[p] = ((w - gx * u1) - gy * u2) - I0[p];
}

// zoom_in of both flow components at fine pixel (i1, j1) (zoom.cpp:61-85)
// times 1/zfactor (tvl1.cpp:213-214).  `uc` is the coarse scale's current
// buffer, `uf` the fine scale's buffer 0.
__device__ __forceinline__ void zoom_in_flow_px(const float* __restrict__ uc,
                                                float* __restrict__ uf, int nx, int ny, int nxx,
                                                int nyy, float factorx, float factory, float invz,
                                                int i1, int j1)
{
    const size_t nc = (size_t)nx * ny, nf = (size_t)nxx * nyy;
    const float i2 = (float)i1 / factory;
    const float j2 = (float)j1 / factorx;
    const BicubicTaps t = bicubic_taps(j2, i2, nx, ny);
    const float a = bicubic_sample(uc, nx, t);
    const float b = bicubic_sample(uc + nc, nx, t);
    uf[(size_t)i1 * nxx + j1] = invz * a;
    uf[nf + (size_t)i1 * nxx + j1] = invz * b;
}

}  // namespace amdflow
