// tvl1_pixel.hpp -- per-pixel bodies of the TV-L1 scale and warping kernels,
// shared by the one-pair kernels (tvl1_kernels.hip) and the batched
// kernels that cover one scale of many pairs per launch (batch_lockstep.hip),
// so both produce the same bits.
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "device_math.hpp"
#include "tvl1_kernels.hpp"

namespace amdflow {

// (I1, I1x, I1y, 0) at pixel (i, j): the centred gradient of mask.cpp:139-196
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

// Warping set-up at pixel (i, j) (tvl1.cpp:198-212): I1, I1x, I1y sampled
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
// times 1/zfactor (tvl1.cpp:171-172).  `uc` is the coarse scale's current
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

// ---------------------------------------------------------------------------
// phase 1: normalisation, Gaussian, pyramid sampling
// ---------------------------------------------------------------------------

// image_normalization bounds (tvl1.cpp:302-329) from the min/max keys of
// both frames: mx = max of the two isamax picks, mn = min of the two isamin
// picks, den = mx - mn.
__device__ __forceinline__ void norm_bounds(const TvCtl* ctl, const float* a, const float* b,
                                            float* mn, float* den)
{
    const unsigned imax0 = 0xFFFFFFFFu - (unsigned)(ctl->maxkey[0] & 0xFFFFFFFFull);
    const unsigned imax1 = 0xFFFFFFFFu - (unsigned)(ctl->maxkey[1] & 0xFFFFFFFFull);
    const unsigned imin0 = (unsigned)(ctl->minkey[0] & 0xFFFFFFFFull);
    const unsigned imin1 = (unsigned)(ctl->minkey[1] & 0xFFFFFFFFull);
    const float a0 = a[imax0], a1 = b[imax1];
    const float b0 = a[imin0], b1 = b[imin1];
    const float mx = (a0 < a1) ? a1 : a0;
    *mn = (b1 < b0) ? b1 : b0;
    *den = mx - *mn;
}

// One normalised value: unchanged when den <= 0, else 255*(x-mn)/den in double.
__device__ __forceinline__ float norm_px(float x, float mn, float den)
{
    return den <= 0 ? x : (float)(255.0 * (double)(x - mn) / (double)den);
}

// cblas_isamax / cblas_isamin of both frames as 64-bit keys: |x| bits in
// the high word, the index (complemented for the max, so that the first
// index wins) in the low word.
__device__ __forceinline__ void minmax_keys(float va, float vb, unsigned i, unsigned long long& kmax0,
                                            unsigned long long& kmax1, unsigned long long& kmin0,
                                            unsigned long long& kmin1)
{
    const unsigned long long ka = (unsigned long long)__float_as_uint(fabsf(va)) << 32;
    const unsigned long long kb = (unsigned long long)__float_as_uint(fabsf(vb)) << 32;
    const unsigned long long hi = 0xFFFFFFFFull - i;
    // a NaN compares neither greater nor smaller: the scan's first element
    // is kept if it is NaN, a NaN anywhere else is never picked
    const bool na = va != va, nb = vb != vb;
    const unsigned long long xa = na ? (i == 0 ? ~0ull : 0ull) : (ka | hi);
    const unsigned long long xb = nb ? (i == 0 ? ~0ull : 0ull) : (kb | hi);
    const unsigned long long ya = na ? (i == 0 ? 0ull : ~0ull) : (ka | i);
    const unsigned long long yb = nb ? (i == 0 ? 0ull : ~0ull) : (kb | i);
    kmax0 = xa > kmax0 ? xa : kmax0;
    kmax1 = xb > kmax1 ? xb : kmax1;
    kmin0 = ya < kmin0 ? ya : kmin0;
    kmin1 = yb < kmin1 ? yb : kmin1;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

// Low side: whole-sample mirror x[-k] = x[k]; high side: half-sample mirror
// x[n-1+k] = x[n-k] (mask.cpp:240-242, 261-263).
__device__ __forceinline__ int mirror_lo(int i) { return i < 0 ? -i : i; }
__device__ __forceinline__ int mirror_hi(int i, int n) { return i >= n ? 2 * n - 1 - i : i; }

// Separable Gaussian (mask.cpp:204-274), double accumulation: the row pass
// and the column pass at pixel (y, x).
__device__ __forceinline__ float gauss_row_px(const float* __restrict__ src, int nx, int x, int y,
                                              const GaussTaps& g)
{
    const float* row = src + (size_t)y * nx;
    double sum = g.B[0] * (double)row[x];
    for (int j = 1; j < g.size; j++) {
        const float pair = row[mirror_lo(x - j)] + row[mirror_hi(x + j, nx)];
        sum += g.B[j] * (double)pair;
    }
    return (float)sum;
}

__device__ __forceinline__ float gauss_col_px(const float* __restrict__ src, int nx, int ny, int x,
                                              int y, const GaussTaps& g)
{
    const float* col = src + x;
    double sum = g.B[0] * (double)col[(size_t)y * nx];
    for (int j = 1; j < g.size; j++) {
        const float pair = col[(size_t)mirror_lo(y - j) * nx] + col[(size_t)mirror_hi(y + j, ny) * nx];
        sum += g.B[j] * (double)pair;
    }
    return (float)sum;
}

// zoom_out's resampling of the blurred image at (j/f, i/f), zoom.cpp:28-52.
__device__ __forceinline__ float zoom_out_px(const float* __restrict__ src, int nx, int ny, int i1,
                                             int j1, float factor)
{
    const float i2 = (float)i1 / factor;
    const float j2 = (float)j1 / factor;
    return bicubic_at(src, j2, i2, nx, ny, false);
}

}  // namespace amdflow
