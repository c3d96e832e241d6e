// hs.hip -- Horn-Schunck on gfx950: kernels, host orchestration and C ABI.
//
// Follows the control flow of ComputeFlow (src/hs/cuda/flow.cu:112-208):
// a /2 pyramid, coarse to fine, nWarpIters x (warp I1, derivatives,
// nSolverIters Jacobi sweeps, u += du), flow upscaled between levels.  The
// reference kernels index images incorrectly (warpingKernel.cuh:38-42,
// upscaleKernel.cuh:33-39, derivativesKernel.cuh:42-46); these kernels
// implement what they intend: bilinear sampling with clamp-to-edge
// addressing and clamp-to-edge stencils.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "amdflow.h"
#include "amdflow_debug.h"
#include "common.hpp"
#include "device_math.hpp"

namespace amdflow {
namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Bilinear sample at (x, y) in pixel-centre coordinates, clamp-to-edge.
__device__ __forceinline__ float bilinear_at(const float* __restrict__ src, int w, int h, int s,
                                             float x, float y)
{
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    // limited to [-1, w] first so that +1 cannot overflow (clamps unchanged)
    const int x0 = clampi(f2i_sat(fx0), -1, w), y0 = clampi(f2i_sat(fy0), -1, h);
    const int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    const int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    const float a = src[(size_t)ya * s + xa], b = src[(size_t)ya * s + xb];
    const float c = src[(size_t)yb * s + xa], d = src[(size_t)yb * s + xb];
    const float top = (1.0f - ax) * a + ax * b;
    const float bot = (1.0f - ax) * c + ax * d;
    return (1.0f - ay) * top + ay * bot;
}

// 2x2 box average (downscaleKernel.cuh:27-40), every output pixel.
__global__ __launch_bounds__(256) void k_hs_downscale(const float* __restrict__ src, int s,
                                                      float* __restrict__ out, int nw, int nh,
                                                      int ns)
{
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ix >= nw || iy >= nh) return;
    const size_t sx = (size_t)ix * 2, sy = (size_t)iy * 2;
    const float a = src[sx + sy * s], b = src[sx + (sy + 1) * s];
    const float c = src[(sx + 1) + sy * s], d = src[(sx + 1) + (sy + 1) * s];
    out[(size_t)ix + (size_t)iy * ns] = 0.25f * (((a + b) + c) + d);
}

// Flow upscale x2 at ((ix-0.5)/2, (iy-0.5)/2) times `scale`
// (upscaleKernel.cuh:26-40).
__global__ __launch_bounds__(256) void k_hs_upscale(const float* __restrict__ src, int w, int h,
                                                    int s, float* __restrict__ out, int nw, int nh,
                                                    int ns, float scale)
{
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ix >= nw || iy >= nh) return;
    const float x = ((float)ix - 0.5f) * 0.5f;
    const float y = ((float)iy - 0.5f) * 0.5f;
    out[(size_t)ix + (size_t)iy * ns] = bilinear_at(src, w, h, s, x, y) * scale;
}

// Warp I1 by (u, v) then the derivatives of derivativesKernel.cuh:29-73:
// 5-tap [1 -8 0 8 -1]/12 on I0 and on the warped I1 (clamp to edge),
// Ix/Iy = mean of the two, Iz = I1w - I0.  The warped image is computed at
// the 9 stencil positions it is needed at.
__device__ __forceinline__ float d5(float m2, float m1, float p1, float p2)
{
    float t = m2;
    t -= m1 * 8.0f;
    t += p1 * 8.0f;
    t -= p2;
    t /= 12.0f;
    return t;
}

__global__ __launch_bounds__(256) void k_hs_warp(const float* __restrict__ I1, int w, int h, int s,
                                                 const float* __restrict__ u,
                                                 const float* __restrict__ v,
                                                 float* __restrict__ out)
{
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ix >= w || iy >= h) return;
    const size_t pos = (size_t)ix + (size_t)iy * s;
    out[pos] = bilinear_at(I1, w, h, s, (float)ix + u[pos], (float)iy + v[pos]);
}

__global__ __launch_bounds__(256) void k_hs_derivatives(const float* __restrict__ I0,
                                                        const float* __restrict__ I1w, int w,
                                                        int h, int s, float* __restrict__ Ix,
                                                        float* __restrict__ Iy,
                                                        float* __restrict__ Iz)
{
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ix >= w || iy >= h) return;
    const size_t pos = (size_t)ix + (size_t)iy * s;
    const size_t row = (size_t)iy * s;
    const int xm2 = clampi(ix - 2, 0, w - 1), xm1 = clampi(ix - 1, 0, w - 1);
    const int xp1 = clampi(ix + 1, 0, w - 1), xp2 = clampi(ix + 2, 0, w - 1);
    const size_t ym2 = (size_t)clampi(iy - 2, 0, h - 1) * s, ym1 = (size_t)clampi(iy - 1, 0, h - 1) * s;
    const size_t yp1 = (size_t)clampi(iy + 1, 0, h - 1) * s, yp2 = (size_t)clampi(iy + 2, 0, h - 1) * s;
    float t0 = d5(I0[row + xm2], I0[row + xm1], I0[row + xp1], I0[row + xp2]);
    float t1 = d5(I1w[row + xm2], I1w[row + xm1], I1w[row + xp1], I1w[row + xp2]);
    Ix[pos] = (t0 + t1) * 0.5f;
    Iz[pos] = I1w[pos] - I0[pos];
    t0 = d5(I0[ym2 + ix], I0[ym1 + ix], I0[yp1 + ix], I0[yp2 + ix]);
    t1 = d5(I1w[ym2 + ix], I1w[ym1 + ix], I1w[yp1 + ix], I1w[yp2 + ix]);
    Iy[pos] = (t0 + t1) * 0.5f;
}

// Waves per workgroup of the Jacobi kernel.  Eight waves of R rows hide the
// latency of a sweep better than four of 2R: at 640x480 with 4 sweeps per
// launch, 8 waves x 4 rows took 0.174 ms for 100 sweeps, 4 waves x 8 rows
// 0.198 ms and 4 waves x 4 rows 0.217 ms.
constexpr int kHsWaves = 8;

// K Jacobi sweeps (solverKernel.cuh:39-153) per launch.  Layout as in the
// TV-L1 iteration kernel, with kHsWaves waves: each lane a column of R rows
// held in registers, horizontal neighbours by lane shuffles, vertical ones
// from registers or one LDS row per wave (double-buffered: one barrier per
// sweep).  Values within K pixels of the tile edge go stale one pixel per
// sweep; the central (64-2K) x (8R-2K) pixels are stored.  `zero_in`: the
// input increments are all zero (first launch of a warping).

template <int K, int R>
__global__ __launch_bounds__(64 * kHsWaves) void k_hs_jacobi(const float* __restrict__ du0,
                                                   const float* __restrict__ dv0,
                                                   const float* __restrict__ Ix,
                                                   const float* __restrict__ Iy,
                                                   const float* __restrict__ Iz, int w, int h,
                                                   int s, float alpha, int kk, int zero_in,
                                                   float* __restrict__ du1,
                                                   float* __restrict__ dv1)
{
    constexpr int NW = kHsWaves;
    constexpr int TW = 64 - 2 * K;
    constexpr int TH = NW * R - 2 * K;
    __shared__ float sh_first[2][2][NW][64];   // [buf][u/v][wave][lane]: row 0 of each wave
    __shared__ float sh_last[2][2][NW][64];    // row R-1 of each wave

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int x = (int)blockIdx.x * TW - K + lane;
    const int ybase = (int)blockIdx.y * TH - K + wave * R;
    const int oy0 = (int)blockIdx.y * TH;
    const bool xin = x >= 0 && x < w;
    const bool xout = xin && lane >= K && lane < K + TW;
    const bool fc = x <= 0, lc = x >= w - 1;

    float u[R], v[R], gx[R], gy[R], gz[R], den[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = ybase + r;
        if (xin && y >= 0 && y < h) {
            const size_t p = (size_t)y * s + x;
            u[r] = zero_in ? 0.0f : du0[p];
            v[r] = zero_in ? 0.0f : dv0[p];
            gx[r] = Ix[p];
            gy[r] = Iy[p];
            gz[r] = Iz[p];
        } else {
            u[r] = v[r] = gx[r] = gy[r] = gz[r] = 0.0f;
        }
        den[r] = (gx[r] * gx[r] + gy[r] * gy[r]) + alpha;
    }

#pragma unroll 1
    for (int t = 0; t < kk; t++) {
        const int b = t & 1;
        sh_first[b][0][wave][lane] = u[0];
        sh_first[b][1][wave][lane] = v[0];
        sh_last[b][0][wave][lane] = u[R - 1];
        sh_last[b][1][wave][lane] = v[R - 1];
        __syncthreads();
        const float upU = wave < NW - 1 ? sh_first[b][0][wave + 1][lane] : u[R - 1];
        const float upV = wave < NW - 1 ? sh_first[b][1][wave + 1][lane] : v[R - 1];
        const float dnU = wave ? sh_last[b][0][wave - 1][lane] : u[0];
        const float dnV = wave ? sh_last[b][1][wave - 1][lane] : v[0];
        float nu[R], nv[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int y = ybase + r;
            const float lu = from_left(u[r]), lv = from_left(v[r]);
            const float ru = from_right(u[r]), rv = from_right(v[r]);
            const float left_u = fc ? u[r] : lu, left_v = fc ? v[r] : lv;
            const float right_u = lc ? u[r] : ru, right_v = lc ? v[r] : rv;
            // "up" is row y+1 and "down" row y-1, clamped at the image edge
            const float up_u = (y >= h - 1) ? u[r] : (r < R - 1 ? u[r + 1] : upU);
            const float up_v = (y >= h - 1) ? v[r] : (r < R - 1 ? v[r + 1] : upV);
            const float dn_u = (y <= 0) ? u[r] : (r > 0 ? u[r - 1] : dnU);
            const float dn_v = (y <= 0) ? v[r] : (r > 0 ? v[r - 1] : dnV);
            const float sumU = (((left_u + right_u) + up_u) + dn_u) * 0.25f;
            const float sumV = (((left_v + right_v) + up_v) + dn_v) * 0.25f;
            const float frac = ((gx[r] * sumU + gy[r] * sumV) + gz[r]) / den[r];
            nu[r] = sumU - gx[r] * frac;
            nv[r] = sumV - gy[r] * frac;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            u[r] = nu[r];
            v[r] = nv[r];
        }
    }
    if (xout) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int y = ybase + r;
            if (y >= oy0 && y < oy0 + TH && y < h) {
                const size_t p = (size_t)y * s + x;
                du1[p] = u[r];
                dv1[p] = v[r];
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_hs_add(float* __restrict__ u, const float* __restrict__ du,
                                                float* __restrict__ v, const float* __restrict__ dv,
                                                int w, int h, int s)
{
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (ix >= w || iy >= h) return;
    const size_t p = (size_t)ix + (size_t)iy * s;
    u[p] = u[p] + du[p];
    v[p] = v[p] + dv[p];
}

__global__ __launch_bounds__(256) void k_hs_zero(float* __restrict__ a, float* __restrict__ b,
                                                 size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        a[i] = 0.0f;
        b[i] = 0.0f;
    }
}

inline dim3 grid2d(int w, int h) { return dim3(cdiv(w, 64), cdiv(h, 4)); }

#define HS_CHECK(name)                                      \
    do {                                                    \
        hipError_t e_ = hipGetLastError();                  \
        if (e_ != hipSuccess) {                             \
            record_hip_error(e_, name, __FILE__, __LINE__); \
            return OF_ERR_HIP;                              \
        }                                                   \
    } while (0)

bool jacobi_shape_ok(int K, int R)
{
    if (!(K == 1 || K == 2 || K == 4 || K == 8)) return false;
    if (!(R == 2 || R == 4 || R == 8)) return false;
    return kHsWaves * R - 2 * K >= 4;
}

template <int K, int R>
int launch_jacobi_kr(const float* du0, const float* dv0, const float* Ix, const float* Iy,
                     const float* Iz, int w, int h, int s, float alpha, int kk, int zero_in,
                     float* du1, float* dv1, hipStream_t st)
{
    const dim3 g(cdiv(w, 64 - 2 * K), cdiv(h, kHsWaves * R - 2 * K));
    hipLaunchKernelGGL((k_hs_jacobi<K, R>), g, dim3(64 * kHsWaves), 0, st, du0, dv0, Ix, Iy, Iz, w, h, s,
                       alpha, kk, zero_in, du1, dv1);
    HS_CHECK("k_hs_jacobi");
    return OF_OK;
}

int launch_jacobi(int K, int R, const float* du0, const float* dv0, const float* Ix,
                  const float* Iy, const float* Iz, int w, int h, int s, float alpha, int kk,
                  int zero_in, float* du1, float* dv1, hipStream_t st)
{
#define HS_CASE(k, r)                                                                           \
    if (K == k && R == r)                                                                       \
        return launch_jacobi_kr<k, r>(du0, dv0, Ix, Iy, Iz, w, h, s, alpha, kk, zero_in, du1, \
                                      dv1, st);
    HS_CASE(1, 2) HS_CASE(2, 2) HS_CASE(1, 4) HS_CASE(2, 4) HS_CASE(4, 4)
    HS_CASE(1, 8) HS_CASE(2, 8) HS_CASE(4, 8) HS_CASE(8, 8)
#undef HS_CASE
    return OF_ERR_UNSUPPORTED;
}

// n Jacobi sweeps from (du0,dv0) (or from zero) with launches of K sweeps;
// returns the buffer index (0: du0/dv0, 1: du1/dv1) holding the result.
int run_jacobi(int K, int R, float* du[2], float* dv[2], const float* Ix, const float* Iy,
               const float* Iz, int w, int h, int s, float alpha, int n, bool zero_first,
               int* result, hipStream_t st)
{
    int cur = 0;
    bool zero = zero_first;
    int done = 0;
    while (done < n) {
        const int kk = (n - done) < K ? (n - done) : K;
        OF_TRY(launch_jacobi(K, R, du[cur], dv[cur], Ix, Iy, Iz, w, h, s, alpha, kk, zero ? 1 : 0,
                             du[cur ^ 1], dv[cur ^ 1], st));
        zero = false;
        cur ^= 1;
        done += kk;
    }
    *result = cur;
    return OF_OK;
}

}  // namespace
}  // namespace amdflow

using namespace amdflow;

struct of_hs {
    int nLevels = 0, stride = 0, W = 0, H = 0;
    hipStream_t stream = nullptr;
    float** pI0 = nullptr;
    float** pI1 = nullptr;
    int pW[32]{}, pH[32]{}, pS[32]{};
    float *u = nullptr, *v = nullptr, *nu = nullptr, *nv = nullptr;
    float *du[2]{}, *dv[2]{};
    float *Ix = nullptr, *Iy = nullptr, *Iz = nullptr, *tmp = nullptr;
    float *hin0 = nullptr, *hin1 = nullptr;   // device staging of host inputs
    int K = 4, R = 4;                         // Jacobi launch shape
};

namespace {

void hs_release(of_hs* h)
{
    if (!h) return;
    for (int l = 0; l < h->nLevels; l++) {
        if (h->pI0) dev_free(h->pI0[l]);
        if (h->pI1) dev_free(h->pI1[l]);
    }
    delete[] h->pI0;
    delete[] h->pI1;
    dev_free(h->u); dev_free(h->v); dev_free(h->nu); dev_free(h->nv);
    dev_free(h->du[0]); dev_free(h->du[1]); dev_free(h->dv[0]); dev_free(h->dv[1]);
    dev_free(h->Ix); dev_free(h->Iy); dev_free(h->Iz); dev_free(h->tmp);
    dev_free(h->hin0); dev_free(h->hin1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

// ComputeFlow, src/hs/cuda/flow.cu:112-208, from device frames (stride s).
int hs_solve(of_hs* h, const float* dI0, const float* dI1, int width, int height, int stride,
             float alpha, int nLevels, int nWarpIters, int nSolverIters, float* du_out,
             float* dv_out, hipStream_t st)
{
    if (width != h->W || height != h->H || stride != h->stride || nLevels != h->nLevels)
        return OF_ERR_ARG;
    if (nWarpIters < 0 || nSolverIters < 0) return OF_ERR_ARG;
    const size_t fsize = (size_t)stride * height;
    int L = nLevels - 1;
    OF_HIP(hipMemcpyAsync(h->pI0[L], dI0, fsize * sizeof(float), hipMemcpyDeviceToDevice, st));
    OF_HIP(hipMemcpyAsync(h->pI1[L], dI1, fsize * sizeof(float), hipMemcpyDeviceToDevice, st));
    for (; L > 0; --L) {
        hipLaunchKernelGGL(k_hs_downscale, grid2d(h->pW[L - 1], h->pH[L - 1]), dim3(256), 0, st,
                           h->pI0[L], h->pS[L], h->pI0[L - 1], h->pW[L - 1], h->pH[L - 1], h->pS[L - 1]);
        HS_CHECK("k_hs_downscale");
        hipLaunchKernelGGL(k_hs_downscale, grid2d(h->pW[L - 1], h->pH[L - 1]), dim3(256), 0, st,
                           h->pI1[L], h->pS[L], h->pI1[L - 1], h->pW[L - 1], h->pH[L - 1], h->pS[L - 1]);
        HS_CHECK("k_hs_downscale");
    }
    hipLaunchKernelGGL(k_hs_zero, dim3(cdiv(fsize, 256) > 2048 ? 2048 : cdiv(fsize, 256)), dim3(256),
                       0, st, h->u, h->v, fsize);
    HS_CHECK("k_hs_zero");
    for (L = 0; L < nLevels; ++L) {
        const int w = h->pW[L], hh = h->pH[L], s = h->pS[L];
        for (int wi = 0; wi < nWarpIters; ++wi) {
            hipLaunchKernelGGL(k_hs_warp, grid2d(w, hh), dim3(256), 0, st, h->pI1[L], w, hh, s, h->u,
                               h->v, h->tmp);
            HS_CHECK("k_hs_warp");
            hipLaunchKernelGGL(k_hs_derivatives, grid2d(w, hh), dim3(256), 0, st, h->pI0[L], h->tmp,
                               w, hh, s, h->Ix, h->Iy, h->Iz);
            HS_CHECK("k_hs_derivatives");
            int res = 0;
            if (nSolverIters > 0) {
                OF_TRY(run_jacobi(h->K, h->R, h->du, h->dv, h->Ix, h->Iy, h->Iz, w, hh, s, alpha,
                                  nSolverIters, true, &res, st));
                hipLaunchKernelGGL(k_hs_add, grid2d(w, hh), dim3(256), 0, st, h->u, h->du[res],
                                   h->v, h->dv[res], w, hh, s);
                HS_CHECK("k_hs_add");
            }
        }
        if (L != nLevels - 1) {
            const float scaleX = (float)h->pW[L + 1] / (float)h->pW[L];
            const float scaleY = (float)h->pH[L + 1] / (float)h->pH[L];
            hipLaunchKernelGGL(k_hs_upscale, grid2d(h->pW[L + 1], h->pH[L + 1]), dim3(256), 0, st,
                               h->u, w, hh, s, h->nu, h->pW[L + 1], h->pH[L + 1], h->pS[L + 1], scaleX);
            HS_CHECK("k_hs_upscale");
            hipLaunchKernelGGL(k_hs_upscale, grid2d(h->pW[L + 1], h->pH[L + 1]), dim3(256), 0, st,
                               h->v, w, hh, s, h->nv, h->pW[L + 1], h->pH[L + 1], h->pS[L + 1], scaleY);
            HS_CHECK("k_hs_upscale");
            float* t = h->u; h->u = h->nu; h->nu = t;
            t = h->v; h->v = h->nv; h->nv = t;
        }
    }
    // the result is in h->u/h->v, which the level loop swaps with nu/nv
    if (du_out) OF_HIP(hipMemcpyAsync(du_out, h->u, fsize * sizeof(float), hipMemcpyDeviceToDevice, st));
    if (dv_out) OF_HIP(hipMemcpyAsync(dv_out, h->v, fsize * sizeof(float), hipMemcpyDeviceToDevice, st));
    OF_HIP(hipStreamSynchronize(st));
    return OF_OK;
}

}  // namespace

extern "C" {

int of_hs_create(int nLevels, int stride, int width, int height, of_hs** out)
{
    if (!out) return OF_ERR_ARG;
    *out = nullptr;
    if (nLevels < 1 || nLevels > 32 || width < 1 || height < 1 || stride < width) return OF_ERR_ARG;
    if ((long long)stride * height >= (1ll << 31)) return OF_ERR_ARG;
    of_hs* h = new (std::nothrow) of_hs();
    if (!h) return OF_ERR_ALLOC;
    h->nLevels = nLevels;
    h->stride = stride;
    h->W = width;
    h->H = height;
    const int top = nLevels - 1;
    h->pW[top] = width; h->pH[top] = height; h->pS[top] = stride;
    for (int L = top; L > 0; --L) {
        h->pW[L - 1] = h->pW[L] / 2;
        h->pH[L - 1] = h->pH[L] / 2;
        h->pS[L - 1] = ((h->pW[L - 1] + 31) / 32) * 32;   // iAlignUp, common.cuh:47-55
        if (h->pW[L - 1] < 1 || h->pH[L - 1] < 1) { delete h; return OF_ERR_ARG; }
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        h->stream = nullptr;
        hs_release(h);
        return OF_ERR_HIP;
    }
    h->pI0 = new (std::nothrow) float*[nLevels]();
    h->pI1 = new (std::nothrow) float*[nLevels]();
    if (!h->pI0 || !h->pI1) { hs_release(h); return OF_ERR_ALLOC; }
    const size_t fsize = (size_t)stride * height;
    int st = OF_OK;
    for (int L = 0; L < nLevels && !st; L++) {
        const size_t n = (size_t)h->pS[L] * h->pH[L];
        st = dev_alloc(&h->pI0[L], n);
        if (!st) st = dev_alloc(&h->pI1[L], n);
    }
    float** bufs[] = {&h->u, &h->v, &h->nu, &h->nv, &h->du[0], &h->du[1], &h->dv[0], &h->dv[1],
                      &h->Ix, &h->Iy, &h->Iz, &h->tmp, &h->hin0, &h->hin1};
    for (float** b : bufs)
        if (!st) st = dev_alloc(b, fsize);
    if (st) { hs_release(h); return st; }
    // padding columns of strided buffers are never read by the kernels;
    // clear everything once so downloads are deterministic
    for (float** b : bufs)
        if (hipMemsetAsync(*b, 0, fsize * sizeof(float), h->stream) != hipSuccess) st = OF_ERR_HIP;
    for (int L = 0; L < nLevels; L++) {
        const size_t n = (size_t)h->pS[L] * h->pH[L];
        if (hipMemsetAsync(h->pI0[L], 0, n * sizeof(float), h->stream) != hipSuccess) st = OF_ERR_HIP;
        if (hipMemsetAsync(h->pI1[L], 0, n * sizeof(float), h->stream) != hipSuccess) st = OF_ERR_HIP;
    }
    if (!st && hipStreamSynchronize(h->stream) != hipSuccess) st = OF_ERR_HIP;
    if (st) { hs_release(h); return st; }
    *out = h;
    return OF_OK;
}

void of_hs_destroy(of_hs* h) { hs_release(h); }

int of_hs_set_jacobi_shape(of_hs* h, int sweeps_per_launch, int rows_per_wave)
{
    if (!h) return OF_ERR_ARG;
    if (!jacobi_shape_ok(sweeps_per_launch, rows_per_wave)) return OF_ERR_UNSUPPORTED;
    h->K = sweeps_per_launch;
    h->R = rows_per_wave;
    return OF_OK;
}

int of_hs_compute_device(of_hs* h, const float* dI0, const float* dI1, int width, int height,
                         int stride, float alpha, int nLevels, int nWarpIters, int nSolverIters,
                         float* du, float* dv, void* stream)
{
    if (!h || !dI0 || !dI1 || !du || !dv) return OF_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;   // NULL: the HIP null stream
    return hs_solve(h, dI0, dI1, width, height, stride, alpha, nLevels, nWarpIters, nSolverIters,
                    du, dv, st);
}

int of_hs_compute(of_hs* h, const float* I0, const float* I1, int width, int height, int stride,
                  float alpha, int nLevels, int nWarpIters, int nSolverIters, float* u, float* v)
{
    if (!h || !I0 || !I1 || !u || !v) return OF_ERR_ARG;
    if (width != h->W || height != h->H || stride != h->stride) return OF_ERR_ARG;
    const size_t fsize = (size_t)stride * height;
    OF_HIP(hipMemcpyAsync(h->hin0, I0, fsize * sizeof(float), hipMemcpyHostToDevice, h->stream));
    OF_HIP(hipMemcpyAsync(h->hin1, I1, fsize * sizeof(float), hipMemcpyHostToDevice, h->stream));
    OF_TRY(hs_solve(h, h->hin0, h->hin1, width, height, stride, alpha, nLevels, nWarpIters,
                    nSolverIters, nullptr, nullptr, h->stream));
    // read the buffers after the solve: every level but the last swaps u
    // with nu (and v with nv), so which one holds the flow depends on the
    // level count
    OF_HIP(hipMemcpyAsync(u, h->u, fsize * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    OF_HIP(hipMemcpyAsync(v, h->v, fsize * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    OF_HIP(hipStreamSynchronize(h->stream));
    return OF_OK;
}

// ---- per-kernel test entry points (amdflow_debug.h) ----

int of_dbg_hs_jacobi(const float* du, const float* dv, const float* Ix, const float* Iy,
                     const float* Iz, int w, int h, float alpha, int iters, int K, int R,
                     float* du_out, float* dv_out)
{
    if (!du || !dv || !Ix || !Iy || !Iz || !du_out || !dv_out || w < 1 || h < 1 || iters < 0)
        return OF_ERR_ARG;
    if (!jacobi_shape_ok(K, R)) return OF_ERR_UNSUPPORTED;
    const size_t n = (size_t)w * h;
    float* b[7] = {};
    int st = OF_OK;
    for (float*& p : b)
        if (!st) st = dev_alloc(&p, n);
    if (!st) {
        const float* src[5] = {du, dv, Ix, Iy, Iz};
        for (int i = 0; i < 5 && !st; i++)
            if (hipMemcpy(b[i], src[i], n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
                st = OF_ERR_HIP;
    }
    int res = 0;
    float* dus[2] = {b[0], b[5]};
    float* dvs[2] = {b[1], b[6]};
    if (!st) st = run_jacobi(K, R, dus, dvs, b[2], b[3], b[4], w, h, w, alpha, iters, false, &res, nullptr);
    if (!st && hipMemcpy(du_out, dus[res], n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) st = OF_ERR_HIP;
    if (!st && hipMemcpy(dv_out, dvs[res], n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) st = OF_ERR_HIP;
    for (float*& p : b) dev_free(p);
    return st;
}

int of_dbg_hs_warp_derivatives(const float* I0, const float* I1, const float* u, const float* v,
                               int w, int h, float* I1w, float* Ix, float* Iy, float* Iz)
{
    if (!I0 || !I1 || !u || !v || !I1w || !Ix || !Iy || !Iz || w < 1 || h < 1) return OF_ERR_ARG;
    const size_t n = (size_t)w * h;
    float* b[8] = {};
    int st = OF_OK;
    for (float*& p : b)
        if (!st) st = dev_alloc(&p, n);
    const float* src[4] = {I0, I1, u, v};
    for (int i = 0; i < 4 && !st; i++)
        if (hipMemcpy(b[i], src[i], n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) st = OF_ERR_HIP;
    if (!st) {
        hipLaunchKernelGGL(k_hs_warp, grid2d(w, h), dim3(256), 0, nullptr, b[1], w, h, w, b[2], b[3], b[4]);
        hipLaunchKernelGGL(k_hs_derivatives, grid2d(w, h), dim3(256), 0, nullptr, b[0], b[4], w, h, w,
                           b[5], b[6], b[7]);
        if (hipGetLastError() != hipSuccess) st = OF_ERR_HIP;
    }
    float* dst[4] = {I1w, Ix, Iy, Iz};
    for (int i = 0; i < 4 && !st; i++)
        if (hipMemcpy(dst[i], b[4 + i], n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) st = OF_ERR_HIP;
    for (float*& p : b) dev_free(p);
    return st;
}

// Runs one pyramid step (down: 2x2 box; up: x2 flow upscale times `scale`)
// on a dense w x h plane into a dense nw x nh plane.
static int dbg_hs_resample(const float* src, int w, int h, float* out, int nw, int nh, bool up,
                           float scale)
{
    if (!src || !out || w < 1 || h < 1 || nw < 1 || nh < 1) return OF_ERR_ARG;
    // the box reads rows/columns 2i and 2i+1 of the source
    if (!up && (2 * (long long)nw > w || 2 * (long long)nh > h)) return OF_ERR_ARG;
    const size_t n = (size_t)w * h, nn = (size_t)nw * nh;
    float *ds = nullptr, *dd = nullptr;
    int st = dev_alloc(&ds, n);
    if (!st) st = dev_alloc(&dd, nn);
    if (!st && hipMemcpy(ds, src, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) st = OF_ERR_HIP;
    if (!st) {
        if (up)
            hipLaunchKernelGGL(k_hs_upscale, grid2d(nw, nh), dim3(256), 0, nullptr, ds, w, h, w, dd,
                               nw, nh, nw, scale);
        else
            hipLaunchKernelGGL(k_hs_downscale, grid2d(nw, nh), dim3(256), 0, nullptr, ds, w, dd, nw,
                               nh, nw);
        if (hipGetLastError() != hipSuccess) st = OF_ERR_HIP;
    }
    if (!st && hipMemcpy(out, dd, nn * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) st = OF_ERR_HIP;
    dev_free(ds);
    dev_free(dd);
    return st;
}

int of_dbg_hs_downscale(const float* src, int w, int h, float* out, int nw, int nh)
{
    return dbg_hs_resample(src, w, h, out, nw, nh, false, 1.0f);
}

int of_dbg_hs_upscale(const float* src, int w, int h, float* out, int nw, int nh, float scale)
{
    return dbg_hs_resample(src, w, h, out, nw, nh, true, scale);
}

}  // extern "C"
