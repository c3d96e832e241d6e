// synth.hip -- seeded synthetic frame pairs of known motion, on the GPU.
//
// The same definition as amdflow/synth.py (a different random stream):
//   frame0 = U(0,1) noise, Gaussian blur sigma 2, rescaled to [0, 255];
//   flow   = sum of 4 sinusoids per component (+ Gaussian bumps), scaled so
//            that max |flow| equals the cap;
//   frame1 = frame0 sampled bilinearly (clamp to edge) at x - flow(x),
//            plus N(0, noise) noise;
//   both frames rounded and clamped to u8.
// The random parameters of the flow are drawn by the caller and passed in.
#include <hip/hip_runtime.h>

#include <cmath>

#include "amdflow.h"
#include "common.hpp"
#include "tvl1_launch.hpp"

namespace amdflow {
namespace {

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    // splitmix64 finaliser
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float unit(unsigned long long seed, unsigned long long i)
{
    return (float)(mix64(seed * 0x100000001B3ull + i) >> 40) * (1.0f / 16777216.0f);
}

__global__ __launch_bounds__(256) void k_noise(float* __restrict__ out, size_t n,
                                               unsigned long long seed)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = unit(seed, i);
}

// min / max / max-norm reductions into red[0..2] (float bits as ordered ints)
__device__ __forceinline__ int ordered(float f)
{
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float unordered(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__global__ void k_red_init(int* red)
{
    red[0] = ordered(INFINITY);
    red[1] = ordered(-INFINITY);
    red[2] = ordered(0.0f);
}

__global__ __launch_bounds__(256) void k_minmax(const float* __restrict__ a, size_t n, int* red)
{
    float lo = INFINITY, hi = -INFINITY;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        lo = fminf(lo, a[i]);
        hi = fmaxf(hi, a[i]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off, 64));
        hi = fmaxf(hi, __shfl_xor(hi, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&red[0], ordered(lo));
        atomicMax(&red[1], ordered(hi));
    }
}

__global__ __launch_bounds__(256) void k_rescale(float* __restrict__ a, size_t n, const int* red)
{
    const float lo = unordered(red[0]), hi = unordered(red[1]);
    const float s = hi > lo ? 255.0f / (hi - lo) : 0.0f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        a[i] = (a[i] - lo) * s;
}

struct FlowParams {
    float period[2][4], cosang[2][4], sinang[2][4], phase[2][4], amp[2][4];
    int nbumps;
    float bx[8], by[8], br[8], bu[8], bv[8];
};

__global__ __launch_bounds__(256) void k_flow(float* __restrict__ fu, float* __restrict__ fv, int w,
                                              int h, FlowParams p, int* red)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    float m = 0.0f;
    if (x < w && y < h) {
        float c[2] = {0.0f, 0.0f};
        for (int k = 0; k < 2; k++)
            for (int j = 0; j < 4; j++)
                c[k] += p.amp[k][j] * sinf(6.2831853f * (p.cosang[k][j] * x + p.sinang[k][j] * y) /
                                               p.period[k][j] + p.phase[k][j]);
        for (int b = 0; b < p.nbumps; b++) {
            const float dx = x - p.bx[b], dy = y - p.by[b];
            const float g = expf(-(dx * dx + dy * dy) / (2.0f * p.br[b] * p.br[b]));
            c[0] += p.bu[b] * g;
            c[1] += p.bv[b] * g;
        }
        fu[(size_t)y * w + x] = c[0];
        fv[(size_t)y * w + x] = c[1];
        m = sqrtf(c[0] * c[0] + c[1] * c[1]);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&red[2], ordered(m));
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// frame1 = frame0(x - s*f) + noise; both frames to u8; the scaled flow is
// written back (ground truth).
__global__ __launch_bounds__(256) void k_warp_quantize(const float* __restrict__ f0,
                                                       float* __restrict__ fu,
                                                       float* __restrict__ fv, int w, int h,
                                                       float cap, float noise, const int* red,
                                                       unsigned long long seed,
                                                       uint8_t* __restrict__ q0,
                                                       uint8_t* __restrict__ q1)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const float mx = unordered(red[2]);
    const float s = mx > 0.0f ? cap / mx : 0.0f;
    const size_t p = (size_t)y * w + x;
    const float u = fu[p] * s, v = fv[p] * s;
    fu[p] = u;
    fv[p] = v;
    const float sx = x - u, sy = y - v;
    const float x0f = floorf(sx), y0f = floorf(sy);
    const float ax = sx - x0f, ay = sy - y0f;
    const int xa = clampi((int)x0f, 0, w - 1), xb = clampi((int)x0f + 1, 0, w - 1);
    const int ya = clampi((int)y0f, 0, h - 1), yb = clampi((int)y0f + 1, 0, h - 1);
    const float top = (1.0f - ax) * f0[(size_t)ya * w + xa] + ax * f0[(size_t)ya * w + xb];
    const float bot = (1.0f - ax) * f0[(size_t)yb * w + xa] + ax * f0[(size_t)yb * w + xb];
    float v1 = (1.0f - ay) * top + ay * bot;
    // Box-Muller
    const float r1 = fmaxf(unit(seed ^ 0xABCDEFull, 2 * p), 1e-7f), r2 = unit(seed ^ 0xABCDEFull, 2 * p + 1);
    v1 += noise * sqrtf(-2.0f * logf(r1)) * cosf(6.2831853f * r2);
    q0[p] = (uint8_t)fminf(fmaxf(rintf(f0[p]), 0.0f), 255.0f);
    q1[p] = (uint8_t)fminf(fmaxf(rintf(v1), 0.0f), 255.0f);
}

unsigned g1(size_t n) { const unsigned g = cdiv(n, 256); return g > 2048 ? 2048 : (g ? g : 1); }

}  // namespace
}  // namespace amdflow

using namespace amdflow;

extern "C" int of_synth_pair(int width, int height, float cap, float noise, unsigned long long seed,
                             const float* params, int nbumps, uint8_t* dI0, uint8_t* dI1,
                             float* dflow, float* dscratch, void* stream)
{
    // params: 2x4 (period, angle, phase, amp) then nbumps x (cx, cy, r, du, dv)
    if (width < 8 || height < 8 || !params || !dI0 || !dI1 || !dflow || !dscratch) return OF_ERR_ARG;
    if (nbumps < 0 || nbumps > 8) return OF_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    FlowParams p{};
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 4; j++) {
            const float* q = params + (k * 4 + j) * 4;
            p.period[k][j] = q[0];
            p.cosang[k][j] = std::cos(q[1]);
            p.sinang[k][j] = std::sin(q[1]);
            p.phase[k][j] = q[2];
            p.amp[k][j] = q[3];
        }
    p.nbumps = nbumps;
    for (int b = 0; b < nbumps; b++) {
        const float* q = params + 32 + b * 5;
        p.bx[b] = q[0]; p.by[b] = q[1]; p.br[b] = q[2]; p.bu[b] = q[3]; p.bv[b] = q[4];
    }
    const size_t n = (size_t)width * height;
    float* f0 = dscratch;            // n
    float* tmp = dscratch + n;       // n
    int* red = (int*)(dscratch + 2 * n);
    GaussTaps g;
    OF_TRY(gaussian_taps(2.0, &g));
    if (g.size > width || g.size > height) return OF_ERR_SIGMA;
    hipLaunchKernelGGL(k_noise, dim3(g1(n)), dim3(256), 0, st, tmp, n, seed);
    OF_TRY(launch_gaussian(tmp, f0, tmp, width, height, g, st));
    // blurred noise is in tmp; rescale to [0,255] into f0
    hipLaunchKernelGGL(k_red_init, dim3(1), dim3(1), 0, st, red);
    hipLaunchKernelGGL(k_minmax, dim3(g1(n)), dim3(256), 0, st, tmp, n, red);
    OF_HIP(hipMemcpyAsync(f0, tmp, n * sizeof(float), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_rescale, dim3(g1(n)), dim3(256), 0, st, f0, n, red);
    const dim3 grid(cdiv(width, 64), cdiv(height, 4));
    hipLaunchKernelGGL(k_flow, grid, dim3(256), 0, st, dflow, dflow + n, width, height, p, red);
    hipLaunchKernelGGL(k_warp_quantize, grid, dim3(256), 0, st, f0, dflow, dflow + n, width, height,
                       cap, noise, red, seed, dI0, dI1);
    OF_HIP(hipGetLastError());
    return OF_OK;
}
