// cnn.hip -- the learned warm start of TV-L1 (no reference counterpart).
//
// A small fully convolutional network predicts the flow at the coarsest
// pyramid scale from that scale's I0, I1, Ix, Iy and It; the solver then
// starts from it instead of zero (the reference starts from zero,
// src/tvl1/sequential/tvl1.cpp:157-158).  Layers: conv3x3 5->C + ReLU,
// `hidden` x (conv3x3 C->C + ReLU), conv3x3 C->2; zero padding.
//
// Every convolution is an implicit GEMM on the bf16 MFMA
// (v_mfma_f32_16x16x32_bf16, fp32 accumulation), with M = pixels:
//   forward        Y[m][n]  = sum_(t,c) X[m + s(t)][c] W[n][t][c]
//   data gradient  dX[m][c] = sum_(t,n) dY[m - s(t)][n] W[n][t][c]
//   weight grad.   dW[n][t][c] = sum_m dY[m][n] X[m + s(t)][c]
// with tap offsets s(t) = (t/3 - 1, t%3 - 1).  Activations are NHWC fp32 in
// memory and rounded to bf16 (nearest even) as they enter an MFMA operand;
// the master weights are fp32 (Adam) with bf16 copies packed for the two
// GEMM orientations.  Every reduction runs in a fixed order, so a training
// run is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "amdflow.h"
#include "amdflow_debug.h"
#include "common.hpp"

namespace amdflow {
namespace {

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ short to_bf16(float f)
{
    unsigned u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (short)(u >> 16);   // inf / nan: truncate
    u += 0x7FFFu + ((u >> 16) & 1u);                                   // nearest even
    return (short)(u >> 16);
}

// gfx950 16x16x32 MFMA: lane l supplies row (l % 16) of A and column
// (l % 16) of B for k = 8*(l/16) .. +7, and receives column (l % 16), rows
// 4*(l/16) .. +3 of the 16x16 product (of_dbg_cnn_mfma checks this through
// k_conv).
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// fp32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ bf16x8 to_bf16x8(f32x8 v) { return __builtin_convertvector(v, bf16x8); }

struct Pix {
    int b, y, x;
};

__device__ __forceinline__ Pix pix_of(long long m, int H, int W)
{
    Pix p;
    const int hw = H * W;
    const int mi = (int)m;   // pixel counts stay below 2^31
    p.b = mi / hw;
    const int r = mi - p.b * hw;
    p.y = r / W;
    p.x = r - p.y * W;
    return p;
}

__device__ __forceinline__ void pix_next(Pix& p, int H, int W)
{
    if (++p.x == W) {
        p.x = 0;
        if (++p.y == H) {
            p.y = 0;
            p.b++;
        }
    }
}

// Eight consecutive GEMM-k entries k .. k+7 of pixel p's im2col row
// (k = t*CX + c), gathered from an NHWC tensor with CX channels; `sgn` = -1
// mirrors the tap offsets (data gradient).
template <int CX>
__device__ __forceinline__ bf16x8 gather8(const float* __restrict__ X, const Pix& p, bool valid,
                                          int k, int K, int H, int W, int sgn)
{
    f32x8 r;
    if constexpr (CX % 8 == 0) {
        // the eight entries share a tap: two 16-byte loads
        const int t = k / CX, c = k - t * CX;
        const int yy = p.y + sgn * (t / 3 - 1), xx = p.x + sgn * (t % 3 - 1);
        float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0;
        if (valid && k < K && yy >= 0 && yy < H && xx >= 0 && xx < W) {
            const float4* q = reinterpret_cast<const float4*>(X + (((long long)p.b * H + yy) * W + xx) * CX + c);
            v0 = q[0];
            v1 = q[1];
        }
        r[0] = v0.x; r[1] = v0.y; r[2] = v0.z; r[3] = v0.w;
        r[4] = v1.x; r[5] = v1.y; r[6] = v1.z; r[7] = v1.w;
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int kk = k + j;
            const int t = kk / CX, c = kk - t * CX;
            const int yy = p.y + sgn * (t / 3 - 1), xx = p.x + sgn * (t % 3 - 1);
            float v = 0.0f;
            if (valid && kk < K && yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = X[(((long long)p.b * H + yy) * W + xx) * CX + c];
            r[j] = v;
        }
    }
    return to_bf16x8(r);
}

// Implicit-GEMM 3x3 convolution.  A block is 4 waves; a wave owns 16
// pixels and NT 16-wide tiles of output channels.  Wp: bf16 [NT*16][Kpad]
// (Kpad a multiple of 32), row n = output channel, zero beyond K and N.
// Epilogue: + bias, ReLU, and/or zero where mask[m][n] <= 0 (the ReLU
// derivative of the layer below).
template <int CX, int NT>
__global__ __launch_bounds__(256) void k_conv(const float* __restrict__ X, int B, int H, int W,
                                              const short* __restrict__ Wp, int Kpad, int N,
                                              const float* __restrict__ bias, int sgn, int relu,
                                              const float* __restrict__ mask,
                                              float* __restrict__ Y)
{
    const int K = 9 * CX;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long M = (long long)B * H * W;
    const long long m0 = ((long long)blockIdx.x * 4 + wave) * 16;
    if (m0 >= M) return;
    const int row = lane & 15, kg = lane >> 4;
    const long long ma = m0 + row;
    const bool valid = ma < M;
    const Pix p = pix_of(valid ? ma : 0, H, W);
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    constexpr int KP = (9 * CX + 31) / 32 * 32;   // == Kpad (checked by the launcher)
#pragma unroll
    for (int k0 = 0; k0 < KP; k0 += 32) {
        const bf16x8 a = gather8<CX>(X, p, valid, k0 + 8 * kg, K, H, W, sgn);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const bf16x8 b = __builtin_bit_cast(
                bf16x8, *reinterpret_cast<const s16x8*>(Wp + (size_t)(t * 16 + row) * KP + k0 + 8 * kg));
            acc[t] = mfma32(a, b, acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int n = t * 16 + row;
        if (n >= N) continue;
        const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long long m = m0 + 4 * kg + j;
            if (m >= M) continue;
            float v = acc[t][j] + bn;
            if (relu) v = v > 0.0f ? v : 0.0f;
            if (mask && !(mask[m * N + n] > 0.0f)) v = 0.0f;
            Y[m * N + n] = v;
        }
    }
}

// The same convolution for CX % 32 == 0 with the operands staged in LDS:
// a block computes an 8x8-pixel tile of one image (waves take 2 rows each)
// from a 10x10xCX bf16 input patch and the bf16 weights, both loaded once,
// so each input pixel and each weight leaves HBM / L2 once per block
// instead of once per tap and per wave.  MFMA operands come from LDS with
// 16-byte reads; accumulation, epilogue and results equal k_conv's.
template <int CX, int NT>
__global__ __launch_bounds__(256) void k_conv_lds(const float* __restrict__ X, int B, int H, int W,
                                                  const short* __restrict__ Wp, int N,
                                                  const float* __restrict__ bias, int sgn, int relu,
                                                  const float* __restrict__ mask,
                                                  float* __restrict__ Y)
{
    static_assert(CX % 32 == 0, "one 32-wide k-step must stay inside one tap");
    constexpr int KP = 9 * CX;           // a multiple of 32
    constexpr int PW = 10, PP = PW * PW;
    __shared__ __align__(16) short patch[PP * CX];
    __shared__ __align__(16) short wts[NT * 16 * KP];
    const int tiles_x = (W + 7) / 8, tiles_y = (H + 7) / 8;
    const int tile = blockIdx.x;
    const int b = tile / (tiles_x * tiles_y);
    const int r = tile - b * tiles_x * tiles_y;
    const int y0 = (r / tiles_x) * 8, x0 = (r % tiles_x) * 8;
    // input patch rows y0-1 .. y0+8, columns x0-1 .. x0+8, zero outside
    for (int e = threadIdx.x; e < PP * (CX / 8); e += 256) {
        const int pix = e / (CX / 8), c8 = (e - pix * (CX / 8)) * 8;
        const int yy = y0 - 1 + pix / PW, xx = x0 - 1 + pix % PW;
        f32x8 v;
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = 0.0f;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
            const float4* q = reinterpret_cast<const float4*>(X + (((long long)b * H + yy) * W + xx) * CX + c8);
            const float4 v0 = q[0], v1 = q[1];
            v[0] = v0.x; v[1] = v0.y; v[2] = v0.z; v[3] = v0.w;
            v[4] = v1.x; v[5] = v1.y; v[6] = v1.z; v[7] = v1.w;
        }
        *reinterpret_cast<bf16x8*>(patch + pix * CX + c8) = to_bf16x8(v);
    }
    for (int e = threadIdx.x; e < NT * 16 * KP / 8; e += 256)
        reinterpret_cast<s16x8*>(wts)[e] = reinterpret_cast<const s16x8*>(Wp)[e];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const int py = wave * 2 + (row >> 3), px = row & 7;   // this lane's A pixel in the tile
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < KP; k0 += 32) {
        const int tap = k0 / CX, c = k0 - tap * CX + 8 * kg;
        const int dy = sgn * (tap / 3 - 1), dx = sgn * (tap % 3 - 1);
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(patch + ((py + 1 + dy) * PW + (px + 1 + dx)) * CX + c);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const bf16x8 bv = *reinterpret_cast<const bf16x8*>(wts + (t * 16 + row) * KP + k0 + 8 * kg);
            acc[t] = mfma32(a, bv, acc[t]);
        }
    }
    // accumulator: column n = t*16 + row, rows (pixels) 4*kg + j of the wave's 16
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int n = t * 16 + row;
        if (n >= N) continue;
        const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int q = 4 * kg + j;
            const int yy = y0 + wave * 2 + (q >> 3), xx = x0 + (q & 7);
            if (yy >= H || xx >= W) continue;
            const long long m = ((long long)b * H + yy) * W + xx;
            float v = acc[t][j] + bn;
            if (relu) v = v > 0.0f ? v : 0.0f;
            if (mask && !(mask[m * N + n] > 0.0f)) v = 0.0f;
            Y[m * N + n] = v;
        }
    }
}

// Weight-gradient partials.  Block (kt, nt, z) forms the 16x16 tile
// dW[16*nt ..][16*kt ..] over the pixels [z*chunk, (z+1)*chunk): the four
// waves take interleaved 32-pixel steps (A = dY^T, B = im2col(X), GEMM-k =
// pixels) and are then added in wave order.  part: [splits][Npad][Kpad].
template <int CX>
__global__ __launch_bounds__(256) void k_wgrad(const float* __restrict__ dY, int N,
                                               const float* __restrict__ X, int B, int H, int W,
                                               int Kpad, long long chunk, float* __restrict__ part)
{
    __shared__ float sh[4][16][17];
    const int K = 9 * CX;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
    const long long M = (long long)B * H * W;
    const long long p_begin = (long long)blockIdx.z * chunk;
    const long long p_end = p_begin + chunk < M ? p_begin + chunk : M;
    const int row = lane & 15, kg = lane >> 4;
    const int k = k0 + row;   // B column
    const bool kvalid = k < K;
    int dy = 0, dx = 0, c = 0;
    if (kvalid) {
        const int t = k / CX;
        c = k - t * CX;
        dy = t / 3 - 1;
        dx = t % 3 - 1;
    }
    const int n = n0 + row;   // A row
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long q0 = p_begin + wave * 32; q0 < p_end; q0 += 128) {
        f32x8 a, b;
        const long long qf = q0 + 8 * kg;
        Pix p = pix_of(qf < M ? qf : 0, H, W);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const long long q = qf + j;
            float va = 0.0f, vb = 0.0f;
            if (q < p_end) {
                if (n < N) va = dY[q * N + n];
                if (kvalid) {
                    const int y = p.y + dy, x = p.x + dx;
                    if (y >= 0 && y < H && x >= 0 && x < W)
                        vb = X[(((long long)p.b * H + y) * W + x) * CX + c];
                }
            }
            a[j] = va;
            b[j] = vb;
            pix_next(p, H, W);
        }
        acc = mfma32(to_bf16x8(a), to_bf16x8(b), acc);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) sh[wave][4 * kg + j][row] = acc[j];   // [n][k]
    __syncthreads();
    const int i = threadIdx.x >> 4, jj = threadIdx.x & 15;
    const float s = ((sh[0][i][jj] + sh[1][i][jj]) + sh[2][i][jj]) + sh[3][i][jj];
    const size_t Npad = (size_t)gridDim.y * 16;
    part[((size_t)blockIdx.z * Npad + (n0 + i)) * Kpad + (k0 + jj)] = s;
}

// dW[n][k] = sum over the splits in order (k < K, n < N).
__global__ __launch_bounds__(256) void k_wgrad_reduce(const float* __restrict__ part, int splits,
                                                      int Npad, int Kpad, int N, int K,
                                                      float* __restrict__ dW)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * K) return;
    const int n = idx / K, k = idx - n * K;
    float s = 0.0f;
    for (int sp = 0; sp < splits; sp++) s += part[((size_t)sp * Npad + n) * Kpad + k];
    dW[idx] = s;
}

// Bias-gradient partials: block z sums dY over its pixel chunk, channel by
// channel, in a fixed tree.
__global__ __launch_bounds__(256) void k_bias_partial(const float* __restrict__ dY, int N, long long M,
                                                      long long chunk, float* __restrict__ part)
{
    __shared__ float sh[256];
    const long long p0 = (long long)blockIdx.x * chunk;
    const long long p1 = p0 + chunk < M ? p0 + chunk : M;
    for (int n = 0; n < N; n++) {
        float s = 0.0f;
        for (long long q = p0 + threadIdx.x; q < p1; q += 256) s += dY[q * N + n];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * N + n] = sh[0];
        __syncthreads();
    }
}

__global__ void k_bias_reduce(const float* __restrict__ part, int splits, int N, float* __restrict__ db)
{
    const int n = threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int sp = 0; sp < splits; sp++) s += part[(size_t)sp * N + n];
    db[n] = s;
}

// Mean-squared flow error against planar targets T[b][2][H*W]:
// L = sum (Y - T)^2 / (2 M), dY = (Y - T) / M (M pixels); per-block double
// partial sums of the squared error.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                             long long M, int HW, float invM, float* __restrict__ dY,
                                             double* __restrict__ part)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < 2 * M; i += (long long)gridDim.x * 256) {
        const long long m = i >> 1;
        const int c = (int)(i & 1);
        const long long b = m / HW;
        const long long r = m - b * HW;
        const float d = Y[i] - T[(b * 2 + c) * HW + r];
        dY[i] = d * invM;
        s += (double)d * (double)d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum_f64(const double* __restrict__ part, int n, double scale, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; i++) s += part[i];
    *out = s * scale;
}

// Adam (Kingma & Ba 2015) with bias-correction factors c1 = 1 - b1^t,
// c2 = 1 - b2^t.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, int n,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 packs of one layer's weights Wm[n][t][c] (N outputs, C inputs):
//   fwd [Npad][Kpad]:  row n, k = t*C + c
//   bwd [Cpad][Kbpad]: row c, k = t*N + n
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ Wm, int N, int C, int Npad,
                                              int Kpad, int Cpad, int Kbpad, short* __restrict__ fwd,
                                              short* __restrict__ bwd)
{
    const int K = 9 * C;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < Npad * Kpad) {
        const int n = idx / Kpad, k = idx - n * Kpad;
        fwd[idx] = (n < N && k < K) ? to_bf16(Wm[n * K + k]) : (short)0;
    }
    if (idx < Cpad * Kbpad) {
        const int c = idx / Kbpad, kb = idx - c * Kbpad;
        short v = 0;
        if (c < C && kb < 9 * N) {
            const int t = kb / N, n = kb - t * N;
            v = to_bf16(Wm[(n * 9 + t) * C + c]);
        }
        bwd[idx] = v;
    }
}

// Network input: I0, I1, Ix, Iy, It of one scale, /255.  Ix, Iy are
// centred differences of I1 (x0.5, clamped neighbours at the border).
__global__ __launch_bounds__(256) void k_features(const float* __restrict__ I0,
                                                  const float* __restrict__ I1, int H, int W,
                                                  float* __restrict__ X)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= H * W) return;
    const int y = idx / W, x = idx - y * W;
    const int xl = x > 0 ? x - 1 : 0, xr = x < W - 1 ? x + 1 : W - 1;
    const int yu = y > 0 ? y - 1 : 0, yd = y < H - 1 ? y + 1 : H - 1;
    const float s = 1.0f / 255.0f;
    const float a = I0[idx], b = I1[idx];
    float* o = X + (size_t)idx * 5;
    o[0] = a * s;
    o[1] = b * s;
    o[2] = 0.5f * (I1[y * W + xr] - I1[y * W + xl]) * s;
    o[3] = 0.5f * (I1[yd * W + x] - I1[yu * W + x]) * s;
    o[4] = (b - a) * s;
}

// NHWC two-channel output -> planar [u1 | u2].
__global__ __launch_bounds__(256) void k_to_planar(const float* __restrict__ Y, int HW,
                                                   float* __restrict__ U)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= HW) return;
    U[i] = Y[2 * (size_t)i];
    U[HW + i] = Y[2 * (size_t)i + 1];
}

inline int pad16(int v) { return (v + 15) / 16 * 16; }
inline int pad32(int v) { return (v + 31) / 32 * 32; }
inline unsigned blocks256(long long n)
{
    const long long g = (n + 255) / 256;
    return (unsigned)std::max<long long>(1, std::min<long long>(g, 1 << 20));
}

#define CNN_CHECK(name)                                     \
    do {                                                    \
        hipError_t e_ = hipGetLastError();                  \
        if (e_ != hipSuccess) {                             \
            record_hip_error(e_, name, __FILE__, __LINE__); \
            return OF_ERR_HIP;                              \
        }                                                   \
    } while (0)

// Dispatch over the instantiated (input channels, output tiles) pairs.
int launch_conv(const float* X, int B, int H, int W, int CX, const short* Wp, int Kpad, int N,
                const float* bias, int sgn, int relu, const float* mask, float* Y, hipStream_t st)
{
    const long long M = (long long)B * H * W;
    const unsigned blocks = (unsigned)((M + 63) / 64);
    const int NT = N <= 16 ? 1 : 2;
    if (N > 32 || Kpad != (9 * CX + 31) / 32 * 32) return OF_ERR_UNSUPPORTED;
    if (CX == 32) {   // LDS-staged tiles of 8x8 pixels
        const unsigned tiles = (unsigned)(B * ((H + 7) / 8) * ((W + 7) / 8));
        if (NT == 2)
            hipLaunchKernelGGL((k_conv_lds<32, 2>), dim3(tiles), dim3(256), 0, st, X, B, H, W, Wp, N, bias,
                               sgn, relu, mask, Y);
        else
            hipLaunchKernelGGL((k_conv_lds<32, 1>), dim3(tiles), dim3(256), 0, st, X, B, H, W, Wp, N, bias,
                               sgn, relu, mask, Y);
        CNN_CHECK("k_conv_lds");
        return OF_OK;
    }
#define CONV_CASE(cx, nt)                                                                          \
    if (CX == cx && NT == nt) {                                                                    \
        hipLaunchKernelGGL((k_conv<cx, nt>), dim3(blocks), dim3(256), 0, st, X, B, H, W, Wp, Kpad, \
                           N, bias, sgn, relu, mask, Y);                                           \
        CNN_CHECK("k_conv");                                                                       \
        return OF_OK;                                                                              \
    }
    CONV_CASE(5, 2)
    CONV_CASE(32, 2)
    CONV_CASE(32, 1)
    CONV_CASE(2, 2)
    CONV_CASE(16, 1)
#undef CONV_CASE
    return OF_ERR_UNSUPPORTED;
}

int launch_wgrad(const float* dY, int N, const float* X, int CX, int B, int H, int W, int Kpad,
                 int Npad, int splits, long long chunk, float* part, hipStream_t st)
{
    const dim3 grid(Kpad / 16, Npad / 16, splits);
#define WG_CASE(cx)                                                                               \
    if (CX == cx) {                                                                               \
        hipLaunchKernelGGL((k_wgrad<cx>), grid, dim3(256), 0, st, dY, N, X, B, H, W, Kpad, chunk, \
                           part);                                                                 \
        CNN_CHECK("k_wgrad");                                                                     \
        return OF_OK;                                                                             \
    }
    WG_CASE(5)
    WG_CASE(32)
#undef WG_CASE
    return OF_ERR_UNSUPPORTED;
}

struct Layer {
    int cin = 0, cout = 0, relu = 0;
    size_t woff = 0, boff = 0;   // offsets into the parameter vector
    int Kpad = 0, Npad = 0, Cpad = 0, Kbpad = 0;
    short* fwd = nullptr;        // bf16 packs
    short* bwd = nullptr;
};

}  // namespace
}  // namespace amdflow

using namespace amdflow;

struct of_cnn {
    int hidden = 0, channels = 0;
    std::vector<Layer> layers;
    size_t nparams = 0;
    float* w = nullptr;   // fp32 master weights: per layer W[n][t][c], then bias[n]
    float* g = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    int step = 0;
    // batch workspace, grown on demand
    long long cap_px = 0;
    std::vector<float*> act;   // output of every layer, NHWC
    float* dA = nullptr;       // gradient ping-pong buffers
    float* dB = nullptr;
    float* part = nullptr;     // weight / bias gradient partials
    double* lpart = nullptr;   // loss partials
    double* loss = nullptr;
    // single-image inference workspace
    long long cap_in = 0;
    float* xin = nullptr;      // 5-channel features
    float* pred = nullptr;     // planar prediction
    // warm-start workspace: the prediction carried to the first solved scale, zero flow
    long long cap_lift = 0;
    float* lift = nullptr;
    float* zero = nullptr;
};

namespace {

constexpr int kLossBlocks = 1024;

int splits_for(long long M)
{
    return (int)std::max<long long>(1, std::min<long long>(256, (M + 4095) / 4096));
}

void free_ws(of_cnn* h)
{
    for (float*& p : h->act) dev_free(p);
    h->act.clear();
    dev_free(h->dA);
    dev_free(h->dB);
    dev_free(h->part);
    h->cap_px = 0;
}

int ensure_ws(of_cnn* h, long long px)
{
    if (px <= h->cap_px) return OF_OK;
    OF_HIP(hipDeviceSynchronize());   // queued work may still use the old buffers
    free_ws(h);
    int maxc = 5;
    size_t need = 1;
    const int splits = splits_for(px);
    for (const Layer& L : h->layers) {
        float* p = nullptr;
        OF_TRY(dev_alloc(&p, (size_t)px * L.cout));
        h->act.push_back(p);
        maxc = std::max(maxc, std::max(L.cin, L.cout));
        need = std::max(need, (size_t)splits * (size_t)std::max(L.Npad * L.Kpad, L.cout));
    }
    OF_TRY(dev_alloc(&h->dA, (size_t)px * maxc));
    OF_TRY(dev_alloc(&h->dB, (size_t)px * maxc));
    OF_TRY(dev_alloc(&h->part, need));
    h->cap_px = px;
    return OF_OK;
}

int ensure_in(of_cnn* h, long long n)
{
    if (n <= h->cap_in) return OF_OK;
    OF_HIP(hipDeviceSynchronize());
    dev_free(h->xin);
    dev_free(h->pred);
    h->cap_in = 0;
    OF_TRY(dev_alloc(&h->xin, (size_t)n * 5));
    OF_TRY(dev_alloc(&h->pred, (size_t)n * 2));
    h->cap_in = n;
    return OF_OK;
}

int pack(of_cnn* h, hipStream_t st)
{
    for (Layer& L : h->layers) {
        const int n = std::max(L.Npad * L.Kpad, L.Cpad * L.Kbpad);
        hipLaunchKernelGGL(k_pack, dim3(blocks256(n)), dim3(256), 0, st, h->w + L.woff, L.cout, L.cin,
                           L.Npad, L.Kpad, L.Cpad, L.Kbpad, L.fwd, L.bwd);
        CNN_CHECK("k_pack");
    }
    return OF_OK;
}

int forward(of_cnn* h, const float* X, int B, int H, int W, hipStream_t st)
{
    const float* in = X;
    for (size_t l = 0; l < h->layers.size(); l++) {
        const Layer& L = h->layers[l];
        OF_TRY(launch_conv(in, B, H, W, L.cin, L.fwd, L.Kpad, L.cout, h->w + L.boff, 1, L.relu,
                           nullptr, h->act[l], st));
        in = h->act[l];
    }
    return OF_OK;
}

// Backward pass from the loss gradient in h->dA; fills h->g.
int backward(of_cnn* h, const float* X, int B, int H, int W, hipStream_t st)
{
    const long long M = (long long)B * H * W;
    const int splits = splits_for(M);
    const long long chunk = (M + splits - 1) / splits;
    float* dcur = h->dA;
    float* dnext = h->dB;
    for (int l = (int)h->layers.size() - 1; l >= 0; l--) {
        const Layer& L = h->layers[l];
        const float* in = l ? h->act[l - 1] : X;
        const int K = 9 * L.cin;
        OF_TRY(launch_wgrad(dcur, L.cout, in, L.cin, B, H, W, L.Kpad, L.Npad, splits, chunk, h->part, st));
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks256(L.cout * K)), dim3(256), 0, st, h->part, splits,
                           L.Npad, L.Kpad, L.cout, K, h->g + L.woff);
        CNN_CHECK("k_wgrad_reduce");
        hipLaunchKernelGGL(k_bias_partial, dim3(splits), dim3(256), 0, st, dcur, L.cout, M, chunk, h->part);
        CNN_CHECK("k_bias_partial");
        hipLaunchKernelGGL(k_bias_reduce, dim3(1), dim3(64), 0, st, h->part, splits, L.cout, h->g + L.boff);
        CNN_CHECK("k_bias_reduce");
        if (l == 0) break;
        // gradient w.r.t. this layer's input, zero where the ReLU below was inactive
        const Layer& P = h->layers[l - 1];
        OF_TRY(launch_conv(dcur, B, H, W, L.cout, L.bwd, L.Kbpad, L.cin, nullptr, -1, 0,
                           P.relu ? h->act[l - 1] : nullptr, dnext, st));
        std::swap(dcur, dnext);
    }
    return OF_OK;
}

int loss_grad(of_cnn* h, const float* T, int B, int H, int W, hipStream_t st)
{
    const long long M = (long long)B * H * W;
    const unsigned blocks = std::min<unsigned>(kLossBlocks, blocks256(2 * M));
    hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, st, h->act.back(), T, M, H * W,
                       1.0f / (float)M, h->dA, h->lpart);
    CNN_CHECK("k_mse");
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(64), 0, st, h->lpart, (int)blocks, 0.5 / (double)M, h->loss);
    CNN_CHECK("k_sum_f64");
    return OF_OK;
}

void release(of_cnn* h)
{
    if (!h) return;
    free_ws(h);
    for (Layer& L : h->layers) {
        dev_free(L.fwd);
        dev_free(L.bwd);
    }
    dev_free(h->w);
    dev_free(h->g);
    dev_free(h->m);
    dev_free(h->v);
    dev_free(h->lpart);
    dev_free(h->loss);
    dev_free(h->xin);
    dev_free(h->pred);
    dev_free(h->lift);
    dev_free(h->zero);
    delete h;
}

bool dims_ok(const of_cnn* h, const void* a, const void* b, int B, int H, int W)
{
    return h && a && b && B >= 1 && H >= 1 && W >= 1 && (long long)B * H * W < (1ll << 31) / 64;
}

}  // namespace

extern "C" {

int of_cnn_create(int hidden_layers, int channels, unsigned long long seed, of_cnn** out)
{
    return guarded([&]() -> int {
        if (!out) return OF_ERR_ARG;
        *out = nullptr;
        if (hidden_layers < 0 || hidden_layers > 16 || channels != 32) return OF_ERR_ARG;
        of_cnn* h = new (std::nothrow) of_cnn();
        if (!h) return OF_ERR_ALLOC;
        h->hidden = hidden_layers;
        h->channels = channels;
        std::vector<std::pair<int, int>> dims;
        dims.push_back({5, channels});
        for (int i = 0; i < hidden_layers; i++) dims.push_back({channels, channels});
        dims.push_back({channels, 2});
        size_t off = 0;
        for (size_t i = 0; i < dims.size(); i++) {
            Layer L;
            L.cin = dims[i].first;
            L.cout = dims[i].second;
            L.relu = i + 1 < dims.size();
            L.woff = off;
            off += (size_t)L.cout * 9 * L.cin;
            L.boff = off;
            off += L.cout;
            L.Kpad = pad32(9 * L.cin);
            L.Npad = pad16(L.cout);
            L.Cpad = pad16(L.cin);
            L.Kbpad = pad32(9 * L.cout);
            h->layers.push_back(L);
        }
        h->nparams = off;
        // He-normal initial weights (the head scaled by 0.1), zero biases
        std::vector<float> w(off, 0.0f);
        std::mt19937_64 rng(seed);
        for (const Layer& L : h->layers) {
            const float sd = std::sqrt(2.0f / (9.0f * (float)L.cin)) * (L.relu ? 1.0f : 0.1f);
            std::normal_distribution<float> nd(0.0f, sd);
            for (size_t i = 0; i < (size_t)L.cout * 9 * L.cin; i++) w[L.woff + i] = nd(rng);
        }
        int st = dev_alloc(&h->w, off);
        if (!st) st = dev_alloc(&h->g, off);
        if (!st) st = dev_alloc(&h->m, off);
        if (!st) st = dev_alloc(&h->v, off);
        if (!st) st = dev_alloc(&h->lpart, kLossBlocks);
        if (!st) st = dev_alloc(&h->loss, 1);
        for (Layer& L : h->layers) {
            if (!st) st = dev_alloc(&L.fwd, (size_t)L.Npad * L.Kpad);
            if (!st) st = dev_alloc(&L.bwd, (size_t)L.Cpad * L.Kbpad);
        }
        if (!st && hipMemcpy(h->w, w.data(), off * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) st = OF_ERR_HIP;
        if (!st && hipMemset(h->g, 0, off * sizeof(float)) != hipSuccess) st = OF_ERR_HIP;
        if (!st && hipMemset(h->m, 0, off * sizeof(float)) != hipSuccess) st = OF_ERR_HIP;
        if (!st && hipMemset(h->v, 0, off * sizeof(float)) != hipSuccess) st = OF_ERR_HIP;
        if (!st) st = pack(h, nullptr);
        if (!st && hipDeviceSynchronize() != hipSuccess) st = OF_ERR_HIP;
        if (st) {
            (void)hipGetLastError();
            release(h);
            return st;
        }
        *out = h;
        return OF_OK;
    });
}

void of_cnn_destroy(of_cnn* h) { release(h); }

long long of_cnn_num_params(const of_cnn* h) { return h ? (long long)h->nparams : 0; }

int of_cnn_get_params(const of_cnn* h, float* host)
{
    if (!h || !host) return OF_ERR_ARG;
    OF_HIP(hipDeviceSynchronize());
    OF_HIP(hipMemcpy(host, h->w, h->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return OF_OK;
}

int of_cnn_set_params(of_cnn* h, const float* host)
{
    if (!h || !host) return OF_ERR_ARG;
    OF_HIP(hipDeviceSynchronize());
    OF_HIP(hipMemcpy(h->w, host, h->nparams * sizeof(float), hipMemcpyHostToDevice));
    OF_HIP(hipMemset(h->m, 0, h->nparams * sizeof(float)));
    OF_HIP(hipMemset(h->v, 0, h->nparams * sizeof(float)));
    h->step = 0;
    OF_TRY(pack(h, nullptr));
    OF_HIP(hipDeviceSynchronize());
    return OF_OK;
}

int of_cnn_features(const float* dI0s, const float* dI1s, int H, int W, float* dX, void* stream)
{
    if (!dI0s || !dI1s || !dX || H < 1 || W < 1 || (long long)H * W >= (1ll << 31) / 8) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_features, dim3(blocks256((long long)H * W)), dim3(256), 0, (hipStream_t)stream,
                       dI0s, dI1s, H, W, dX);
    CNN_CHECK("k_features");
    return OF_OK;
}

int of_cnn_forward(of_cnn* h, const float* dX, int B, int H, int W, float* dY, void* stream)
{
    return guarded([&]() -> int {
        if (!dims_ok(h, dX, dY, B, H, W)) return OF_ERR_ARG;
        hipStream_t st = (hipStream_t)stream;
        OF_TRY(ensure_ws(h, (long long)B * H * W));
        OF_TRY(forward(h, dX, B, H, W, st));
        OF_HIP(hipMemcpyAsync(dY, h->act.back(), (size_t)B * H * W * 2 * sizeof(float),
                              hipMemcpyDeviceToDevice, st));
        return OF_OK;
    });
}

int of_cnn_predict(of_cnn* h, const float* dI0s, const float* dI1s, int H, int W, float* dU, void* stream)
{
    return guarded([&]() -> int {
        if (!dims_ok(h, dI0s, dI1s, 1, H, W) || !dU) return OF_ERR_ARG;
        hipStream_t st = (hipStream_t)stream;
        const long long n = (long long)H * W;
        OF_TRY(ensure_in(h, n));
        OF_TRY(ensure_ws(h, n));
        OF_TRY(of_cnn_features(dI0s, dI1s, H, W, h->xin, st));
        OF_TRY(forward(h, h->xin, 1, H, W, st));
        hipLaunchKernelGGL(k_to_planar, dim3(blocks256(n)), dim3(256), 0, st, h->act.back(), (int)n, dU);
        CNN_CHECK("k_to_planar");
        return OF_OK;
    });
}

int of_cnn_gradients(of_cnn* h, const float* dX, const float* dT, int B, int H, int W, float* host_grads,
                     double* loss)
{
    return guarded([&]() -> int {
        if (!dims_ok(h, dX, dT, B, H, W)) return OF_ERR_ARG;
        OF_TRY(ensure_ws(h, (long long)B * H * W));
        OF_TRY(forward(h, dX, B, H, W, nullptr));
        OF_TRY(loss_grad(h, dT, B, H, W, nullptr));
        OF_TRY(backward(h, dX, B, H, W, nullptr));
        OF_HIP(hipDeviceSynchronize());
        if (host_grads) OF_HIP(hipMemcpy(host_grads, h->g, h->nparams * sizeof(float), hipMemcpyDeviceToHost));
        if (loss) OF_HIP(hipMemcpy(loss, h->loss, sizeof(double), hipMemcpyDeviceToHost));
        return OF_OK;
    });
}

int of_cnn_train_step(of_cnn* h, const float* dX, const float* dT, int B, int H, int W, float lr,
                      double* loss)
{
    return guarded([&]() -> int {
        if (!dims_ok(h, dX, dT, B, H, W) || !(lr > 0.0f)) return OF_ERR_ARG;
        OF_TRY(ensure_ws(h, (long long)B * H * W));
        OF_TRY(forward(h, dX, B, H, W, nullptr));
        OF_TRY(loss_grad(h, dT, B, H, W, nullptr));
        OF_TRY(backward(h, dX, B, H, W, nullptr));
        h->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)h->step), c2 = 1.0f - std::pow(b2, (float)h->step);
        hipLaunchKernelGGL(k_adam, dim3(blocks256((long long)h->nparams)), dim3(256), 0, nullptr, h->w, h->g,
                           h->m, h->v, (int)h->nparams, lr, b1, b2, 1e-8f, c1, c2);
        CNN_CHECK("k_adam");
        OF_TRY(pack(h, nullptr));
        if (loss) OF_HIP(hipMemcpy(loss, h->loss, sizeof(double), hipMemcpyDeviceToHost));
        else OF_HIP(hipDeviceSynchronize());
        return OF_OK;
    });
}

// Warm-started solve of one pair: prepare the pyramids, predict the
// coarsest-scale flow, carry it to `first_scale`, keep it only if its TV-L1
// energy there is not above that of zero flow, then solve scales
// first_scale .. 0 with at most `warps` warpings and `max_iterations`
// iterations per warping (0: the handle's parameters).  When zero flow wins,
// the solve is the plain one: zero initial flow at the coarsest scale and
// the handle's own parameters.
int of_tvl1_warm_run_device_u8(of_tvl1* tv, of_cnn* net, const uint8_t* dI0, const uint8_t* dI1,
                               int first_scale, int warps, int max_iterations, float* dU,
                               int* fell_back, double* energies, void* stream)
{
    return guarded([&]() -> int {
        if (!tv || !net || !dI0 || !dI1 || !dU || warps < 0 || max_iterations < 0) return OF_ERR_ARG;
        const int ns = of_tvl1_nscales(tv);
        if (first_scale < -1 || first_scale >= ns) return OF_ERR_ARG;
        if (first_scale < 0) first_scale = ns - 1;
        int nx = 0, ny = 0, fx = 0, fy = 0;
        OF_TRY(of_tvl1_scale_size(tv, ns - 1, &nx, &ny));
        OF_TRY(of_tvl1_scale_size(tv, first_scale, &fx, &fy));
        const long long n = (long long)nx * ny, nf = (long long)fx * fy;
        OF_TRY(ensure_in(net, n));
        if (2 * nf > net->cap_lift) {
            OF_HIP(hipDeviceSynchronize());
            dev_free(net->lift);
            dev_free(net->zero);
            net->cap_lift = 0;
            OF_TRY(dev_alloc(&net->lift, 2 * (size_t)nf));
            OF_TRY(dev_alloc(&net->zero, 2 * (size_t)nf));
            net->cap_lift = 2 * nf;
        }
        hipStream_t st = (hipStream_t)stream;
        OF_TRY(of_tvl1_prepare_device_u8(tv, dI0, dI1, stream));
        const float* I0c = nullptr;
        const float* I1c = nullptr;
        OF_TRY(of_tvl1_scale_images(tv, ns - 1, &I0c, &I1c));
        OF_TRY(of_cnn_predict(net, I0c, I1c, ny, nx, net->pred, stream));
        OF_TRY(of_tvl1_lift_flow(tv, ns - 1, net->pred, first_scale, net->lift, stream));
        OF_HIP(hipMemsetAsync(net->zero, 0, sizeof(float) * 2 * (size_t)nf, st));
        double ep[3], ez[3];
        OF_TRY(of_tvl1_energy(tv, first_scale, net->lift, ep, stream));
        OF_TRY(of_tvl1_energy(tv, first_scale, net->zero, ez, stream));
        if (energies) {
            energies[0] = ep[0];
            energies[1] = ez[0];
        }
        const bool keep = ep[0] <= ez[0];
        if (fell_back) *fell_back = keep ? 0 : 1;
        if (keep) return of_tvl1_solve_prepared(tv, net->lift, first_scale, warps, max_iterations, dU, stream);
        return of_tvl1_solve_prepared(tv, nullptr, -1, 0, 0, dU, stream);
    });
}

// ---- test entry point: MFMA operand layout through k_conv ----

// D = A (16x16) * B (16x16), row-major fp32 (operands rounded to bf16),
// computed as a convolution of a 1x16 image with 16 channels whose only
// non-zero tap is the centre one.
int of_dbg_cnn_mfma(const float* A, const float* Bm, float* D)
{
    return guarded([&]() -> int {
        if (!A || !Bm || !D) return OF_ERR_ARG;
        float* dX = nullptr;
        float* dY = nullptr;
        short* dW = nullptr;
        const int Kpad = 160;
        int st = dev_alloc(&dX, 256);
        if (!st) st = dev_alloc(&dY, 256);
        if (!st) st = dev_alloc(&dW, 16 * Kpad);
        std::vector<short> w(16 * Kpad, 0);
        for (int n = 0; n < 16; n++)
            for (int c = 0; c < 16; c++) w[n * Kpad + 4 * 16 + c] = to_bf16(Bm[c * 16 + n]);
        if (!st && hipMemcpy(dX, A, 256 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) st = OF_ERR_HIP;
        if (!st && hipMemcpy(dW, w.data(), w.size() * sizeof(short), hipMemcpyHostToDevice) != hipSuccess)
            st = OF_ERR_HIP;
        if (!st) st = launch_conv(dX, 1, 1, 16, 16, dW, Kpad, 16, nullptr, 1, 0, nullptr, dY, nullptr);
        if (!st && hipDeviceSynchronize() != hipSuccess) st = OF_ERR_HIP;
        if (!st && hipMemcpy(D, dY, 256 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) st = OF_ERR_HIP;
        dev_free(dX);
        dev_free(dY);
        dev_free(dW);
        return st;
    });
}

}  // extern "C"
