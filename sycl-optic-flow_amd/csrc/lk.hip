// lk.hip -- Lucas-Kanade on gfx950: kernels, host orchestration and C ABI.
//
// Single-scale LK over a ring of T u8 frames, as LucasKanade::lucas_kanade
// (src/lk/cuda/lucaskanade.cu:54-71, kernels/kernels.cu:14-181): temporal
// FIR It, zero-padded spatial correlation Ix, Iy of frame (iframe+1)%T,
// w x w window sums and a Cramer 2x2 solve where det != 0.  The reference's
// y-border test against nx (kernels.cu:106) is not reproduced and the
// border ring it leaves unwritten is 0.
//
// Pyramidal iterative LK (no reference counterpart): Gaussian pyramid as in
// TV-L1's zoom_out with the given sigma, per level the template gradient
// (centred differences) and its structure tensor G summed over the window,
// then `iters` x (bilinear warp of I1, b = window sums of (Ix It, Iy It),
// u -= G^-1 b where det G > det_eps), flow carried up with zoom_in x 1/zf.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "amdflow.h"
#include "amdflow_debug.h"
#include "common.hpp"
#include "device_math.hpp"
#include "tvl1_launch.hpp"
#include "tvl1_pixel.hpp"

namespace amdflow {
namespace {

constexpr int kMaxFilt = 9;

struct Filt {
    float f[kMaxFilt];
};

// lucaskanade.cu:79-161
bool lk_filter(int size, bool temporal, Filt* out)
{
    float* f = out->f;
    for (int i = 0; i < kMaxFilt; i++) f[i] = 0.0f;
    switch (size) {
    case 1: f[0] = 1.0f; return true;
    case 2:
        if (!temporal) return false;
        f[0] = -1.0f; f[1] = 1.0f; return true;
    case 3: f[0] = -1.0f / 2.0f; f[1] = 0.0f; f[2] = 1.0f / 2.0f; return true;
    case 5:
        f[0] = 1.0f / 12.0f; f[1] = -2.0f / 3.0f; f[2] = 0.0f; f[3] = 2.0f / 3.0f; f[4] = -1.0f / 12.0f;
        return true;
    case 7:
        f[0] = -1.0f / 60.0f; f[1] = 3.0f / 20.0f; f[2] = -3.0f / 4.0f; f[3] = 0.0f;
        f[4] = 3.0f / 4.0f; f[5] = -3.0f / 20.0f; f[6] = 1.0f / 60.0f;
        return true;
    case 9:
        f[0] = 1.0f / 280.0f; f[1] = -4.0f / 105.0f; f[2] = 1.0f / 5.0f; f[3] = -4.0f / 5.0f;
        f[4] = 0.0f; f[5] = 4.0f / 5.0f; f[6] = -1.0f / 5.0f; f[7] = 4.0f / 105.0f;
        f[8] = -1.0f / 280.0f;
        return true;
    default: return false;
    }
}

inline dim3 grid2d(int w, int h) { return dim3(cdiv(w, 64), cdiv(h, 4)); }

#define LK_CHECK(name)                                      \
    do {                                                    \
        hipError_t e_ = hipGetLastError();                  \
        if (e_ != hipSuccess) {                             \
            record_hip_error(e_, name, __FILE__, __LINE__); \
            return OF_ERR_HIP;                              \
        }                                                   \
    } while (0)

// It, Ix, Iy of the single-scale LK (kernels.cu:14-121).
__global__ __launch_bounds__(256) void k_lk_derivs(const uint8_t* __restrict__ ring, int T,
                                                   int iframe, int nx, int ny, Filt ft, int spac,
                                                   Filt fs, float* __restrict__ It,
                                                   float* __restrict__ Ix, float* __restrict__ Iy)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= nx || i >= ny) return;
    const size_t n = (size_t)nx * ny, p = (size_t)i * nx + j;
    float t = 0.0f;
    for (int k = 0; k < T; k++) t += ft.f[k] * (float)ring[(size_t)((k + iframe + 1) % T) * n + p];
    It[p] = t;
    const uint8_t* fr = ring + (size_t)((iframe + 1) % T) * n;
    const int fc = (spac - 1) / 2;
    float gx = 0.0f, gy = 0.0f;
    for (int k = 0; k < spac; k++) {
        const int jj = j + k - fc, ii = i + k - fc;
        if (jj >= 0 && jj < nx) gx += (float)fr[(size_t)i * nx + jj] * fs.f[k];
        if (ii >= 0 && ii < ny) gy += (float)fr[(size_t)ii * nx + j] * fs.f[k];
    }
    Ix[p] = gx;
    Iy[p] = gy;
}

// Window sums and Cramer solve (kernels.cu:134-181); the ring of width
// (win-1)/2 is written as 0.
__global__ __launch_bounds__(256) void k_lk_solve_single(const float* __restrict__ Ix,
                                                         const float* __restrict__ Iy,
                                                         const float* __restrict__ It, int nx,
                                                         int ny, int win, float* __restrict__ Vx,
                                                         float* __restrict__ Vy)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= nx || i >= ny) return;
    const size_t p = (size_t)i * nx + j;
    const int wc = (win - 1) / 2;
    float vx = 0.0f, vy = 0.0f;
    if (i >= wc && j >= wc && i < ny - wc && j < nx - wc) {
        float sxx = 0.0f, sxy = 0.0f, syy = 0.0f, sxt = 0.0f, syt = 0.0f;
        for (int ii = -wc; ii <= wc; ii++)
            for (int jj = -wc; jj <= wc; jj++) {
                const size_t q = (size_t)(i + ii) * nx + (j + jj);
                const float gx = Ix[q], gy = Iy[q], gt = It[q];
                sxx += gx * gx;
                sxy += gx * gy;
                syy += gy * gy;
                sxt += gx * gt;
                syt += gy * gt;
            }
        const float det = (sxx * syy - sxy * sxy);
        if (det != 0.0f) {
            vx = 1.0f / det * (syy * (-sxt) + (-sxy) * (-syt));
            vy = 1.0f / det * ((-sxy) * (-sxt) + sxx * (-syt));
        }
    }
    Vx[p] = vx;
    Vy[p] = vy;
}

// ---- pyramidal LK ----

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ float bilinear_at(const float* __restrict__ src, int w, int h, float x,
                                             float y)
{
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float ax = x - fx0, ay = y - fy0;
    // limited to [-1, w] first so that +1 cannot overflow (clamps unchanged)
    const int x0 = clampi(f2i_sat(fx0), -1, w), y0 = clampi(f2i_sat(fy0), -1, h);
    const int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    const int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    const float a = src[(size_t)ya * w + xa], b = src[(size_t)ya * w + xb];
    const float c = src[(size_t)yb * w + xa], d = src[(size_t)yb * w + xb];
    const float top = (1.0f - ax) * a + ax * b;
    const float bot = (1.0f - ax) * c + ax * d;
    return (1.0f - ay) * top + ay * bot;
}

// Products of the template gradient: Ix^2, Ix*Iy, Iy^2 (centred differences
// with one-sided x0.5 borders, as the TV-L1 gradient).
__global__ __launch_bounds__(256) void k_lk_tensor_products(const float* __restrict__ P0, int w,
                                                            int h, float* __restrict__ Ix,
                                                            float* __restrict__ Iy,
                                                            float* __restrict__ t3)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= w || i >= h) return;
    const size_t p = (size_t)i * w + j, n = (size_t)w * h;
    const int jl = j > 0 ? j - 1 : 0, jr = j < w - 1 ? j + 1 : w - 1;
    const int iu = i > 0 ? i - 1 : 0, id = i < h - 1 ? i + 1 : h - 1;
    const float ddx = P0[(size_t)i * w + jr] - P0[(size_t)i * w + jl];
    const float ddy = P0[(size_t)id * w + j] - P0[(size_t)iu * w + j];
    const float gx = (float)(0.5 * (double)ddx), gy = (float)(0.5 * (double)ddy);
    Ix[p] = gx;
    Iy[p] = gy;
    t3[p] = gx * gx;
    t3[n + p] = gx * gy;
    t3[2 * n + p] = gy * gy;
}

// Output rows of the fused LK tiles (64 columns wide).  8 rows give twice
// the workgroups of 16 for the same work, which the coarse levels need: the
// 1280x720 3-level solve took 0.260 ms with 8 everywhere, 0.265 ms with 8
// below 300,000 pixels and 16 above, 0.290 ms with 16 everywhere.
constexpr int kLkTileRows = 8;

// The gradient, its products and their window sums in one launch, on a
// 64 x kLkTileRows output tile: the products on the tile plus a WC-pixel halo (LDS),
// their row sums, then the column sums into G (Gxx, Gxy, Gyy planes); Ix and
// Iy are written for the output pixels.  Same expressions and summation
// order as k_lk_tensor_products followed by k_lk_boxsum along rows and then
// columns (windows clipped to the image, ascending order).
template <int WC, int TH = kLkTileRows>
__global__ __launch_bounds__(256) void k_lk_tensor_fused(const float* __restrict__ P0, int w,
                                                         int h, float* __restrict__ Ix,
                                                         float* __restrict__ Iy,
                                                         float* __restrict__ G)
{
    constexpr int TW = 64, EW = TW + 2 * WC, EH = TH + 2 * WC;
    constexpr int NE = (EW * EH + 255) / 256;
    constexpr int NO = TW * TH / 256;
    __shared__ float pr[3][EH][EW];
    __shared__ float rs[3][EH][TW];
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const size_t n = (size_t)w * h;
    // the four neighbour loads of every element are issued without branches
    // (outside the image: clamped, in-range addresses, result discarded)
    float nl[NE], nr[NE], nu[NE], nd[NE];
#pragma unroll
    for (int k = 0; k < NE; k++) {
        const int e = threadIdx.x + 256 * k;
        const int ey = e / EW, ex = e - ey * EW;
        const int j = min(max(x0 - WC + ex, 0), w - 1), i = min(max(y0 - WC + ey, 0), h - 1);
        const int jl = j > 0 ? j - 1 : 0, jr = j < w - 1 ? j + 1 : w - 1;
        const int iu = i > 0 ? i - 1 : 0, id = i < h - 1 ? i + 1 : h - 1;
        nl[k] = P0[(size_t)i * w + jl];
        nr[k] = P0[(size_t)i * w + jr];
        nu[k] = P0[(size_t)iu * w + j];
        nd[k] = P0[(size_t)id * w + j];
    }
#pragma unroll
    for (int k = 0; k < NE; k++) {
        const int e = threadIdx.x + 256 * k;
        if (e >= EW * EH) break;
        const int ey = e / EW, ex = e - ey * EW;
        const int j = x0 - WC + ex, i = y0 - WC + ey;
        float a = 0.0f, b = 0.0f, c = 0.0f;
        if (j >= 0 && j < w && i >= 0 && i < h) {
            const float ddx = nr[k] - nl[k];
            const float ddy = nd[k] - nu[k];
            const float gx = (float)(0.5 * (double)ddx), gy = (float)(0.5 * (double)ddy);
            a = gx * gx;
            b = gx * gy;
            c = gy * gy;
            const int ox = ex - WC, oy = ey - WC;
            if (ox >= 0 && ox < TW && oy >= 0 && oy < TH) {
                const size_t p = (size_t)i * w + j;
                Ix[p] = gx;
                Iy[p] = gy;
            }
        }
        pr[0][ey][ex] = a;
        pr[1][ey][ex] = b;
        pr[2][ey][ex] = c;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < TW * EH; e += 256) {
        const int ey = e / TW, ox = e - ey * TW;
        // products outside the image are +0.0, and adding +0.0 leaves a
        // sum that starts at +0.0 unchanged: every tap is added, with the
        // bits of the image-clipped window sum
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int k = -WC; k <= WC; k++) {
            s0 += pr[0][ey][ox + WC + k];
            s1 += pr[1][ey][ox + WC + k];
            s2 += pr[2][ey][ox + WC + k];
        }
        rs[0][ey][ox] = s0;
        rs[1][ey][ox] = s1;
        rs[2][ey][ox] = s2;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NO; k++) {
        const int e = threadIdx.x + 256 * k;
        const int oy = e / TW, ox = e - oy * TW;
        const int gx = x0 + ox, gy = y0 + oy;
        if (gx >= w || gy >= h) continue;
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int kk = -WC; kk <= WC; kk++) {
            s0 += rs[0][oy + WC + kk][ox];
            s1 += rs[1][oy + WC + kk][ox];
            s2 += rs[2][oy + WC + kk][ox];
        }
        const size_t p = (size_t)gy * w + gx;
        G[p] = s0;
        G[n + p] = s1;
        G[2 * n + p] = s2;
    }
}

// Row (dir 0) or column (dir 1) window sums of `planes` planes, window
// clipped to the image, ascending order.
__global__ __launch_bounds__(256) void k_lk_boxsum(const float* __restrict__ in,
                                                   float* __restrict__ out, int w, int h,
                                                   int planes, int wc, int dir)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= w || i >= h) return;
    const size_t n = (size_t)w * h;
    for (int c = 0; c < planes; c++) {
        const float* src = in + c * n;
        float s = 0.0f;
        for (int k = -wc; k <= wc; k++) {
            if (dir == 0) {
                const int jj = j + k;
                if (jj >= 0 && jj < w) s += src[(size_t)i * w + jj];
            } else {
                const int ii = i + k;
                if (ii >= 0 && ii < h) s += src[(size_t)ii * w + j];
            }
        }
        out[c * n + (size_t)i * w + j] = s;
    }
}

// It = I1(x+u) - I0 and the two products Ix*It, Iy*It.
__global__ __launch_bounds__(256) void k_lk_it_products(const float* __restrict__ P0,
                                                        const float* __restrict__ P1,
                                                        const float* __restrict__ Ix,
                                                        const float* __restrict__ Iy,
                                                        const float* __restrict__ u,
                                                        const float* __restrict__ v, int w, int h,
                                                        float* __restrict__ t2)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= w || i >= h) return;
    const size_t p = (size_t)i * w + j, n = (size_t)w * h;
    const float wv = bilinear_at(P1, w, h, (float)j + u[p], (float)i + v[p]);
    const float itv = wv - P0[p];
    t2[p] = Ix[p] * itv;
    t2[n + p] = Iy[p] * itv;
}

// Column sums of the two row-summed products, then the 2x2 update.
__global__ __launch_bounds__(256) void k_lk_update(const float* __restrict__ rows2,
                                                   const float* __restrict__ G, int w, int h,
                                                   int wc, float det_eps, float* __restrict__ u,
                                                   float* __restrict__ v)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= w || i >= h) return;
    const size_t p = (size_t)i * w + j, n = (size_t)w * h;
    float bx = 0.0f, by = 0.0f;
    for (int k = -wc; k <= wc; k++) {
        const int ii = i + k;
        if (ii >= 0 && ii < h) {
            bx += rows2[(size_t)ii * w + j];
            by += rows2[n + (size_t)ii * w + j];
        }
    }
    const float gxx = G[p], gxy = G[n + p], gyy = G[2 * n + p];
    const float det = gxx * gyy - gxy * gxy;
    if (det > det_eps) {
        const float du = (gxy * by - gyy * bx) / det;
        const float dv = (gxy * bx - gxx * by) / det;
        u[p] = u[p] + du;
        v[p] = v[p] + dv;
    }
}

// kk whole iterations on a 64 x 64 tile in one launch (temporal blocking):
// the iteration at a pixel reads the flow of every pixel of its window, so
// after t iterations a tile's flow is exact at distance >= t*WC from its
// edge, and the tile stores its central (64 - 2*kk*WC)^2 pixels.  Wave r of
// the 16 holds tile rows 4r .. 4r+3 (lane = column) with their flow, frame,
// gradient and tensor values in registers.  A row of window products lies
// in one wave, so the row sums take their neighbours through DPP lane
// shifts; the column sums read the 2*WC rows next to a thread's four from
// LDS (row sums, double-buffered: one barrier per iteration).  The
// per-pixel arithmetic and both summation orders are those of the
// three-kernel iteration k_lk_it_products / k_lk_boxsum / k_lk_update (the
// windows clipped to the image), so the flow is the same bits as kk rounds
// of it.
constexpr int kLkBlockEdge = 64;

// Whole-wave lane shifts for the row sums: lane i gets lane i-1's (from the
// left) or lane i+1's value; the lane without a source gets 0.  Those lanes
// (0 and 63) lie outside every square the blocked kernel stores.
__device__ __forceinline__ float lane_from_left0(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138 /* wave_shr:1 */, 0xf, 0xf, true));
}

__device__ __forceinline__ float lane_from_right0(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130 /* wave_shl:1 */, 0xf, 0xf, true));
}

// clampi(f2i_sat(f), -1, n) for an integer-valued f (a floor) and 0 < n <
// 2^24, without branches: NaN gives 0, everything else is clamped in float
// (where -1 and n are exact) and converted.
__device__ __forceinline__ int floor_index(float f, int n)
{
    const int c = (int)fminf(fmaxf(f, -1.0f), (float)n);
    return f != f ? 0 : c;
}

// img[idx] through a 32-bit byte offset (the blocked kernel runs on planes
// of at most 2^30 floats).
__device__ __forceinline__ float load_at(const float* __restrict__ img, int idx)
{
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(img) + (unsigned)idx * 4u);
}

template <int WC>
__global__ __launch_bounds__(1024) void k_lk_iters_blocked(const float* __restrict__ P0,
                                                           const float* __restrict__ P1,
                                                           const float* __restrict__ Ix,
                                                           const float* __restrict__ Iy,
                                                           const float* __restrict__ G,
                                                           const float* __restrict__ u,
                                                           const float* __restrict__ v, int w,
                                                           int h, float det_eps, int kk,
                                                           float* __restrict__ uo,
                                                           float* __restrict__ vo)
{
    constexpr int E = kLkBlockEdge;
    constexpr int NP = E * E / 1024;   // rows per thread
    __shared__ float rs[2][2][E][E];   // [buffer][plane][row][column]
    const int halo = kk * WC;
    const int O = E - 2 * halo;
    const int x0 = (int)blockIdx.x * O - halo, y0 = (int)blockIdx.y * O - halo;
    const int ex = threadIdx.x & 63, ey0 = (threadIdx.x >> 6) * NP;
    const size_t n = (size_t)w * h;
    const int gx = x0 + ex;
    const bool xin = gx >= 0 && gx < w;

    float uu[NP], vv[NP], i0[NP], gxv[NP], gyv[NP], gxx[NP], gxy[NP], gyy[NP];
    bool in[NP];
    const float* __restrict__ Gxy = G + n;
    const float* __restrict__ Gyy = G + 2 * n;
#pragma unroll
    for (int k = 0; k < NP; k++) {
        const int gy = y0 + ey0 + k;
        in[k] = xin && gy >= 0 && gy < h;
        const int p = in[k] ? gy * w + gx : 0;
        uu[k] = in[k] ? load_at(u, p) : 0.0f;
        vv[k] = in[k] ? load_at(v, p) : 0.0f;
        i0[k] = load_at(P0, p);
        gxv[k] = load_at(Ix, p);
        gyv[k] = load_at(Iy, p);
        gxx[k] = load_at(G, p);
        gxy[k] = load_at(Gxy, p);
        gyy[k] = load_at(Gyy, p);
    }
    for (int it = 0; it < kk; it++) {
        const int lo1 = (it + 1) * WC, hi1 = E - (it + 1) * WC;   // flow exact here after it
        const int b = it & 1;
        // the window products of all four rows (k_lk_it_products'); the taps
        // are clamped into the frame, so every gather is in bounds and all of
        // them are issued before the first is used.  Values outside the
        // square where the flow is exact are computed and never used.
        float ta[NP], tb[NP], tc[NP], td[NP], ax[NP], ay[NP];
#pragma unroll
        for (int k = 0; k < NP; k++) {
            const int gy = y0 + ey0 + k;
            const float x = (float)gx + uu[k], y = (float)gy + vv[k];
            const float fx0 = floorf(x), fy0 = floorf(y);
            ax[k] = x - fx0;
            ay[k] = y - fy0;
            const int xi = floor_index(fx0, w), yi = floor_index(fy0, h);
            const int xa = clampi(xi, 0, w - 1), xb = clampi(xi + 1, 0, w - 1);
            const int ya = clampi(yi, 0, h - 1), yb = clampi(yi + 1, 0, h - 1);
            const int ra = (int)__umul24((unsigned)ya, (unsigned)w);
            const int rb = (int)__umul24((unsigned)yb, (unsigned)w);
            ta[k] = load_at(P1, ra + xa);
            tb[k] = load_at(P1, ra + xb);
            tc[k] = load_at(P1, rb + xa);
            td[k] = load_at(P1, rb + xb);
        }
        float r0[NP], r1[NP];   // row sums of this thread's rows
#pragma unroll
        for (int k = 0; k < NP; k++) {
            const float top = (1.0f - ax[k]) * ta[k] + ax[k] * tb[k];
            const float bot = (1.0f - ax[k]) * tc[k] + ax[k] * td[k];
            const float wv = (1.0f - ay[k]) * top + ay[k] * bot;
            const float itv = wv - i0[k];
            float a[2 * WC + 1], c[2 * WC + 1];   // products at columns ex - WC .. ex + WC
            a[WC] = in[k] ? gxv[k] * itv : 0.0f;
            c[WC] = in[k] ? gyv[k] * itv : 0.0f;
#pragma unroll
            for (int d = 1; d <= WC; d++) {
                a[WC + d] = lane_from_right0(a[WC + d - 1]);
                c[WC + d] = lane_from_right0(c[WC + d - 1]);
                a[WC - d] = lane_from_left0(a[WC - d + 1]);
                c[WC - d] = lane_from_left0(c[WC - d + 1]);
            }
            // every tap is added: a tap outside the image carries +0.0 (its
            // lane's `in` is false), and s + (+0.0) is s for every s but
            // -0.0, which a sum that starts at +0.0 never is; so the sum has
            // the bits of the image-clipped window sum
            float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
            for (int q = 0; q <= 2 * WC; q++) {
                s0 += a[q];
                s1 += c[q];
            }
            r0[k] = s0;
            r1[k] = s1;
            rs[b][0][ey0 + k][ex] = s0;
            rs[b][1][ey0 + k][ex] = s1;
        }
        __syncthreads();
        // column sums of rows ey0 - WC .. ey0 + NP - 1 + WC: the thread's own
        // from registers, the others from LDS (rows clamped into the tile are
        // read only for pixels outside the exact square)
        float y0s[NP + 2 * WC], y1s[NP + 2 * WC];
#pragma unroll
        for (int j = 0; j < WC; j++) {
            const int ra = ey0 - WC + j < 0 ? 0 : ey0 - WC + j;
            const int rb = ey0 + NP + j > E - 1 ? E - 1 : ey0 + NP + j;
            y0s[j] = rs[b][0][ra][ex];
            y1s[j] = rs[b][1][ra][ex];
            y0s[WC + NP + j] = rs[b][0][rb][ex];
            y1s[WC + NP + j] = rs[b][1][rb][ex];
        }
#pragma unroll
        for (int k = 0; k < NP; k++) {
            y0s[WC + k] = r0[k];
            y1s[WC + k] = r1[k];
        }
        // rows outside the image hold row sums of +0.0, so the column sums
        // need no clipping either
#pragma unroll
        for (int k = 0; k < NP; k++) {
            const int ey = ey0 + k;
            if (!in[k] || ey < lo1 || ey >= hi1 || ex < lo1 || ex >= hi1) continue;
            float bx = 0.0f, by = 0.0f;
#pragma unroll
            for (int q = -WC; q <= WC; q++) {
                bx += y0s[WC + k + q];
                by += y1s[WC + k + q];
            }
            const float det = gxx[k] * gyy[k] - gxy[k] * gxy[k];
            if (det > det_eps) {
                const float du = (gxy[k] * by - gyy[k] * bx) / det;
                const float dv = (gxy[k] * bx - gxx[k] * by) / det;
                uu[k] = uu[k] + du;
                vv[k] = vv[k] + dv;
            }
        }
        // rs[b] is written again two iterations on, after the barrier of
        // the next iteration, which every thread reaches only once its reads
        // of rs[b] above are done
    }
#pragma unroll
    for (int k = 0; k < NP; k++) {
        const int ey = ey0 + k, gy = y0 + ey;
        if (!in[k] || ey < halo || ey >= E - halo || ex < halo || ex >= E - halo) continue;
        const size_t p = (size_t)gy * w + gx;
        uo[p] = uu[k];
        vo[p] = vv[k];
    }
}

// Iterations per launch of k_lk_iters_blocked.  A launch stores the centre
// (64 - 2*kk*WC)^2 of each 64 x 64 tile, so (64 / (64 - 2*kk*WC))^2 pixels
// are read and iterated per stored pixel.  Levels of at most kLkSmallLevel
// pixels have too few tiles to fill the GPU and take the deepest blocks (a
// centre of at least 32 pixels per side); larger levels keep the halo within
// 9 pixels (1.9 pixels iterated per stored one at WC = 3, kk = 3).  On 1280 x
// 720 frames, 3 levels and 7 x 7 windows the solve took 0.241 ms with one
// launch per iteration on every level, 0.226 ms with blocks on the two small
// levels only, 0.212 ms with level 0 in blocks of 3 and 0.244 ms in blocks
// of 5 (blocks of 2 were slower than blocks of 3 in an earlier version of
// the kernel).
constexpr long long kLkSmallLevel = 300000;
constexpr int lk_block_iters(int wc, long long npix)
{
    return npix <= kLkSmallLevel ? (kLkBlockEdge - 32) / (2 * wc) : (9 / wc > 1 ? 9 / wc : 1);
}

// One pyramid step of both frames per launch (blockIdx.z = frame): the
// Gaussian rows, the Gaussian columns, then zoom_out's resampling, with the
// pixel bodies of the TV-L1 pyramid (mask.cpp:204-274, zoom.cpp:28-52).
__global__ __launch_bounds__(256) void k_lk_pyr_rows(const float* __restrict__ a,
                                                     const float* __restrict__ b,
                                                     float* __restrict__ oa, float* __restrict__ ob,
                                                     int nx, int ny, GaussTaps g)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= nx || y >= ny) return;
    (blockIdx.z ? ob : oa)[(size_t)y * nx + x] = gauss_row_px(blockIdx.z ? b : a, nx, x, y, g);
}

__global__ __launch_bounds__(256) void k_lk_pyr_cols(const float* __restrict__ a,
                                                     const float* __restrict__ b,
                                                     float* __restrict__ oa, float* __restrict__ ob,
                                                     int nx, int ny, GaussTaps g)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= nx || y >= ny) return;
    (blockIdx.z ? ob : oa)[(size_t)y * nx + x] = gauss_col_px(blockIdx.z ? b : a, nx, ny, x, y, g);
}

__global__ __launch_bounds__(256) void k_lk_pyr_zoom(const float* __restrict__ a,
                                                     const float* __restrict__ b,
                                                     float* __restrict__ oa, float* __restrict__ ob,
                                                     int nx, int ny, int nxx, int nyy, float factor)
{
    const int j1 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i1 = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j1 >= nxx || i1 >= nyy) return;
    (blockIdx.z ? ob : oa)[(size_t)i1 * nxx + j1] = zoom_out_px(blockIdx.z ? b : a, nx, ny, i1, j1, factor);
}

__global__ __launch_bounds__(256) void k_lk_zero2(float* __restrict__ a, float* __restrict__ b,
                                                  size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        a[i] = 0.0f;
        b[i] = 0.0f;
    }
}

unsigned g1(size_t n) { const unsigned g = cdiv(n, 256); return g > 4096 ? 4096 : (g ? g : 1); }

}  // namespace

// Device workspace of the pyramidal LK (the entry points take no handle).
// Workspaces live in a process-wide pool, one list per device: a call takes
// one for its duration and gives it back idle (the call has synchronised its
// stream, also on its error paths), so growing one needs no device-wide wait.
// The pool is never destroyed: nothing in the library calls HIP while the
// process exits (static and thread-local destructors run after the runtime
// may have begun its own teardown); the driver reclaims the memory.
// Geometry of one pyramidal LK solve: level sizes, and the offsets of
// levels 1.. inside the pyramid planes P0/P1 (level 0 is read from the
// caller's frames in place).
struct LkPlan {
    int levels = 0;
    int nx[16], ny[16];
    size_t off[16];
    size_t pyr = 0;   // floats of levels 1.. per frame
    size_t n0 = 0;    // pixels of level 0
    GaussTaps g;
};

int lk_plan(int W, int H, int levels, float zf, float sigma, int win, int iters, LkPlan* P)
{
    if (W < 2 || H < 2) return OF_ERR_ARG;
    if (levels < 1 || levels > 16 || win < 1 || (win & 1) == 0 || iters < 0) return OF_ERR_ARG;
    if (!(zf > 0 && zf < 1) || !(sigma > 0)) return OF_ERR_ARG;
    OF_TRY(gaussian_taps((double)sigma, &P->g));
    P->levels = levels;
    P->nx[0] = W;
    P->ny[0] = H;
    P->off[0] = 0;
    P->pyr = 0;
    for (int l = 0; l < levels; l++) {
        if (l) zoom_size(P->nx[l - 1], P->ny[l - 1], zf, &P->nx[l], &P->ny[l]);
        if (P->nx[l] < 2 || P->ny[l] < 2) return OF_ERR_ARG;
        if (l + 1 < levels && (P->g.size > P->nx[l] || P->g.size > P->ny[l])) return OF_ERR_SIGMA;
        if (l) {
            P->off[l] = P->pyr;
            P->pyr += (size_t)P->nx[l] * P->ny[l];
        }
    }
    P->n0 = (size_t)W * H;
    return OF_OK;
}

// Device buffers of one solve, carved from one block of lk_block_floats().
struct LkBufs {
    float *P0, *P1;            // pyramid levels 1.. of each frame (LkPlan::off)
    float *Ix, *Iy, *u, *v, *nu, *nv;   // one plane of level 0's size each
    float *PROD, *ROWS, *GT;   // three planes each: per-pixel products, their
                               // row sums, the structure tensor Gxx, Gxy, Gyy
};

size_t lk_round(size_t n) { return (n + 63) & ~(size_t)63; }   // 256-byte multiples

size_t lk_block_floats(const LkPlan& P)
{
    return 2 * lk_round(P.pyr) + 6 * lk_round(P.n0) + 3 * lk_round(3 * P.n0);
}

LkBufs lk_carve(float* base, const LkPlan& P)
{
    LkBufs B;
    float* q = base;
    auto take = [&](size_t n) { float* r = q; q += lk_round(n); return r; };
    B.P0 = take(P.pyr);
    B.P1 = take(P.pyr);
    B.Ix = take(P.n0);
    B.Iy = take(P.n0);
    B.u = take(P.n0);
    B.v = take(P.n0);
    B.nu = take(P.n0);
    B.nv = take(P.n0);
    B.PROD = take(3 * P.n0);
    B.ROWS = take(3 * P.n0);
    B.GT = take(3 * P.n0);
    return B;
}

// The solve proper: pyramid, then coarse to fine the structure tensor and
// `iters` iterations per level, zoom_in between levels.  Every buffer of B
// is written before it is read (u, v are cleared at the coarsest level), so
// a block that held an earlier call's data gives the same result.  Returns
// after synchronising `st`.
int lk_pyr_run(const LkPlan& P, const LkBufs& B, const float* dI0, const float* dI1, float zf,
               int win, int iters, float det_eps, float* du, float* dv, hipStream_t st)
{
    const int levels = P.levels;
    const int* nx = P.nx;
    const int* ny = P.ny;
    const size_t n0 = P.n0;
    const GaussTaps& g = P.g;
    float* P0 = B.P0;
    float* P1 = B.P1;
    float* Ix = B.Ix;
    float* Iy = B.Iy;
    float* PROD = B.PROD;
    float* ROWS = B.ROWS;
    float* GT = B.GT;
    float* u = B.u;
    float* v = B.v;
    float* nu = B.nu;
    float* nv = B.nv;
    size_t off[16];
    for (int l = 0; l < levels; l++) off[l] = P.off[l];
    // level 0 is read from the caller's frames; levels 1.. live in P0/P1.
    // Each pyramid step treats both frames in one launch, with PROD and
    // ROWS (free until the structure tensor) as the two blur buffers.
    const float* L0[16];
    const float* L1[16];
    for (int l = 0; l < levels; l++) {
        L0[l] = l ? P0 + off[l] : dI0;
        L1[l] = l ? P1 + off[l] : dI1;
    }
    for (int l = 1; l < levels; l++) {
        const int pw = nx[l - 1], ph = ny[l - 1];
        if (zf == 0.5f) {   // one tiled launch for both frames
            const int sz = launch_zoom_out_half_pair(L0[l - 1], L1[l - 1], P0 + off[l], P1 + off[l],
                                                     pw, ph, nx[l], ny[l], g, st);
            if (sz != OF_ERR_UNSUPPORTED) {
                OF_TRY(sz);
                continue;
            }
        }
        const dim3 gb(cdiv(pw, 64), cdiv(ph, 4), 2), gz(cdiv(nx[l], 64), cdiv(ny[l], 4), 2);
        hipLaunchKernelGGL(k_lk_pyr_rows, gb, dim3(256), 0, st, L0[l - 1], L1[l - 1], PROD,
                           PROD + n0, pw, ph, g);
        LK_CHECK("k_lk_pyr_rows");
        hipLaunchKernelGGL(k_lk_pyr_cols, gb, dim3(256), 0, st, PROD, PROD + n0, ROWS, ROWS + n0,
                           pw, ph, g);
        LK_CHECK("k_lk_pyr_cols");
        hipLaunchKernelGGL(k_lk_pyr_zoom, gz, dim3(256), 0, st, ROWS, ROWS + n0, P0 + off[l],
                           P1 + off[l], pw, ph, nx[l], ny[l], zf);
        LK_CHECK("k_lk_pyr_zoom");
    }
    const int wc = (win - 1) / 2;
    // the last iteration launch at level 0 writes the caller's du/dv
    // directly unless they overlap the frames or each other
    // address ranges of len floats, compared as integers (the buffers
    // are separate allocations)
    auto overlaps = [](const float* a, const float* b, size_t len) {
        const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b, n = len * sizeof(float);
        return x < y + n && y < x + n;
    };
    // the blocked iterations index planes with 32-bit byte offsets and
    // 24-bit row products; level 0 is the largest level
    const bool blocked = wc >= 1 && wc <= 4 && nx[0] < (1 << 24) && ny[0] < (1 << 24) &&
                         n0 <= ((size_t)1 << 30);
    const bool direct = blocked && iters >= 1 && !overlaps(du, dv, n0) &&
                        !overlaps(du, dI0, n0) && !overlaps(du, dI1, n0) &&
                        !overlaps(dv, dI0, n0) && !overlaps(dv, dI1, n0);
    const float invZ = 1 / zf;
    {
        const size_t nc = (size_t)nx[levels - 1] * ny[levels - 1];
        hipLaunchKernelGGL(k_lk_zero2, dim3(g1(nc)), dim3(256), 0, st, u, v, nc);
        LK_CHECK("k_lk_zero2");
    }
    for (int l = levels - 1; l >= 0; l--) {
        const int w = nx[l], h = ny[l];
        const float* p0 = L0[l];
        const float* p1 = L1[l];
        const dim3 fgrid(cdiv(w, 64), cdiv(h, kLkTileRows));
        if (wc >= 1 && wc <= 4) {
            // gradient, products and both window sums in one launch
            if (wc == 1) hipLaunchKernelGGL((k_lk_tensor_fused<1>), fgrid, dim3(256), 0, st, p0, w, h, Ix, Iy, GT);
            if (wc == 2) hipLaunchKernelGGL((k_lk_tensor_fused<2>), fgrid, dim3(256), 0, st, p0, w, h, Ix, Iy, GT);
            if (wc == 3) hipLaunchKernelGGL((k_lk_tensor_fused<3>), fgrid, dim3(256), 0, st, p0, w, h, Ix, Iy, GT);
            if (wc == 4) hipLaunchKernelGGL((k_lk_tensor_fused<4>), fgrid, dim3(256), 0, st, p0, w, h, Ix, Iy, GT);
            LK_CHECK("k_lk_tensor_fused");
        } else {
            hipLaunchKernelGGL(k_lk_tensor_products, grid2d(w, h), dim3(256), 0, st, p0, w, h, Ix,
                               Iy, PROD);
            LK_CHECK("k_lk_tensor_products");
            hipLaunchKernelGGL(k_lk_boxsum, grid2d(w, h), dim3(256), 0, st, PROD, ROWS, w, h, 3, wc,
                               0);
            LK_CHECK("k_lk_boxsum");
            hipLaunchKernelGGL(k_lk_boxsum, grid2d(w, h), dim3(256), 0, st, ROWS, GT, w, h, 3, wc,
                               1);
            LK_CHECK("k_lk_boxsum");
        }
        // windows up to 9 x 9: blocks of iterations per launch (k_lk_iters_blocked)
        for (int it = 0; it < iters && blocked;) {
            const int kb = lk_block_iters(wc, (long long)w * h);
            const int kk = iters - it < kb ? iters - it : kb;
            const int O = kLkBlockEdge - 2 * kk * wc;
            const dim3 bgrid(cdiv(w, O), cdiv(h, O));
            if (direct && l == 0 && it + kk == iters) {
                nu = du;
                nv = dv;
            }
            if (wc == 1) hipLaunchKernelGGL((k_lk_iters_blocked<1>), bgrid, dim3(1024), 0, st, p0, p1, Ix, Iy, GT, u, v, w, h, det_eps, kk, nu, nv);
            if (wc == 2) hipLaunchKernelGGL((k_lk_iters_blocked<2>), bgrid, dim3(1024), 0, st, p0, p1, Ix, Iy, GT, u, v, w, h, det_eps, kk, nu, nv);
            if (wc == 3) hipLaunchKernelGGL((k_lk_iters_blocked<3>), bgrid, dim3(1024), 0, st, p0, p1, Ix, Iy, GT, u, v, w, h, det_eps, kk, nu, nv);
            if (wc == 4) hipLaunchKernelGGL((k_lk_iters_blocked<4>), bgrid, dim3(1024), 0, st, p0, p1, Ix, Iy, GT, u, v, w, h, det_eps, kk, nu, nv);
            LK_CHECK("k_lk_iters_blocked");
            float* t = u; u = nu; nu = t;
            t = v; v = nv; nv = t;
            it += kk;
        }
        for (int it = 0; it < iters && !blocked; it++) {
            hipLaunchKernelGGL(k_lk_it_products, grid2d(w, h), dim3(256), 0, st, p0, p1, Ix, Iy,
                               u, v, w, h, PROD);
            LK_CHECK("k_lk_it_products");
            hipLaunchKernelGGL(k_lk_boxsum, grid2d(w, h), dim3(256), 0, st, PROD, ROWS, w, h, 2, wc, 0);
            LK_CHECK("k_lk_boxsum");
            hipLaunchKernelGGL(k_lk_update, grid2d(w, h), dim3(256), 0, st, ROWS, GT, w, h, wc,
                               det_eps, u, v);
            LK_CHECK("k_lk_update");
        }
        if (l > 0) {
            // zoom_in (zoom.cpp:61-85) of both components, times 1/zf
            OF_TRY(launch_zoom_in_planes(u, v, nu, nv, w, h, nx[l - 1], ny[l - 1], invZ, st));
            float* t = u; u = nu; nu = t;
            t = v; v = nv; nv = t;
        }
    }
    if (!direct) {
        OF_HIP(hipMemcpyAsync(du, u, n0 * sizeof(float), hipMemcpyDeviceToDevice, st));
        OF_HIP(hipMemcpyAsync(dv, v, n0 * sizeof(float), hipMemcpyDeviceToDevice, st));
    }
    OF_HIP(hipStreamSynchronize(st));
    return OF_OK;
}

// Device workspace of the pyramidal LK (the entry points take no handle):
// one block of lk_block_floats().  Workspaces live in a process-wide pool,
// one list per device: a call takes one for its duration and gives it back
// idle (the call has synchronised its stream, also on its error paths), so
// growing one needs no device-wide wait.  The pool is never destroyed:
// nothing in the library calls HIP while the process exits (static and
// thread-local destructors run after the runtime may have begun its own
// teardown); the driver reclaims the memory.
struct LkWorkspace {
    int device = -1;
    size_t cap = 0;
    float* block = nullptr;
    void release()
    {
        dev_free(block);
        cap = 0;
    }
    // idle workspace: no kernel of an earlier call still uses the block
    int ensure(size_t floats)
    {
        if (floats <= cap) return OF_OK;
        release();
        OF_TRY(dev_alloc(&block, floats));
        cap = floats;
        return OF_OK;
    }
};

struct LkPool {
    std::mutex m;
    std::vector<LkWorkspace*> idle;
};

LkPool& lk_pool()
{
    static LkPool* pool = new LkPool();   // never deleted (see above)
    return *pool;
}

// Takes an idle workspace of `device` for one call; gives it back when the
// call ends, after draining `st` if the call returned early with work queued.
class LkLease {
public:
    LkLease(int device, hipStream_t st) : st_(st)
    {
        LkPool& p = lk_pool();
        std::lock_guard<std::mutex> g(p.m);
        for (size_t i = 0; i < p.idle.size(); i++)
            if (p.idle[i]->device == device) {
                ws_ = p.idle[i];
                p.idle.erase(p.idle.begin() + i);
                return;
            }
        ws_ = new (std::nothrow) LkWorkspace();
        if (ws_) ws_->device = device;
    }
    ~LkLease()
    {
        if (!ws_) return;
        if (!drained_) {
            (void)hipStreamSynchronize(st_);
            (void)hipGetLastError();
        }
        LkPool& p = lk_pool();
        std::lock_guard<std::mutex> g(p.m);
        try {
            p.idle.push_back(ws_);
        } catch (...) {
            ws_->release();
            delete ws_;
        }
    }
    LkWorkspace* get() const { return ws_; }
    void drained() { drained_ = true; }

private:
    hipStream_t st_;
    LkWorkspace* ws_ = nullptr;
    bool drained_ = false;
};

// Pyramidal LK from device frames, on a pooled workspace.
int lk_pyr_solve(const float* dI0, const float* dI1, int W, int H, int levels, float zf,
                 float sigma, int win, int iters, float det_eps, float* du, float* dv,
                 hipStream_t st)
{
    if (!dI0 || !dI1 || !du || !dv) return OF_ERR_ARG;
    LkPlan P;
    OF_TRY(lk_plan(W, H, levels, zf, sigma, win, iters, &P));
    int device = 0;
    OF_HIP(hipGetDevice(&device));
    LkLease lease(device, st);
    if (!lease.get()) return OF_ERR_ALLOC;
    OF_TRY(lease.get()->ensure(lk_block_floats(P)));
    const int status = lk_pyr_run(P, lk_carve(lease.get()->block, P), dI0, dI1, zf, win, iters,
                                  det_eps, du, dv, st);
    if (status == OF_OK) lease.drained();
    return status;
}

}  // namespace amdflow

// Scratch of of_lk_pyr_compute_device_noalloc: one fixed block per GPU in
// the code object's data, so that entry point never allocates.  Host code
// never uses this name as an address (it is a host-side stand-in there); it
// uses the device address that hipGetSymbolAddress returns for each device.
constexpr size_t kLkScratchFloats = (size_t)32 << 20;   // 128 MiB
static __device__ __align__(256) float scratch[kLkScratchFloats];

namespace {

std::mutex g_scratch_mu;
float* g_scratch_ptr[64];   // device address of `scratch` per device id

// Device address of `scratch` on the current device `device`, looked up once
// per device with that device current.
int lk_scratch(int device, float** out)
{
    if (device < 0 || device >= 64) return OF_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        OF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return OF_ERR_HIP;
        g_scratch_ptr[device] = (float*)ptr;
    }
    *out = g_scratch_ptr[device];
    return OF_OK;
}

}  // namespace

using namespace amdflow;

struct of_lk {
    int spac = 3, temp = 2, win = 5, nx = 0, ny = 0;
    Filt fs{}, ft{};
    hipStream_t stream = nullptr;
    uint8_t* ring = nullptr;
    float *It = nullptr, *Ix = nullptr, *Iy = nullptr, *Vx = nullptr, *Vy = nullptr;
};

namespace {
void lk_release(of_lk* h)
{
    if (!h) return;
    dev_free(h->ring);
    dev_free(h->It); dev_free(h->Ix); dev_free(h->Iy); dev_free(h->Vx); dev_free(h->Vy);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}
}  // namespace

extern "C" {

int of_lk_create(int spac_filt_size, int temp_filt_size, int window_size, int nx, int ny, of_lk** out)
{
    if (!out) return OF_ERR_ARG;
    *out = nullptr;
    if (nx < 1 || ny < 1 || window_size < 1 || (window_size & 1) == 0) return OF_ERR_ARG;
    if ((long long)nx * ny >= (1ll << 31)) return OF_ERR_ARG;
    Filt fs, ft;
    if (!lk_filter(spac_filt_size, false, &fs) || !lk_filter(temp_filt_size, true, &ft))
        return OF_ERR_UNSUPPORTED;
    of_lk* h = new (std::nothrow) of_lk();
    if (!h) return OF_ERR_ALLOC;
    h->spac = spac_filt_size;
    h->temp = temp_filt_size;
    h->win = window_size;
    h->nx = nx;
    h->ny = ny;
    h->fs = fs;
    h->ft = ft;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        h->stream = nullptr;
        lk_release(h);
        return OF_ERR_HIP;
    }
    const size_t n = (size_t)nx * ny;
    int st = dev_alloc(&h->ring, (size_t)temp_filt_size * n);
    float** bufs[] = {&h->It, &h->Ix, &h->Iy, &h->Vx, &h->Vy};
    for (float** b : bufs)
        if (!st) st = dev_alloc(b, n);
    if (!st && hipMemsetAsync(h->ring, 0, (size_t)temp_filt_size * n, h->stream) != hipSuccess) st = OF_ERR_HIP;
    if (!st && hipStreamSynchronize(h->stream) != hipSuccess) st = OF_ERR_HIP;
    if (st) { lk_release(h); return st; }
    *out = h;
    return OF_OK;
}

void of_lk_destroy(of_lk* h) { lk_release(h); }

int of_lk_push_u8(of_lk* h, const uint8_t* frame, int temp_size, int fr, int frame_size)
{
    if (!h || !frame || temp_size != h->temp || fr < 0) return OF_ERR_ARG;
    if ((size_t)frame_size != (size_t)h->nx * h->ny) return OF_ERR_ARG;
    const int idx = fr % temp_size;
    OF_HIP(hipMemcpyAsync(h->ring + (size_t)idx * frame_size, frame, (size_t)frame_size,
                          hipMemcpyHostToDevice, h->stream));
    OF_HIP(hipStreamSynchronize(h->stream));
    return OF_OK;
}

int of_lk_compute(of_lk* h, float* Vx, float* Vy, int iframe)
{
    if (!h || !Vx || !Vy || iframe < 0) return OF_ERR_ARG;
    const size_t n = (size_t)h->nx * h->ny;
    hipLaunchKernelGGL(k_lk_derivs, grid2d(h->nx, h->ny), dim3(256), 0, h->stream, h->ring, h->temp,
                       iframe, h->nx, h->ny, h->ft, h->spac, h->fs, h->It, h->Ix, h->Iy);
    LK_CHECK("k_lk_derivs");
    hipLaunchKernelGGL(k_lk_solve_single, grid2d(h->nx, h->ny), dim3(256), 0, h->stream, h->Ix,
                       h->Iy, h->It, h->nx, h->ny, h->win, h->Vx, h->Vy);
    LK_CHECK("k_lk_solve_single");
    OF_HIP(hipMemcpyAsync(Vx, h->Vx, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    OF_HIP(hipMemcpyAsync(Vy, h->Vy, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    OF_HIP(hipStreamSynchronize(h->stream));
    return OF_OK;
}

int of_lk_pyr_compute_device(const float* dI0, const float* dI1, int width, int height, int levels,
                             float zf, float sigma, int window, int iters, float det_eps, float* du,
                             float* dv, void* stream)
{
    return lk_pyr_solve(dI0, dI1, width, height, levels, zf, sigma, window, iters, det_eps, du, dv,
                        (hipStream_t)stream);
}

int of_lk_pyr_compute_device_noalloc(const float* dI0, const float* dI1, int width, int height,
                                     int levels, float zf, float sigma, int window, int iters,
                                     float det_eps, float* du, float* dv, void* stream)
{
    if (!dI0 || !dI1 || !du || !dv) return OF_ERR_ARG;
    LkPlan P;
    OF_TRY(lk_plan(width, height, levels, zf, sigma, window, iters, &P));
    if (lk_block_floats(P) > kLkScratchFloats) return OF_ERR_CAPACITY;
    int device = 0;
    OF_HIP(hipGetDevice(&device));
    float* base = nullptr;
    OF_TRY(lk_scratch(device, &base));
    const hipStream_t st = (hipStream_t)stream;
    const int status = lk_pyr_run(P, lk_carve(base, P), dI0, dI1, zf, window, iters, det_eps, du,
                                  dv, st);
    if (status != OF_OK) {
        // the next call reuses the scratch: let this one's queued work end first
        (void)hipStreamSynchronize(st);
        (void)hipGetLastError();
    }
    return status;
}

int of_lk_pyr_compute(const float* I0, const float* I1, int width, int height, int levels, float zf,
                      float sigma, int window, int iters, float det_eps, float* u, float* v)
{
    if (!I0 || !I1 || !u || !v || width < 2 || height < 2) return OF_ERR_ARG;
    const size_t n = (size_t)width * height;
    float *d0 = nullptr, *d1 = nullptr, *du = nullptr, *dv = nullptr;
    int st = dev_alloc(&d0, n);
    if (!st) st = dev_alloc(&d1, n);
    if (!st) st = dev_alloc(&du, n);
    if (!st) st = dev_alloc(&dv, n);
    if (!st && hipMemcpy(d0, I0, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) st = OF_ERR_HIP;
    if (!st && hipMemcpy(d1, I1, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) st = OF_ERR_HIP;
    if (!st) st = lk_pyr_solve(d0, d1, width, height, levels, zf, sigma, window, iters, det_eps, du, dv, nullptr);
    if (!st && hipMemcpy(u, du, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) st = OF_ERR_HIP;
    if (!st && hipMemcpy(v, dv, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) st = OF_ERR_HIP;
    dev_free(d0); dev_free(d1); dev_free(du); dev_free(dv);
    return st;
}

}  // extern "C"
