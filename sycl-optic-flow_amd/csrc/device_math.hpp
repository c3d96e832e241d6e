// device_math.hpp -- per-pixel arithmetic shared by the TV-L1, HS and LK
// kernels.  Every expression is written with the operand types and the
// evaluation order of the reference's sequential path so that, compiled with
// -ffp-contract=off (no fused multiply-add) and IEEE fp32 division/sqrt, the
// GPU produces the same bits as the CPU.
#pragma once
#include <climits>
#include <hip/hip_runtime.h>

namespace amdflow {

// Clamp used by the bicubic sampler: bicubic_interpolation.cpp:17-30
// (Neumann boundary, flags the tap as outside).
__device__ __forceinline__ int neumann_clamp(int x, int n, bool& out)
{
    if (x < 0) { out = true; return 0; }
    if (x >= n) { out = true; return n - 1; }
    return x;
}

// Keys cubic convolution (a = -1/2) over four samples at fraction x:
// bicubic_interpolation.cpp:38-46.  The literals are double in the
// reference, the two bare differences are float, the result is rounded to
// float.
// The products of a float by 3, 5 and 4 are exact in double, so each
// product-then-sum below is one fused multiply-add with the same rounding
// (and the same infinities, NaNs and signed zeros).
__device__ __forceinline__ float cubic_cell(float v0, float v1, float v2, float v3, float x)
{
    const double xd = (double)x;
    const float d20 = v2 - v0;
    const float d12 = v1 - v2;
    const double c3 = __builtin_fma(3.0, (double)d12, (double)v3) - (double)v0;
    const double c2 = __builtin_fma(4.0, (double)v2, __builtin_fma(-5.0, (double)v1, 2.0 * (double)v0)) -
                      (double)v3;
    const double inner = (double)d20 + xd * (c2 + xd * c3);
    return (float)((double)v1 + (0.5 * xd) * inner);
}

// Tap geometry of one bicubic sample: bicubic_interpolation.cpp:74-99.
// Row `my` uses the x direction sign, as the reference does (:92).
struct BicubicTaps {
    int cols[4];
    int rows[4];
    float fx, fy;
    bool out;
};

// float -> int truncation defined for every input: NaN gives 0, values
// beyond the int range saturate (the conversion instruction's own results,
// spelled out because the language leaves them undefined).
__device__ __forceinline__ int f2i_sat(float x)
{
    return x != x ? 0 : (x >= 2147483647.0f ? INT_MAX : (x <= -2147483648.0f ? INT_MIN : (int)x));
}

// A tap base index limited to [-4, n+4]: every tap within 2 of it clamps to
// the same pixel and sets `out` as without the limit, and the offsets
// cannot overflow.
__device__ __forceinline__ int tap_base(int i, int n) { return i < -4 ? -4 : (i > n + 4 ? n + 4 : i); }

__device__ __forceinline__ BicubicTaps bicubic_taps(float uu, float vv, int nx, int ny)
{
    BicubicTaps t;
    const int sx = (uu < 0) ? -1 : 1;
    const int sy = (vv < 0) ? -1 : 1;
    bool out = false;
    const int iu = tap_base(f2i_sat(uu), nx), iv = tap_base(f2i_sat(vv), ny);
    const int x = neumann_clamp(iu, nx, out);
    const int y = neumann_clamp(iv, ny, out);
    const int mx = neumann_clamp(iu - sx, nx, out);
    const int my = neumann_clamp(iv - sx, ny, out);
    const int dx = neumann_clamp(iu + sx, nx, out);
    const int dy = neumann_clamp(iv + sy, ny, out);
    const int ddx = neumann_clamp(iu + 2 * sx, nx, out);
    const int ddy = neumann_clamp(iv + 2 * sy, ny, out);
    t.cols[0] = mx; t.cols[1] = x; t.cols[2] = dx; t.cols[3] = ddx;
    t.rows[0] = my; t.rows[1] = y; t.rows[2] = dy; t.rows[3] = ddy;
    t.fx = uu - (float)x;
    t.fy = vv - (float)y;
    t.out = out;
    return t;
}

// Sample `img` with precomputed taps: columns first along y, then along x
// (bicubic_interpolation_cell, bicubic_interpolation.cpp:53-65).
__device__ __forceinline__ float bicubic_sample(const float* __restrict__ img, int nx,
                                                const BicubicTaps& t)
{
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float* col = img + t.cols[c];
        v[c] = cubic_cell(col[(size_t)nx * t.rows[0]], col[(size_t)nx * t.rows[1]],
                          col[(size_t)nx * t.rows[2]], col[(size_t)nx * t.rows[3]], t.fy);
    }
    return cubic_cell(v[0], v[1], v[2], v[3], t.fx);
}

// The same sample of three images stored interleaved, one float4 per pixel
// (x, y, z used): one 16-byte gather per tap instead of three; each channel
// is evaluated exactly as bicubic_sample evaluates it.
__device__ __forceinline__ void bicubic_sample3(const float4* __restrict__ img, int nx,
                                                const BicubicTaps& t, float& a, float& b,
                                                float& c)
{
    float va[4], vb[4], vc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float4* col = img + t.cols[k];
        const float4 r0 = col[(size_t)nx * t.rows[0]], r1 = col[(size_t)nx * t.rows[1]];
        const float4 r2 = col[(size_t)nx * t.rows[2]], r3 = col[(size_t)nx * t.rows[3]];
        va[k] = cubic_cell(r0.x, r1.x, r2.x, r3.x, t.fy);
        vb[k] = cubic_cell(r0.y, r1.y, r2.y, r3.y, t.fy);
        vc[k] = cubic_cell(r0.z, r1.z, r2.z, r3.z, t.fy);
    }
    a = cubic_cell(va[0], va[1], va[2], va[3], t.fx);
    b = cubic_cell(vb[0], vb[1], vb[2], vb[3], t.fx);
    c = cubic_cell(vc[0], vc[1], vc[2], vc[3], t.fx);
}

__device__ __forceinline__ float bicubic_at(const float* __restrict__ img, float uu, float vv,
                                            int nx, int ny, bool border_out)
{
    const BicubicTaps t = bicubic_taps(uu, vv, nx, ny);
    if (t.out && border_out) return 0.0f;
    return bicubic_sample(img, nx, t);
}

// Whole-wave lane shifts through DPP (no LDS traffic).  from_left(v) gives
// lane i the value of lane i-1 and from_right(v) the value of lane i+1; the
// lane without a source keeps its own value.
__device__ __forceinline__ float from_left(float v)
{
    const int i = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(i, i, 0x138 /* wave_shr:1 */, 0xf, 0xf, false));
}

__device__ __forceinline__ float from_right(float v)
{
    const int i = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(i, i, 0x130 /* wave_shl:1 */, 0xf, 0xf, false));
}

// hypotf as glibc 2.35 evaluates it for finite arguments: the double sum of
// the exact double squares, square-rooted in double, rounded to float.
__device__ __forceinline__ float hypot_ref(float a, float b)
{
    return (float)__builtin_sqrt((double)a * (double)a + (double)b * (double)b);
}

// Correctly rounded double square root of s = 0 or s in [2^-767, 2^1023]:
// the hardware reciprocal square root refined by a Goldschmidt step and two
// Newton corrections with fused multiply-adds, the same iteration the
// compiler emits for sqrt (which first scales s below 2^-767 by 2^256 and
// passes 0 and infinity through, two steps that cannot apply here).
__device__ __forceinline__ double sqrt_unscaled(double s)
{
    const double r = __builtin_amdgcn_rsq(s);
    double g = s * r;
    double h = 0.5 * r;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    g = __builtin_fma(__builtin_fma(-g, g, s), h, g);
    g = __builtin_fma(__builtin_fma(-g, g, s), h, g);
    return s == 0.0 ? 0.0 : g;
}

// hypot_ref for finite a, b: their double sum of squares is 0 or lies in
// [2^-298, 2^257], inside sqrt_unscaled's range.
__device__ __forceinline__ float hypot_fin(float a, float b)
{
    return (float)sqrt_unscaled((double)a * (double)a + (double)b * (double)b);
}

}  // namespace amdflow
