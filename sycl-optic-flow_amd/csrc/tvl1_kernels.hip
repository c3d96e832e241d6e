// tvl1_kernels.hip -- gfx950 kernels of the TV-L1 solver.
//
// Arithmetic follows src/tvl1/sequential/{tvl1.cpp,mask/mask.cpp,zoom/zoom.cpp,
// bicubic_interpolation/bicubic_interpolation.cpp} operation by operation;
// the library is built with -ffp-contract=off so no multiply-add is fused.
#include "device_math.hpp"
#include "tvl1_kernels.hpp"
#include "tvl1_launch.hpp"

namespace amdflow {

// ---------------------------------------------------------------------------
// control-block set-up
// ---------------------------------------------------------------------------

// Start of a solve: every scale's current buffer is 0, min/max keys reset.
__global__ void k_ctl_init(TvCtl* ctl)
{
    if (threadIdx.x < kMaxScales) ctl->cur[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        ctl->n = 0;
        ctl->stop = 1;
        ctl->redo = 0;
        ctl->counter = 0;
        ctl->launches = 0;
        ctl->wlaunches = 0;
        ctl->maxkey[0] = ctl->maxkey[1] = 0ull;
        ctl->minkey[0] = ctl->minkey[1] = ~0ull;
    }
}

// Start of a warping without the warp set-up kernel (test entry points).
__global__ void k_ctl_begin(TvCtl* ctl)
{
    if (threadIdx.x == 0) {
        ctl->n = 0;
        ctl->stop = 0;
        ctl->redo = 0;
        ctl->counter = 0;
        ctl->wlaunches = 0;
    }
}

// ---------------------------------------------------------------------------
// input conversion + joint min/max (tvl1.cpp:160-165, 352-361)
// ---------------------------------------------------------------------------

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

// Converts both frames to float and finds, per frame, the first index of the
// largest and of the smallest |x| (cblas_isamax / cblas_isamin semantics).
template <typename T>
__global__ __launch_bounds__(256) void k_load_minmax(const T* __restrict__ a,
                                                     const T* __restrict__ b,
                                                     float* __restrict__ fa,
                                                     float* __restrict__ fb, unsigned n,
                                                     TvCtl* ctl)
{
    unsigned long long kmax0 = 0ull, kmax1 = 0ull, kmin0 = ~0ull, kmin1 = ~0ull;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float va = (float)a[i], vb = (float)b[i];
        fa[i] = va;
        fb[i] = vb;
        const unsigned long long ka = (unsigned long long)__float_as_uint(fabsf(va)) << 32;
        const unsigned long long kb = (unsigned long long)__float_as_uint(fabsf(vb)) << 32;
        const unsigned long long hi = 0xFFFFFFFFull - i;
        kmax0 = (ka | hi) > kmax0 ? (ka | hi) : kmax0;
        kmax1 = (kb | hi) > kmax1 ? (kb | hi) : kmax1;
        kmin0 = (ka | i) < kmin0 ? (ka | i) : kmin0;
        kmin1 = (kb | i) < kmin1 ? (kb | i) : kmin1;
    }
    kmax0 = wave_max_u64(kmax0);
    kmax1 = wave_max_u64(kmax1);
    kmin0 = wave_min_u64(kmin0);
    kmin1 = wave_min_u64(kmin1);
    // one atomic per key per workgroup: the grid is small (see launcher)
    __shared__ unsigned long long sk[4][4];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sk[0][wave] = kmax0;
        sk[1][wave] = kmax1;
        sk[2][wave] = kmin0;
        sk[3][wave] = kmin1;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        unsigned long long v = sk[k][0];
        for (int w = 1; w < 4; w++) v = (k < 2) ? (sk[k][w] > v ? sk[k][w] : v) : (sk[k][w] < v ? sk[k][w] : v);
        if (k == 0) atomicMax(&ctl->maxkey[0], v);
        if (k == 1) atomicMax(&ctl->maxkey[1], v);
        if (k == 2) atomicMin(&ctl->minkey[0], v);
        if (k == 3) atomicMin(&ctl->minkey[1], v);
    }
}

// image_normalization, tvl1.cpp:344-371: max = std::max of the two isamax
// picks, min = std::min of the two isamin picks, den = max - min; when
// den <= 0 the frames are copied unchanged, else 255*(I-min)/den in double.
__global__ __launch_bounds__(256) void k_normalize(const float* __restrict__ a,
                                                   const float* __restrict__ b,
                                                   float* __restrict__ oa,
                                                   float* __restrict__ ob, unsigned n,
                                                   const TvCtl* ctl)
{
    const unsigned imax0 = 0xFFFFFFFFu - (unsigned)(ctl->maxkey[0] & 0xFFFFFFFFull);
    const unsigned imax1 = 0xFFFFFFFFu - (unsigned)(ctl->maxkey[1] & 0xFFFFFFFFull);
    const unsigned imin0 = (unsigned)(ctl->minkey[0] & 0xFFFFFFFFull);
    const unsigned imin1 = (unsigned)(ctl->minkey[1] & 0xFFFFFFFFull);
    const float a0 = a[imax0], a1 = b[imax1];
    const float b0 = a[imin0], b1 = b[imin1];
    const float mx = (a0 < a1) ? a1 : a0;
    const float mn = (b1 < b0) ? b1 : b0;
    const float den = mx - mn;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (den <= 0) {
            oa[i] = a[i];
            ob[i] = b[i];
        } else {
            oa[i] = (float)(255.0 * (double)(a[i] - mn) / (double)den);
            ob[i] = (float)(255.0 * (double)(b[i] - mn) / (double)den);
        }
    }
}

// ---------------------------------------------------------------------------
// separable Gaussian (mask.cpp:217-287)
// ---------------------------------------------------------------------------

// Low side: whole-sample mirror x[-k] = x[k]; high side: half-sample mirror
// x[n-1+k] = x[n-k] (mask.cpp:253-255, 274-276).
__device__ __forceinline__ int mirror_lo(int i) { return i < 0 ? -i : i; }
__device__ __forceinline__ int mirror_hi(int i, int n) { return i >= n ? 2 * n - 1 - i : i; }

__global__ __launch_bounds__(256) void k_gauss_rows(const float* __restrict__ src,
                                                    float* __restrict__ dst, int nx, int ny,
                                                    GaussTaps g)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= nx || y >= ny) return;
    const float* row = src + (size_t)y * nx;
    double sum = g.B[0] * (double)row[x];
    for (int j = 1; j < g.size; j++) {
        const float pair = row[mirror_lo(x - j)] + row[mirror_hi(x + j, nx)];
        sum += g.B[j] * (double)pair;
    }
    dst[(size_t)y * nx + x] = (float)sum;
}

__global__ __launch_bounds__(256) void k_gauss_cols(const float* __restrict__ src,
                                                    float* __restrict__ dst, int nx, int ny,
                                                    GaussTaps g)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= nx || y >= ny) return;
    const float* col = src + x;
    double sum = g.B[0] * (double)col[(size_t)y * nx];
    for (int j = 1; j < g.size; j++) {
        const float pair = col[(size_t)mirror_lo(y - j) * nx] + col[(size_t)mirror_hi(y + j, ny) * nx];
        sum += g.B[j] * (double)pair;
    }
    dst[(size_t)y * nx + x] = (float)sum;
}

// ---------------------------------------------------------------------------
// resampling (zoom.cpp:20-85)
// ---------------------------------------------------------------------------

// zoom_out's resampling of the blurred image at (j/f, i/f), zoom.cpp:28-52.
__global__ __launch_bounds__(256) void k_zoom_out_sample(const float* __restrict__ src,
                                                         float* __restrict__ dst, int nx, int ny,
                                                         int nxx, int nyy, float factor)
{
    const int j1 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i1 = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j1 >= nxx || i1 >= nyy) return;
    const float i2 = (float)i1 / factor;
    const float j2 = (float)j1 / factor;
    dst[(size_t)i1 * nxx + j1] = bicubic_at(src, j2, i2, nx, ny, false);
}

// zoom_in of both flow components (zoom.cpp:61-85) fused with the
// cblas_sscal by 1/zfactor (tvl1.cpp:213-214).  Reads the coarse scale's
// current buffer, writes buffer 0 of the finer scale.
__global__ __launch_bounds__(256) void k_zoom_in_flow(const float* __restrict__ uc0,
                                                      const float* __restrict__ uc1,
                                                      float* __restrict__ uf, int nx, int ny,
                                                      int nxx, int nyy, float factorx,
                                                      float factory, float invz,
                                                      const TvCtl* ctl, int coarse_level)
{
    const int j1 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i1 = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j1 >= nxx || i1 >= nyy) return;
    const float* uc = ctl->cur[coarse_level] ? uc1 : uc0;
    const size_t nc = (size_t)nx * ny, nf = (size_t)nxx * nyy;
    const float i2 = (float)i1 / factory;
    const float j2 = (float)j1 / factorx;
    const BicubicTaps t = bicubic_taps(j2, i2, nx, ny);
    const float a = bicubic_sample(uc, nx, t);
    const float b = bicubic_sample(uc + nc, nx, t);
    uf[(size_t)i1 * nxx + j1] = invz * a;
    uf[nf + (size_t)i1 * nxx + j1] = invz * b;
}

// zoom_in of two separate planes times `scale` (used by the pyramidal LK).
__global__ __launch_bounds__(256) void k_zoom_in_planes(const float* __restrict__ a,
                                                        const float* __restrict__ b,
                                                        float* __restrict__ oa,
                                                        float* __restrict__ ob, int nx, int ny,
                                                        int nxx, int nyy, float factorx,
                                                        float factory, float scale)
{
    const int j1 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i1 = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j1 >= nxx || i1 >= nyy) return;
    const float i2 = (float)i1 / factory;
    const float j2 = (float)j1 / factorx;
    const BicubicTaps t = bicubic_taps(j2, i2, nx, ny);
    oa[(size_t)i1 * nxx + j1] = scale * bicubic_sample(a, nx, t);
    ob[(size_t)i1 * nxx + j1] = scale * bicubic_sample(b, nx, t);
}

// ---------------------------------------------------------------------------
// centred gradient (mask.cpp:152-209)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_centered_gradient(const float* __restrict__ in,
                                                           float* __restrict__ dx,
                                                           float* __restrict__ dy, int nx, int ny)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= nx || i >= ny) return;
    const int jl = j > 0 ? j - 1 : 0, jr = j < nx - 1 ? j + 1 : nx - 1;
    const int iu = i > 0 ? i - 1 : 0, id = i < ny - 1 ? i + 1 : ny - 1;
    const float ddx = in[(size_t)i * nx + jr] - in[(size_t)i * nx + jl];
    const float ddy = in[(size_t)id * nx + j] - in[(size_t)iu * nx + j];
    dx[(size_t)i * nx + j] = (float)(0.5 * (double)ddx);
    dy[(size_t)i * nx + j] = (float)(0.5 * (double)ddy);
}

__global__ __launch_bounds__(256) void k_fill(float* __restrict__ p, unsigned n, float v)
{
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// warping set-up (tvl1.cpp:237-254)
// ---------------------------------------------------------------------------

// Warps I1, I1x, I1y by the current flow with border_out = true (one shared
// tap geometry, three gathers) and stores I1wx, I1wy and
// rho_c = I1w - I1wx*u1 - I1wy*u2 - I0.  |grad I1w|^2 is recomputed by the
// iteration kernel with the same expression.  Resets the per-warping control.
__global__ __launch_bounds__(256) void k_warp_setup(const float* __restrict__ I0,
                                                    const float* __restrict__ I1,
                                                    const float* __restrict__ I1x,
                                                    const float* __restrict__ I1y,
                                                    const float* __restrict__ ub0,
                                                    const float* __restrict__ ub1,
                                                    float* __restrict__ wx, float* __restrict__ wy,
                                                    float* __restrict__ rhoc, int nx, int ny,
                                                    TvCtl* ctl, int level)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        ctl->n = 0;
        ctl->stop = 0;
        ctl->redo = 0;
        ctl->counter = 0;
        ctl->wlaunches = 0;
    }
    if (j >= nx || i >= ny) return;
    const size_t plane = (size_t)nx * ny;
    const float* ub = ctl->cur[level] ? ub1 : ub0;
    const size_t p = (size_t)i * nx + j;
    const float u1 = ub[p], u2 = ub[plane + p];
    const float uu = (float)j + u1;
    const float vv = (float)i + u2;
    const BicubicTaps t = bicubic_taps(uu, vv, nx, ny);
    float w = 0.0f, gx = 0.0f, gy = 0.0f;
    if (!t.out) {
        w = bicubic_sample(I1, nx, t);
        gx = bicubic_sample(I1x, nx, t);
        gy = bicubic_sample(I1y, nx, t);
    }
    wx[p] = gx;
    wy[p] = gy;
    rhoc[p] = ((w - gx * u1) - gy * u2) - I0[p];
}

// Copies the finest scale's current flow into the caller's planar buffer.
__global__ __launch_bounds__(256) void k_copy_flow(const float* __restrict__ ub0,
                                                   const float* __restrict__ ub1,
                                                   float* __restrict__ out, unsigned n2,
                                                   const TvCtl* ctl)
{
    const float* ub = ctl->cur[0] ? ub1 : ub0;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += gridDim.x * blockDim.x)
        out[i] = ub[i];
}

// ---------------------------------------------------------------------------
// primal-dual iterations (tvl1.cpp:256-334), K iterations per launch
// ---------------------------------------------------------------------------

// Divergence at one pixel from the current value and the left/up neighbours
// with the border rules of mask.cpp:44-90 (first/last column cases
// associate left to right as written there).
// Written with selects instead of branches; every candidate is the exact
// expression of its reference case, so the selected value is bit-identical.
__device__ __forceinline__ float div_rule(float v1, float v1l, float v2, float v2u, bool fr,
                                          bool lr, bool fc, bool lc)
{
    const float a = fc ? v1 : (lc ? -v1l : (v1 - v1l));
    const float t_first_row = a + v2;               // :70, corners :86-87
    const float t_last_row = a - v2u;               // :71, corners :88-89
    const float t_edge_col = (a + v2) - v2u;        // :81-82
    const float t_body = a + (v2 - v2u);            // :59
    return fr ? t_first_row : (lr ? t_last_row : ((fc || lc) ? t_edge_col : t_body));
}

// Thresholding operator TH, tvl1.cpp:263-287, branch-free: the three
// reference cases multiply the warped gradient by l_t, -l_t or -rho/grad,
// and the `grad < 1e-10` case stores +0 exactly as the reference does.
__device__ __forceinline__ void th_step(float rc, float gx, float gy, float u1, float u2,
                                        float l_t, float& v1, float& v2)
{
    const float grad = (gx * gx) + (gy * gy);
    const float rho = rc + ((gx * u1) + (gy * u2));
    const float fi = -rho / grad;
    const bool c1 = rho >= -l_t * grad;
    const bool c2 = rho > l_t * grad;
    const bool c3 = grad < 1E-10f;
    const bool zero = c1 && !c2 && c3;
    const float k = c1 ? (c2 ? -l_t : fi) : l_t;
    const float d1 = zero ? 0.0f : k * gx;
    const float d2 = zero ? 0.0f : k * gy;
    v1 = u1 + d1;
    v2 = u2 + d2;
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One workgroup = 4 waves; the workgroup's extended tile is 64 columns (one
// per lane) by 4*R rows (R consecutive rows per wave, held in registers).
// Horizontal neighbours come from lane shuffles, vertical ones from the same
// lane's registers or, across waves, from one LDS row.  A tile computes K
// iterations; values within K pixels of the tile edge go stale one pixel per
// iteration, so only the central (64-2K) x (4R-2K) pixels are stored.
//
// Stopping rule, evaluated by the last workgroup to finish: the warping stops
// after the first iteration whose mean |u^n - u^(n-1)|^2 is not above
// epsilon^2, or after max_iter iterations (tvl1.cpp:256-258).  When that
// first iteration falls inside the launch before its last iteration, the
// launch's output is discarded and the next launch re-runs exactly that many
// iterations from the same input buffer.
template <int K, int R>
__global__ __launch_bounds__(256) void k_tvl1_iter(IterArgs a)
{
    constexpr int NW = 4;
    constexpr int TW = 64 - 2 * K;
    constexpr int TH = NW * R - 2 * K;
    __shared__ float sh_a[NW][64], sh_b[NW][64], sh_c[NW][64], sh_d[NW][64];
    __shared__ double sh_err[NW][K];
    __shared__ double sh_tot[K];
    __shared__ int sh_last;

    TvCtl* ctl = a.ctl;
    if (ctl->stop) return;
    const int cur = ctl->cur[a.level];
    const int redo = ctl->redo;
    const int n0 = ctl->n;
    int kk = redo ? redo : a.max_iter - n0;
    if (kk > K) kk = K;
    if (kk <= 0) return;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
    const int nx = a.nx, ny = a.ny;
    const int x = (int)blockIdx.x * TW - K + lane;
    const int ybase = (int)blockIdx.y * TH - K + wave * R;
    const int oy0 = (int)blockIdx.y * TH;
    const bool xin = x >= 0 && x < nx;
    const bool xout = xin && lane >= K && lane < K + TW;
    const bool fc = x == 0, lc = x == nx - 1;
    // waves holding no first/last image column take the body formula of the
    // divergence on every row away from the first/last image row
    const bool edge_cols = __any(fc || lc);

    const float* uin = a.u[cur];
    const float* pin = a.p[cur];
    const size_t plane = (size_t)a.plane;

    float u1[R], u2[R], p11[R], p12[R], p21[R], p22[R], gx[R], gy[R], rc[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = ybase + r;
        if (xin && y >= 0 && y < ny) {
            const size_t idx = (size_t)y * nx + x;
            u1[r] = uin[idx];
            u2[r] = uin[plane + idx];
            p11[r] = pin[idx];
            p12[r] = pin[plane + idx];
            p21[r] = pin[2 * plane + idx];
            p22[r] = pin[3 * plane + idx];
            gx[r] = a.wx[idx];
            gy[r] = a.wy[idx];
            rc[r] = a.rhoc[idx];
        } else {
            u1[r] = u2[r] = p11[r] = p12[r] = p21[r] = p22[r] = 0.0f;
            gx[r] = gy[r] = rc[r] = 0.0f;
        }
    }

    double acc[K];
#pragma unroll
    for (int t = 0; t < K; t++) acc[t] = 0.0;

    const float l_t = a.l_t, theta = a.theta, taut = a.taut;
#pragma unroll
    for (int t = 0; t < K; t++) {
        if (t < kk) {
            // ---- u update: TH, divergence of p, u = v + theta*div p ----
            sh_a[wave][lane] = p12[R - 1];
            sh_b[wave][lane] = p22[R - 1];
            __syncthreads();
            const float up12 = wave ? sh_a[wave - 1][lane] : p12[0];
            const float up22 = wave ? sh_b[wave - 1][lane] : p22[0];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int y = ybase + r;
                const float p11l = from_left(p11[r]);
                const float p21l = from_left(p21[r]);
                const float p12u = r ? p12[r - 1] : up12;
                const float p22u = r ? p22[r - 1] : up22;
                const bool fr = y == 0, lr = y == ny - 1;
                float div1, div2;
                if (!edge_cols && !fr && !lr) {   // uniform branch: interior rows of interior waves
                    div1 = (p11[r] - p11l) + (p12[r] - p12u);
                    div2 = (p21[r] - p21l) + (p22[r] - p22u);
                } else {
                    div1 = div_rule(p11[r], p11l, p12[r], p12u, fr, lr, fc, lc);
                    div2 = div_rule(p21[r], p21l, p22[r], p22u, fr, lr, fc, lc);
                }
                float v1, v2;
                th_step(rc[r], gx[r], gy[r], u1[r], u2[r], l_t, v1, v2);
                const float nu1 = v1 + theta * div1;
                const float nu2 = v2 + theta * div2;
                const float e1 = nu1 - u1[r], e2 = nu2 - u2[r];
                const float e = e1 * e1 + e2 * e2;
                const bool mine = xout && y >= oy0 && y < oy0 + TH && y < ny;
                acc[t] += mine ? (double)e : 0.0;
                u1[r] = nu1;
                u2[r] = nu2;
            }
            // ---- p update: forward gradient of u, g, p = (p + taut*grad)/g ----
            sh_c[wave][lane] = u1[0];
            sh_d[wave][lane] = u2[0];
            __syncthreads();
            const float dn1 = wave < NW - 1 ? sh_c[wave + 1][lane] : u1[R - 1];
            const float dn2 = wave < NW - 1 ? sh_d[wave + 1][lane] : u2[R - 1];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int y = ybase + r;
                const float u1r = from_right(u1[r]);
                const float u2r = from_right(u2[r]);
                const float u1d = r < R - 1 ? u1[r + 1] : dn1;
                const float u2d = r < R - 1 ? u2[r + 1] : dn2;
                const bool xl = x < nx - 1, yl = y < ny - 1;
                const float fx1 = xl ? u1r - u1[r] : 0.0f;
                const float fy1 = yl ? u1d - u1[r] : 0.0f;
                const float fx2 = xl ? u2r - u2[r] : 0.0f;
                const float fy2 = yl ? u2d - u2[r] : 0.0f;
                const float g1 = 1.0f + taut * hypot_ref(fx1, fy1);
                const float g2 = 1.0f + taut * hypot_ref(fx2, fy2);
                p11[r] = (p11[r] + taut * fx1) / g1;
                p12[r] = (p12[r] + taut * fy1) / g1;
                p21[r] = (p21[r] + taut * fx2) / g2;
                p22[r] = (p22[r] + taut * fy2) / g2;
            }
        }
    }

    // ---- store the central tile into the other buffer ----
    float* uout = a.u[cur ^ 1];
    float* pout = a.p[cur ^ 1];
    if (xout) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int y = ybase + r;
            if (y >= oy0 && y < oy0 + TH && y < ny) {
                const size_t idx = (size_t)y * nx + x;
                uout[idx] = u1[r];
                uout[plane + idx] = u2[r];
                pout[idx] = p11[r];
                pout[plane + idx] = p12[r];
                pout[2 * plane + idx] = p21[r];
                pout[3 * plane + idx] = p22[r];
            }
        }
    }

    // ---- per-workgroup error partials, fixed reduction order ----
    const unsigned nblocks = gridDim.x * gridDim.y;
    const unsigned bid = blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int t = 0; t < K; t++) {
        const double v = wave_sum_f64(acc[t]);
        if (lane == 0) sh_err[wave][t] = v;
    }
    __syncthreads();
    // The partials are published with device-scope atomic stores, which are
    // written through to the point of coherence of all XCDs, and the wave
    // waits for their completion (s_waitcnt vmcnt(0)) before the counter
    // increment; the last workgroup reads them with device-scope atomic
    // loads.  A device-scope release fence would instead write back the
    // XCD's whole L2 at the end of every workgroup.
    if (threadIdx.x < (unsigned)kk)
        __hip_atomic_store(a.partials + (size_t)threadIdx.x * nblocks + bid,
                           ((sh_err[0][threadIdx.x] + sh_err[1][threadIdx.x]) +
                            sh_err[2][threadIdx.x]) + sh_err[3][threadIdx.x],
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    // ---- the last workgroup reduces the partials and updates the control ----
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (threadIdx.x == 0) sh_last = atomicAdd(&ctl->counter, 1u) == nblocks - 1;
    __syncthreads();
    if (!sh_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (int t = 0; t < kk; t++) {
        double s = 0.0;
        const double* col = a.partials + (size_t)t * nblocks;
        for (unsigned j = threadIdx.x; j < nblocks; j += blockDim.x)
            s += __hip_atomic_load(col + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s = wave_sum_f64(s);
        if (lane == 0) sh_err[wave][0] = s;
        __syncthreads();
        if (threadIdx.x == 0) sh_tot[t] = ((sh_err[0][0] + sh_err[1][0]) + sh_err[2][0]) + sh_err[3][0];
        __syncthreads();
    }
    if (a.err_log && threadIdx.x < (unsigned)kk) a.err_log[n0 + threadIdx.x] = sh_tot[threadIdx.x];
    if (threadIdx.x == 0) {
        int n = n0, stop = 0, nredo = 0, flip = 0;
        if (redo) {
            n = n0 + redo;
            stop = 1;
            flip = 1;
        } else {
            int st = -1;
            for (int t = 0; t < kk; t++) {
                const double err = sh_tot[t] / (double)a.plane;
                if (!(err > a.eps2)) { st = t; break; }
            }
            if (st == kk - 1 || (st < 0 && n0 + kk >= a.max_iter)) {
                n = n0 + kk;
                stop = 1;
                flip = 1;
            } else if (st >= 0) {
                nredo = st + 1;
            } else {
                n = n0 + kk;
                flip = 1;
            }
        }
        ctl->n = n;
        ctl->redo = nredo;
        ctl->stop = stop;
        if (flip) ctl->cur[a.level] = cur ^ 1;
        if (stop) *a.iters_slot = n;
        ctl->counter = 0;
        ctl->launches = ctl->launches + 1;
        const unsigned wl = ctl->wlaunches + 1;
        ctl->wlaunches = wl;
        const unsigned long long word = ((unsigned long long)a.epoch << 32) |
                                        ((unsigned long long)stop << 31) |
                                        (unsigned long long)(wl & 0x7FFFFFFFu);
        // the host only reads this word itself: no ordering with other data
        if (a.host_status)
            __hip_atomic_store(a.host_status, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ __launch_bounds__(256) void k_hypot(const float* __restrict__ a,
                                               const float* __restrict__ b,
                                               float* __restrict__ o, unsigned n)
{
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        o[i] = hypot_ref(a[i], b[i]);
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

bool iter_config_ok(int K, int R)
{
    if (!(K == 1 || K == 2 || K == 4 || K == 8)) return false;
    if (!(R == 2 || R == 4 || R == 8)) return false;
    return 4 * R - 2 * K >= 4;
}

void iter_grid(int K, int R, int nx, int ny, unsigned* gx, unsigned* gy)
{
    *gx = cdiv((size_t)nx, (size_t)(64 - 2 * K));
    *gy = cdiv((size_t)ny, (size_t)(4 * R - 2 * K));
}

template <int K, int R>
static hipError_t launch_kr(const IterArgs& a, hipStream_t s)
{
    unsigned gx, gy;
    iter_grid(K, R, a.nx, a.ny, &gx, &gy);
    hipLaunchKernelGGL((k_tvl1_iter<K, R>), dim3(gx, gy), dim3(256), 0, s, a);
    return hipGetLastError();
}

int launch_iter(int K, int R, const IterArgs& a, hipStream_t s)
{
    hipError_t e = hipErrorInvalidValue;
#define OF_CASE(k, r) \
    if (K == k && R == r) e = launch_kr<k, r>(a, s);
    OF_CASE(1, 2) OF_CASE(2, 2)
    OF_CASE(1, 4) OF_CASE(2, 4) OF_CASE(4, 4)
    OF_CASE(1, 8) OF_CASE(2, 8) OF_CASE(4, 8) OF_CASE(8, 8)
#undef OF_CASE
    if (e != hipSuccess) {
        record_hip_error(e, "k_tvl1_iter launch", __FILE__, __LINE__);
        return OF_ERR_HIP;
    }
    return OF_OK;
}

static inline dim3 grid2d(int nx, int ny) { return dim3(cdiv(nx, 64), cdiv(ny, 4)); }
static inline unsigned grid1d(size_t n) { unsigned g = cdiv(n, 256); return g > 4096 ? 4096 : (g ? g : 1); }

#define OF_LAUNCH_CHECK(name)                                       \
    do {                                                            \
        hipError_t e_ = hipGetLastError();                          \
        if (e_ != hipSuccess) {                                     \
            record_hip_error(e_, name, __FILE__, __LINE__);         \
            return OF_ERR_HIP;                                      \
        }                                                           \
    } while (0)

int launch_ctl_init(TvCtl* ctl, hipStream_t s)
{
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, s, ctl);
    OF_LAUNCH_CHECK("k_ctl_init");
    return OF_OK;
}

int launch_ctl_begin(TvCtl* ctl, hipStream_t s)
{
    hipLaunchKernelGGL(k_ctl_begin, dim3(1), dim3(64), 0, s, ctl);
    OF_LAUNCH_CHECK("k_ctl_begin");
    return OF_OK;
}

int launch_load_minmax_u8(const uint8_t* a, const uint8_t* b, float* fa, float* fb, size_t n,
                          TvCtl* ctl, hipStream_t s)
{
    hipLaunchKernelGGL(k_load_minmax<uint8_t>, dim3(std::min(grid1d(n), 1024u)), dim3(256), 0, s, a, b, fa, fb,
                       (unsigned)n, ctl);
    OF_LAUNCH_CHECK("k_load_minmax<u8>");
    return OF_OK;
}

int launch_load_minmax_f32(const float* a, const float* b, float* fa, float* fb, size_t n,
                           TvCtl* ctl, hipStream_t s)
{
    hipLaunchKernelGGL(k_load_minmax<float>, dim3(std::min(grid1d(n), 1024u)), dim3(256), 0, s, a, b, fa, fb,
                       (unsigned)n, ctl);
    OF_LAUNCH_CHECK("k_load_minmax<f32>");
    return OF_OK;
}

int launch_normalize(const float* a, const float* b, float* oa, float* ob, size_t n,
                     const TvCtl* ctl, hipStream_t s)
{
    hipLaunchKernelGGL(k_normalize, dim3(grid1d(n)), dim3(256), 0, s, a, b, oa, ob, (unsigned)n,
                       ctl);
    OF_LAUNCH_CHECK("k_normalize");
    return OF_OK;
}

int launch_gaussian(const float* src, float* tmp, float* dst, int nx, int ny, const GaussTaps& g,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_gauss_rows, grid2d(nx, ny), dim3(256), 0, s, src, tmp, nx, ny, g);
    OF_LAUNCH_CHECK("k_gauss_rows");
    hipLaunchKernelGGL(k_gauss_cols, grid2d(nx, ny), dim3(256), 0, s, tmp, dst, nx, ny, g);
    OF_LAUNCH_CHECK("k_gauss_cols");
    return OF_OK;
}

int launch_zoom_out_sample(const float* src, float* dst, int nx, int ny, int nxx, int nyy,
                           float factor, hipStream_t s)
{
    hipLaunchKernelGGL(k_zoom_out_sample, grid2d(nxx, nyy), dim3(256), 0, s, src, dst, nx, ny,
                       nxx, nyy, factor);
    OF_LAUNCH_CHECK("k_zoom_out_sample");
    return OF_OK;
}

int launch_zoom_in_flow(const float* uc0, const float* uc1, float* uf, int nx, int ny, int nxx,
                        int nyy, float invz, const TvCtl* ctl, int coarse_level, hipStream_t s)
{
    const float factorx = ((float)nxx / (float)nx);
    const float factory = ((float)nyy / (float)ny);
    hipLaunchKernelGGL(k_zoom_in_flow, grid2d(nxx, nyy), dim3(256), 0, s, uc0, uc1, uf, nx, ny,
                       nxx, nyy, factorx, factory, invz, ctl, coarse_level);
    OF_LAUNCH_CHECK("k_zoom_in_flow");
    return OF_OK;
}

int launch_zoom_in_planes(const float* a, const float* b, float* oa, float* ob, int nx, int ny,
                          int nxx, int nyy, float scale, hipStream_t s)
{
    const float factorx = ((float)nxx / (float)nx);
    const float factory = ((float)nyy / (float)ny);
    hipLaunchKernelGGL(k_zoom_in_planes, grid2d(nxx, nyy), dim3(256), 0, s, a, b, oa, ob, nx, ny,
                       nxx, nyy, factorx, factory, scale);
    OF_LAUNCH_CHECK("k_zoom_in_planes");
    return OF_OK;
}

int launch_centered_gradient(const float* in, float* dx, float* dy, int nx, int ny, hipStream_t s)
{
    hipLaunchKernelGGL(k_centered_gradient, grid2d(nx, ny), dim3(256), 0, s, in, dx, dy, nx, ny);
    OF_LAUNCH_CHECK("k_centered_gradient");
    return OF_OK;
}

int launch_fill(float* p, size_t n, float v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(grid1d(n)), dim3(256), 0, s, p, (unsigned)n, v);
    OF_LAUNCH_CHECK("k_fill");
    return OF_OK;
}

int launch_warp_setup(const float* I0, const float* I1, const float* I1x, const float* I1y,
                      const float* ub0, const float* ub1, float* wx, float* wy, float* rhoc,
                      int nx, int ny, TvCtl* ctl, int level, hipStream_t s)
{
    hipLaunchKernelGGL(k_warp_setup, grid2d(nx, ny), dim3(256), 0, s, I0, I1, I1x, I1y, ub0, ub1,
                       wx, wy, rhoc, nx, ny, ctl, level);
    OF_LAUNCH_CHECK("k_warp_setup");
    return OF_OK;
}

int launch_copy_flow(const float* ub0, const float* ub1, float* out, size_t n2, const TvCtl* ctl,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_flow, dim3(grid1d(n2)), dim3(256), 0, s, ub0, ub1, out,
                       (unsigned)n2, ctl);
    OF_LAUNCH_CHECK("k_copy_flow");
    return OF_OK;
}

int launch_hypot(const float* a, const float* b, float* o, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_hypot, dim3(grid1d(n)), dim3(256), 0, s, a, b, o, (unsigned)n);
    OF_LAUNCH_CHECK("k_hypot");
    return OF_OK;
}

}  // namespace amdflow
