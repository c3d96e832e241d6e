// tvl1_kernels.hip -- gfx950 kernels of the TV-L1 solver.
//
// Arithmetic follows src/tvl1/sequential/{tvl1.cpp,mask/mask.cpp,zoom/zoom.cpp,
// bicubic_interpolation/bicubic_interpolation.cpp} operation by operation;
// the library is built with -ffp-contract=off so no multiply-add is fused.
#include "device_math.hpp"
#include "tvl1_kernels.hpp"
#include "tvl1_pixel.hpp"
#include "tvl1_launch.hpp"

namespace amdflow {

// ---------------------------------------------------------------------------
// control-block set-up
// ---------------------------------------------------------------------------

// Start of a solve: every scale's current buffer is 0, min/max keys reset.
// Also clears `iters[0..n_iters)` (the per-warping iteration counts), when
// given.
__global__ void k_ctl_init(TvCtl* ctl, int* iters, int n_iters)
{
    for (int k = threadIdx.x; k < n_iters; k += blockDim.x) iters[k] = 0;
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
        ctl->bar_count = 0;
        for (int g = 0; g < kBarGroups; g++) ctl->grp_count[g].v = 0;
        ctl->chunk_count[0].v = ctl->chunk_count[1].v = 0;
        ctl->decision[0].v = ctl->decision[1].v = 0ull;
        ctl->error = 0;
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
        ctl->chunk_count[0].v = ctl->chunk_count[1].v = 0;
    }
}

// ---------------------------------------------------------------------------
// input conversion + joint min/max (tvl1.cpp:118-123, 310-319)
// ---------------------------------------------------------------------------

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
        minmax_keys(va, vb, i, kmax0, kmax1, kmin0, kmin1);
    }
    kmax0 = wave_max_u64(kmax0);
    kmax1 = wave_max_u64(kmax1);
    kmin0 = wave_min_u64(kmin0);
    kmin1 = wave_min_u64(kmin1);
    // per-workgroup keys, reduced by k_minmax_final (no atomics on 4 words
    // from a thousand workgroups)
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
        ctl->mmpart[k][blockIdx.x] = v;
    }
}

// Combines the per-workgroup keys of k_load_minmax (one workgroup).
__global__ __launch_bounds__(256) void k_minmax_final(TvCtl* ctl, unsigned nblocks)
{
    unsigned long long kmax0 = 0ull, kmax1 = 0ull, kmin0 = ~0ull, kmin1 = ~0ull;
    for (unsigned i = threadIdx.x; i < nblocks; i += 256) {
        kmax0 = ctl->mmpart[0][i] > kmax0 ? ctl->mmpart[0][i] : kmax0;
        kmax1 = ctl->mmpart[1][i] > kmax1 ? ctl->mmpart[1][i] : kmax1;
        kmin0 = ctl->mmpart[2][i] < kmin0 ? ctl->mmpart[2][i] : kmin0;
        kmin1 = ctl->mmpart[3][i] < kmin1 ? ctl->mmpart[3][i] : kmin1;
    }
    kmax0 = wave_max_u64(kmax0);
    kmax1 = wave_max_u64(kmax1);
    kmin0 = wave_min_u64(kmin0);
    kmin1 = wave_min_u64(kmin1);
    __shared__ unsigned long long sk[4][4];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sk[0][wave] = kmax0;
        sk[1][wave] = kmax1;
        sk[2][wave] = kmin0;
        sk[3][wave] = kmin1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long v[4];
        for (int k = 0; k < 4; k++) {
            v[k] = sk[k][0];
            for (int w = 1; w < 4; w++)
                v[k] = (k < 2) ? (sk[k][w] > v[k] ? sk[k][w] : v[k]) : (sk[k][w] < v[k] ? sk[k][w] : v[k]);
        }
        ctl->maxkey[0] = v[0];
        ctl->maxkey[1] = v[1];
        ctl->minkey[0] = v[2];
        ctl->minkey[1] = v[3];
    }
}

// image_normalization, tvl1.cpp:302-329: max = std::max of the two isamax
// picks, min = std::min of the two isamin picks, den = max - min; when
// den <= 0 the frames are copied unchanged, else 255*(I-min)/den in double.
__global__ __launch_bounds__(256) void k_normalize(const float* __restrict__ a,
                                                   const float* __restrict__ b,
                                                   float* __restrict__ oa,
                                                   float* __restrict__ ob, unsigned n,
                                                   const TvCtl* ctl)
{
    float mn, den;
    norm_bounds(ctl, a, b, &mn, &den);
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        oa[i] = norm_px(a[i], mn, den);
        ob[i] = norm_px(b[i], mn, den);
    }
}

// ---------------------------------------------------------------------------
// separable Gaussian (mask.cpp:204-274)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_gauss_rows(const float* __restrict__ src,
                                                    float* __restrict__ dst, int nx, int ny,
                                                    GaussTaps g)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= nx || y >= ny) return;
    dst[(size_t)y * nx + x] = gauss_row_px(src, nx, x, y, g);
}

__global__ __launch_bounds__(256) void k_gauss_cols(const float* __restrict__ src,
                                                    float* __restrict__ dst, int nx, int ny,
                                                    GaussTaps g)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= nx || y >= ny) return;
    dst[(size_t)y * nx + x] = gauss_col_px(src, nx, ny, x, y, g);
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
    dst[(size_t)i1 * nxx + j1] = zoom_out_px(src, nx, ny, i1, j1, factor);
}

// Tiled separable Gaussian of two frames per launch (blockIdx.z = frame).
// Output tile: 64 columns x TR rows of an output image whose pixel (oj, oi)
// is the blurred input at (STEP*oj, STEP*oi):
//   STEP 1: the blurred image (the pre-smoothing, tvl1.cpp:143-144); with
//           NORM the input is first normalised as image_normalization does
//           (tvl1.cpp:302-329), sample by sample;
//   STEP 2: zoom_out by 1/2 (zoom.cpp:20-53): the bicubic at an integer
//           position returns the sample there, so the kept samples are the
//           blurred image at even positions and only those are blurred.
// Every source sample the tile reads (rows and columns outside the image
// hold the mirrored sample, as mask.cpp:240-242, 261-263 read them) is
// staged in LDS first, all loads of a thread issued before the first is
// stored; then the row pass runs at the STEP*oj columns of every staged
// row, and the column pass at the STEP*oi rows: the per-pixel arithmetic of
// k_gauss_rows, k_gauss_cols and k_zoom_out_sample, so the same values (an
// exact zero may differ in sign from the bicubic's, which only matters for
// inputs holding -0).  H is the largest half-width the instance stages
// (taps of at most H + 1).
constexpr int kBlurHalo = 15;

template <int TR, int STEP, bool NORM, int H>
__global__ __launch_bounds__(256) void k_blur_tile(const float* __restrict__ a,
                                                   const float* __restrict__ b,
                                                   float* __restrict__ oa, float* __restrict__ ob,
                                                   int nx, int ny, int onx, int ony, GaussTaps g,
                                                   const TvCtl* ctl)
{
    constexpr int LR = STEP * (TR - 1) + 1 + 2 * H;    // staged rows
    constexpr int SW = STEP * 63 + 1 + 2 * H;          // staged columns
    constexpr int KR = (LR + 3) / 4, KC = (SW + 63) / 64;
    __shared__ float S[LR][SW];
    __shared__ float R[LR][64];
    const float* src = blockIdx.z ? b : a;
    float* dst = blockIdx.z ? ob : oa;
    const int oj0 = blockIdx.x * 64, oi0 = blockIdx.y * TR;
    const int hw = g.size - 1;
    const int oilast = min(oi0 + TR, ony) - 1;
    const int ojlast = min(oj0 + 64, onx) - 1;
    const int rlo = STEP * oi0 - hw, nrows = STEP * (oilast - oi0) + 1 + 2 * hw;
    const int xlo = STEP * oj0 - hw, ncols = STEP * (ojlast - oj0) + 1 + 2 * hw;
    // thread t stages rows 4*kr + t/64 and columns 64*kc + t%64; a slot
    // beyond the window loads its clamped neighbour and stores nothing
    const int tr = (int)threadIdx.x >> 6, tc = (int)threadIdx.x & 63;
    float v[KR][KC];
#pragma unroll
    for (int kr = 0; kr < KR; kr++) {
        const int lr = min(4 * kr + tr, nrows - 1);
        const int r = rlo + lr;
        const int y = r < 0 ? -r : (r >= ny ? 2 * ny - 1 - r : r);
        const float* row = src + (size_t)y * nx;
#pragma unroll
        for (int kc = 0; kc < KC; kc++) {
            const int x = xlo + min(64 * kc + tc, ncols - 1);
            v[kr][kc] = row[x < 0 ? -x : (x >= nx ? 2 * nx - 1 - x : x)];
        }
    }
    float mn = 0.0f, den = 0.0f;
    if constexpr (NORM) norm_bounds(ctl, a, b, &mn, &den);
#pragma unroll
    for (int kr = 0; kr < KR; kr++) {
        const int lr = 4 * kr + tr;
#pragma unroll
        for (int kc = 0; kc < KC; kc++) {
            const int c = 64 * kc + tc;
            if (lr < nrows && c < ncols) S[lr][c] = NORM ? norm_px(v[kr][kc], mn, den) : v[kr][kc];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nrows * 64; e += 256) {
        const int lr = e >> 6, c = e & 63;
        if (oj0 + c > ojlast) continue;
        const int sc = STEP * c + hw;   // source column STEP*oj in S
        double sum = g.B[0] * (double)S[lr][sc];
        for (int j = 1; j < g.size; j++) {
            const float pair = S[lr][sc - j] + S[lr][sc + j];
            sum += g.B[j] * (double)pair;
        }
        R[lr][c] = (float)sum;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < TR * 64; e += 256) {
        const int oi = oi0 + (e >> 6), oj = oj0 + (e & 63);
        if (oi > oilast || oj >= onx) continue;
        const int c = e & 63;
        const int y = STEP * oi - rlo;   // LDS row of the source row STEP*oi
        double sum = g.B[0] * (double)R[y][c];
        for (int j = 1; j < g.size; j++) {
            const float pair = R[y - j][c] + R[y + j][c];
            sum += g.B[j] * (double)pair;
        }
        dst[(size_t)oi * onx + oj] = (float)sum;
    }
}

// Rows per tile: 16, or fewer when the output image is small, so that the
// launch still has a few hundred workgroups; the staged half-width 7 when
// the taps allow it (the default TV-L1 and LK sigmas), else kBlurHalo.
template <int STEP, bool NORM, int H>
static void launch_blur_tile_h(const float* a, const float* b, float* oa, float* ob, int nx,
                               int ny, int onx, int ony, const GaussTaps& g, const TvCtl* ctl,
                               hipStream_t s)
{
    const unsigned gx = cdiv(onx, 64);
    int tr = 16;
    while (tr > 2 && (size_t)gx * cdiv(ony, tr) * 2 < 512) tr /= 2;
    const dim3 grid(gx, cdiv(ony, tr), 2);
    if (tr == 16)
        hipLaunchKernelGGL((k_blur_tile<16, STEP, NORM, H>), grid, dim3(256), 0, s, a, b, oa, ob,
                           nx, ny, onx, ony, g, ctl);
    else if (tr == 8)
        hipLaunchKernelGGL((k_blur_tile<8, STEP, NORM, H>), grid, dim3(256), 0, s, a, b, oa, ob,
                           nx, ny, onx, ony, g, ctl);
    else if (tr == 4)
        hipLaunchKernelGGL((k_blur_tile<4, STEP, NORM, H>), grid, dim3(256), 0, s, a, b, oa, ob,
                           nx, ny, onx, ony, g, ctl);
    else
        hipLaunchKernelGGL((k_blur_tile<2, STEP, NORM, H>), grid, dim3(256), 0, s, a, b, oa, ob,
                           nx, ny, onx, ony, g, ctl);
}

template <int STEP, bool NORM>
static void launch_blur_tile(const float* a, const float* b, float* oa, float* ob, int nx, int ny,
                             int onx, int ony, const GaussTaps& g, const TvCtl* ctl, hipStream_t s)
{
    if (g.size - 1 <= 7)
        launch_blur_tile_h<STEP, NORM, 7>(a, b, oa, ob, nx, ny, onx, ony, g, ctl, s);
    else
        launch_blur_tile_h<STEP, NORM, kBlurHalo>(a, b, oa, ob, nx, ny, onx, ony, g, ctl, s);
}

// zoom_in of both flow components (zoom.cpp:61-85) fused with the
// cblas_sscal by 1/zfactor (tvl1.cpp:171-172).  Reads the coarse scale's
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
    zoom_in_flow_px(uc, uf, nx, ny, nxx, nyy, factorx, factory, invz, i1, j1);
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
// centred gradient (mask.cpp:139-196)
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

// The centred gradient of I1 stored with I1 itself as one float4 per pixel
// (I1, I1x, I1y, 0): the source of the warping's three bicubic samples.
// With `pzero`, also clears the dual variables of the scale (4 floats per
// pixel, tvl1.cpp:190-193).
__global__ __launch_bounds__(256) void k_centered_gradient_packed(const float* __restrict__ in,
                                                                  float4* __restrict__ out, int nx,
                                                                  int ny, float4* __restrict__ pzero)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= nx || i >= ny) return;
    gradient_packed_px(in, out, nx, ny, i, j);
    if (pzero) pzero[(size_t)i * nx + j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ __launch_bounds__(256) void k_pack3(const float* __restrict__ a, const float* __restrict__ b,
                                               const float* __restrict__ c, float4* __restrict__ out,
                                               unsigned n)
{
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = make_float4(a[i], b[i], c[i], 0.0f);
}

__global__ __launch_bounds__(256) void k_scale(float* __restrict__ p, unsigned n, float f)
{
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = f * p[i];
}

__global__ __launch_bounds__(256) void k_fill(float* __restrict__ p, unsigned n, float v)
{
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// warping set-up (tvl1.cpp:198-212)
// ---------------------------------------------------------------------------

// Warps I1, I1x, I1y (interleaved, one float4 per pixel) by the current flow
// with border_out = true (one shared tap geometry, one 16-byte gather per
// tap) and stores I1wx, I1wy and
// rho_c = I1w - I1wx*u1 - I1wy*u2 - I0.  |grad I1w|^2 is recomputed by the
// iteration kernel with the same expression.  Resets the per-warping control.
__global__ __launch_bounds__(256) void k_warp_setup(const float* __restrict__ I0,
                                                    const float4* __restrict__ src,
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
        ctl->chunk_count[0].v = ctl->chunk_count[1].v = 0;
    }
    if (j >= nx || i >= ny) return;
    warp_setup_px(I0, src, ctl->cur[level] ? ub1 : ub0, wx, wy, rhoc, nx, ny, i, j);
}

// Copies the finest scale's current flow into the caller's planar buffer.
// With `report`, one thread also writes the solve's per-warping iteration
// counts (`iters`, n_iters of them) and then its error word to
// report[0..n_iters], host-mapped memory the host reads after its stream
// synchronisation (no copy back).
__global__ __launch_bounds__(256) void k_copy_flow(const float* __restrict__ ub0,
                                                   const float* __restrict__ ub1,
                                                   float* __restrict__ out, unsigned n2,
                                                   const TvCtl* ctl, int level,
                                                   const int* __restrict__ iters, int n_iters,
                                                   int* __restrict__ report)
{
    if (report && blockIdx.x == 0 && threadIdx.x == 0) {
        for (int k = 0; k < n_iters; k++) report[k] = iters[k];
        report[n_iters] = ctl->error;
    }
    if (ctl->error) return;   // a failed solve leaves the destination as it was
    const float* ub = ctl->cur[level] ? ub1 : ub0;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += gridDim.x * blockDim.x)
        out[i] = ub[i];
}

__global__ __launch_bounds__(256) void k_hypot(const float* __restrict__ a,
                                               const float* __restrict__ b,
                                               float* __restrict__ o, unsigned n)
{
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        o[i] = hypot_fin(a[i], b[i]);
}

// TV-L1 energy of a flow at one scale (no reference counterpart; it decides
// between a predicted initial flow and zero):
//   E(u) = lambda * sum_x |I1(x + u(x)) - I0(x)| + sum_x (|grad u1| + |grad u2|)
// I1 is sampled with the Keys bicubic and a clamped border, the gradient is
// the forward difference (zero across the last column / row), |.| of a
// gradient is hypot.  Per-pixel terms are float; the two sums are double,
// reduced per block in a fixed tree, partials [2*block] = (data, tv).
__global__ __launch_bounds__(256) void k_energy(const float* __restrict__ I0,
                                                const float* __restrict__ I1,
                                                const float* __restrict__ u1,
                                                const float* __restrict__ u2, int nx, int ny,
                                                double* __restrict__ part)
{
    __shared__ double sd[256], st[256];
    const long long n = (long long)nx * ny;
    double data = 0.0, tv = 0.0;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int y = (int)(i / nx), x = (int)(i - (long long)y * nx);
        const float a = u1[i], b = u2[i];
        const float w = bicubic_at(I1, (float)x + a, (float)y + b, nx, ny, false);
        data += (double)fabsf(w - I0[i]);
        const float ax = x < nx - 1 ? u1[i + 1] - a : 0.0f;
        const float ay = y < ny - 1 ? u1[i + nx] - a : 0.0f;
        const float bx = x < nx - 1 ? u2[i + 1] - b : 0.0f;
        const float by = y < ny - 1 ? u2[i + nx] - b : 0.0f;
        tv += (double)hypot_fin(ax, ay) + (double)hypot_fin(bx, by);
    }
    sd[threadIdx.x] = data;
    st[threadIdx.x] = tv;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sd[threadIdx.x] += sd[threadIdx.x + w];
            st[threadIdx.x] += st[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sd[0];
        part[2 * blockIdx.x + 1] = st[0];
    }
}

// One block of 256: each thread sums the partials b = tid, tid+256, ... in
// order, then a fixed-order tree, so the result does not depend on timing.
__global__ __launch_bounds__(256) void k_energy_final(const double* __restrict__ part,
                                                      int nblocks, double lambda,
                                                      double* __restrict__ out)
{
    __shared__ double sd[256], st[256];
    double d = 0.0, t = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) {
        d += part[2 * b];
        t += part[2 * b + 1];
    }
    sd[threadIdx.x] = d;
    st[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sd[threadIdx.x] += sd[threadIdx.x + w];
            st[threadIdx.x] += st[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = lambda * sd[0] + st[0];
        out[1] = sd[0];
        out[2] = st[0];
    }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

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

int launch_ctl_init(TvCtl* ctl, hipStream_t s, int* iters, int n_iters)
{
    hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(64), 0, s, ctl, iters, n_iters);
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
    const unsigned nb = std::min(grid1d(n), (unsigned)kMinMaxBlocks);
    hipLaunchKernelGGL(k_load_minmax<uint8_t>, dim3(nb), dim3(256), 0, s, a, b, fa, fb, (unsigned)n, ctl);
    OF_LAUNCH_CHECK("k_load_minmax<u8>");
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, s, ctl, nb);
    OF_LAUNCH_CHECK("k_minmax_final");
    return OF_OK;
}

int launch_load_minmax_f32(const float* a, const float* b, float* fa, float* fb, size_t n,
                           TvCtl* ctl, hipStream_t s)
{
    const unsigned nb = std::min(grid1d(n), (unsigned)kMinMaxBlocks);
    hipLaunchKernelGGL(k_load_minmax<float>, dim3(nb), dim3(256), 0, s, a, b, fa, fb, (unsigned)n, ctl);
    OF_LAUNCH_CHECK("k_load_minmax<f32>");
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, s, ctl, nb);
    OF_LAUNCH_CHECK("k_minmax_final");
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

int launch_smooth_pair(const float* a, const float* b, float* oa, float* ob, int nx, int ny,
                       const GaussTaps& g, const TvCtl* ctl, hipStream_t s)
{
    if (g.size < 1 || g.size - 1 > kBlurHalo || g.size > nx || g.size > ny) return OF_ERR_UNSUPPORTED;
    if (ctl) launch_blur_tile<1, true>(a, b, oa, ob, nx, ny, nx, ny, g, ctl, s);
    else launch_blur_tile<1, false>(a, b, oa, ob, nx, ny, nx, ny, g, ctl, s);
    OF_LAUNCH_CHECK("k_blur_tile<1>");
    return OF_OK;
}

int launch_zoom_out_half_pair(const float* a, const float* b, float* oa, float* ob, int nx, int ny,
                              int nxx, int nyy, const GaussTaps& g, hipStream_t s)
{
    if (g.size < 1 || g.size - 1 > kBlurHalo || g.size > nx || g.size > ny) return OF_ERR_UNSUPPORTED;
    if (2 * (nxx - 1) > nx - 1 || 2 * (nyy - 1) > ny - 1) return OF_ERR_UNSUPPORTED;
    launch_blur_tile<2, false>(a, b, oa, ob, nx, ny, nxx, nyy, g, nullptr, s);
    OF_LAUNCH_CHECK("k_blur_tile<2>");
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

int launch_scale(float* p, size_t n, float f, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale, dim3(grid1d(n)), dim3(256), 0, s, p, (unsigned)n, f);
    OF_LAUNCH_CHECK("k_scale");
    return OF_OK;
}

int launch_fill(float* p, size_t n, float v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(grid1d(n)), dim3(256), 0, s, p, (unsigned)n, v);
    OF_LAUNCH_CHECK("k_fill");
    return OF_OK;
}

int launch_centered_gradient_packed(const float* in, float* out4, int nx, int ny, hipStream_t s,
                                    float* pzero)
{
    hipLaunchKernelGGL(k_centered_gradient_packed, grid2d(nx, ny), dim3(256), 0, s, in,
                       reinterpret_cast<float4*>(out4), nx, ny, reinterpret_cast<float4*>(pzero));
    OF_LAUNCH_CHECK("k_centered_gradient_packed");
    return OF_OK;
}

int launch_pack3(const float* a, const float* b, const float* c, float* out4, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack3, dim3(grid1d(n)), dim3(256), 0, s, a, b, c,
                       reinterpret_cast<float4*>(out4), (unsigned)n);
    OF_LAUNCH_CHECK("k_pack3");
    return OF_OK;
}

int launch_warp_setup(const float* I0, const float* src4, const float* ub0, const float* ub1,
                      float* wx, float* wy, float* rhoc, int nx, int ny, TvCtl* ctl, int level,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_warp_setup, grid2d(nx, ny), dim3(256), 0, s, I0,
                       reinterpret_cast<const float4*>(src4), ub0, ub1, wx, wy, rhoc, nx, ny, ctl,
                       level);
    OF_LAUNCH_CHECK("k_warp_setup");
    return OF_OK;
}

int launch_copy_flow(const float* ub0, const float* ub1, float* out, size_t n2, const TvCtl* ctl,
                     int level, hipStream_t s, const int* iters, int n_iters, int* report)
{
    hipLaunchKernelGGL(k_copy_flow, dim3(grid1d(n2)), dim3(256), 0, s, ub0, ub1, out,
                       (unsigned)n2, ctl, level, iters, n_iters, report);
    OF_LAUNCH_CHECK("k_copy_flow");
    return OF_OK;
}

int launch_hypot(const float* a, const float* b, float* o, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_hypot, dim3(grid1d(n)), dim3(256), 0, s, a, b, o, (unsigned)n);
    OF_LAUNCH_CHECK("k_hypot");
    return OF_OK;
}

// out[0] = E, out[1] = data sum, out[2] = TV sum; `partials` holds
// 2*max_partials doubles.
int launch_energy(const float* I0, const float* I1, const float* u1, const float* u2, int nx,
                  int ny, float lambda, double* partials, int max_partials, double* out,
                  hipStream_t s)
{
    const unsigned blocks = std::min(grid1d((size_t)nx * ny), (unsigned)max_partials);
    hipLaunchKernelGGL(k_energy, dim3(blocks), dim3(256), 0, s, I0, I1, u1, u2, nx, ny, partials);
    OF_LAUNCH_CHECK("k_energy");
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, s, partials, (int)blocks,
                       (double)lambda, out);
    OF_LAUNCH_CHECK("k_energy_final");
    return OF_OK;
}

}  // namespace amdflow
