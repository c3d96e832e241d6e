// tvl1_iter.hip -- the TV-L1 primal-dual iterations (tvl1.cpp:214-292).
//
// Two drivers share one tile core:
//   k_tvl1_iter<K,R>     one launch = K iterations; the last workgroup to
//                        finish evaluates the stopping rule; the host keeps a
//                        few launches queued (large scales);
//   k_tvl1_persist<K,R>  one launch of a co-resident grid = a whole warping;
//                        chunks of K iterations separated by a grid barrier, every
//                        workgroup evaluating the stopping rule on the same
//                        data (small, launch-latency-bound scales).
// Arithmetic follows the reference operation by operation; the library is
// built with -ffp-contract=off.
#include <algorithm>
#include <atomic>
#include <cstdlib>

#include "device_math.hpp"
#include "lockstep.hpp"
#include "tvl1_kernels.hpp"

namespace amdflow {

// Divergence at one pixel from the current value and the left/up neighbours
// with the border rules of mask.cpp:31-77.  Written with selects; every
// candidate is the exact expression of its reference case (the first/last
// column cases associate left to right as written at :81-82).
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

// Thresholding operator TH, tvl1.cpp:221-245, branch-free: the three
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

// Loads and stores of the iteration state.  COH: device-scope atomic
// accesses, coherent across XCDs without cache maintenance (used when other
// workgroups of the same launch read the values).
template <bool COH>
__device__ __forceinline__ float ld(const float* p)
{
    if constexpr (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}

template <bool COH>
__device__ __forceinline__ void st(float* p, float v)
{
    if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

// Called by every thread of a tile's workgroup between its iterations and
// its stores: false discards the tile's results (the neighbour-synchronised
// warping checks its speculative chunk there).  The default stores always.
struct StoreAlways {
    __device__ bool operator()() const { return true; }
};

// One LDS row per wave and exchanged quantity; NW waves per workgroup.
template <int NW>
struct TileSharedN {
    float a[NW][64], b[NW][64], c[NW][64], d[NW][64];
};
using TileShared = TileSharedN<4>;

// kk (<= K) iterations on tile (bx, by): reads buffer `cur`, writes the
// central tile into buffer cur^1, adds each iteration's per-lane error sum
// over the central tile to acc[t].
//
// One workgroup = NW waves; the extended tile is 64 columns (one per lane) by
// NW*R rows (R consecutive rows per wave, in registers).  Horizontal
// neighbours come from DPP lane shifts, vertical ones from the lane's own
// registers or, across waves, from one LDS row.  Values within K pixels of
// the tile edge go stale one pixel per iteration; the central (64-2K) x
// (NW*R-2K) pixels are exact.
template <int K, int R, bool COH, bool INNER, int NW, class Pre>
__device__ __forceinline__ void tile_iterate_impl(const IterArgs& a, int bx, int by, int cur, int kk,
                                                  double (&acc)[K], TileSharedN<NW>& sh, Pre& pre)
{
    constexpr int TW = 64 - 2 * K;
    constexpr int TH = NW * R - 2 * K;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
    const int nx = a.nx, ny = a.ny;
    const int x = bx * TW - K + lane;
    const int ybase = by * TH - K + wave * R;
    const int oy0 = by * TH;
    // INNER: the extended tile lies within columns 1..nx-2 and rows 1..ny-2,
    // so every border test below is known and folds away
    const bool xin = INNER || (x >= 0 && x < nx);
    const bool xout = xin && lane >= K && lane < K + TW;
    const bool fc = !INNER && x == 0, lc = !INNER && x == nx - 1;
    // waves holding no first/last image column take the body formula of the
    // divergence on every row away from the first/last image row
    const bool edge_cols = !INNER && __any(fc || lc);

    const float* uin = a.u[cur];
    const float* pin = a.p[cur];
    const size_t plane = (size_t)a.plane;

    float u1[R], u2[R], p11[R], p12[R], p21[R], p22[R], gx[R], gy[R], rc[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = ybase + r;
        if (INNER || (xin && y >= 0 && y < ny)) {
            const size_t idx = (size_t)y * nx + x;
            u1[r] = ld<COH>(uin + idx);
            u2[r] = ld<COH>(uin + plane + idx);
            p11[r] = ld<COH>(pin + idx);
            p12[r] = ld<COH>(pin + plane + idx);
            p21[r] = ld<COH>(pin + 2 * plane + idx);
            p22[r] = ld<COH>(pin + 3 * plane + idx);
            gx[r] = a.wx[idx];
            gy[r] = a.wy[idx];
            rc[r] = a.rhoc[idx];
        } else {
            u1[r] = u2[r] = p11[r] = p12[r] = p21[r] = p22[r] = 0.0f;
            gx[r] = gy[r] = rc[r] = 0.0f;
        }
    }

    const float l_t = a.l_t, theta = a.theta, taut = a.taut;
#pragma unroll
    for (int t = 0; t < K; t++) {
        if (t < kk) {
            // ---- u update: TH, divergence of p, u = v + theta*div p ----
            sh.a[wave][lane] = p12[R - 1];
            sh.b[wave][lane] = p22[R - 1];
            __syncthreads();
            const float up12 = wave ? sh.a[wave - 1][lane] : p12[0];
            const float up22 = wave ? sh.b[wave - 1][lane] : p22[0];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int y = ybase + r;
                const float p11l = from_left(p11[r]);
                const float p21l = from_left(p21[r]);
                const float p12u = r ? p12[r - 1] : up12;
                const float p22u = r ? p22[r - 1] : up22;
                const bool fr = !INNER && y == 0, lr = !INNER && y == ny - 1;
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
                const bool mine = xout && y >= oy0 && y < oy0 + TH && (INNER || y < ny);
                acc[t] += mine ? (double)e : 0.0;
                u1[r] = nu1;
                u2[r] = nu2;
            }
            // ---- p update: forward gradient of u, g, p = (p + taut*grad)/g ----
            sh.c[wave][lane] = u1[0];
            sh.d[wave][lane] = u2[0];
            __syncthreads();
            const float dn1 = wave < NW - 1 ? sh.c[wave + 1][lane] : u1[R - 1];
            const float dn2 = wave < NW - 1 ? sh.d[wave + 1][lane] : u2[R - 1];
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int y = ybase + r;
                const float u1r = from_right(u1[r]);
                const float u2r = from_right(u2[r]);
                const float u1d = r < R - 1 ? u1[r + 1] : dn1;
                const float u2d = r < R - 1 ? u2[r + 1] : dn2;
                const bool xl = INNER || x < nx - 1, yl = INNER || y < ny - 1;
                const float fx1 = xl ? u1r - u1[r] : 0.0f;
                const float fy1 = yl ? u1d - u1[r] : 0.0f;
                const float fx2 = xl ? u2r - u2[r] : 0.0f;
                const float fy2 = yl ? u2d - u2[r] : 0.0f;
                const float g1 = 1.0f + taut * hypot_fin(fx1, fy1);
                const float g2 = 1.0f + taut * hypot_fin(fx2, fy2);
                // cblas_saxpy (p += taut*grad) then the division, tvl1.cpp:280-291
                p11[r] = (p11[r] + taut * fx1) / g1;
                p12[r] = (p12[r] + taut * fy1) / g1;
                p21[r] = (p21[r] + taut * fx2) / g2;
                p22[r] = (p22[r] + taut * fy2) / g2;
            }
        }
    }

    if (!pre()) return;
    float* uout = a.u[cur ^ 1];
    float* pout = a.p[cur ^ 1];
    if (xout) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int y = ybase + r;
            if (y >= oy0 && y < oy0 + TH && (INNER || y < ny)) {
                const size_t idx = (size_t)y * nx + x;
                st<COH>(uout + idx, u1[r]);
                st<COH>(uout + plane + idx, u2[r]);
                st<COH>(pout + idx, p11[r]);
                st<COH>(pout + plane + idx, p12[r]);
                st<COH>(pout + 2 * plane + idx, p21[r]);
                st<COH>(pout + 3 * plane + idx, p22[r]);
            }
        }
    }
}

// Tiles whose extended area (64 columns x 4R rows from (bx*TW-K, by*TH-K))
// stays clear of the image border run the INNER variant.
template <int K, int R, bool COH, int NW, class Pre = StoreAlways>
__device__ __forceinline__ void tile_iterate(const IterArgs& a, int bx, int by, int cur, int kk,
                                             double (&acc)[K], TileSharedN<NW>& sh, Pre pre = Pre{})
{
    constexpr int TW = 64 - 2 * K;
    constexpr int TH = NW * R - 2 * K;
    const int x0 = bx * TW - K, y0 = by * TH - K;
    const bool inner = x0 >= 1 && x0 + 63 <= a.nx - 2 && y0 >= 1 && y0 + NW * R - 1 <= a.ny - 2;
    if (inner) tile_iterate_impl<K, R, COH, true, NW>(a, bx, by, cur, kk, acc, sh, pre);
    else tile_iterate_impl<K, R, COH, false, NW>(a, bx, by, cur, kk, acc, sh, pre);
}

// Square tiles for the small scales: one workgroup of 256 lanes covers a
// 16 x 16 pixel tile, one pixel per lane, neighbours exchanged through LDS.
// With K iterations per launch the central (16-2K) x (16-2K) pixels are
// exact.  One pixel per lane and more workgroups hide less latency per
// lane but keep many more CUs busy than 64-column tiles at scales of a few
// ten thousand pixels (120x68: 1.2 instead of 3.3 us per iteration).
constexpr int kSqTile = 16;

// p exchange in a..d, u exchange in e, f: the two halves of an iteration
// use different arrays, so each needs one barrier, between its writes and
// its reads.  A half's next writes come after the other half's barrier,
// which every thread reaches only after its reads of this half.
struct SqShared {
    float a[kSqTile][kSqTile + 1], b[kSqTile][kSqTile + 1], c[kSqTile][kSqTile + 1],
        d[kSqTile][kSqTile + 1], e[kSqTile][kSqTile + 1], f[kSqTile][kSqTile + 1];
};

// kk (<= K) iterations on square tile (bx, by): reads buffer `cur`, writes
// the central tile into buffer cur^1, adds each iteration's per-lane error
// term over the central tile to acc[t].  Same per-pixel expressions as
// tile_iterate_impl; pixels outside the image never feed pixels inside (the
// border rules do not read across the border).
template <int K, bool COH, class Pre = StoreAlways>
__device__ __forceinline__ void tile_iterate_sq(const IterArgs& a, int bx, int by, int cur, int kk,
                                                double (&acc)[K], SqShared& sh, Pre pre = Pre{})
{
    constexpr int T = kSqTile, C = T - 2 * K;
    const int lx = threadIdx.x % T, ly = threadIdx.x / T;
    const int nx = a.nx, ny = a.ny;
    const int x = bx * C - K + lx, y = by * C - K + ly;
    const bool in = x >= 0 && x < nx && y >= 0 && y < ny;
    const bool mine = in && lx >= K && lx < K + C && ly >= K && ly < K + C;
    const size_t plane = (size_t)a.plane;
    const size_t idx = (size_t)y * nx + x;
    const float* uin = a.u[cur];
    const float* pin = a.p[cur];
    float u1 = 0.0f, u2 = 0.0f, p11 = 0.0f, p12 = 0.0f, p21 = 0.0f, p22 = 0.0f;
    float gx = 0.0f, gy = 0.0f, rc = 0.0f;
    if (in) {
        u1 = ld<COH>(uin + idx);
        u2 = ld<COH>(uin + plane + idx);
        p11 = ld<COH>(pin + idx);
        p12 = ld<COH>(pin + plane + idx);
        p21 = ld<COH>(pin + 2 * plane + idx);
        p22 = ld<COH>(pin + 3 * plane + idx);
        gx = a.wx[idx];
        gy = a.wy[idx];
        rc = a.rhoc[idx];
    }
    const bool fc = x == 0, lc = x == nx - 1, fr = y == 0, lr = y == ny - 1;
    // neighbour slots, clamped to the tile: the edge lanes read their own
    // slot (their results are outside the exact centre)
    const int xm = lx > 0 ? lx - 1 : lx, xp = lx < T - 1 ? lx + 1 : lx;
    const int ym = ly > 0 ? ly - 1 : ly, yp = ly < T - 1 ? ly + 1 : ly;
    const bool xl = x < nx - 1, yl = y < ny - 1;
    const float l_t = a.l_t, theta = a.theta, taut = a.taut;
#pragma unroll
    for (int t = 0; t < K; t++) {
        if (t < kk) {
            // ---- u update: TH, divergence of p, u = v + theta*div p ----
            sh.a[ly][lx] = p11;
            sh.b[ly][lx] = p21;
            sh.c[ly][lx] = p12;
            sh.d[ly][lx] = p22;
            __syncthreads();
            const float p11l = sh.a[ly][xm], p21l = sh.b[ly][xm];
            const float p12u = sh.c[ym][lx], p22u = sh.d[ym][lx];
            const float div1 = div_rule(p11, p11l, p12, p12u, fr, lr, fc, lc);
            const float div2 = div_rule(p21, p21l, p22, p22u, fr, lr, fc, lc);
            float v1, v2;
            th_step(rc, gx, gy, u1, u2, l_t, v1, v2);
            const float nu1 = v1 + theta * div1;
            const float nu2 = v2 + theta * div2;
            const float e1 = nu1 - u1, e2 = nu2 - u2;
            const float e = e1 * e1 + e2 * e2;
            acc[t] += mine ? (double)e : 0.0;
            u1 = nu1;
            u2 = nu2;
            // ---- p update: forward gradient of u, g, p = (p + taut*grad)/g ----
            sh.e[ly][lx] = u1;
            sh.f[ly][lx] = u2;
            __syncthreads();
            const float u1r = sh.e[ly][xp], u2r = sh.f[ly][xp];
            const float u1d = sh.e[yp][lx], u2d = sh.f[yp][lx];
            const float fx1 = xl ? u1r - u1 : 0.0f;
            const float fy1 = yl ? u1d - u1 : 0.0f;
            const float fx2 = xl ? u2r - u2 : 0.0f;
            const float fy2 = yl ? u2d - u2 : 0.0f;
            const float g1 = 1.0f + taut * hypot_fin(fx1, fy1);
            const float g2 = 1.0f + taut * hypot_fin(fx2, fy2);
            // cblas_saxpy (p += taut*grad) then the division, tvl1.cpp:280-291
            p11 = (p11 + taut * fx1) / g1;
            p12 = (p12 + taut * fy1) / g1;
            p21 = (p21 + taut * fx2) / g2;
            p22 = (p22 + taut * fy2) / g2;
        }
    }
    if (!pre()) return;
    if (mine) {
        float* uout = a.u[cur ^ 1];
        float* pout = a.p[cur ^ 1];
        st<COH>(uout + idx, u1);
        st<COH>(uout + plane + idx, u2);
        st<COH>(pout + idx, p11);
        st<COH>(pout + plane + idx, p12);
        st<COH>(pout + 2 * plane + idx, p21);
        st<COH>(pout + 3 * plane + idx, p22);
    }
}

// Workgroup sum of the per-lane partials, fixed order: wave tree, then
// waves 0..NW-1 left to right.  Result valid in thread t for t < K (thread t
// holds sum t).
template <int K, int NW = 4>
__device__ __forceinline__ double block_sum(const double (&acc)[K], double (*sh_err)[K], int t_out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < K; t++) {
        const double v = wave_sum_f64(acc[t]);
        if (lane == 0) sh_err[wave][t] = v;
    }
    __syncthreads();
    double s = 0.0;
    if (t_out < K) {
        s = sh_err[0][t_out];
#pragma unroll
        for (int w = 1; w < NW; w++) s += sh_err[w][t_out];
    }
    __syncthreads();
    return s;
}

// Sums over all workgroups of columns t < kk of `part` (column t holds nb
// entries at part + t*nb), fixed order per column: thread j sums entries j,
// j+64*NW, ...; then wave trees and waves 0..NW-1 left to right.  All
// columns are read in one pass, so the sums cost one memory round trip.
// Identical in every workgroup that calls it on the same data; out[t] is
// valid in every thread after the call (sh: K*NW doubles of LDS).
template <int K, int NW>
__device__ __forceinline__ void grid_sums(const double* part, unsigned nb, int kk, double* sh,
                                          double* out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double s[K];
#pragma unroll
    for (int t = 0; t < K; t++) s[t] = 0.0;
    for (unsigned j = threadIdx.x; j < nb; j += 64 * NW) {
#pragma unroll
        for (int t = 0; t < K; t++)
            if (t < kk)
                s[t] += __hip_atomic_load(part + (size_t)t * nb + j, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int t = 0; t < K; t++) {
        const double v = wave_sum_f64(s[t]);
        if (lane == 0) sh[t * NW + wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)K) {
        double tot = sh[threadIdx.x * NW];
#pragma unroll
        for (int w = 1; w < NW; w++) tot += sh[threadIdx.x * NW + w];
        out[threadIdx.x] = tot;
    }
    __syncthreads();
}

// The stopping rule applied to kk error sums after iteration n0 (1-based
// iterations n0+1 .. n0+kk): returns the number of iterations to keep when
// the warping stops inside or at the end of this chunk, 0 to continue.
// tvl1.cpp:214-216: continue while error > epsilon^2 and n < max.
__device__ __forceinline__ int stop_after(const double* tot, int kk, int n0, const IterArgs& a)
{
    for (int t = 0; t < kk; t++) {
        const double err = tot[t] / (double)a.plane;
        if (!(err > a.eps2)) return t + 1;
    }
    if (n0 + kk >= a.max_iter) return kk;
    return 0;
}

// Workgroups of 256 per CU the 4-wave iteration kernels are compiled for:
// 3, i.e. 3 waves per SIMD, which caps them at 168 VGPRs (allocated in steps
// of 8 out of 512).  Without the cap the largest shapes took 170 registers
// and fell to 2 waves per SIMD, and the 1920x1080 warping slowed by 10%.
constexpr int kIterBlocksPerCU = 3;

// Waves per workgroup of the one-pair iteration kernels.  A launch shape
// (K, R) names tiles of 64 x 4R pixels; they run as 8 waves of R/2 rows,
// which hides the latency of an iteration better than 4 waves of R rows:
// host-paced, every shape was faster at every scale of a 1920x1080 pyramid
// (e.g. (4,4) at 480x270: 5.22 -> 4.87 us per iteration; (4,8) at 960x540:
// 9.29 -> 8.18 us).  OF_ITER_BOUNDS asks for at least 2 waves per SIMD
// (its second argument counts waves per SIMD, not workgroups per CU); the
// 8-wave kernels hold 2 workgroups per CU wherever they stay within 128
// VGPRs, which all shapes but (8,8) do.
constexpr int kIterWaves = 8;
// The lockstep batch kernel keeps 4 waves of R rows: with 8 waves of R/2 the
// batch of 64 mixed-size pairs took 79.0 ms instead of 75.7 ms (same box).
constexpr int kLsWaves = 4;
// __launch_bounds__(threads, minimum waves per SIMD)
#define OF_ITER_BOUNDS(NW) __launch_bounds__(64 * (NW), (NW) == 4 ? kIterBlocksPerCU : 2)

// ---------------------------------------------------------------------------
// launch-per-chunk driver
// ---------------------------------------------------------------------------

// Stopping rule, evaluated by the last workgroup to finish: when the first
// iteration that meets it falls inside the launch before its last
// iteration, the launch's output is discarded and the next launch re-runs
// exactly that many iterations from the same input buffer.
template <int K, int R, int NW = 4, bool SQ = false>
__global__ OF_ITER_BOUNDS(NW) void k_tvl1_iter(IterArgs a)
{
    __shared__ TileSharedN<NW> sh;
    __shared__ SqShared sh_sq;
    __shared__ double sh_err[NW][K];
    __shared__ double sh_tot[K];
    __shared__ double sh4[K * NW];
    __shared__ int sh_last;

    TvCtl* ctl = a.ctl;
    if (ctl->stop) return;
    const int cur = ctl->cur[a.level];
    const int redo = ctl->redo;
    const int n0 = ctl->n;
    int kk = redo ? redo : a.max_iter - n0;
    if (kk > K) kk = K;
    if (kk <= 0) return;

    double acc[K];
#pragma unroll
    for (int t = 0; t < K; t++) acc[t] = 0.0;
    if constexpr (SQ) tile_iterate_sq<K, false>(a, blockIdx.x, blockIdx.y, cur, kk, acc, sh_sq);
    else tile_iterate<K, R, false, NW>(a, blockIdx.x, blockIdx.y, cur, kk, acc, sh);

    const unsigned nblocks = gridDim.x * gridDim.y;
    const unsigned bid = blockIdx.y * gridDim.x + blockIdx.x;
    const double part = block_sum<K, NW>(acc, sh_err, threadIdx.x);
    // The partials are published with device-scope atomic stores, which are
    // written through to the point of coherence of all XCDs, and every wave
    // waits for their completion (s_waitcnt vmcnt(0)) before the counter
    // increment; the last workgroup reads them with device-scope atomic
    // loads.  A device-scope release fence would instead write back the
    // XCD's whole L2 at the end of every workgroup.
    if (threadIdx.x < (unsigned)kk)
        __hip_atomic_store(a.partials + (size_t)threadIdx.x * nblocks + bid, part, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (threadIdx.x == 0) sh_last = atomicAdd(&ctl->counter, 1u) == nblocks - 1;
    __syncthreads();
    if (!sh_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    grid_sums<K, NW>(a.partials, nblocks, kk, sh4, sh_tot);
    if (a.err_log && threadIdx.x < (unsigned)kk) a.err_log[n0 + threadIdx.x] = sh_tot[threadIdx.x];
    if (threadIdx.x == 0) {
        int n = n0, stop = 0, nredo = 0, flip = 0;
        if (redo) {
            n = n0 + redo;
            stop = 1;
            flip = 1;
        } else {
            const int keep = stop_after(sh_tot, kk, n0, a);
            if (keep == kk || keep == 0) {
                n = n0 + kk;
                flip = 1;
                stop = keep != 0;
            } else {
                nredo = keep;
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

// ---------------------------------------------------------------------------
// persistent driver
// ---------------------------------------------------------------------------

// Residency bound of a first grid barrier: `ticks` > 0 is the wall-clock
// time (ticks of the device's constant-rate clock) that a workgroup waits for the others
// before it reports kBarNotResident; `force` reports it at once (test hook).
struct FirstBound {
    unsigned long long ticks;
    int force;
};

__device__ __forceinline__ bool past_first_bound(const FirstBound& fb, unsigned long long t_end,
                                                 unsigned spins, TvCtl* ctl)
{
    if (!fb.ticks || (spins & 15u) != 0 || (unsigned long long)wall_clock64() < t_end) return false;
    __hip_atomic_store(&ctl->error, kBarNotResident, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// Grid barrier over `nb` co-resident workgroups.  Every wave first waits for
// its own memory operations to complete, so all state written by the
// workgroup (device-scope stores) is visible at the point of coherence before
// the arrival.  The wait is bounded: past ~1 s of polls it flags kBarTimeout
// and returns false, and the kernel ends; with a FirstBound it also ends
// after fb.ticks of wall-clock time with kBarNotResident.
__device__ __forceinline__ bool grid_barrier(TvCtl* ctl, unsigned nb, unsigned& gen, int* sh_ok,
                                             FirstBound fb = {0ull, 0})
{
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned g = gen;
        if (fb.force)
            __hip_atomic_store(&ctl->error, kBarNotResident, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long t_end = fb.ticks ? (unsigned long long)wall_clock64() + fb.ticks : 0ull;
        const unsigned old =
            __hip_atomic_fetch_add(&ctl->bar_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == nb - 1) {
            __hip_atomic_store(&ctl->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __hip_atomic_store(&ctl->bar_gen, g + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned spins = 0;
            while (__hip_atomic_load(&ctl->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > (1u << 21)) {
                    __hip_atomic_store(&ctl->error, kBarTimeout, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                if (past_first_bound(fb, t_end, spins, ctl)) {
                    ok = 0;
                    break;
                }
                // another workgroup gave up: stop waiting too
                if ((spins & 1023u) == 0 &&
                    __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    ok = 0;
                    break;
                }
            }
        }
        if (__hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
        gen = g + 1;
        *sh_ok = ok;
    }
    __syncthreads();
    return *sh_ok != 0;
}

// Two-level grid barrier, for grids of many workgroups (the square-tile
// warping kernel: one workgroup per 8x8 centre).  Workgroup b arrives on the
// counter of group b % 8; the last arrival of a group arrives on the top
// counter; the last top arrival releases the group leaders, each leader its
// group.  With all workgroups arriving together a flat counter took ~10 us
// per barrier at 768 workgroups, the two levels ~3 us.  Each last arriver
// resets its counter and waits for the reset before it releases anyone; the
// release words only grow, and `ep` holds their values read at kernel start
// plus the barriers passed.  Bounded as grid_barrier.
struct BarEpoch {
    unsigned top, grp;
};

__device__ __forceinline__ bool bar_wait(const unsigned* w, unsigned target, TvCtl* ctl,
                                         const FirstBound& fb, unsigned long long t_end)
{
    unsigned spins = 0;
    while ((int)(__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > (1u << 21)) {
            __hip_atomic_store(&ctl->error, kBarTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        if (past_first_bound(fb, t_end, spins, ctl)) return false;
        if ((spins & 1023u) == 0 &&
            __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
    }
    return true;
}

__device__ __forceinline__ bool grid_barrier2(TvCtl* ctl, unsigned nb, BarEpoch& ep, int* sh_ok,
                                              FirstBound fb = {0ull, 0})
{
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (fb.force)
            __hip_atomic_store(&ctl->error, kBarNotResident, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long t_end = fb.ticks ? (unsigned long long)wall_clock64() + fb.ticks : 0ull;
        const unsigned g = blockIdx.x % kBarGroups;
        const unsigned members = nb / kBarGroups + (g < nb % kBarGroups ? 1u : 0u);
        const unsigned groups = nb < (unsigned)kBarGroups ? nb : (unsigned)kBarGroups;
        unsigned* gc = &ctl->grp_count[g].v;
        unsigned* gg = &ctl->grp_gen[g].v;
        int ok = 1;
        const unsigned old = __hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == members - 1) {
            __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned t =
                __hip_atomic_fetch_add(&ctl->bar_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t == groups - 1) {
                __hip_atomic_store(&ctl->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_s_waitcnt(0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __hip_atomic_store(&ctl->bar_gen, ep.top + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                ok = bar_wait(&ctl->bar_gen, ep.top + 1, ctl, fb, t_end);
            }
            if (ok) {
                __builtin_amdgcn_s_waitcnt(0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __hip_atomic_store(gg, ep.grp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else {
            ok = bar_wait(gg, ep.grp + 1, ctl, fb, t_end);
        }
        ep.top++;
        ep.grp++;
        if (fb.force) ok = 0;
        *sh_ok = ok;
    }
    __syncthreads();
    return *sh_ok != 0;
}

// A whole warping: chunks of K iterations over all tiles (workgroup b takes
// tiles b, b+nb, ...), a grid barrier, then every workgroup sums the chunk's
// per-workgroup error partials in the same fixed order and takes the same
// stopping decision.  A stop inside a chunk re-runs exactly that many
// iterations from the chunk's input buffer.  The state is read and written
// with device-scope atomic accesses so that halos written by workgroups on
// other XCDs are seen without cache maintenance.  Partials are
// double-buffered by chunk parity.
template <int K, int R, int NW = 4, bool SQ = false>
__global__ OF_ITER_BOUNDS(NW) void k_tvl1_persist(IterArgs a, unsigned tiles_x, unsigned ntiles)
{
    __shared__ TileSharedN<NW> sh;
    __shared__ SqShared sh_sq;
    __shared__ double sh_err[NW][K];
    __shared__ double sh_tot[K];
    __shared__ double sh4[K * NW];
    __shared__ int sh_ok;

    TvCtl* ctl = a.ctl;
    const unsigned nb = gridDim.x;
    unsigned gen = 0;
    BarEpoch ep{0u, 0u};
    if (threadIdx.x == 0) {
        gen = __hip_atomic_load(&ctl->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ep.top = gen;
        ep.grp = __hip_atomic_load(&ctl->grp_gen[blockIdx.x % kBarGroups].v, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    }
    int cur = ctl->cur[a.level];
    int n = 0;
    int redo = 0;
    int parity = 0;
    // the first barrier of a plain launch also checks that the grid is resident
    FirstBound fb{a.resident_ticks, a.force_not_resident};
    while (true) {
        int kk = redo ? redo : a.max_iter - n;
        if (kk > K) kk = K;
        if (kk <= 0) break;
        double acc[K];
#pragma unroll
        for (int t = 0; t < K; t++) acc[t] = 0.0;
        for (unsigned tile = blockIdx.x; tile < ntiles; tile += nb)
            if constexpr (SQ)
                tile_iterate_sq<K, true>(a, (int)(tile % tiles_x), (int)(tile / tiles_x), cur, kk, acc, sh_sq);
            else
                tile_iterate<K, R, true, NW>(a, (int)(tile % tiles_x), (int)(tile / tiles_x), cur, kk, acc, sh);
        if (redo) {   // the re-run chunk is final: no decision needed
            n += redo;
            cur ^= 1;
            break;
        }
        const double part = block_sum<K, NW>(acc, sh_err, threadIdx.x);
        double* pp = a.partials + (size_t)parity * K * nb;
        if (threadIdx.x < (unsigned)kk)
            __hip_atomic_store(pp + (size_t)threadIdx.x * nb + blockIdx.x, part, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        // square tiles run one workgroup per small tile: many workgroups
        if (!(SQ ? grid_barrier2(ctl, nb, ep, &sh_ok, fb) : grid_barrier(ctl, nb, gen, &sh_ok, fb)))
            return;
        fb = FirstBound{0ull, 0};
        grid_sums<K, NW>(pp, nb, kk, sh4, sh_tot);
        const int keep = stop_after(sh_tot, kk, n, a);
        if (a.err_log && blockIdx.x == 0 && threadIdx.x < (unsigned)kk)
            a.err_log[n + threadIdx.x] = sh_tot[threadIdx.x];
        parity ^= 1;
        if (keep == 0) {
            n += kk;
            cur ^= 1;
            continue;
        }
        if (keep == kk) {
            n += kk;
            cur ^= 1;
            break;
        }
        redo = keep;   // state of buffer `cur` is unchanged: re-run `keep` iterations
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->cur[a.level] = cur;
        ctl->n = n;
        ctl->stop = 1;
        *a.iters_slot = n;
    }
}

// ---------------------------------------------------------------------------
// neighbour-synchronised persistent driver
// ---------------------------------------------------------------------------

// Bounded wait of one thread until the 64-bit word at `w` reaches `want`
// (the words only grow within a warping).  Returns false after flagging
// kBarTimeout past ~1 s of polls, after the first bound (kBarNotResident),
// or once another workgroup has flagged an error.
__device__ __forceinline__ bool word_wait_ge(const unsigned long long* w, unsigned long long want,
                                             TvCtl* ctl, const FirstBound& fb)
{
    const unsigned long long t_end = fb.ticks ? (unsigned long long)wall_clock64() + fb.ticks : 0ull;
    if (fb.force) {
        __hip_atomic_store(&ctl->error, kBarNotResident, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
    }
    unsigned spins = 0;
    while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1u << 21)) {
            __hip_atomic_store(&ctl->error, kBarTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        if (past_first_bound(fb, t_end, spins, ctl)) return false;
        if ((spins & 1023u) == 0 &&
            __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
    }
    return true;
}

// The referee's bounded wait for a chunk's counter to reach `want`.
__device__ __forceinline__ bool count_wait(const unsigned* w, unsigned want, TvCtl* ctl)
{
    unsigned spins = 0;
    while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1u << 21)) {
            __hip_atomic_store(&ctl->error, kBarTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        if ((spins & 1023u) == 0 &&
            __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
    }
    return true;
}

// Tile (bx, by) may start chunk c once its (up to) 8 neighbours have
// completed chunk c-1: its extended tile reads their centres, and its
// stores overwrite values their chunk c-1 read.  Threads 0..8 poll one
// neighbour each; the workgroup learns whether every wait ended normally.
__device__ __forceinline__ bool wait_neighbours(const IterArgs& a, unsigned tiles_x, unsigned tiles_y,
                                                int bx, int by, int c, const FirstBound& fb)
{
    int ok = 1;
    const unsigned t = threadIdx.x;
    if (t < 9 && t != 4) {
        const int x = bx + (int)(t % 3) - 1, y = by + (int)(t / 3) - 1;
        if (x >= 0 && y >= 0 && x < (int)tiles_x && y < (int)tiles_y)
            ok = word_wait_ge(a.tile_flags + (size_t)y * tiles_x + x, (a.seq << 20) | (unsigned)c,
                              a.ctl, fb);
    }
    return __syncthreads_and(ok) != 0;
}

// The stopping decision of chunk c as published by the referee workgroup:
// keep (0: continue) in every thread, or -1 if the wait failed.
__device__ __forceinline__ int wait_decision(const IterArgs& a, int c, int parity, const FirstBound& fb,
                                             int* sh_keep)
{
    if (threadIdx.x == 0) {
        const unsigned long long want = (a.seq << 24) | ((unsigned long long)(c + 1) << 4);
        const unsigned long long* w = &a.ctl->decision[parity].v;
        *sh_keep = word_wait_ge(w, want, a.ctl, fb)
                       ? (int)(__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 0xFull)
                       : -1;
    }
    __syncthreads();
    const int keep = *sh_keep;
    __syncthreads();
    return keep;
}

// A whole warping without grid barriers.  Workgroups 0..nbw-1 iterate tiles
// (workgroup b takes tiles b, b+nbw, ...); a tile starts chunk c as soon as
// its neighbours have completed chunk c-1 (per-tile flags), so no workgroup
// waits for the whole grid.  Workgroup nbw is the referee: it waits for the
// error partials of chunk c from every workgroup, sums them in the fixed
// order of the barrier kernel (grid_sums), applies the stopping rule and
// publishes the decision.  Tiles do not wait for it: they iterate chunk c+1
// from chunk c's output and check decision c before their first store of
// chunk c+1.  "Continue": they store as usual.  "Stop after keep
// iterations": they discard chunk c+1, and if keep is short of chunk c's
// length they re-run keep iterations from chunk c's input, which no store of
// chunk c+1 has touched.  Same arithmetic, partial sums, order and
// decisions as k_tvl1_persist, so the same bits and iteration counts.
template <int K, int R, int NW = 4, bool SQ = false>
__global__ OF_ITER_BOUNDS(NW) void k_tvl1_persist_nbr(IterArgs a, unsigned tiles_x, unsigned ntiles)
{
    static_assert(SQ ? (kSqTile - 2 * K >= K) : (NW * R - 2 * K >= K && 64 - 2 * K >= K),
                  "the halo of a tile must lie within its neighbours");
    __shared__ TileSharedN<NW> sh;
    __shared__ SqShared sh_sq;
    __shared__ double sh_err[NW][K];
    __shared__ double sh_tot[K];
    __shared__ double sh4[K * NW];
    __shared__ int sh_keep;

    TvCtl* ctl = a.ctl;
    const unsigned nbw = gridDim.x - 1;   // tile workgroups; the last one is the referee
    const unsigned tiles_y = ntiles / tiles_x;
    int cur = ctl->cur[a.level];
    int n = 0, c = 0, parity = 0;
    // the first wait of a plain launch also checks that the grid is resident
    FirstBound fb{a.resident_ticks, a.force_not_resident};

    if (blockIdx.x == nbw) {   // referee
        while (true) {
            int kk = a.max_iter - n;
            if (kk > K) kk = K;
            if (kk <= 0) break;
            if (threadIdx.x == 0) {
                // every tile workgroup's first wait checks residency; the
                // referee waits for chunk 0 of all of them without that bound
                sh_keep = count_wait(&ctl->chunk_count[parity].v, nbw, ctl) ? 1 : 0;
            }
            __syncthreads();
            if (!sh_keep) return;
            const double* pp = a.partials + (size_t)parity * K * nbw;
            grid_sums<K, NW>(pp, nbw, kk, sh4, sh_tot);
            const int keep = stop_after(sh_tot, kk, n, a);
            if (a.err_log && threadIdx.x < (unsigned)kk) a.err_log[n + threadIdx.x] = sh_tot[threadIdx.x];
            if (threadIdx.x == 0) {
                // the counter is reused two chunks on, by increments that
                // follow this decision
                __hip_atomic_store(&ctl->chunk_count[parity].v, 0u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_s_waitcnt(0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __hip_atomic_store(&ctl->decision[parity].v,
                                   (a.seq << 24) | ((unsigned long long)(c + 1) << 4) |
                                       (unsigned long long)keep,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            n += keep ? keep : kk;
            cur ^= 1;
            parity ^= 1;
            c++;
            if (keep) break;
        }
        if (threadIdx.x == 0) {
            ctl->cur[a.level] = cur;
            ctl->n = n;
            ctl->stop = 1;
            *a.iters_slot = n;
        }
        return;
    }

    int kk_prev = 0;
    int keep_prev = 0;
    while (true) {
        int kk = a.max_iter - n;
        if (kk > K) kk = K;
        if (kk <= 0) {   // past the last chunk: its decision says whether to re-run part of it
            if (c > 0) keep_prev = wait_decision(a, c - 1, parity ^ 1, fb, &sh_keep);
            break;
        }
        double acc[K];
#pragma unroll
        for (int t = 0; t < K; t++) acc[t] = 0.0;
        bool decided = c == 0;
        bool failed = false;
        // decision c-1 before the chunk's first store: false discards
        auto check = [&]() -> bool {
            if (decided) return true;
            decided = true;
            keep_prev = wait_decision(a, c - 1, parity ^ 1, fb, &sh_keep);
            fb = FirstBound{0ull, 0};
            if (keep_prev < 0) failed = true;
            return keep_prev == 0;
        };
        for (unsigned tile = blockIdx.x; tile < ntiles; tile += nbw) {
            const int bx = (int)(tile % tiles_x), by = (int)(tile / tiles_x);
            if (c > 0) {
                if (!wait_neighbours(a, tiles_x, tiles_y, bx, by, c, fb)) return;
                fb = FirstBound{0ull, 0};
            }
            if constexpr (SQ) tile_iterate_sq<K, true>(a, bx, by, cur, kk, acc, sh_sq, check);
            else tile_iterate<K, R, true, NW>(a, bx, by, cur, kk, acc, sh, check);
            if (failed) return;
            if (keep_prev != 0) break;
            // after the workgroup's last tile of the chunk, its error partials
            // go out with the tile's stores, so one wait for completion
            // covers both
            const bool last = tile + nbw >= ntiles;
            if (last) {
                const double part = block_sum<K, NW>(acc, sh_err, threadIdx.x);
                double* pp = a.partials + (size_t)parity * K * nbw;
                if (threadIdx.x < (unsigned)kk)
                    __hip_atomic_store(pp + (size_t)threadIdx.x * nbw + blockIdx.x, part,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            // publish the tile's chunk (and the partials) once the stores
            // have completed
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __syncthreads();
            if (threadIdx.x == 0) {
                __hip_atomic_store(a.tile_flags + tile, (a.seq << 20) | (unsigned)(c + 1),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (last)
                    __hip_atomic_fetch_add(&ctl->chunk_count[parity].v, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (keep_prev != 0) break;
        kk_prev = kk;
        n += kk;
        cur ^= 1;
        parity ^= 1;
        c++;
    }
    if (keep_prev < 0) return;
    // the warping stopped after keep_prev iterations of chunk c-1: re-run
    // them from its input (buffer cur^1) into its output (buffer cur)
    if (keep_prev > 0 && keep_prev < kk_prev) {
        double acc[K];
#pragma unroll
        for (int t = 0; t < K; t++) acc[t] = 0.0;
        for (unsigned tile = blockIdx.x; tile < ntiles; tile += nbw) {
            const int bx = (int)(tile % tiles_x), by = (int)(tile / tiles_x);
            if constexpr (SQ) tile_iterate_sq<K, true>(a, bx, by, cur ^ 1, keep_prev, acc, sh_sq);
            else tile_iterate<K, R, true, NW>(a, bx, by, cur ^ 1, keep_prev, acc, sh);
        }
    }
}

// ---------------------------------------------------------------------------
// lockstep driver: one scale of many pairs per launch
// ---------------------------------------------------------------------------

// Inclusive prefix sums of v[0..n), n <= kLsMaxPairs, by one workgroup of
// NT threads (Hillis-Steele over LDS).
template <int NT>
__device__ __forceinline__ void ls_block_scan(unsigned* v, int n)
{
    static_assert(kLsMaxPairs % NT == 0, "pairs per thread");
    constexpr int PER = kLsMaxPairs / NT;
    for (int off = 1; off < n; off <<= 1) {
        unsigned t[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int e = (int)threadIdx.x + NT * k;
            t[k] = (e < n && e >= off) ? v[e - off] : 0u;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int e = (int)threadIdx.x + NT * k;
            if (e < n) v[e] += t[k];
        }
        __syncthreads();
    }
}

// Iterations of one pair's tile `bid` (kk <= K of them) and, in the
// workgroup that finishes the pair's last tile of this launch, the pair's
// stopping decision: the same partial sums, order and rule as k_tvl1_iter.
template <int K, int R, int NW>
__device__ __forceinline__ void ls_tile(const LsPhase& ph, const LsPair* P, unsigned bid,
                                        TileSharedN<NW>& sh, double (*sh_err)[K], double* sh_tot,
                                        double* sh4, int* sh_last)
{
    // the table entry is read in place: a private copy would be indexed by
    // `cur` at run time and live in scratch memory
    const IterArgs& a = P->a;
    TvCtl* ctl = a.ctl;
    const int cur = ctl->cur[a.level];
    const int redo = ctl->redo;
    const int n0 = ctl->n;
    int kk = redo ? redo : a.max_iter - n0;
    if (kk > K) kk = K;
    const unsigned nblocks = P->ntiles;
    double acc[K];
#pragma unroll
    for (int t = 0; t < K; t++) acc[t] = 0.0;
    if (kk > 0) tile_iterate<K, R, false, NW>(a, (int)(bid % P->gx), (int)(bid / P->gx), cur, kk, acc, sh);
    const double part = block_sum<K, NW>(acc, sh_err, threadIdx.x);
    if ((int)threadIdx.x < kk)
        __hip_atomic_store(a.partials + (size_t)threadIdx.x * nblocks + bid, part, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (threadIdx.x == 0) *sh_last = atomicAdd(&ctl->counter, 1u) == nblocks - 1;
    __syncthreads();
    if (*sh_last) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        grid_sums<K, NW>(a.partials, nblocks, kk, sh4, sh_tot);
        if (threadIdx.x == 0) {
            int n = n0, stop = 0, nredo = 0, flip = 0;
            if (kk <= 0) {   // nothing left to do: only reached if the list was stale
                stop = 1;
            } else if (redo) {
                n = n0 + redo;
                stop = 1;
                flip = 1;
            } else {
                const int keep = stop_after(sh_tot, kk, n0, a);
                if (keep == kk || keep == 0) {
                    n = n0 + kk;
                    flip = 1;
                    stop = keep != 0;
                } else {
                    nredo = keep;
                }
            }
            ctl->n = n;
            ctl->redo = nredo;
            if (flip) ctl->cur[a.level] = cur ^ 1;
            if (stop)
                ph.iters[((size_t)P->pair_id * kMaxScales + a.level) * ph.warps + ph.warp] = n;
            ctl->counter = 0;
            ctl->launches = ctl->launches + 1;
            // read in this launch by the workgroup that rebuilds the work list
            __hip_atomic_store(&ctl->stop, stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
}

template <int K, int R, int NW>
__global__ OF_ITER_BOUNDS(NW) void k_ls_iter(LsPhase ph, const LsPair* __restrict__ pairs,
                                             LsCtl* mc, unsigned long long* host_status,
                                             unsigned epoch)
{
    // `pairs` is ph.pairs as a restrict kernel argument: the compiler may
    // then read the (wave-uniform) table entries with scalar loads
    constexpr int NT = 64 * NW;
    __shared__ TileSharedN<NW> sh;
    __shared__ double sh_err[NW][K];
    __shared__ double sh_tot[K];
    __shared__ double sh4[K * NW];
    __shared__ int sh_last;
    __shared__ unsigned s_pair[kLsMaxPairs], s_base[kLsMaxPairs];

    const unsigned nlist = mc->nlist;
    const unsigned ntot = mc->ntiles;
    for (unsigned k = threadIdx.x; k < nlist; k += NT) {
        s_pair[k] = mc->list_pair[k];
        s_base[k] = mc->list_base[k];
    }
    __syncthreads();
    for (unsigned t = blockIdx.x; t < ntot; t += gridDim.x) {
        int lo = 0, hi = (int)nlist - 1;   // last listed pair whose first entry is <= t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_base[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        // the owner is the same in every lane: keep it (and the table entry's
        // fields) in scalar registers
        const unsigned pk = __builtin_amdgcn_readfirstlane(s_pair[lo]);
        const unsigned bid = __builtin_amdgcn_readfirstlane(t - s_base[lo]);
        ls_tile<K, R, NW>(ph, pairs + pk, bid, sh, sh_err, sh_tot, sh4, &sh_last);
    }

    // launch completion: the last workgroup rebuilds the work list
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (threadIdx.x == 0) sh_last = atomicAdd(&mc->counter, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!sh_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int np = ph.npairs;
    for (int k = threadIdx.x; k < np; k += NT) {
        const int stop =
            __hip_atomic_load(&ph.pairs[k].a.ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_pair[k] = stop ? 0u : 1u;
        s_base[k] = stop ? 0u : ph.pairs[k].ntiles;
    }
    __syncthreads();
    ls_block_scan<NT>(s_pair, np);
    ls_block_scan<NT>(s_base, np);
    for (int k = threadIdx.x; k < np; k += NT) {
        const unsigned before = k ? s_pair[k - 1] : 0u;
        if (s_pair[k] != before) {   // pair k still iterates
            mc->list_pair[before] = (unsigned)k;
            mc->list_base[before] = k ? s_base[k - 1] : 0u;
        }
    }
    if (threadIdx.x == 0) {
        const unsigned left = np ? s_pair[np - 1] : 0u;
        mc->nlist = left;
        mc->ntiles = np ? s_base[np - 1] : 0u;
        mc->counter = 0;
        const unsigned done = mc->launches + 1;
        mc->launches = done;
        const unsigned long long word = ((unsigned long long)epoch << 32) |
                                        ((unsigned long long)(left == 0) << 31) |
                                        (unsigned long long)(done & 0x7FFFFFFFu);
        if (host_status)
            __hip_atomic_store(host_status, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

bool iter_config_ok(int K, int R)
{
    if (R == 0) return K == 1 || K == 2 || K == 4;   // square tiles
    if (!(K == 1 || K == 2 || K == 4 || K == 8)) return false;
    if (!(R == 2 || R == 4 || R == 8)) return false;
    return 4 * R - 2 * K >= 4;
}

void iter_grid(int K, int R, int nx, int ny, unsigned* gx, unsigned* gy)
{
    if (R == 0) {   // square tiles
        *gx = cdiv((size_t)nx, (size_t)(kSqTile - 2 * K));
        *gy = cdiv((size_t)ny, (size_t)(kSqTile - 2 * K));
        return;
    }
    *gx = cdiv((size_t)nx, (size_t)(64 - 2 * K));
    *gy = cdiv((size_t)ny, (size_t)(4 * R - 2 * K));
}

template <int K, int R>
static hipError_t launch_kr(const IterArgs& a, hipStream_t s)
{
    unsigned gx, gy;
    iter_grid(K, R, a.nx, a.ny, &gx, &gy);
    if constexpr (R == 0)
        hipLaunchKernelGGL((k_tvl1_iter<K, 0, 4, true>), dim3(gx, gy), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_tvl1_iter<K, R * 4 / kIterWaves, kIterWaves>), dim3(gx, gy),
                           dim3(64 * kIterWaves), 0, s, a);
    return hipGetLastError();
}

int launch_iter(int K, int R, const IterArgs& a, hipStream_t s)
{
    hipError_t e = hipErrorInvalidValue;
#define OF_CASE(k, r) \
    if (K == k && R == r) e = launch_kr<k, r>(a, s);
    OF_CASE(1, 0) OF_CASE(2, 0) OF_CASE(4, 0)
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

// The persistent warping kernel of launch shape (K, R): 64 x 4R tiles as
// kIterWaves waves, or (R = 0) square tiles as 4 waves.
template <int K, int R, bool NBR>
static const void* persist_kernel()
{
    if constexpr (NBR) {
        if constexpr (R == 0) return (const void*)k_tvl1_persist_nbr<K, 0, 4, true>;
        else return (const void*)k_tvl1_persist_nbr<K, R * 4 / kIterWaves, kIterWaves>;
    } else {
        if constexpr (R == 0) return (const void*)k_tvl1_persist<K, 0, 4, true>;
        else return (const void*)k_tvl1_persist<K, R * 4 / kIterWaves, kIterWaves>;
    }
}

unsigned iter_tiles(int K, int R, int nx, int ny)
{
    unsigned gx, gy;
    iter_grid(K, R, nx, ny, &gx, &gy);
    return gx * gy;
}

template <int R>
static constexpr int persist_threads()
{
    return R == 0 ? 256 : 64 * kIterWaves;
}

// Launch mode of the persistent kernel.  It is launched as a plain kernel of
// a grid that fits the device at once (occupancy x CUs), which starts every
// workgroup when nothing else holds the CUs.  A cooperative launch would
// guarantee co-residency, but the runtime brackets it with about 11 us of
// idle device time on each side (kernel trace of the 1920x1080 solve, 20
// such launches per solve: 6.07 vs 5.67 ms per solve).  Instead the first
// grid barrier of each launch checks residency with a wall-clock bound: if
// the grid did not fully start (another persistent kernel on another stream
// or process holding CUs, a CU-masked stream), the kernel ends with
// kBarNotResident, the host switches the process to host-paced launches for
// the warpings that would be persistent, and repeats the solve.  Host-paced
// launches need no co-residency and give the same bits.
static std::atomic<int> g_host_paced{-1};   // -1: not read from the environment yet
static std::atomic<int> g_force_not_resident{0};
constexpr double kResidentBoundUs = 2000.0;   // first-barrier wait of a launch

bool persist_host_paced()
{
    int m = g_host_paced.load(std::memory_order_relaxed);
    if (m < 0) {
        const char* e = std::getenv("AMDFLOW_HOST_PACED");
        const int want = (e && e[0] == '1') ? 1 : 0;
        g_host_paced.compare_exchange_strong(m, want);
        m = g_host_paced.load(std::memory_order_relaxed);
    }
    return m == 1;
}

void set_persist_host_paced(bool on) { g_host_paced.store(on ? 1 : 0); }

void set_persist_force_not_resident(bool on) { g_force_not_resident.store(on ? 1 : 0); }

static std::atomic<unsigned long long> g_warp_seq{0};

unsigned long long next_warp_seq() { return g_warp_seq.fetch_add(1) + 1; }

// NBR: the neighbour-synchronised kernel, whose grid is the tile
// workgroups plus one referee workgroup.
template <int K, int R, bool NBR>
static int launch_persist_kr(const IterArgs& a, int nblocks, hipStream_t s)
{
    static int resident = -1;   // co-resident workgroups of this kernel on the device
    static unsigned long long bound_ticks = 0;
    if (resident < 0) {
        int dev = 0, per_cu = 0, cus = 0, clock_khz = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_kernel<K, R, NBR>(),
                                                         persist_threads<R>(), 0) != hipSuccess) {
            (void)hipGetLastError();
            return OF_ERR_HIP;
        }
        // the residency bound in wall-clock ticks (at least 1 ms at any rate)
        bound_ticks = (unsigned long long)(kResidentBoundUs * 1e-3 * std::max(clock_khz, 1000));
        // at most 4 workgroups per CU: 1024 arrivals on one counter and
        // pollers of one word per barrier are already plenty (the compiled
        // bound allows 2 of 8 waves)
        resident = std::min(per_cu, 4) * cus;
    }
    if (resident <= (NBR ? 1 : 0)) return OF_ERR_UNSUPPORTED;
    unsigned gx, gy;
    iter_grid(K, R, a.nx, a.ny, &gx, &gy);
    const unsigned ntiles = gx * gy;
    unsigned nb = nblocks > 0 ? (unsigned)nblocks : ntiles;
    if (nb > ntiles) nb = ntiles;
    if (nb > (unsigned)resident - (NBR ? 1u : 0u)) nb = (unsigned)resident - (NBR ? 1u : 0u);
    if (NBR) {
        if (!a.tile_flags || a.seq == 0 || a.seq >= kNbrMaxSeq) return OF_ERR_ARG;
        if ((long long)a.max_iter / K + 1 >= kNbrMaxChunks) return OF_ERR_UNSUPPORTED;
        nb += 1;   // the referee
    }
    IterArgs args = a;
    args.resident_ticks = bound_ticks;
    args.force_not_resident = g_force_not_resident.load(std::memory_order_relaxed);
    unsigned tx = gx, nt = ntiles;
    void* kargs[] = {(void*)&args, (void*)&tx, (void*)&nt};
    const hipError_t e =
        hipLaunchKernel(persist_kernel<K, R, NBR>(), dim3(nb), dim3(persist_threads<R>()), kargs, 0, s);
    if (e != hipSuccess) {
        record_hip_error(e, "k_tvl1_persist launch", __FILE__, __LINE__);
        return OF_ERR_HIP;
    }
    return OF_OK;
}

int launch_persist(int K, int R, const IterArgs& a, int nblocks, hipStream_t s, int mode)
{
    if (mode != kPersistBarrier && mode != kPersistNeighbour) return OF_ERR_ARG;
#define OF_PCASE(k, r)                                                            \
    if (K == k && R == r)                                                         \
        return mode == kPersistNeighbour ? launch_persist_kr<k, r, true>(a, nblocks, s) \
                                         : launch_persist_kr<k, r, false>(a, nblocks, s);
    OF_PCASE(1, 0) OF_PCASE(2, 0) OF_PCASE(4, 0)
    OF_PCASE(1, 2) OF_PCASE(2, 2)
    OF_PCASE(1, 4) OF_PCASE(2, 4) OF_PCASE(4, 4)
    OF_PCASE(1, 8) OF_PCASE(2, 8) OF_PCASE(4, 8) OF_PCASE(8, 8)
#undef OF_PCASE
    return OF_ERR_UNSUPPORTED;
}

template <int K, int R>
static hipError_t launch_ls_kr(const LsPhase& ph, LsCtl* mc, unsigned grid,
                               unsigned long long* host_status, unsigned epoch, hipStream_t s)
{
    hipLaunchKernelGGL((k_ls_iter<K, R * 4 / kLsWaves, kLsWaves>), dim3(grid),
                       dim3(64 * kLsWaves), 0, s, ph, ph.pairs, mc, host_status, epoch);
    return hipGetLastError();
}

int launch_ls_iter(int K, int R, const LsPhase& ph, LsCtl* mc, unsigned grid,
                   unsigned long long* host_status, unsigned epoch, hipStream_t s)
{
    if (grid == 0 || ph.npairs < 1 || ph.npairs > kLsMaxPairs) return OF_ERR_ARG;
    hipError_t e = hipErrorInvalidValue;
#define OF_CASE(k, r) \
    if (K == k && R == r) e = launch_ls_kr<k, r>(ph, mc, grid, host_status, epoch, s);
    OF_CASE(1, 2) OF_CASE(2, 2)
    OF_CASE(1, 4) OF_CASE(2, 4) OF_CASE(4, 4)
    OF_CASE(1, 8) OF_CASE(2, 8) OF_CASE(4, 8) OF_CASE(8, 8)
#undef OF_CASE
    if (e != hipSuccess) {
        record_hip_error(e, "k_ls_iter launch", __FILE__, __LINE__);
        return OF_ERR_HIP;
    }
    return OF_OK;
}

template <int K, int R>
static int ls_resident_kr(unsigned* blocks)
{
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, (const void*)k_ls_iter<K, R * 4 / kLsWaves, kLsWaves>, 64 * kLsWaves,
            0) != hipSuccess) {
        (void)hipGetLastError();
        return OF_ERR_HIP;
    }
    *blocks = (unsigned)std::max(1, per_cu) * (unsigned)std::max(1, cus);
    return OF_OK;
}

int ls_iter_resident(int K, int R, unsigned* blocks)
{
#define OF_CASE(k, r) \
    if (K == k && R == r) return ls_resident_kr<k, r>(blocks);
    OF_CASE(1, 2) OF_CASE(2, 2)
    OF_CASE(1, 4) OF_CASE(2, 4) OF_CASE(4, 4)
    OF_CASE(1, 8) OF_CASE(2, 8) OF_CASE(4, 8) OF_CASE(8, 8)
#undef OF_CASE
    return OF_ERR_ARG;
}

}  // namespace amdflow
