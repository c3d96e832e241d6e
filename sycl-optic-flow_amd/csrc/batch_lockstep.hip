// batch_lockstep.hip -- the per-pixel kernels of the lockstep batch solver:
// one launch covers one scale of every pair in the group.  Each launch's
// work is the concatenation of the pairs' 64x4 pixel blocks (prefix sums
// in LsPhase); a workgroup takes blocks round robin and finds the owning
// pair by binary search in a copy of the prefix table in LDS.  The
// per-pixel bodies are the one-pair kernels' (tvl1_pixel.hpp).
#include <algorithm>

#include "common.hpp"
#include "lockstep.hpp"
#include "tvl1_pixel.hpp"

namespace amdflow {
namespace {

constexpr unsigned kLsPixelGrid = 2048;   // workgroups of a pixel launch, at most (8 per CU)

__device__ __forceinline__ unsigned cdiv_dev(int a, int b) { return (unsigned)((a + b - 1) / b); }

// Loads base[0..n] into LDS (n+1 entries).
__device__ __forceinline__ void ls_load_bases(const unsigned* __restrict__ base, int n,
                                              unsigned* sb)
{
    for (int k = threadIdx.x; k <= n; k += blockDim.x) sb[k] = base[k];
    __syncthreads();
}

// Index of the pair owning block b: the last k with sb[k] <= b.
__device__ __forceinline__ int ls_owner(const unsigned* sb, int n, unsigned b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sb[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Calls f(k, local) for this workgroup's contiguous share of the items
// 0 .. sb[n]-1, k being the owning pair (wave-uniform) and `local` the
// item's index within the pair: one search per workgroup, then a walk.
template <typename F>
__device__ __forceinline__ void ls_for_items(const unsigned* sb, int n, F&& f)
{
    const unsigned total = sb[n];
    const unsigned per = (total + gridDim.x - 1) / gridDim.x;
    unsigned b = blockIdx.x * per;
    const unsigned e = min(total, b + per);
    if (b >= e) return;
    int k = ls_owner(sb, n, b);
    for (; b < e; b++) {
        while (sb[k + 1] <= b) k++;
        const int ku = __builtin_amdgcn_readfirstlane(k);
        f(ku, b - sb[ku]);
    }
}

__global__ __launch_bounds__(256) void k_ls_level_begin(LsPhase ph, const LsPair* __restrict__ pairs)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    ls_load_bases(ph.blk_base, ph.npairs, sb);
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    ls_for_items(sb, ph.npairs, [&](int k, unsigned local) {
        const LsPair& P = pairs[k];
        const unsigned lb = local;
        const int nx = P.a.nx, ny = P.a.ny;
        const int j = (int)(lb % P.bx) * 64 + lane;
        const int i = (int)(lb / P.bx) * 4 + row;
        if (j >= nx || i >= ny) return;
        gradient_packed_px(P.I1s, P.wsrc, nx, ny, i, j);
        // p = 0 in buffer 0 (tvl1.cpp:190-193)
        const size_t N = (size_t)P.a.plane, p = (size_t)i * nx + j;
        P.p0[p] = 0.0f;
        P.p0[N + p] = 0.0f;
        P.p0[2 * N + p] = 0.0f;
        P.p0[3 * N + p] = 0.0f;
        if (P.ucoarse) {   // zero initial flow at the pair's coarsest scale
            P.ucoarse[p] = 0.0f;
            P.ucoarse[N + p] = 0.0f;
        }
    });
}

__global__ __launch_bounds__(256) void k_ls_warp_setup(LsPhase ph, const LsPair* __restrict__ pairs, LsCtl* mc)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    if (blockIdx.x == 0) {
        // the work list of the warping's first iteration launch: every pair
        for (int k = threadIdx.x; k < ph.npairs; k += blockDim.x) {
            mc->list_pair[k] = (unsigned)k;
            mc->list_base[k] = ph.tile_base[k];
        }
        if (threadIdx.x == 0) {
            mc->nlist = (unsigned)ph.npairs;
            mc->ntiles = ph.ntiles;
            mc->counter = 0;
            mc->launches = 0;
        }
    }
    ls_load_bases(ph.blk_base, ph.npairs, sb);
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    ls_for_items(sb, ph.npairs, [&](int k, unsigned local) {
        const LsPair& P = pairs[k];
        const unsigned lb = local;
        TvCtl* ctl = P.a.ctl;
        if (lb == 0 && threadIdx.x == 0) {   // per-warping control, as k_warp_setup
            ctl->n = 0;
            ctl->stop = 0;
            ctl->redo = 0;
            ctl->counter = 0;
            ctl->wlaunches = 0;
        }
        const int nx = P.a.nx, ny = P.a.ny;
        const int j = (int)(lb % P.bx) * 64 + lane;
        const int i = (int)(lb / P.bx) * 4 + row;
        if (j >= nx || i >= ny) return;
        const float* ub = ctl->cur[P.a.level] ? P.a.u[1] : P.a.u[0];
        warp_setup_px(P.I0s, P.wsrc, ub, const_cast<float*>(P.a.wx), const_cast<float*>(P.a.wy),
                      const_cast<float*>(P.a.rhoc), nx, ny, i, j);
    });
}

__global__ __launch_bounds__(256) void k_ls_zoom_in(LsPhase ph, const LsPair* __restrict__ pairs)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    ls_load_bases(ph.fblk_base, ph.npairs, sb);
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    ls_for_items(sb, ph.npairs, [&](int k, unsigned local) {
        const LsPair& P = pairs[k];
        const unsigned lb = local;
        const int j1 = (int)(lb % P.bxf) * 64 + lane;
        const int i1 = (int)(lb / P.bxf) * 4 + row;
        if (j1 >= P.nxf || i1 >= P.nyf) return;
        const float* uc = P.a.ctl->cur[P.a.level] ? P.a.u[1] : P.a.u[0];
        zoom_in_flow_px(uc, P.ufine, P.a.nx, P.a.ny, P.nxf, P.nyf, P.fx, P.fy, P.invz, i1, j1);
    });
}

__global__ __launch_bounds__(256) void k_ls_copy_flow(LsPhase ph, const LsPair* __restrict__ pairs)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    ls_load_bases(ph.blk_base, ph.npairs, sb);
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    ls_for_items(sb, ph.npairs, [&](int k, unsigned local) {
        const LsPair& P = pairs[k];
        const unsigned lb = local;
        const int nx = P.a.nx, ny = P.a.ny;
        const int j = (int)(lb % P.bx) * 64 + lane;
        const int i = (int)(lb / P.bx) * 4 + row;
        if (j >= nx || i >= ny) return;
        const float* ub = P.a.ctl->cur[P.a.level] ? P.a.u[1] : P.a.u[0];
        const size_t N = (size_t)P.a.plane, p = (size_t)i * nx + j;
        P.dst[p] = ub[p];
        P.dst[N + p] = ub[N + p];
    });
}

// ---------------------------------------------------------------------------
// batched preparation (frames to pyramids, tvl1.cpp:137-154)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_lp_load(LsPrepStep st, const LsPrep* __restrict__ preps)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    __shared__ unsigned long long sk[4][4];
    ls_load_bases(st.base, st.npart, sb);
    const int wave = threadIdx.x >> 6;
    ls_for_items(sb, st.npart, [&](int k, unsigned local) {
        const LsPrep& P = preps[st.part[k]];
        const unsigned q = local;
        unsigned long long kmax0 = 0ull, kmax1 = 0ull, kmin0 = ~0ull, kmin1 = ~0ull;
        for (unsigned i = q * 256u + threadIdx.x; i < P.n0; i += P.nmm * 256u) {
            const float va = (float)P.s0[i], vb = (float)P.s1[i];
            P.A0[i] = va;
            P.A1[i] = vb;
            minmax_keys(va, vb, i, kmax0, kmax1, kmin0, kmin1);
        }
        kmax0 = wave_max_u64(kmax0);
        kmax1 = wave_max_u64(kmax1);
        kmin0 = wave_min_u64(kmin0);
        kmin1 = wave_min_u64(kmin1);
        if ((threadIdx.x & 63) == 0) {
            sk[0][wave] = kmax0;
            sk[1][wave] = kmax1;
            sk[2][wave] = kmin0;
            sk[3][wave] = kmin1;
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            const int c = threadIdx.x;
            unsigned long long v = sk[c][0];
            for (int w = 1; w < 4; w++)
                v = (c < 2) ? (sk[c][w] > v ? sk[c][w] : v) : (sk[c][w] < v ? sk[c][w] : v);
            P.ctl->mmpart[c][q] = v;
        }
        __syncthreads();
    });
}

// One workgroup per pair: the min/max keys of both frames and the control
// block of a new solve (every scale's current buffer 0, as k_ctl_init).
__global__ __launch_bounds__(256) void k_lp_final(const LsPrep* __restrict__ preps)
{
    const LsPrep& P = preps[blockIdx.x];
    TvCtl* ctl = P.ctl;
    unsigned long long kmax0 = 0ull, kmax1 = 0ull, kmin0 = ~0ull, kmin1 = ~0ull;
    for (unsigned i = threadIdx.x; i < P.nmm; i += 256) {
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
    if (threadIdx.x < kMaxScales) ctl->cur[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        unsigned long long v[4];
        for (int c = 0; c < 4; c++) {
            v[c] = sk[c][0];
            for (int w = 1; w < 4; w++)
                v[c] = (c < 2) ? (sk[c][w] > v[c] ? sk[c][w] : v[c]) : (sk[c][w] < v[c] ? sk[c][w] : v[c]);
        }
        ctl->maxkey[0] = v[0];
        ctl->maxkey[1] = v[1];
        ctl->minkey[0] = v[2];
        ctl->minkey[1] = v[3];
        ctl->n = 0;
        ctl->stop = 1;
        ctl->redo = 0;
        ctl->counter = 0;
        ctl->launches = 0;
        ctl->wlaunches = 0;
        ctl->bar_count = 0;
        for (int g = 0; g < kBarGroups; g++) ctl->grp_count[g].v = 0;
        ctl->error = 0;
    }
}

__global__ __launch_bounds__(256) void k_lp_normalize(LsPrepStep st, const LsPrep* __restrict__ preps)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    ls_load_bases(st.base, st.npart, sb);
    ls_for_items(sb, st.npart, [&](int k, unsigned local) {
        const LsPrep& P = preps[st.part[k]];
        const unsigned i = local * 256u + threadIdx.x;
        if (i >= P.n0) return;
        float mn, den;
        norm_bounds(P.ctl, P.A0, P.A1, &mn, &den);
        P.I0p[i] = norm_px(P.A0[i], mn, den);
        P.I1p[i] = norm_px(P.A1[i], mn, den);
    });
}

// Items: for each pair, the 64x4 blocks of frame 0 then of frame 1 at the
// blurred scale (0 for step 0, s-1 for step s).
template <bool COLS>
__global__ __launch_bounds__(256) void k_lp_blur(LsPrepStep st, const LsPrep* __restrict__ preps,
                                                 GaussTaps g)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    ls_load_bases(st.base, st.npart, sb);
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const int sc = st.s > 0 ? st.s - 1 : 0;
    ls_for_items(sb, st.npart, [&](int k, unsigned local) {
        const LsPrep& P = preps[st.part[k]];
        const int nx = P.nx[sc], ny = P.ny[sc];
        const unsigned bx = cdiv_dev(nx, 64), nb = bx * cdiv_dev(ny, 4);
        unsigned lb = local;
        const int img = lb >= nb;
        if (img) lb -= nb;
        const int x = (int)(lb % bx) * 64 + lane;
        const int y = (int)(lb / bx) * 4 + row;
        if (x >= nx || y >= ny) return;
        float* tmp = img ? P.B0 : P.A0;
        float* pyr = img ? P.I1p : P.I0p;
        if (!COLS) {
            tmp[(size_t)y * nx + x] = gauss_row_px(pyr + P.off[sc], nx, x, y, g);
        } else {
            // step 0 blurs scale 0 in place; step s writes the blurred copy
            float* dst = st.s == 0 ? pyr : (img ? P.B1 : P.A1);
            dst[(size_t)y * nx + x] = gauss_col_px(tmp, nx, ny, x, y, g);
        }
    });
}

// Items: for each pair, the 64x4 blocks of frame 0 then of frame 1 at scale s.
__global__ __launch_bounds__(256) void k_lp_zoom_out(LsPrepStep st, const LsPrep* __restrict__ preps,
                                                     float zf)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    ls_load_bases(st.base, st.npart, sb);
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const int s = st.s;
    ls_for_items(sb, st.npart, [&](int k, unsigned local) {
        const LsPrep& P = preps[st.part[k]];
        const int nxx = P.nx[s], nyy = P.ny[s];
        const unsigned bx = cdiv_dev(nxx, 64), nb = bx * cdiv_dev(nyy, 4);
        unsigned lb = local;
        const int img = lb >= nb;
        if (img) lb -= nb;
        const int j1 = (int)(lb % bx) * 64 + lane;
        const int i1 = (int)(lb / bx) * 4 + row;
        if (j1 >= nxx || i1 >= nyy) return;
        const float* src = img ? P.B1 : P.A1;
        float* dst = (img ? P.I1p : P.I0p) + P.off[s];
        dst[(size_t)i1 * nxx + j1] = zoom_out_px(src, P.nx[s - 1], P.ny[s - 1], i1, j1, zf);
    });
}

inline unsigned pixel_grid(unsigned blocks) { return std::max(1u, std::min(blocks, kLsPixelGrid)); }

#define OF_LS_CHECK(name)                                           \
    do {                                                            \
        hipError_t e_ = hipGetLastError();                          \
        if (e_ != hipSuccess) {                                     \
            record_hip_error(e_, name, __FILE__, __LINE__);         \
            return OF_ERR_HIP;                                      \
        }                                                           \
    } while (0)

bool phase_ok(const LsPhase& ph) { return ph.npairs >= 1 && ph.npairs <= kLsMaxPairs; }

bool step_ok(const LsPrepStep& st) { return st.npart >= 1 && st.npart <= kLsMaxPairs; }

}  // namespace

int launch_lp_load(const LsPrepStep& st, hipStream_t s)
{
    if (!step_ok(st)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_lp_load, dim3(pixel_grid(st.nitems)), dim3(256), 0, s, st, st.preps);
    OF_LS_CHECK("k_lp_load");
    return OF_OK;
}

int launch_lp_final(const LsPrep* preps, int npairs, hipStream_t s)
{
    if (npairs < 1) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_lp_final, dim3(npairs), dim3(256), 0, s, preps);
    OF_LS_CHECK("k_lp_final");
    return OF_OK;
}

int launch_lp_normalize(const LsPrepStep& st, hipStream_t s)
{
    if (!step_ok(st)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_lp_normalize, dim3(pixel_grid(st.nitems)), dim3(256), 0, s, st, st.preps);
    OF_LS_CHECK("k_lp_normalize");
    return OF_OK;
}

int launch_lp_blur(const LsPrepStep& st, const GaussTaps& g, hipStream_t s)
{
    if (!step_ok(st)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_lp_blur<false>, dim3(pixel_grid(st.nitems)), dim3(256), 0, s, st, st.preps, g);
    OF_LS_CHECK("k_lp_blur rows");
    hipLaunchKernelGGL(k_lp_blur<true>, dim3(pixel_grid(st.nitems)), dim3(256), 0, s, st, st.preps, g);
    OF_LS_CHECK("k_lp_blur cols");
    return OF_OK;
}

int launch_lp_zoom_out(const LsPrepStep& st, float zfactor, hipStream_t s)
{
    if (!step_ok(st) || st.s < 1) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_lp_zoom_out, dim3(pixel_grid(st.nitems)), dim3(256), 0, s, st, st.preps,
                       zfactor);
    OF_LS_CHECK("k_lp_zoom_out");
    return OF_OK;
}

int launch_ls_level_begin(const LsPhase& ph, hipStream_t s)
{
    if (!phase_ok(ph)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_ls_level_begin, dim3(pixel_grid(ph.nblk)), dim3(256), 0, s, ph, ph.pairs);
    OF_LS_CHECK("k_ls_level_begin");
    return OF_OK;
}

int launch_ls_warp_setup(const LsPhase& ph, LsCtl* mc, hipStream_t s)
{
    if (!phase_ok(ph)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_ls_warp_setup, dim3(pixel_grid(ph.nblk)), dim3(256), 0, s, ph, ph.pairs, mc);
    OF_LS_CHECK("k_ls_warp_setup");
    return OF_OK;
}

int launch_ls_zoom_in(const LsPhase& ph, hipStream_t s)
{
    if (!phase_ok(ph)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_ls_zoom_in, dim3(pixel_grid(ph.nfblk)), dim3(256), 0, s, ph, ph.pairs);
    OF_LS_CHECK("k_ls_zoom_in");
    return OF_OK;
}

int launch_ls_copy_flow(const LsPhase& ph, hipStream_t s)
{
    if (!phase_ok(ph)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_ls_copy_flow, dim3(pixel_grid(ph.nblk)), dim3(256), 0, s, ph, ph.pairs);
    OF_LS_CHECK("k_ls_copy_flow");
    return OF_OK;
}

}  // namespace amdflow
