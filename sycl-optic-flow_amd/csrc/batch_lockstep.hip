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

constexpr unsigned kLsPixelGrid = 4096;   // workgroups of a pixel launch, at most

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

__global__ __launch_bounds__(256) void k_ls_level_begin(LsPhase ph)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    ls_load_bases(ph.blk_base, ph.npairs, sb);
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    for (unsigned b = blockIdx.x; b < ph.nblk; b += gridDim.x) {
        const int k = ls_owner(sb, ph.npairs, b);
        const LsPair& P = ph.pairs[k];
        const unsigned lb = b - sb[k];
        const int nx = P.a.nx, ny = P.a.ny;
        const int j = (int)(lb % P.bx) * 64 + lane;
        const int i = (int)(lb / P.bx) * 4 + row;
        if (j >= nx || i >= ny) continue;
        gradient_packed_px(P.I1s, P.wsrc, nx, ny, i, j);
        // p = 0 in buffer 0 (tvl1.cpp:232-235)
        const size_t N = (size_t)P.a.plane, p = (size_t)i * nx + j;
        P.p0[p] = 0.0f;
        P.p0[N + p] = 0.0f;
        P.p0[2 * N + p] = 0.0f;
        P.p0[3 * N + p] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_ls_warp_setup(LsPhase ph, LsCtl* mc)
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
    for (unsigned b = blockIdx.x; b < ph.nblk; b += gridDim.x) {
        const int k = ls_owner(sb, ph.npairs, b);
        const LsPair& P = ph.pairs[k];
        const unsigned lb = b - sb[k];
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
        if (j >= nx || i >= ny) continue;
        const float* ub = ctl->cur[P.a.level] ? P.a.u[1] : P.a.u[0];
        warp_setup_px(P.I0s, P.wsrc, ub, const_cast<float*>(P.a.wx), const_cast<float*>(P.a.wy),
                      const_cast<float*>(P.a.rhoc), nx, ny, i, j);
    }
}

__global__ __launch_bounds__(256) void k_ls_zoom_in(LsPhase ph)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    ls_load_bases(ph.fblk_base, ph.npairs, sb);
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    for (unsigned b = blockIdx.x; b < ph.nfblk; b += gridDim.x) {
        const int k = ls_owner(sb, ph.npairs, b);
        const LsPair& P = ph.pairs[k];
        const unsigned lb = b - sb[k];
        const int j1 = (int)(lb % P.bxf) * 64 + lane;
        const int i1 = (int)(lb / P.bxf) * 4 + row;
        if (j1 >= P.nxf || i1 >= P.nyf) continue;
        const float* uc = P.a.ctl->cur[P.a.level] ? P.a.u[1] : P.a.u[0];
        zoom_in_flow_px(uc, P.ufine, P.a.nx, P.a.ny, P.nxf, P.nyf, P.fx, P.fy, P.invz, i1, j1);
    }
}

__global__ __launch_bounds__(256) void k_ls_copy_flow(LsPhase ph)
{
    __shared__ unsigned sb[kLsMaxPairs + 1];
    ls_load_bases(ph.blk_base, ph.npairs, sb);
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    for (unsigned b = blockIdx.x; b < ph.nblk; b += gridDim.x) {
        const int k = ls_owner(sb, ph.npairs, b);
        const LsPair& P = ph.pairs[k];
        const unsigned lb = b - sb[k];
        const int nx = P.a.nx, ny = P.a.ny;
        const int j = (int)(lb % P.bx) * 64 + lane;
        const int i = (int)(lb / P.bx) * 4 + row;
        if (j >= nx || i >= ny) continue;
        const float* ub = P.a.ctl->cur[P.a.level] ? P.a.u[1] : P.a.u[0];
        const size_t N = (size_t)P.a.plane, p = (size_t)i * nx + j;
        P.dst[p] = ub[p];
        P.dst[N + p] = ub[N + p];
    }
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

}  // namespace

int launch_ls_level_begin(const LsPhase& ph, hipStream_t s)
{
    if (!phase_ok(ph)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_ls_level_begin, dim3(pixel_grid(ph.nblk)), dim3(256), 0, s, ph);
    OF_LS_CHECK("k_ls_level_begin");
    return OF_OK;
}

int launch_ls_warp_setup(const LsPhase& ph, LsCtl* mc, hipStream_t s)
{
    if (!phase_ok(ph)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_ls_warp_setup, dim3(pixel_grid(ph.nblk)), dim3(256), 0, s, ph, mc);
    OF_LS_CHECK("k_ls_warp_setup");
    return OF_OK;
}

int launch_ls_zoom_in(const LsPhase& ph, hipStream_t s)
{
    if (!phase_ok(ph)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_ls_zoom_in, dim3(pixel_grid(ph.nfblk)), dim3(256), 0, s, ph);
    OF_LS_CHECK("k_ls_zoom_in");
    return OF_OK;
}

int launch_ls_copy_flow(const LsPhase& ph, hipStream_t s)
{
    if (!phase_ok(ph)) return OF_ERR_ARG;
    hipLaunchKernelGGL(k_ls_copy_flow, dim3(pixel_grid(ph.nblk)), dim3(256), 0, s, ph);
    OF_LS_CHECK("k_ls_copy_flow");
    return OF_OK;
}

}  // namespace amdflow
