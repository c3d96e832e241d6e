// tvl1_handle.hpp -- the TV-L1 handle (of_tvl1) shared by the solver and
// the batched solver: every device buffer of one frame size, the per-scale
// geometry and the launch rules.  Private to the library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "amdflow.h"
#include "common.hpp"
#include "tvl1_kernels.hpp"

namespace amdflow {

constexpr int kInFlight = 3;   // iteration launches kept queued ahead of the device

// Launch shape for scales of at least `min_pixels` pixels.
struct LaunchRule {
    long long min_pixels;
    int K;
    int R;
    int persistent;   // kHostPaced, kPersistBarrier or kPersistNeighbour
};

// Enqueues the first phase of a solve on `st`: frames to float (u8 or
// float, `is_u8`), joint normalisation, pre-smoothing and both pyramids;
// resets the handle's control block.
int tvl1_prepare(of_tvl1* h, const void* src0, const void* src1, bool is_u8, hipStream_t st);

}  // namespace amdflow

struct of_tvl1 {
    of_tvl1_params prm{};
    int W = 0, H = 0, ns = 0;
    int nx[amdflow::kMaxScales]{}, ny[amdflow::kMaxScales]{};
    size_t off[amdflow::kMaxScales]{};
    size_t total = 0;

    hipStream_t stream = nullptr;
    float* I0p = nullptr;
    float* I1p = nullptr;
    float* ubase = nullptr;     // per scale s at 4*off[s]: [buf0 u1|u2][buf1 u1|u2]
    float* pbase = nullptr;     // [buf0 p11|p12|p21|p22][buf1 ...], buffer stride 4*N0
    float* wsrc = nullptr;      // 4*N0: (I1, I1x, I1y, 0) per pixel of the active scale
    float* wx = nullptr;
    float* wy = nullptr;
    float* rhoc = nullptr;
    float* A0 = nullptr;
    float* A1 = nullptr;
    uint8_t* in8 = nullptr;     // 2*N0 bytes of upload staging
    float* inf = nullptr;       // 2*N0 floats of upload staging
    float* prev = nullptr;      // previous frame of the streaming API
    bool have_prev = false;
    float* dflow = nullptr;     // 2*N0
    float* init = nullptr;      // 2*N0, allocated on first use: of_tvl1_set_init_flow's flow
    int init_level = -1;        // scale the next solve starts from with `init`; -1: none
    float* hflow = nullptr;     // pinned host copy returned by of_tvl1_flow
    amdflow::TvCtl* ctl = nullptr;
    double* partials = nullptr;
    unsigned long long* tile_flags = nullptr;   // one word per tile (neighbour-synchronised warpings)
    double* energy = nullptr;   // 2*kEnergyBlocks partials, then (E, data, TV)
    bool prepared = false;      // pyramids hold the last prepared pair
    bool solved = false;        // per-scale flows hold the last solve
    int run_warps = 0, run_max = 0;   // limits of the scale being solved
    int run_warps_s[amdflow::kMaxScales]{}, run_max_s[amdflow::kMaxScales]{};   // per scale
    int run_first = -1;               // coarsest scale solved in the last solve
    int* d_iters = nullptr;
    unsigned long long* hstatus = nullptr;   // host-mapped
    int* hreport = nullptr;   // host-mapped: a solve's iteration counts, then its error word
    unsigned epoch = 0;
    std::vector<int> h_iters;
    amdflow::GaussTaps g_pre{}, g_zoom{};
    std::vector<amdflow::LaunchRule> rules;
    long long launches_issued = 0;
    int inflight = amdflow::kInFlight;
    double host_launch_s = 0.0, host_wait_s = 0.0;

    // per-launch event timing (of_tvl1_profile_enable)
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    struct Timed {
        int ev;          // index of the start event; end event is ev+1
        int K, R, persistent;
    };
    std::vector<Timed> timed;
    std::vector<of_tvl1_kernel_profile> prof;
    int scale_K[amdflow::kMaxScales]{}, scale_R[amdflow::kMaxScales]{}, scale_P[amdflow::kMaxScales]{};

    // per-scale event timing (autotuner, of_tvl1_scale_timing_enable)
    bool scale_timing = false;
    std::vector<hipEvent_t> scale_ev;
    double scale_ms[amdflow::kMaxScales]{};
    int scale_solves = 0;
};

