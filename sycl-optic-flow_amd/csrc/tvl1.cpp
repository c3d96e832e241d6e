// tvl1.cpp -- host orchestration of the TV-L1 solver and its C ABI.
//
// Mirrors TV_L1 of src/tvl1/sequential/tvl1.{hpp,cpp}: the constructor fixes
// the pyramid, runDualTVL1Multiscale builds it and solves coarse to fine.
// All work is enqueued on one HIP stream.  The only host involvement inside
// a solve is pacing the iteration launches of each warping: the device
// decides when a warping has converged and publishes its progress in a
// host-mapped word that the host polls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "amdflow.h"
#include "amdflow_debug.h"
#include "common.hpp"
#include "tvl1_kernels.hpp"
#include "tvl1_handle.hpp"
#include "tvl1_launch.hpp"

using namespace amdflow;

namespace {

// Built-in launch shapes per scale size (pixels): large scales are bound by
// memory traffic and use short temporal blocking with host-paced launches;
// small scales are bound by launch latency and run each warping in one
// persistent launch.
const LaunchRule kDefaultRules[] = {
    {200000, 2, 8, 0},
    {0, 4, 0, 0},   // square tiles (rows_per_wave 0), host-paced
};

constexpr int kEnergyBlocks = 1024;

}  // namespace

namespace {

// The handle's own stream, for the host-pointer calls and the autotuner,
// created on first use: every stream holds one of the device's few hardware
// queues (4 by default), and idle streams created before a batch's lanes
// make the lanes share queues.
int own_stream(of_tvl1* h, hipStream_t* out)
{
    if (!h->stream && hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        h->stream = nullptr;
        record_hip_error(hipErrorOutOfMemory, "hipStreamCreateWithFlags", __FILE__, __LINE__);
        return OF_ERR_HIP;
    }
    *out = h->stream;
    return OF_OK;
}

int pick_rule(const of_tvl1* h, long long npx, int* K, int* R, int* persistent)
{
    const LaunchRule* best = nullptr;
    for (const LaunchRule& r : h->rules)
        if (npx >= r.min_pixels && (!best || r.min_pixels > best->min_pixels)) best = &r;
    if (!best)
        for (const LaunchRule& r : kDefaultRules)
            if (npx >= r.min_pixels) { best = &r; break; }
    if (!best) best = &kDefaultRules[sizeof(kDefaultRules) / sizeof(kDefaultRules[0]) - 1];
    *K = best->K;
    *R = best->R;
    *persistent = best->persistent;
    return OF_OK;
}

void release(of_tvl1* h)
{
    if (!h) return;
    dev_free(h->I0p);
    dev_free(h->I1p);
    dev_free(h->ubase);
    dev_free(h->pbase);
    dev_free(h->wsrc);
    dev_free(h->wx);
    dev_free(h->wy);
    dev_free(h->rhoc);
    dev_free(h->A0);
    dev_free(h->A1);
    dev_free(h->in8);
    dev_free(h->inf);
    dev_free(h->prev);
    dev_free(h->dflow);
    dev_free(h->init);
    dev_free(h->ctl);
    dev_free(h->partials);
    dev_free(h->tile_flags);
    dev_free(h->energy);
    dev_free(h->d_iters);
    host_free(h->hflow);
    host_free(h->hstatus);
    host_free(h->hreport);
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : h->scale_ev) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

// Launch bracketed by two events from the handle's pool.
int timed_launch(of_tvl1* h, int K, int R, int persistent, const IterArgs& a, hipStream_t st)
{
    const int ev = 2 * (int)h->timed.size();
    while ((int)h->ev_pool.size() < ev + 2) {
        hipEvent_t e;
        OF_HIP(hipEventCreate(&e));
        h->ev_pool.push_back(e);
    }
    OF_HIP(hipEventRecord(h->ev_pool[ev], st));
    if (persistent) OF_TRY(launch_persist(K, R, a, 0, st, persistent));
    else OF_TRY(launch_iter(K, R, a, st));
    OF_HIP(hipEventRecord(h->ev_pool[ev + 1], st));
    h->timed.push_back({ev, K, R, persistent});
    return OF_OK;
}

of_tvl1_kernel_profile& prof_entry(of_tvl1* h, int K, int R, int persistent)
{
    for (of_tvl1_kernel_profile& p : h->prof)
        if (p.iters_per_launch == K && p.rows_per_wave == R && p.persistent == persistent) return p;
    of_tvl1_kernel_profile p{};
    p.iters_per_launch = K;
    p.rows_per_wave = R;
    p.persistent = persistent;
    h->prof.push_back(p);
    return h->prof.back();
}

// After a synchronised solve: fold the timed launches and the completed
// iterations into the per-kernel accumulators.
int collect_profile(of_tvl1* h)
{
    for (const of_tvl1::Timed& t : h->timed) {
        float ms = 0.0f;
        OF_HIP(hipEventElapsedTime(&ms, h->ev_pool[t.ev], h->ev_pool[t.ev + 1]));
        of_tvl1_kernel_profile& p = prof_entry(h, t.K, t.R, t.persistent);
        p.launches++;
        p.milliseconds += ms;
    }
    h->timed.clear();
    for (int s = 0; s < h->ns; s++) {
        if (!h->scale_K[s]) continue;
        of_tvl1_kernel_profile& p = prof_entry(h, h->scale_K[s], h->scale_R[s], h->scale_P[s]);
        const long long N = (long long)h->nx[s] * h->ny[s];
        for (int w = 0; w < h->prm.warps; w++) {
            const long long pi = (long long)h->h_iters[(size_t)s * h->prm.warps + w] * N;
            p.pixel_iterations += pi;
            p.algorithmic_bytes += 60.0 * (double)pi;
        }
    }
    return OF_OK;
}

// Runs the iterations of one warping: keeps up to kInFlight launches queued
// and stops issuing once the device reports that the warping has stopped.
int run_warping(of_tvl1* h, int s, int w, hipStream_t st)
{
    const of_tvl1_params& P = h->prm;
    if (h->run_max <= 0) return OF_OK;
    int K, R, persistent;
    pick_rule(h, (long long)h->nx[s] * h->ny[s], &K, &R, &persistent);
    h->scale_K[s] = K;
    h->scale_R[s] = R;
    h->scale_P[s] = persistent;
    const size_t N = (size_t)h->nx[s] * h->ny[s];
    IterArgs a{};
    a.wx = h->wx;
    a.wy = h->wy;
    a.rhoc = h->rhoc;
    float* ub = h->ubase + 4 * h->off[s];
    a.u[0] = ub;
    a.u[1] = ub + 2 * N;
    a.p[0] = h->pbase;
    a.p[1] = h->pbase + 4 * (size_t)h->W * h->H;
    a.nx = h->nx[s];
    a.ny = h->ny[s];
    a.plane = (int)N;
    a.l_t = P.lambda * P.theta;
    a.theta = P.theta;
    a.taut = P.tau / P.theta;
    a.eps2 = (double)(P.epsilon * P.epsilon);
    a.max_iter = h->run_max;
    a.level = s;
    a.ctl = h->ctl;
    a.iters_slot = h->d_iters + (size_t)s * P.warps + w;
    a.partials = h->partials;
    a.host_status = h->hstatus;
    a.epoch = ++h->epoch;
    a.tile_flags = h->tile_flags;
    a.seq = next_warp_seq();

    if (persistent && persist_host_paced()) {
        persistent = 0;   // the process gave up persistent launches
        h->scale_P[s] = 0;
    }
    if (persistent) {
        // the whole warping in one launch: no host involvement
        if (h->profiling) return timed_launch(h, K, R, persistent, a, st);
        h->launches_issued++;
        return launch_persist(K, R, a, 0, st, persistent);
    }

    const long long max_launches = (h->run_max + K - 1) / K + 1;
    long long issued = 0, completed = 0;
    bool stopped = false;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto tw = t0;
    bool waiting = false;
    unsigned spins = 0;
    const int inflight = h->inflight;
    while (!stopped && issued < max_launches) {
        if (waiting && issued - completed < inflight) {
            h->host_wait_s += std::chrono::duration<double>(clk::now() - tw).count();
            waiting = false;
        }
        const auto tl = clk::now();
        bool launched = false;
        while (issued < max_launches && issued - completed < inflight) {
            if (h->profiling) OF_TRY(timed_launch(h, K, R, 0, a, st));
            else OF_TRY(launch_iter(K, R, a, st));
            issued++;
            h->launches_issued++;
            launched = true;
        }
        if (launched) h->host_launch_s += std::chrono::duration<double>(clk::now() - tl).count();
        if (issued >= max_launches) break;
        if (!waiting) {
            waiting = true;
            tw = clk::now();
        }
        const unsigned long long word = __atomic_load_n(h->hstatus, __ATOMIC_ACQUIRE);
        if ((unsigned)(word >> 32) == a.epoch) {
            completed = (long long)(word & 0x7FFFFFFFull);
            stopped = ((word >> 31) & 1ull) != 0;
            if (stopped || issued - completed < inflight) continue;
        }
        if (++spins % 4096 == 0) {
            const hipError_t q = hipStreamQuery(st);
            if (q != hipSuccess && q != hipErrorNotReady) {
                record_hip_error(q, "hipStreamQuery", __FILE__, __LINE__);
                return OF_ERR_HIP;
            }
            if (q == hipSuccess) {
                // Idle stream: every issued launch has reported.
                const unsigned long long w2 = __atomic_load_n(h->hstatus, __ATOMIC_ACQUIRE);
                if ((unsigned)(w2 >> 32) == a.epoch) {
                    completed = (long long)(w2 & 0x7FFFFFFFull);
                    stopped = ((w2 >> 31) & 1ull) != 0;
                }
                if (!stopped && completed < issued) return OF_ERR_STATE;
            }
            const double secs =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (secs > 60.0) return OF_ERR_HIP;
        } else {
            __builtin_ia32_pause();
        }
    }
    if (waiting) h->host_wait_s += std::chrono::duration<double>(clk::now() - tw).count();
    return OF_OK;
}

}  // namespace

int amdflow::tvl1_prepare(of_tvl1* h, const void* src0, const void* src1, bool is_u8, hipStream_t st)
{
    const of_tvl1_params& P = h->prm;
    const size_t N0 = (size_t)h->W * h->H;
    OF_TRY(launch_ctl_init(h->ctl, st));
    if (is_u8)
        OF_TRY(launch_load_minmax_u8((const uint8_t*)src0, (const uint8_t*)src1, h->A0, h->A1, N0,
                                     h->ctl, st));
    else
        OF_TRY(launch_load_minmax_f32((const float*)src0, (const float*)src1, h->A0, h->A1, N0,
                                      h->ctl, st));
    float* I00 = h->I0p;
    float* I10 = h->I1p;
    // normalisation (tvl1.cpp:140) and pre-smoothing (:143-144) of both frames
    // in one tiled launch; the two-pass form where the taps do not fit it
    int st0 = launch_smooth_pair(h->A0, h->A1, I00, I10, h->W, h->H, h->g_pre, h->ctl, st);
    if (st0 == OF_ERR_UNSUPPORTED) {
        OF_TRY(launch_normalize(h->A0, h->A1, I00, I10, N0, h->ctl, st));
        OF_TRY(launch_gaussian(I00, h->A0, I00, h->W, h->H, h->g_pre, st));
        OF_TRY(launch_gaussian(I10, h->A0, I10, h->W, h->H, h->g_pre, st));
    } else {
        OF_TRY(st0);
    }
    // pyramid, tvl1.cpp:146-154 (zoom_out: blur a copy, then resample): at
    // zfactor 1/2 one launch per scale for both frames, which keeps only the
    // blurred samples at even positions.  That equals the bicubic resampling
    // of every finite frame without negative zeros, as u8 frames normalise
    // to (finite values >= +0); there bicubic_at(2j, 2i) adds (0.5*0)*inner
    // to the sample, and a non-finite neighbour turns that into NaN, a -0
    // sample may turn into +0.  Float frames therefore take the two-step path.
    for (int s = 1; s < h->ns; s++) {
        int sz = OF_ERR_UNSUPPORTED;
        if (P.zfactor == 0.5f && is_u8)
            sz = launch_zoom_out_half_pair(h->I0p + h->off[s - 1], h->I1p + h->off[s - 1],
                                           h->I0p + h->off[s], h->I1p + h->off[s], h->nx[s - 1],
                                           h->ny[s - 1], h->nx[s], h->ny[s], h->g_zoom, st);
        if (sz != OF_ERR_UNSUPPORTED) {
            OF_TRY(sz);
            continue;
        }
        for (int k = 0; k < 2; k++) {
            float* P0 = (k ? h->I1p : h->I0p);
            OF_TRY(launch_gaussian(P0 + h->off[s - 1], h->A0, h->A1, h->nx[s - 1], h->ny[s - 1],
                                   h->g_zoom, st));
            OF_TRY(launch_zoom_out_sample(h->A1, P0 + h->off[s], h->nx[s - 1], h->ny[s - 1],
                                          h->nx[s], h->ny[s], P.zfactor, st));
        }
    }
    h->prepared = true;
    h->solved = false;
    return OF_OK;
}

namespace {

// Returned by solve_levels_once when a persistent launch found its grid not
// fully resident: nothing was written to `dst`.
constexpr int kRetryHostPaced = -1;

// Second phase: the coarse-to-fine solve from the prepared pyramids, with
// `init` (planar [u1|u2] at the coarsest scale, device) as the initial flow,
// or zero when null (tvl1.cpp:157-158); at most warps_s[s] warpings and
// max_s[s] iterations per warping at scale s.  Waits for the result.
int solve_levels_once(of_tvl1* h, const float* init, int first, const int* warps_s,
                      const int* max_s, float* dst, hipStream_t st)
{
    const of_tvl1_params& P = h->prm;
    const size_t N0 = (size_t)h->W * h->H;
    for (int s = 0; s < h->ns; s++) {
        h->run_warps_s[s] = s <= first ? warps_s[s] : 0;
        h->run_max_s[s] = max_s[s];
    }
    h->run_first = first;
    // every scale's current buffer back to 0 (the pyramids stay as prepared)
    OF_TRY(launch_ctl_init(h->ctl, st, h->d_iters, (int)h->h_iters.size()));
    {
        const size_t n2 = 2 * (size_t)h->nx[first] * h->ny[first];
        float* uc = h->ubase + 4 * h->off[first];   // buffer 0
        if (init) OF_HIP(hipMemcpyAsync(uc, init, sizeof(float) * n2, hipMemcpyDeviceToDevice, st));
        else OF_TRY(launch_fill(uc, n2, 0.0f, st));
    }
    const float invZ = 1 / P.zfactor;
    for (int s = first; s >= 0; s--) {
        const int nx = h->nx[s], ny = h->ny[s];
        const size_t N = (size_t)nx * ny;
        const int warps = warps_s[s];
        h->run_warps = warps;
        h->run_max = max_s[s];
        float* I0s = h->I0p + h->off[s];
        float* I1s = h->I1p + h->off[s];
        float* ub = h->ubase + 4 * h->off[s];
        if (h->scale_timing) OF_HIP(hipEventRecord(h->scale_ev[2 * s], st));
        // the packed gradient of I1, and p = 0 in buffer 0 (tvl1.cpp:190-193)
        OF_TRY(launch_centered_gradient_packed(I1s, h->wsrc, nx, ny, st, h->pbase));
        for (int w = 0; w < warps; w++) {
            OF_TRY(launch_warp_setup(I0s, h->wsrc, ub, ub + 2 * N, h->wx, h->wy, h->rhoc, nx, ny,
                                     h->ctl, s, st));
            OF_TRY(run_warping(h, s, w, st));
        }
        if (s > 0) {
            float* uf = h->ubase + 4 * h->off[s - 1];
            OF_TRY(launch_zoom_in_flow(ub, ub + 2 * N, uf, nx, ny, h->nx[s - 1], h->ny[s - 1],
                                       invZ, h->ctl, s, st));
        }
        if (h->scale_timing) OF_HIP(hipEventRecord(h->scale_ev[2 * s + 1], st));
    }
    // the final copy also writes the iteration counts and the error word to
    // host-mapped memory: one synchronisation, no copies back
    const int n_iters = (int)h->h_iters.size();
    OF_TRY(launch_copy_flow(h->ubase, h->ubase + 2 * N0, dst, 2 * N0, h->ctl, 0, st, h->d_iters,
                            n_iters, h->hreport));
    OF_HIP(hipStreamSynchronize(st));
    std::copy(h->hreport, h->hreport + n_iters, h->h_iters.begin());
    {
        // a persistent warping whose grid barrier ran past its bound flags
        // it; so does one whose grid did not fully start (the final copy
        // then skipped `dst`)
        const int derr = h->hreport[n_iters];
        if (derr == kBarNotResident) return kRetryHostPaced;
        if (derr) {
            record_hip_error(hipErrorLaunchTimeOut, "k_tvl1_persist grid barrier", __FILE__, __LINE__);
            return OF_ERR_HIP;
        }
    }
    if (h->profiling) OF_TRY(collect_profile(h));
    if (h->scale_timing) {
        for (int s = first; s >= 0; s--) {
            float ms = 0.0f;
            OF_HIP(hipEventElapsedTime(&ms, h->scale_ev[2 * s], h->scale_ev[2 * s + 1]));
            h->scale_ms[s] += ms;
        }
        h->scale_solves++;
    }
    h->solved = true;
    return OF_OK;
}

// solve_levels_once, repeated once with host-paced launches (from then on,
// for the whole process) if a persistent launch was not co-resident.  The
// repeat starts from the same pyramids and `init`, which a solve does not
// write.
int solve_levels_schedule(of_tvl1* h, const float* init, int first, const int* warps_s,
                          const int* max_s, float* dst, hipStream_t st)
{
    const int rc = solve_levels_once(h, init, first, warps_s, max_s, dst, st);
    if (rc != kRetryHostPaced) return rc;
    set_persist_host_paced(true);
    h->timed.clear();   // launch events of the abandoned solve
    const int rc2 = solve_levels_once(h, init, first, warps_s, max_s, dst, st);
    if (rc2 == kRetryHostPaced) {   // the mode was set back to persistent meanwhile
        record_hip_error(hipErrorLaunchTimeOut, "k_tvl1_persist residency", __FILE__, __LINE__);
        return OF_ERR_HIP;
    }
    return rc2;
}

// The same limits at every scale.
int solve_levels(of_tvl1* h, const float* init, int first, int warps, int max_iter, float* dst,
                 hipStream_t st)
{
    int w[kMaxScales], m[kMaxScales];
    for (int s = 0; s < kMaxScales; s++) {
        w[s] = warps;
        m[s] = max_iter;
    }
    return solve_levels_schedule(h, init, first, w, m, dst, st);
}

// One full multiscale solve with the handle's parameters and zero initial flow.
int solve(of_tvl1* h, const void* src0, const void* src1, bool is_u8, float* dst, hipStream_t st)
{
    OF_TRY(tvl1_prepare(h, src0, src1, is_u8, st));
    if (h->init_level >= 0) {   // one-shot initial flow of of_tvl1_set_init_flow
        const int first = h->init_level;
        h->init_level = -1;
        return solve_levels(h, h->init, first, h->prm.warps, h->prm.max_iterations, dst, st);
    }
    return solve_levels(h, nullptr, h->ns - 1, h->prm.warps, h->prm.max_iterations, dst, st);
}

// Launch-shape autotuner: every supported (iterations per launch, rows per
// wave) pair is applied to all scales in turn; each scale's wall time on
// the device (events around its gradient, warpings and prolongation) is
// averaged over `reps` solves of the given pair, and each scale then keeps
// its fastest shape.  The iteration counts, and the flow, do not depend on
// the shape (the kernels are bit-exact for every shape), so the timings
// compare the same work.
int autotune(of_tvl1* h, const uint8_t* d0, const uint8_t* d1, int reps, int* chosen,
             double* table)
{
    static const int kShapes[][2] = {{1, 2}, {2, 2}, {1, 4}, {2, 4}, {4, 4},
                                     {1, 8}, {2, 8}, {4, 8}, {8, 8}};
    constexpr int nshape = (int)(sizeof(kShapes) / sizeof(kShapes[0]));
    static const int kSquareK[] = {1, 2, 4};   // square tiles
    constexpr int nsquare = (int)(sizeof(kSquareK) / sizeof(kSquareK[0]));
    if (reps < 1) reps = 1;
    if (h->scale_ev.empty()) {
        h->scale_ev.resize(2 * kMaxScales);
        for (hipEvent_t& e : h->scale_ev) OF_HIP(hipEventCreate(&e));
    }
    hipStream_t hs = nullptr;
    OF_TRY(own_stream(h, &hs));
    std::vector<double> best(h->ns, 1e30);
    std::vector<int> bestK(h->ns, 2), bestR(h->ns, 8), bestP(h->ns, 0);
    const std::vector<LaunchRule> saved = h->rules;
    // modes: 64-column tiles host-paced, persistent with grid barriers;
    // square tiles host-paced, persistent with grid barriers; persistent
    // with neighbour flags, 64-column tiles and square tiles
    static const int kModeP[] = {kHostPaced, kPersistBarrier, kHostPaced, kPersistBarrier,
                                 kPersistNeighbour, kPersistNeighbour};
    static const bool kModeSquare[] = {false, false, true, true, false, true};
    int row = 0;
    for (int mode = 0; mode < 6; mode++) {
        const bool square = kModeSquare[mode];
        for (int c = 0; c < (square ? nsquare : nshape); c++, row++) {
            const int K = square ? kSquareK[c] : kShapes[c][0];
            const int R = square ? 0 : kShapes[c][1];
            const int P = kModeP[mode];
            h->rules.assign(1, LaunchRule{0, K, R, P});
            int st = solve(h, d0, d1, true, h->dflow, hs);   // warm-up of this variant
            // a persistent variant whose grid barrier ran past its bound is
            // skipped (the kernels exit and the device stays usable), and so
            // is every persistent variant once the process runs host-paced
            if (P && (persist_host_paced() ||
                      (st != OF_OK && hipDeviceSynchronize() == hipSuccess))) {
                if (table)
                    for (int s = 0; s < h->ns; s++) table[(size_t)row * h->ns + s] = -1.0;
                continue;
            }
            for (int s = 0; s < h->ns; s++) h->scale_ms[s] = 0.0;
            h->scale_timing = true;
            for (int r = 0; r < reps && st == OF_OK; r++) st = solve(h, d0, d1, true, h->dflow, hs);
            h->scale_timing = false;
            if (st != OF_OK) {
                h->rules = saved;
                return st;
            }
            for (int s = 0; s < h->ns; s++) {
                const double ms = h->scale_ms[s] / reps;
                if (table) table[(size_t)row * h->ns + s] = ms;
                if (ms < best[s]) {
                    best[s] = ms;
                    bestK[s] = K;
                    bestR[s] = R;
                    bestP[s] = P;
                }
            }
        }
    }
    h->rules.clear();
    for (int s = 0; s < h->ns; s++) {
        h->rules.push_back(LaunchRule{(long long)h->nx[s] * h->ny[s], bestK[s], bestR[s], bestP[s]});
        if (chosen) {
            chosen[3 * s] = bestK[s];
            chosen[3 * s + 1] = bestR[s];
            chosen[3 * s + 2] = bestP[s];
        }
    }
    return OF_OK;
}

int validate(const of_tvl1_params& p, int W, int H)
{
    if (W < 2 || H < 2) return OF_ERR_ARG;
    if ((long long)W * H >= (1ll << 31)) return OF_ERR_ARG;
    if (!(p.tau > 0) || !(p.theta > 0) || !(p.lambda >= 0)) return OF_ERR_ARG;
    if (!(p.zfactor > 0 && p.zfactor < 1)) return OF_ERR_ARG;
    if (p.warps < 1 || p.nscales < 1 || p.max_iterations < 0 || !(p.epsilon >= 0)) return OF_ERR_ARG;
    return OF_OK;
}

}  // namespace

extern "C" {

void of_tvl1_default_params(of_tvl1_params* p)
{
    if (!p) return;
    p->tau = 0.25f;
    p->lambda = 0.15f;
    p->theta = 0.3f;
    p->nscales = 100;
    p->zfactor = 0.5f;
    p->warps = 5;
    p->epsilon = 0.01f;
    p->max_iterations = 300;
}

int of_tvl1_create(const of_tvl1_params* params, int width, int height, of_tvl1** out)
{
    return guarded([&]() -> int {
        if (!out) return OF_ERR_ARG;
        *out = nullptr;
        of_tvl1_params p;
        if (params) p = *params;
        else of_tvl1_default_params(&p);
        OF_TRY(validate(p, width, height));
    
        // scale count, tvl1.cpp:42-43, clamped to >= 1
        const float logz = std::log(1.0f / p.zfactor);
        const double q = std::log(std::hypot((double)width, (double)height) / 16.0) / (double)logz;
        const float Nf = (float)(1.0 + q);
        int ns = (int)((Nf < (float)p.nscales) ? Nf : (float)p.nscales);
        if (ns < 1) ns = 1;
        if (ns > kMaxScales) ns = kMaxScales;
    
        of_tvl1* h = new (std::nothrow) of_tvl1();
        if (!h) return OF_ERR_ALLOC;
        h->prm = p;
        h->W = width;
        h->H = height;
        h->ns = ns;
        h->nx[0] = width;
        h->ny[0] = height;
        for (int s = 1; s < ns; s++) zoom_size(h->nx[s - 1], h->ny[s - 1], p.zfactor, &h->nx[s], &h->ny[s]);
        size_t off = 0;
        for (int s = 0; s < ns; s++) {
            if (h->nx[s] < 2 || h->ny[s] < 2) { delete h; return OF_ERR_ARG; }
            h->off[s] = off;
            off += (size_t)h->nx[s] * h->ny[s];
        }
        h->total = off;
    
        // Gaussian taps: pre-smoothing sigma 0.8f (tvl1.hpp:7) and the zoom_out
        // sigma 0.6f*sqrt(1/zf^2 - 1) rounded to float (zoom.cpp:38).
        int st = gaussian_taps((double)0.8f, &h->g_pre);
        const float zsig = (float)((double)0.6f * std::sqrt(1.0 / (double)(p.zfactor * p.zfactor) - 1.0));
        if (!st) st = gaussian_taps((double)zsig, &h->g_zoom);
        if (!st && (h->g_pre.size > width || h->g_pre.size > height)) st = OF_ERR_SIGMA;
        for (int s = 0; s + 1 < ns && !st; s++)
            if (h->g_zoom.size > h->nx[s] || h->g_zoom.size > h->ny[s]) st = OF_ERR_SIGMA;
        if (st) { delete h; return st; }
    
        const size_t N0 = (size_t)width * height;
        size_t maxpart = 1, maxtiles = 1;
        const int Ks[] = {1, 2, 4, 8}, Rs[] = {0, 2, 4, 8};
        for (int K : Ks)
            for (int R : Rs)
                if (iter_config_ok(K, R)) {
                    const size_t t = iter_tiles(K, R, width, height);
                    maxpart = std::max(maxpart, t * K);
                    maxtiles = std::max(maxtiles, t);
                }
    
        st = OF_OK;
        if (!st) st = dev_alloc(&h->I0p, h->total);
        if (!st) st = dev_alloc(&h->I1p, h->total);
        if (!st) st = dev_alloc(&h->ubase, 4 * h->total);
        if (!st) st = dev_alloc(&h->pbase, 8 * N0);
        if (!st) st = dev_alloc(&h->wsrc, 4 * N0);
        if (!st) st = dev_alloc(&h->wx, N0);
        if (!st) st = dev_alloc(&h->wy, N0);
        if (!st) st = dev_alloc(&h->rhoc, N0);
        if (!st) st = dev_alloc(&h->A0, N0);
        if (!st) st = dev_alloc(&h->A1, N0);
        if (!st) st = dev_alloc(&h->in8, 2 * N0);
        if (!st) st = dev_alloc(&h->inf, 2 * N0);
        if (!st) st = dev_alloc(&h->prev, N0);
        if (!st) st = dev_alloc(&h->dflow, 2 * N0);
        if (!st) st = dev_alloc(&h->ctl, 1);
        // x2: the persistent kernel double-buffers its partials by chunk parity
        if (!st) st = dev_alloc(&h->partials, 2 * maxpart);
        if (!st) st = dev_alloc(&h->tile_flags, maxtiles);
        if (!st) st = dev_alloc(&h->energy, 2 * kEnergyBlocks + 3);
        if (!st) st = dev_alloc(&h->d_iters, (size_t)ns * p.warps);
        if (!st) st = host_alloc(&h->hflow, 2 * N0, hipHostMallocDefault);
        if (!st) st = host_alloc(&h->hstatus, 1, hipHostMallocMapped | hipHostMallocCoherent);
        if (!st)
            st = host_alloc(&h->hreport, (size_t)ns * p.warps + 1,
                            hipHostMallocMapped | hipHostMallocCoherent);
        if (st) { release(h); return st; }
        *h->hstatus = 0ull;
        std::memset(h->hreport, 0, sizeof(int) * ((size_t)ns * p.warps + 1));
        std::memset(h->hflow, 0, sizeof(float) * 2 * N0);
        h->h_iters.assign((size_t)ns * p.warps, 0);
        // flags carry the warping's seq (from 1): zero means never set
        if (hipMemset(h->d_iters, 0, sizeof(int) * (size_t)ns * p.warps) != hipSuccess ||
            hipMemset(h->tile_flags, 0, sizeof(unsigned long long) * maxtiles) != hipSuccess) {
            release(h);
            return OF_ERR_HIP;
        }
        *out = h;
        return OF_OK;
    });
}

void of_tvl1_destroy(of_tvl1* h) { release(h); }

int of_tvl1_nscales(const of_tvl1* h) { return h ? h->ns : 0; }

int of_tvl1_scale_size(const of_tvl1* h, int s, int* nx, int* ny)
{
    if (!h || s < 0 || s >= h->ns) return OF_ERR_ARG;
    if (nx) *nx = h->nx[s];
    if (ny) *ny = h->ny[s];
    return OF_OK;
}

int of_tvl1_run_device_u8(of_tvl1* h, const uint8_t* dI0, const uint8_t* dI1, float* dU, void* stream)
{
    return guarded([&]() -> int {
        if (!h || !dI0 || !dI1 || !dU) return OF_ERR_ARG;
        hipStream_t st = (hipStream_t)stream;   // NULL: the HIP null stream
        return solve(h, dI0, dI1, true, dU, st);
    });
}

int of_tvl1_run_device_f32(of_tvl1* h, const float* dI0, const float* dI1, float* dU, void* stream)
{
    return guarded([&]() -> int {
        if (!h || !dI0 || !dI1 || !dU) return OF_ERR_ARG;
        hipStream_t st = (hipStream_t)stream;   // NULL: the HIP null stream
        return solve(h, dI0, dI1, false, dU, st);
    });
}

int of_tvl1_prepare_device_u8(of_tvl1* h, const uint8_t* dI0, const uint8_t* dI1, void* stream)
{
    return guarded([&]() -> int {
        if (!h || !dI0 || !dI1) return OF_ERR_ARG;
        hipStream_t st = (hipStream_t)stream;   // NULL: the HIP null stream
        return tvl1_prepare(h, dI0, dI1, true, st);
    });
}

int of_tvl1_scale_images(const of_tvl1* h, int s, const float** dI0s, const float** dI1s)
{
    if (!h || s < 0 || s >= h->ns || !dI0s || !dI1s) return OF_ERR_ARG;
    if (!h->prepared) return OF_ERR_STATE;
    *dI0s = h->I0p + h->off[s];
    *dI1s = h->I1p + h->off[s];
    return OF_OK;
}

int of_tvl1_set_init_flow(of_tvl1* h, int level, const float* u, const float* v)
{
    return guarded([&]() -> int {
        if (!h) return OF_ERR_ARG;
        if (!u && !v) {
            h->init_level = -1;
            return OF_OK;
        }
        if (!u || !v || level < 0 || level >= h->ns) return OF_ERR_ARG;
        if (!h->init) OF_TRY(dev_alloc(&h->init, 2 * (size_t)h->W * h->H));
        const size_t n = (size_t)h->nx[level] * h->ny[level];
        // host or device pointers alike (one address space); synchronous
        OF_HIP(hipMemcpy(h->init, u, sizeof(float) * n, hipMemcpyDefault));
        OF_HIP(hipMemcpy(h->init + n, v, sizeof(float) * n, hipMemcpyDefault));
        h->init_level = level;
        return OF_OK;
    });
}

int of_tvl1_solve_prepared(of_tvl1* h, const float* dUinit, int first_scale, int warps,
                           int max_iterations, float* dU, void* stream)
{
    return guarded([&]() -> int {
        if (!h || !dU || warps < 0 || warps > h->prm.warps || max_iterations < 0) return OF_ERR_ARG;
        if (first_scale < -1 || first_scale >= h->ns) return OF_ERR_ARG;
        if (!h->prepared) return OF_ERR_STATE;
        hipStream_t st = (hipStream_t)stream;   // NULL: the HIP null stream
        if (first_scale < 0) first_scale = h->ns - 1;
        if (warps == 0) warps = h->prm.warps;
        if (max_iterations == 0) max_iterations = h->prm.max_iterations;
        return solve_levels(h, dUinit, first_scale, warps, max_iterations, dU, st);
    });
}

int of_tvl1_solve_prepared_schedule(of_tvl1* h, const float* dUinit, int first_scale,
                                    const int* warps, const int* max_iterations, float* dU,
                                    void* stream)
{
    return guarded([&]() -> int {
        if (!h || !dU || !warps || !max_iterations) return OF_ERR_ARG;
        if (first_scale < -1 || first_scale >= h->ns) return OF_ERR_ARG;
        if (!h->prepared) return OF_ERR_STATE;
        if (first_scale < 0) first_scale = h->ns - 1;
        int w[kMaxScales], m[kMaxScales];
        for (int s = 0; s < h->ns; s++) {
            if (warps[s] < 0 || warps[s] > h->prm.warps || max_iterations[s] < 0) return OF_ERR_ARG;
            w[s] = warps[s] ? warps[s] : h->prm.warps;
            m[s] = max_iterations[s] ? max_iterations[s] : h->prm.max_iterations;
        }
        return solve_levels_schedule(h, dUinit, first_scale, w, m, dU, (hipStream_t)stream);
    });
}

int of_dbg_tvl1_dual(const float* I0, const float* I1, float* u1, float* u2, int width, int height,
                     const of_tvl1_params* params, int* iters)
{
    return guarded([&]() -> int {
        if (!I0 || !I1 || !u1 || !u2 || !params || !iters) return OF_ERR_ARG;
        of_tvl1_params p = *params;
        p.nscales = 1;
        of_tvl1* h = nullptr;
        OF_TRY(of_tvl1_create(&p, width, height, &h));
        const size_t n = (size_t)width * height;
        float* dinit = nullptr;
        hipStream_t st = nullptr;
        int rc = own_stream(h, &st);
        if (!rc) rc = dev_alloc(&dinit, 2 * n);
        // the frames go straight into scale 0 of the pyramids (off[0] = 0)
        auto up = [&](float* dst, const float* src) {
            if (!rc && hipMemcpy(dst, src, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
                rc = OF_ERR_HIP;
        };
        up(h->I0p, I0);
        up(h->I1p, I1);
        up(dinit, u1);
        up(dinit + n, u2);
        if (!rc) {
            h->prepared = true;
            rc = solve_levels(h, dinit, 0, p.warps, p.max_iterations, h->dflow, st);
        }
        if (!rc && hipMemcpy(u1, h->dflow, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = OF_ERR_HIP;
        if (!rc && hipMemcpy(u2, h->dflow + n, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = OF_ERR_HIP;
        if (!rc) std::copy(h->h_iters.begin(), h->h_iters.begin() + p.warps, iters);
        dev_free(dinit);
        release(h);
        return rc;
    });
}

int of_tvl1_lift_flow(of_tvl1* h, int from_scale, const float* dUfrom, int to_scale, float* dUto,
                      void* stream)
{
    if (!h || !dUfrom || !dUto || from_scale < 0 || from_scale >= h->ns || to_scale < 0 ||
        to_scale > from_scale)
        return OF_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const float invZ = 1 / h->prm.zfactor;
    const float* src = dUfrom;
    float* bufs[2] = {h->A0, h->A1};   // N0 floats each: two planes of any scale >= 1
    if (from_scale == to_scale) {
        OF_HIP(hipMemcpyAsync(dUto, dUfrom, sizeof(float) * 2 * (size_t)h->nx[to_scale] * h->ny[to_scale],
                              hipMemcpyDeviceToDevice, st));
        return OF_OK;
    }
    for (int s = from_scale; s > to_scale; s--) {
        const size_t nc = (size_t)h->nx[s] * h->ny[s], nf = (size_t)h->nx[s - 1] * h->ny[s - 1];
        float* out = (s - 1 == to_scale) ? dUto : bufs[s & 1];
        // zoom_in of both components times 1/zfactor, as between solved scales
        OF_TRY(launch_zoom_in_planes(src, src + nc, out, out + nf, h->nx[s], h->ny[s], h->nx[s - 1],
                                     h->ny[s - 1], invZ, st));
        src = out;
    }
    return OF_OK;
}

int of_tvl1_flow_to_scale(of_tvl1* h, const float* dU, int s, float* dUs, void* stream)
{
    if (!h || !dU || !dUs || s < 0 || s >= h->ns) return OF_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const size_t N0 = (size_t)h->W * h->H;
    float scale = 1.0f;
    for (int c = 0; c < 2; c++) {
        const float* src = dU + c * N0;
        float* ping[2] = {h->wx, h->wy};   // N0 floats each
        for (int l = 0; l < s; l++) {
            float* out = (l + 1 == s) ? dUs + c * (size_t)h->nx[s] * h->ny[s] : ping[l & 1];
            // zoom_out (zoom.cpp:20-53): blur, then sample at (j/zf, i/zf)
            OF_TRY(launch_gaussian(src, h->A0, h->A1, h->nx[l], h->ny[l], h->g_zoom, st));
            OF_TRY(launch_zoom_out_sample(h->A1, out, h->nx[l], h->ny[l], h->nx[l + 1], h->ny[l + 1],
                                          h->prm.zfactor, st));
            src = out;
        }
        if (s == 0)
            OF_HIP(hipMemcpyAsync(dUs + c * N0, src, sizeof(float) * N0, hipMemcpyDeviceToDevice, st));
    }
    for (int l = 0; l < s; l++) scale *= h->prm.zfactor;
    OF_TRY(launch_scale(dUs, 2 * (size_t)h->nx[s] * h->ny[s], scale, st));
    OF_HIP(hipStreamSynchronize(st));
    return OF_OK;
}

int of_tvl1_scale_flow(of_tvl1* h, int s, float* dUs, void* stream)
{
    if (!h || s < 0 || s >= h->ns || !dUs) return OF_ERR_ARG;
    if (!h->solved) return OF_ERR_STATE;
    hipStream_t st = (hipStream_t)stream;   // NULL: the HIP null stream
    const size_t N = (size_t)h->nx[s] * h->ny[s];
    float* ub = h->ubase + 4 * h->off[s];
    OF_TRY(launch_copy_flow(ub, ub + 2 * N, dUs, 2 * N, h->ctl, s, st));
    OF_HIP(hipStreamSynchronize(st));
    return OF_OK;
}

int of_tvl1_energy(of_tvl1* h, int s, const float* dUs, double* energy, void* stream)
{
    if (!h || s < 0 || s >= h->ns || !dUs) return OF_ERR_ARG;
    if (!h->prepared) return OF_ERR_STATE;
    hipStream_t st = (hipStream_t)stream;   // NULL: the HIP null stream
    const size_t N = (size_t)h->nx[s] * h->ny[s];
    OF_TRY(launch_energy(h->I0p + h->off[s], h->I1p + h->off[s], dUs, dUs + N, h->nx[s], h->ny[s],
                         h->prm.lambda, h->energy, kEnergyBlocks, h->energy + 2 * kEnergyBlocks, st));
    double e[3];
    OF_HIP(hipMemcpyAsync(e, h->energy + 2 * kEnergyBlocks, sizeof(e), hipMemcpyDeviceToHost, st));
    OF_HIP(hipStreamSynchronize(st));
    if (energy) {
        energy[0] = e[0];
        energy[1] = e[1];
        energy[2] = e[2];
    }
    return OF_OK;
}

int of_tvl1_run_u8(of_tvl1* h, const uint8_t* I0, const uint8_t* I1)
{
    return guarded([&]() -> int {
        if (!h || !I0 || !I1) return OF_ERR_ARG;
        hipStream_t hs = nullptr;
        OF_TRY(own_stream(h, &hs));
        const size_t N0 = (size_t)h->W * h->H;
        OF_HIP(hipMemcpyAsync(h->in8, I0, N0, hipMemcpyHostToDevice, hs));
        OF_HIP(hipMemcpyAsync(h->in8 + N0, I1, N0, hipMemcpyHostToDevice, hs));
        OF_TRY(solve(h, h->in8, h->in8 + N0, true, h->dflow, hs));
        OF_HIP(hipMemcpyAsync(h->hflow, h->dflow, sizeof(float) * 2 * N0, hipMemcpyDeviceToHost, hs));
        OF_HIP(hipStreamSynchronize(hs));
        return OF_OK;
    });
}

int of_tvl1_run_f32(of_tvl1* h, const float* I0, const float* I1)
{
    return guarded([&]() -> int {
        if (!h || !I0 || !I1) return OF_ERR_ARG;
        hipStream_t hs = nullptr;
        OF_TRY(own_stream(h, &hs));
        const size_t N0 = (size_t)h->W * h->H;
        OF_HIP(hipMemcpyAsync(h->inf, I0, sizeof(float) * N0, hipMemcpyHostToDevice, hs));
        OF_HIP(hipMemcpyAsync(h->inf + N0, I1, sizeof(float) * N0, hipMemcpyHostToDevice, hs));
        OF_TRY(solve(h, h->inf, h->inf + N0, false, h->dflow, hs));
        OF_HIP(hipMemcpyAsync(h->hflow, h->dflow, sizeof(float) * 2 * N0, hipMemcpyDeviceToHost, hs));
        OF_HIP(hipStreamSynchronize(hs));
        return OF_OK;
    });
}

int of_tvl1_push_f32(of_tvl1* h, const float* I)
{
    return guarded([&]() -> int {
        if (!h || !I) return OF_ERR_ARG;
        hipStream_t hs = nullptr;
        OF_TRY(own_stream(h, &hs));
        const size_t N0 = (size_t)h->W * h->H;
        if (!h->have_prev) {
            OF_HIP(hipMemcpyAsync(h->prev, I, sizeof(float) * N0, hipMemcpyHostToDevice, hs));
            OF_HIP(hipStreamSynchronize(hs));
            std::memset(h->hflow, 0, sizeof(float) * 2 * N0);
            h->have_prev = true;
            return OF_OK;
        }
        // I0 = previous frame, I1 = this frame (src/tvl1/cuda/tvl1.cu:126-130)
        OF_HIP(hipMemcpyAsync(h->inf, h->prev, sizeof(float) * N0, hipMemcpyDeviceToDevice, hs));
        OF_HIP(hipMemcpyAsync(h->inf + N0, I, sizeof(float) * N0, hipMemcpyHostToDevice, hs));
        OF_TRY(solve(h, h->inf, h->inf + N0, false, h->dflow, hs));
        OF_HIP(hipMemcpyAsync(h->prev, h->inf + N0, sizeof(float) * N0, hipMemcpyDeviceToDevice, hs));
        OF_HIP(hipMemcpyAsync(h->hflow, h->dflow, sizeof(float) * 2 * N0, hipMemcpyDeviceToHost, hs));
        OF_HIP(hipStreamSynchronize(hs));
        return OF_OK;
    });
}

const float* of_tvl1_flow(of_tvl1* h) { return h ? h->hflow : nullptr; }

int of_tvl1_iterations(const of_tvl1* h, int* counts, int cap)
{
    if (!h || !counts) return OF_ERR_ARG;
    const int n = (int)h->h_iters.size();
    for (int i = 0; i < n && i < cap; i++) counts[i] = h->h_iters[i];
    return OF_OK;
}

int of_tvl1_get_stats(const of_tvl1* h, of_tvl1_stats* st)
{
    if (!h || !st) return OF_ERR_ARG;
    st->nscales = h->ns;
    st->warps = h->prm.warps;
    long long pi = 0, sumN = 0;
    for (int s = 0; s < h->ns; s++) {
        const long long N = (long long)h->nx[s] * h->ny[s];
        sumN += N;
        for (int w = 0; w < h->prm.warps; w++) pi += (long long)h->h_iters[(size_t)s * h->prm.warps + w] * N;
    }
    st->pixel_iterations = pi;
    long long pw = 0;
    for (int s = 0; s < h->ns; s++)
        pw += (long long)(h->solved ? h->run_warps_s[s] : h->prm.warps) * h->nx[s] * h->ny[s];
    st->pixel_warps = pw;
    st->pixels_all_scales = sumN;
    st->launches = h->launches_issued;
    st->host_launch_seconds = h->host_launch_s;
    st->host_wait_seconds = h->host_wait_s;
    const long long N0 = (long long)h->W * h->H;
    st->algorithmic_bytes = 60.0 * (double)pi + 36.0 * (double)st->pixel_warps + 64.0 * (double)sumN +
                            48.0 * (double)N0;
    return OF_OK;
}

int of_tvl1_autotune(of_tvl1* h, const uint8_t* dI0, const uint8_t* dI1, int reps, int* chosen,
                     double* ms_table)
{
    return guarded([&]() -> int {
        if (!h || !dI0 || !dI1) return OF_ERR_ARG;
        return autotune(h, dI0, dI1, reps, chosen, ms_table);
    });
}

int of_tvl1_set_inflight(of_tvl1* h, int launches)
{
    if (!h || launches < 1 || launches > 64) return OF_ERR_ARG;
    h->inflight = launches;
    return OF_OK;
}

int of_tvl1_profile_enable(of_tvl1* h, int enable)
{
    return guarded([&]() -> int {
        if (!h) return OF_ERR_ARG;
        h->profiling = enable != 0;
        h->prof.clear();
        h->timed.clear();
        return OF_OK;
    });
}

int of_tvl1_scale_timing_enable(of_tvl1* h, int enable)
{
    return guarded([&]() -> int {
        if (!h) return OF_ERR_ARG;
        if (enable && h->scale_ev.empty()) {
            h->scale_ev.resize(2 * kMaxScales);
            for (hipEvent_t& e : h->scale_ev) OF_HIP(hipEventCreate(&e));
        }
        h->scale_timing = enable != 0;
        for (double& m : h->scale_ms) m = 0.0;
        h->scale_solves = 0;
        return OF_OK;
    });
}

int of_tvl1_scale_times(const of_tvl1* h, double* ms, int cap, int* solves)
{
    if (!h || (!ms && cap > 0)) return OF_ERR_ARG;
    for (int s = 0; s < h->ns && s < cap; s++) ms[s] = h->scale_ms[s];
    if (solves) *solves = h->scale_solves;
    return OF_OK;
}

int of_tvl1_profile_get(const of_tvl1* h, of_tvl1_kernel_profile* out, int cap, int* count)
{
    if (!h || (!out && cap > 0)) return OF_ERR_ARG;
    const int n = (int)h->prof.size();
    for (int i = 0; i < n && i < cap; i++) out[i] = h->prof[i];
    if (count) *count = n;
    return OF_OK;
}

int of_tvl1_set_launch_rule(of_tvl1* h, long long min_pixels, int iters_per_launch, int rows_per_wave,
                            int persistent)
{
    return guarded([&]() -> int {
        if (!h) return OF_ERR_ARG;
        if (min_pixels < 0) {
            h->rules.clear();
            return OF_OK;
        }
        if (!iter_config_ok(iters_per_launch, rows_per_wave)) return OF_ERR_UNSUPPORTED;
        if (persistent < kHostPaced || persistent > kPersistNeighbour) return OF_ERR_ARG;
        const int p = persistent;
        for (LaunchRule& r : h->rules)
            if (r.min_pixels == min_pixels) {
                r.K = iters_per_launch;
                r.R = rows_per_wave;
                r.persistent = p;
                return OF_OK;
            }
        h->rules.push_back({min_pixels, iters_per_launch, rows_per_wave, p});
        return OF_OK;
    });
}

int of_persist_launch_mode(void)
{
    return persist_host_paced() ? OF_PERSIST_HOST_PACED : OF_PERSIST_PLAIN;
}

int of_set_persist_launch_mode(int mode)
{
    if (mode != OF_PERSIST_PLAIN && mode != OF_PERSIST_HOST_PACED) return OF_ERR_ARG;
    set_persist_host_paced(mode == OF_PERSIST_HOST_PACED);
    return OF_OK;
}

}  // extern "C"
