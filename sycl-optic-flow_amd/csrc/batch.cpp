// batch.cpp -- many TV-L1 frame pairs of mixed sizes solved together.
//
// No reference counterpart (the reference solves one pair per call).  Two
// modes, both with results identical to solving each pair alone:
//
//   lockstep (default): every pair gets a TV-L1 handle of its own (pooled
//     by frame size).  All work runs on one stream, and every launch covers
//     all pairs: the preparation (frames to float, normalisation,
//     pre-smoothing, pyramids) step by step, then the solve scale by scale
//     and warping by warping.
//     Iteration launches walk a device work list of (pair, tile) entries;
//     each pair stops on its own rule and drops out of the list, and the
//     host only paces launches until every pair of the warping has
//     stopped.  A coarse scale of 64 pairs is then one launch of a few
//     thousand tiles instead of 64 latency-bound launches of a few dozen.
//   workers: `streams` host threads, each with its own stream and one
//     handle per frame size, take pairs longest first and solve them one
//     by one with host-paced launches; the workers' launches interleave on
//     the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

#include "amdflow.h"
#include "common.hpp"
#include "lockstep.hpp"
#include "tvl1_handle.hpp"
#include "tvl1_launch.hpp"

// One lockstep lane: a group of pairs solved on its own stream, with its
// own device control block and tables.  The calling thread advances all
// lanes in turn without blocking, so the launch-latency-bound coarse scales
// of one lane overlap the bandwidth-bound fine scales of another.
struct LsLane {
    hipStream_t stream = nullptr;             // one of the batch's streams
    amdflow::LsCtl* mc = nullptr;             // device control of the lane
    unsigned long long* hstatus = nullptr;    // host-mapped progress word
    unsigned epoch = 0;
    amdflow::LsPrep* h_preps = nullptr;       // pinned tables, uploaded once per group
    amdflow::LsPrep* d_preps = nullptr;
    size_t preps_cap = 0;
    amdflow::LsPair* h_pairs = nullptr;
    amdflow::LsPair* d_pairs = nullptr;
    size_t pairs_cap = 0;
    unsigned* h_bases = nullptr;
    unsigned* d_bases = nullptr;
    size_t bases_cap = 0;
    int* h_iters = nullptr;
    int* d_iters = nullptr;
    size_t iters_cap = 0;
    std::map<std::pair<int, int>, unsigned> resident;   // (K, R) -> co-resident workgroups

    // the group in flight (see ls_begin / ls_step)
    std::vector<int> members;                 // batch indices of its pairs
    std::vector<amdflow::LsPhase> phases;     // per scale
    std::vector<int> shapeK, shapeR;
    size_t niters = 0;
    int s = -1, w = 0;                        // running scale and warping
    bool done = true;
    long long issued = 0, completed = 0, max_launches = 0;
    unsigned grid = 0, run_epoch = 0, spins = 0;
    std::chrono::steady_clock::time_point t0;
};

struct of_tvl1_batch {
    of_tvl1_params prm{};
    int streams = 1;
    int device = 0;
    // per worker: its stream and its handles (frame size -> handle)
    std::vector<hipStream_t> wstreams;
    std::vector<std::map<std::pair<int, int>, of_tvl1*>> handles;
    // launch rules applied to every handle: (min_pixels, K, R)
    std::vector<std::tuple<long long, int, int>> rules;
    int mode = OF_BATCH_LOCKSTEP;
    // lockstep mode
    std::map<std::pair<int, int>, std::vector<of_tvl1*>> pool;   // handles by frame size
    std::vector<LsLane> lanes;
};

// Batch autotuner: the candidate shapes, and the scale size (pixels) that
// separates "large" from "small" scales when two shapes are combined.
constexpr int kBatchTuneShapes = OF_BATCH_TUNE_SHAPES;
constexpr long long kBatchTuneSplit = 200000;

namespace {

int worker_handle(of_tvl1_batch* b, int w, int W, int H, of_tvl1** out)
{
    auto& m = b->handles[w];
    auto it = m.find({W, H});
    if (it != m.end()) {
        *out = it->second;
        return OF_OK;
    }
    of_tvl1* h = nullptr;
    OF_TRY(of_tvl1_create(&b->prm, W, H, &h));
    // host-paced launches at every scale
    for (int s = 0; s < of_tvl1_nscales(h); s++) {
        int nx, ny;
        of_tvl1_scale_size(h, s, &nx, &ny);
        const long long npx = (long long)nx * ny;
        int K = npx >= 200000 ? 2 : 4, R = npx >= 200000 ? 8 : 4;
        for (const auto& r : b->rules)
            if (npx >= std::get<0>(r)) {
                K = std::get<1>(r);
                R = std::get<2>(r);
                break;
            }
        const int st = of_tvl1_set_launch_rule(h, npx, K, R, 0);
        if (st) {
            of_tvl1_destroy(h);
            return st;
        }
    }
    m[{W, H}] = h;
    *out = h;
    return OF_OK;
}

// The batch itself: workers take pairs longest first until the list is
// empty or one of them fails.
int run_batch(of_tvl1_batch* b, int npairs, const int* widths, const int* heights,
              const uint8_t* const* dI0, const uint8_t* const* dI1, float* const* dU,
              int* iterations_total)
{
    // longest first: the largest frames start early and small ones fill in
    std::vector<int> order(npairs);
    for (int i = 0; i < npairs; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return (long long)widths[x] * heights[x] > (long long)widths[y] * heights[y];
    });
    std::atomic<int> next{0};
    std::atomic<int> status{OF_OK};
    std::vector<long long> iters(npairs, 0);
    auto fail = [&](int st) {
        int expected = OF_OK;
        status.compare_exchange_strong(expected, st);
    };
    auto work = [&](int w) noexcept {
        try {
            if (hipSetDevice(b->device) != hipSuccess) {
                fail(OF_ERR_HIP);
                return;
            }
            while (status.load() == OF_OK) {
                const int k = next.fetch_add(1);
                if (k >= npairs) break;
                const int i = order[k];
                of_tvl1* h = nullptr;
                int st = worker_handle(b, w, widths[i], heights[i], &h);
                if (!st) st = of_tvl1_run_device_u8(h, dI0[i], dI1[i], dU[i], b->wstreams[w]);
                if (!st && iterations_total) {
                    std::vector<int> c((size_t)of_tvl1_nscales(h) * b->prm.warps);
                    of_tvl1_iterations(h, c.data(), (int)c.size());
                    long long sum = 0;
                    for (int v : c) sum += v;
                    iters[i] = sum;
                }
                if (st) fail(st);
            }
        } catch (...) {
            fail(OF_ERR_ALLOC);
        }
    };
    const int nw = std::min(b->streams, std::max(npairs, 1));
    std::vector<std::thread> threads;
    try {
        threads.reserve(nw);
        for (int w = 1; w < nw; w++) threads.emplace_back(work, w);
    } catch (...) {
        fail(OF_ERR_ALLOC);   // the workers that did start stop at their next pair
    }
    work(0);
    for (auto& t : threads) t.join();
    if (iterations_total)
        for (int i = 0; i < npairs; i++) iterations_total[i] = (int)iters[i];
    return status.load();
}

// ---------------------------------------------------------------------------
// lockstep mode
// ---------------------------------------------------------------------------

void ls_release(of_tvl1_batch* b)
{
    for (auto& kv : b->pool)
        for (of_tvl1* h : kv.second) of_tvl1_destroy(h);
    b->pool.clear();
    for (LsLane& L : b->lanes) {
        amdflow::dev_free(L.mc);
        amdflow::dev_free(L.d_preps);
        amdflow::dev_free(L.d_pairs);
        amdflow::dev_free(L.d_bases);
        amdflow::dev_free(L.d_iters);
        amdflow::host_free(L.hstatus);
        amdflow::host_free(L.h_preps);
        amdflow::host_free(L.h_pairs);
        amdflow::host_free(L.h_bases);
        amdflow::host_free(L.h_iters);
    }
    b->lanes.clear();
}

// Pinned host table and its device twin of at least n elements.  Only
// called between groups, when no copy from the old buffers is pending.
template <typename T>
int ls_reserve(T** host, T** dev, size_t* cap, size_t n)
{
    if (n <= *cap) return OF_OK;
    amdflow::dev_free(*dev);
    amdflow::host_free(*host);
    *cap = 0;
    const size_t want = std::max(n, (size_t)64);
    OF_TRY(amdflow::host_alloc(host, want, hipHostMallocDefault));
    OF_TRY(amdflow::dev_alloc(dev, want));
    *cap = want;
    return OF_OK;
}

int ls_setup(LsLane& L)
{
    if (L.mc) return OF_OK;
    OF_TRY(amdflow::dev_alloc(&L.mc, 1));
    OF_HIP(hipMemset(L.mc, 0, sizeof(amdflow::LsCtl)));
    OF_TRY(amdflow::host_alloc(&L.hstatus, 1, hipHostMallocMapped | hipHostMallocCoherent));
    *L.hstatus = 0ull;
    return OF_OK;
}

constexpr size_t kLsMaxPairsPerGroup = amdflow::kLsMaxPairs;

// Launch shape of a scale: the batch's rule for the largest participating
// scale size, else 4 iterations per launch with 8 rows per wave at every
// scale.  (A batch's working set is far beyond the 256 MB Infinity Cache,
// so deeper temporal blocking pays: on the 64-pair mixed batch (4,8) beat
// the one-pair choice (2,8) at the fine scales by 6%, and (4,4) at the
// coarse ones by 2%.)
void ls_shape(const of_tvl1_batch* b, long long npx, int* K, int* R)
{
    *K = 4;
    *R = 8;
    for (const auto& r : b->rules)
        if (npx >= std::get<0>(r)) {
            *K = std::get<1>(r);
            *R = std::get<2>(r);
            return;
        }
}

// Starts the iteration launches of warping L.w of scale L.s.
int ls_start_warping(const of_tvl1_batch* b, LsLane& L)
{
    using namespace amdflow;
    const int K = L.shapeK[L.s], R = L.shapeR[L.s];
    auto key = std::make_pair(K, R);
    auto it = L.resident.find(key);
    if (it == L.resident.end()) {
        unsigned blocks = 0;
        OF_TRY(ls_iter_resident(K, R, &blocks));
        it = L.resident.emplace(key, blocks).first;
    }
    LsPhase ph = L.phases[L.s];
    ph.warp = L.w;
    L.phases[L.s].warp = L.w;
    OF_TRY(launch_ls_warp_setup(ph, L.mc, L.stream));
    L.grid = std::max(1u, std::min(ph.ntiles, it->second));
    L.run_epoch = ++L.epoch;
    L.max_launches = (b->prm.max_iterations + K - 1) / K + 1;
    L.issued = L.completed = 0;
    L.spins = 0;
    L.t0 = std::chrono::steady_clock::now();
    return OF_OK;
}

// One lockstep group of pairs `members` (at most kLsMaxPairs, with their
// handles hs[]) on lane L: builds and uploads the tables, enqueues the
// batched preparation and starts the coarsest scale.  ls_step does the rest.
int ls_begin(const of_tvl1_batch* b, LsLane& L, const std::vector<int>& members,
             of_tvl1* const* hall, const uint8_t* const* dI0, const uint8_t* const* dI1,
             float* const* dU)
{
    using namespace amdflow;
    OF_TRY(ls_setup(L));
    hipStream_t main = L.stream;
    const of_tvl1_params& P = b->prm;
    const int n = (int)members.size();
    L.members = members;
    std::vector<of_tvl1*> hs(n);
    for (int i = 0; i < n; i++) hs[i] = hall[members[i]];

    // tables: per scale, the pairs that have it
    int maxns = 0;
    for (int i = 0; i < n; i++) maxns = std::max(maxns, hs[i]->ns);
    std::vector<std::vector<int>> part(maxns);
    size_t nent = 0, nbase = 0;
    for (int s = 0; s < maxns; s++) {
        for (int i = 0; i < n; i++)
            if (hs[i]->ns > s) part[s].push_back(i);
        nent += part[s].size();
        // solve tables (3 prefix arrays) and preparation tables (participants
        // and 2 prefix arrays at s >= 1; participants and 3 at s = 0)
        nbase += 3 * (part[s].size() + 1) + part[s].size() + 3 * (part[s].size() + 1);
    }
    L.niters = (size_t)n * kMaxScales * P.warps;
    OF_TRY(ls_reserve(&L.h_preps, &L.d_preps, &L.preps_cap, (size_t)n));
    OF_TRY(ls_reserve(&L.h_pairs, &L.d_pairs, &L.pairs_cap, nent));
    OF_TRY(ls_reserve(&L.h_bases, &L.d_bases, &L.bases_cap, nbase));
    OF_TRY(ls_reserve(&L.h_iters, &L.d_iters, &L.iters_cap, L.niters));

    L.phases.assign(maxns, LsPhase{});
    L.shapeK.assign(maxns, 0);
    L.shapeR.assign(maxns, 0);
    size_t e = 0, q = 0;
    for (int s = maxns - 1; s >= 0; s--) {
        const std::vector<int>& ps = part[s];
        const int np = (int)ps.size();
        long long maxpx = 0;
        for (int i : ps) maxpx = std::max(maxpx, (long long)hs[i]->nx[s] * hs[i]->ny[s]);
        int K, R;
        ls_shape(b, maxpx, &K, &R);
        L.shapeK[s] = K;
        L.shapeR[s] = R;
        LsPhase& ph = L.phases[s];
        ph.pairs = L.d_pairs + e;
        ph.blk_base = L.d_bases + q;
        ph.fblk_base = L.d_bases + q + (np + 1);
        ph.tile_base = L.d_bases + q + 2 * (np + 1);
        ph.iters = L.d_iters;
        ph.npairs = np;
        ph.level = s;
        ph.warps = P.warps;
        ph.warp = 0;
        unsigned* blk = L.h_bases + q;
        unsigned* fblk = blk + (np + 1);
        unsigned* tiles = fblk + (np + 1);
        blk[0] = fblk[0] = tiles[0] = 0;
        for (int k = 0; k < np; k++) {
            const int i = ps[k];
            of_tvl1* h = hs[i];
            const int nx = h->nx[s], ny = h->ny[s];
            const size_t N = (size_t)nx * ny;
            LsPair& E = L.h_pairs[e + k];
            E = LsPair{};
            IterArgs& a = E.a;
            a.wx = h->wx;
            a.wy = h->wy;
            a.rhoc = h->rhoc;
            float* ub = h->ubase + 4 * h->off[s];
            a.u[0] = ub;
            a.u[1] = ub + 2 * N;
            a.p[0] = h->pbase;
            a.p[1] = h->pbase + 4 * (size_t)h->W * h->H;
            a.nx = nx;
            a.ny = ny;
            a.plane = (int)N;
            a.l_t = P.lambda * P.theta;
            a.theta = P.theta;
            a.taut = P.tau / P.theta;
            a.eps2 = (double)(P.epsilon * P.epsilon);
            a.max_iter = P.max_iterations;
            a.level = s;
            a.ctl = h->ctl;
            a.partials = h->partials;
            E.I0s = h->I0p + h->off[s];
            E.I1s = h->I1p + h->off[s];
            E.wsrc = reinterpret_cast<float4*>(h->wsrc);
            E.p0 = h->pbase;
            E.dst = dU[members[i]];
            if (s == h->ns - 1) E.ucoarse = ub;   // zero flow, tvl1.cpp:157-158
            E.invz = 1 / P.zfactor;
            E.pair_id = (unsigned)i;
            unsigned gx, gy;
            iter_grid(K, R, nx, ny, &gx, &gy);
            E.gx = gx;
            E.ntiles = gx * gy;
            E.bx = (unsigned)cdiv((size_t)nx, 64);
            const unsigned nb = E.bx * (unsigned)cdiv((size_t)ny, 4);
            unsigned nfb = 0;
            if (s > 0) {
                E.ufine = h->ubase + 4 * h->off[s - 1];
                E.nxf = h->nx[s - 1];
                E.nyf = h->ny[s - 1];
                E.fx = ((float)E.nxf / (float)nx);
                E.fy = ((float)E.nyf / (float)ny);
                E.bxf = (unsigned)cdiv((size_t)E.nxf, 64);
                nfb = E.bxf * (unsigned)cdiv((size_t)E.nyf, 4);
            }
            blk[k + 1] = blk[k] + nb;
            fblk[k + 1] = fblk[k] + nfb;
            tiles[k + 1] = tiles[k] + E.ntiles;
        }
        ph.nblk = blk[np];
        ph.nfblk = fblk[np];
        ph.ntiles = tiles[np];
        e += np;
        q += 3 * (size_t)(np + 1);
    }

    // preparation tables
    for (int i = 0; i < n; i++) {
        of_tvl1* h = hs[i];
        LsPrep& R = L.h_preps[i];
        R = LsPrep{};
        R.s0 = dI0[members[i]];
        R.s1 = dI1[members[i]];
        R.A0 = h->A0;
        R.A1 = h->A1;
        R.B0 = h->wx;
        R.B1 = h->wy;
        R.I0p = h->I0p;
        R.I1p = h->I1p;
        R.ctl = h->ctl;
        R.n0 = (unsigned)((size_t)h->W * h->H);
        R.nmm = std::min((unsigned)cdiv(R.n0, 256), (unsigned)kMinMaxBlocks);
        R.ns = h->ns;
        for (int s = 0; s < h->ns; s++) {
            R.nx[s] = h->nx[s];
            R.ny[s] = h->ny[s];
            R.off[s] = (unsigned)h->off[s];
        }
    }
    // one step: participants and the prefix sums of their item counts
    auto add_step = [&](const std::vector<int>& ps, const unsigned* part_dev, int s,
                        auto&& items) -> LsPrepStep {
        LsPrepStep st{};
        st.preps = L.d_preps;
        st.part = part_dev;
        st.base = L.d_bases + q;
        st.npart = (int)ps.size();
        st.s = s;
        unsigned* base = L.h_bases + q;
        base[0] = 0;
        for (size_t k = 0; k < ps.size(); k++) base[k + 1] = base[k] + items(hs[ps[k]]);
        st.nitems = base[ps.size()];
        q += ps.size() + 1;
        return st;
    };
    auto blocks = [](int nx, int ny) { return (unsigned)(cdiv((size_t)nx, 64) * cdiv((size_t)ny, 4)); };
    std::vector<const unsigned*> part_dev(maxns);
    for (int s = 0; s < maxns; s++) {
        part_dev[s] = L.d_bases + q;
        for (size_t k = 0; k < part[s].size(); k++) L.h_bases[q + k] = (unsigned)part[s][k];
        q += part[s].size();
    }
    const LsPrepStep st_load = add_step(part[0], part_dev[0], 0, [&](of_tvl1* h) {
        return std::min((unsigned)cdiv((size_t)h->W * h->H, 256), (unsigned)kMinMaxBlocks);
    });
    const LsPrepStep st_norm = add_step(part[0], part_dev[0], 0,
                                        [&](of_tvl1* h) { return (unsigned)cdiv((size_t)h->W * h->H, 256); });
    std::vector<LsPrepStep> st_blur(maxns), st_zoom(maxns);
    for (int s = 0; s < maxns; s++) {
        const int sc = s > 0 ? s - 1 : 0;
        st_blur[s] = add_step(part[s], part_dev[s], s,
                              [&](of_tvl1* h) { return 2 * blocks(h->nx[sc], h->ny[sc]); });
        if (s > 0)
            st_zoom[s] = add_step(part[s], part_dev[s], s,
                                  [&](of_tvl1* h) { return 2 * blocks(h->nx[s], h->ny[s]); });
    }
    if (q > nbase) return OF_ERR_STATE;   // table sizing above is wrong

    OF_HIP(hipMemcpyAsync(L.d_preps, L.h_preps, sizeof(LsPrep) * n, hipMemcpyHostToDevice, main));
    OF_HIP(hipMemcpyAsync(L.d_pairs, L.h_pairs, sizeof(LsPair) * nent, hipMemcpyHostToDevice, main));
    OF_HIP(hipMemcpyAsync(L.d_bases, L.h_bases, sizeof(unsigned) * q, hipMemcpyHostToDevice, main));
    OF_HIP(hipMemsetAsync(L.d_iters, 0, sizeof(int) * L.niters, main));

    // phase 1 of all pairs (tvl1.cpp:137-154): frames to float with the
    // min/max keys, control blocks, normalisation, pre-smoothing, pyramids
    const GaussTaps& g_pre = hs[0]->g_pre;    // the same for every handle of the batch
    const GaussTaps& g_zoom = hs[0]->g_zoom;
    OF_TRY(launch_lp_load(st_load, main));
    OF_TRY(launch_lp_final(L.d_preps, n, main));
    OF_TRY(launch_lp_normalize(st_norm, main));
    OF_TRY(launch_lp_blur(st_blur[0], g_pre, main));
    for (int s = 1; s < maxns; s++) {
        OF_TRY(launch_lp_blur(st_blur[s], g_zoom, main));
        OF_TRY(launch_lp_zoom_out(st_zoom[s], P.zfactor, main));
    }

    // phase 2 starts at the coarsest scale (tvl1.cpp:157-174)
    L.done = false;
    L.s = maxns - 1;
    L.w = 0;
    OF_TRY(launch_ls_level_begin(L.phases[L.s], main));
    return ls_start_warping(b, L);
}

// Advances lane L without waiting: keeps kInFlight iteration launches
// queued; when the device reports that every pair of the warping stopped
// (or the launch bound is reached), enqueues the next warping, scale or the
// group's end.  *progress is set when something was enqueued or completed.
int ls_step(const of_tvl1_batch* b, LsLane& L, bool* progress)
{
    using namespace amdflow;
    if (L.done) return OF_OK;
    const int K = L.shapeK[L.s], R = L.shapeR[L.s];
    bool stopped = false;
    const unsigned long long word = __atomic_load_n(L.hstatus, __ATOMIC_ACQUIRE);
    if ((unsigned)(word >> 32) == L.run_epoch) {
        const long long c = (long long)(word & 0x7FFFFFFFull);
        if (c != L.completed) *progress = true;
        L.completed = c;
        stopped = ((word >> 31) & 1ull) != 0;
    }
    if (!stopped) {
        while (L.issued < L.max_launches && L.issued - L.completed < kInFlight) {
            OF_TRY(launch_ls_iter(K, R, L.phases[L.s], L.mc, L.grid, L.hstatus, L.run_epoch, L.stream));
            L.issued++;
            *progress = true;
        }
        if (L.issued < L.max_launches) {
            if (++L.spins % 4096 == 0) {
                const hipError_t q = hipStreamQuery(L.stream);
                if (q != hipSuccess && q != hipErrorNotReady) {
                    record_hip_error(q, "hipStreamQuery", __FILE__, __LINE__);
                    return OF_ERR_HIP;
                }
                if (q == hipSuccess) {
                    // idle stream: every issued launch has reported
                    const unsigned long long w2 = __atomic_load_n(L.hstatus, __ATOMIC_ACQUIRE);
                    const bool ok = (unsigned)(w2 >> 32) == L.run_epoch &&
                                    (((w2 >> 31) & 1ull) || (long long)(w2 & 0x7FFFFFFFull) >= L.issued);
                    if (!ok) return OF_ERR_STATE;
                }
                if (std::chrono::duration<double>(std::chrono::steady_clock::now() - L.t0).count() > 60.0)
                    return OF_ERR_HIP;
            }
            return OF_OK;   // the warping is still running
        }
    }
    // the warping is over (launches still queued after the stop are no-ops)
    *progress = true;
    const of_tvl1_params& P = b->prm;
    if (++L.w < P.warps) return ls_start_warping(b, L);
    if (L.s > 0) OF_TRY(launch_ls_zoom_in(L.phases[L.s], L.stream));
    if (--L.s >= 0) {
        L.w = 0;
        OF_TRY(launch_ls_level_begin(L.phases[L.s], L.stream));
        return ls_start_warping(b, L);
    }
    OF_TRY(launch_ls_copy_flow(L.phases[0], L.stream));
    OF_HIP(hipMemcpyAsync(L.h_iters, L.d_iters, sizeof(int) * L.niters, hipMemcpyDeviceToHost, L.stream));
    L.done = true;
    return OF_OK;
}

int run_lockstep(of_tvl1_batch* b, int npairs, const int* widths, const int* heights,
                 const uint8_t* const* dI0, const uint8_t* const* dI1, float* const* dU,
                 int* iterations_total)
{
    if (npairs == 0) return OF_OK;
    if (hipSetDevice(b->device) != hipSuccess) return OF_ERR_HIP;
    const int nl = std::max(1, std::min(b->streams, npairs));
    while ((int)b->lanes.size() < nl) {
        b->lanes.emplace_back();
        b->lanes.back().stream = b->wstreams[b->lanes.size() - 1];
    }
    // a handle per pair, pooled by frame size
    std::vector<of_tvl1*> hs(npairs, nullptr);
    {
        std::map<std::pair<int, int>, size_t> used;
        for (int i = 0; i < npairs; i++) {
            const auto key = std::make_pair(widths[i], heights[i]);
            auto& v = b->pool[key];
            const size_t k = used[key]++;
            while (v.size() <= k) {
                of_tvl1* h = nullptr;
                OF_TRY(of_tvl1_create(&b->prm, widths[i], heights[i], &h));
                v.push_back(h);
            }
            hs[i] = v[k];
        }
    }
    // pairs to lanes, largest first onto the lane with the fewest pixels
    std::vector<int> order(npairs);
    for (int i = 0; i < npairs; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return (long long)widths[x] * heights[x] > (long long)widths[y] * heights[y];
    });
    std::vector<std::vector<int>> mine(nl);
    std::vector<long long> load(nl, 0);
    for (int i : order) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        mine[l].push_back(i);
        load[l] += (long long)widths[i] * heights[i];
    }
    for (auto& m : mine) std::sort(m.begin(), m.end());
    // If a step fails (or throws), kernels already queued on other lanes
    // still write into the caller's buffers and this batch's tables: wait for
    // every lane before returning.
    struct Drain {
        of_tvl1_batch* b;
        int nl;
        bool armed = true;
        ~Drain()
        {
            if (!armed) return;
            for (int l = 0; l < nl; l++) (void)hipStreamSynchronize(b->lanes[l].stream);
            (void)hipGetLastError();
        }
    } drain{b, nl};
    // rounds of at most kLsMaxPairs pairs per lane, all lanes advanced in turn
    for (size_t g = 0;; g += kLsMaxPairsPerGroup) {
        bool any = false;
        for (int l = 0; l < nl; l++) {
            LsLane& L = b->lanes[l];
            L.done = true;
            if (g >= mine[l].size()) continue;
            const size_t n = std::min(mine[l].size() - g, kLsMaxPairsPerGroup);
            const std::vector<int> members(mine[l].begin() + g, mine[l].begin() + g + n);
            OF_TRY(ls_begin(b, L, members, hs.data(), dI0, dI1, dU));
            any = true;
        }
        if (!any) break;
        for (;;) {
            bool all = true, progress = false;
            for (int l = 0; l < nl; l++) {
                LsLane& L = b->lanes[l];
                OF_TRY(ls_step(b, L, &progress));
                all = all && L.done;
            }
            if (all) break;
            if (!progress) __builtin_ia32_pause();
        }
        for (int l = 0; l < nl; l++) {
            LsLane& L = b->lanes[l];
            if (L.members.empty() || g >= mine[l].size()) continue;
            OF_HIP(hipStreamSynchronize(L.stream));
            if (iterations_total)
                for (size_t k = 0; k < L.members.size(); k++) {
                    long long sum = 0;
                    for (size_t j = 0; j < (size_t)amdflow::kMaxScales * b->prm.warps; j++)
                        sum += L.h_iters[k * amdflow::kMaxScales * b->prm.warps + j];
                    iterations_total[L.members[k]] = (int)sum;
                }
        }
    }
    drain.armed = false;   // every lane was synchronised above
    return OF_OK;
}

}  // namespace

extern "C" {

int of_tvl1_batch_create(const of_tvl1_params* params, int streams, of_tvl1_batch** out)
{
    if (!out || streams < 1 || streams > 64) return OF_ERR_ARG;
    *out = nullptr;
    of_tvl1_batch* b = new (std::nothrow) of_tvl1_batch();
    if (!b) return OF_ERR_ALLOC;
    if (params) b->prm = *params;
    else of_tvl1_default_params(&b->prm);
    b->streams = streams;
    try {
        b->handles.resize(streams);
        b->wstreams.assign(streams, nullptr);
    } catch (...) {
        delete b;
        return OF_ERR_ALLOC;
    }
    bool ok = hipGetDevice(&b->device) == hipSuccess;
    for (int w = 0; w < streams && ok; w++)
        ok = hipStreamCreateWithFlags(&b->wstreams[w], hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        of_tvl1_batch_destroy(b);
        return OF_ERR_HIP;
    }
    *out = b;
    return OF_OK;
}

void of_tvl1_batch_destroy(of_tvl1_batch* b)
{
    if (!b) return;
    for (auto& m : b->handles)
        for (auto& kv : m) of_tvl1_destroy(kv.second);
    ls_release(b);
    for (hipStream_t s : b->wstreams)
        if (s) (void)hipStreamDestroy(s);
    delete b;
}

int of_tvl1_batch_set_shape(of_tvl1_batch* b, long long min_pixels, int iters_per_launch,
                            int rows_per_wave)
{
    if (!b) return OF_ERR_ARG;
    // the batch kernels run the 64-column tiles only
    if (rows_per_wave == 0 || !amdflow::iter_config_ok(iters_per_launch, rows_per_wave))
        return OF_ERR_UNSUPPORTED;
    try {
        b->rules.emplace_back(min_pixels, iters_per_launch, rows_per_wave);
    } catch (...) {
        return OF_ERR_ALLOC;
    }
    std::sort(b->rules.begin(), b->rules.end(),
              [](const auto& x, const auto& y) { return std::get<0>(x) > std::get<0>(y); });
    for (auto& m : b->handles)
        for (auto& kv : m) of_tvl1_destroy(kv.second);
    for (auto& m : b->handles) m.clear();
    return OF_OK;
}

static int batch_autotune(of_tvl1_batch* b, int npairs, const int* widths, const int* heights,
                          const uint8_t* const* dI0, const uint8_t* const* dI1, float* const* dU,
                          int reps, int* chosen, double* ms_table, double* best_ms)
{
    // candidate shapes (iterations per launch, rows per wave), the default first
    static const int kCand[kBatchTuneShapes][2] = {{4, 8}, {2, 8}, {8, 8}, {4, 4}, {2, 4}};
    using Rules = std::vector<std::tuple<long long, int, int>>;
    const Rules saved = b->rules;
    auto install = [&](const Rules& r) {
        b->rules = r;
        for (auto& m : b->handles)
            for (auto& kv : m) of_tvl1_destroy(kv.second);
        for (auto& m : b->handles) m.clear();
    };
    // one untimed run (first launches of a shape load its code), then `reps`
    // timed ones; the runs are synchronous, so the host clock brackets them
    auto time_rules = [&](const Rules& r, double* ms) -> int {
        try {
            install(r);
        } catch (...) {
            return OF_ERR_ALLOC;
        }
        OF_TRY(of_tvl1_batch_run_u8(b, npairs, widths, heights, dI0, dI1, dU, nullptr));
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < reps; k++)
            OF_TRY(of_tvl1_batch_run_u8(b, npairs, widths, heights, dI0, dI1, dU, nullptr));
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
              reps;
        return OF_OK;
    };
    double t[kBatchTuneShapes];
    int st = OF_OK;
    for (int c = 0; c < kBatchTuneShapes && !st; c++)
        st = time_rules(Rules{{0, kCand[c][0], kCand[c][1]}}, &t[c]);
    if (st) {
        install(saved);
        return st;
    }
    int order[kBatchTuneShapes];
    for (int c = 0; c < kBatchTuneShapes; c++) order[c] = c;
    std::sort(order, order + kBatchTuneShapes, [&](int x, int y) { return t[x] < t[y]; });
    // then the three fastest uniform shapes paired: one at scales of at least
    // kBatchTuneSplit pixels, another below
    int big = order[0], small = order[0];
    double best = t[order[0]];
    for (int i = 0; i < 3 && !st; i++)
        for (int j = 0; j < 3 && !st; j++) {
            if (i == j) continue;
            const int x = order[i], y = order[j];
            double ms = 0;
            st = time_rules(Rules{{kBatchTuneSplit, kCand[x][0], kCand[x][1]},
                                  {0, kCand[y][0], kCand[y][1]}},
                            &ms);
            if (!st && ms < best) {
                best = ms;
                big = x;
                small = y;
            }
        }
    if (st) {
        install(saved);
        return st;
    }
    install(Rules{{kBatchTuneSplit, kCand[big][0], kCand[big][1]}, {0, kCand[small][0], kCand[small][1]}});
    chosen[0] = kCand[big][0];
    chosen[1] = kCand[big][1];
    chosen[2] = kCand[small][0];
    chosen[3] = kCand[small][1];
    if (ms_table)
        for (int c = 0; c < kBatchTuneShapes; c++) ms_table[c] = t[c];
    if (best_ms) *best_ms = best;
    return OF_OK;
}

int of_tvl1_batch_autotune(of_tvl1_batch* b, int npairs, const int* widths, const int* heights,
                           const uint8_t* const* dI0, const uint8_t* const* dI1, float* const* dU,
                           int reps, int* chosen, double* ms_table, double* best_ms)
{
    if (!b || !chosen || reps < 1 || npairs < 1) return OF_ERR_ARG;
    // the C ABI never throws: allocation failures become OF_ERR_ALLOC
    try {
        return batch_autotune(b, npairs, widths, heights, dI0, dI1, dU, reps, chosen, ms_table,
                              best_ms);
    } catch (...) {
        return OF_ERR_ALLOC;
    }
}

int of_tvl1_batch_set_mode(of_tvl1_batch* b, int mode)
{
    if (!b || (mode != OF_BATCH_LOCKSTEP && mode != OF_BATCH_WORKERS)) return OF_ERR_ARG;
    b->mode = mode;
    return OF_OK;
}

int of_tvl1_run_batch(const of_tvl1_params* params, int npairs, const int* widths,
                      const int* heights, const uint8_t* const* dI0, const uint8_t* const* dI1,
                      float* const* dU, int* iterations_total)
{
    of_tvl1_batch* b = nullptr;
    int st = of_tvl1_batch_create(params, 3, &b);
    if (!st) st = of_tvl1_batch_run_u8(b, npairs, widths, heights, dI0, dI1, dU, iterations_total);
    of_tvl1_batch_destroy(b);
    return st;
}

int of_tvl1_batch_run_u8(of_tvl1_batch* b, int npairs, const int* widths, const int* heights,
                         const uint8_t* const* dI0, const uint8_t* const* dI1, float* const* dU,
                         int* iterations_total)
{
    if (!b || npairs < 0 || (npairs && (!widths || !heights || !dI0 || !dI1 || !dU)))
        return OF_ERR_ARG;
    for (int i = 0; i < npairs; i++)
        if (!dI0[i] || !dI1[i] || !dU[i] || widths[i] < 2 || heights[i] < 2) return OF_ERR_ARG;
    // the C ABI never throws: allocation failures become OF_ERR_ALLOC
    try {
        if (b->mode == OF_BATCH_LOCKSTEP)
            return run_lockstep(b, npairs, widths, heights, dI0, dI1, dU, iterations_total);
        return run_batch(b, npairs, widths, heights, dI0, dI1, dU, iterations_total);
    } catch (...) {
        return OF_ERR_ALLOC;
    }
}

}  // extern "C"
