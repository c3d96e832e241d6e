// batch.cpp -- many TV-L1 frame pairs of mixed sizes solved together.
//
// No reference counterpart (the reference solves one pair per call).  Two
// modes, both with results identical to solving each pair alone:
//
//   lockstep (default): every pair gets a TV-L1 handle of its own (pooled
//     by frame size).  The pairs are prepared (normalisation, pyramids) on
//     the batch's streams, then solved scale by scale in lockstep on one
//     stream: each launch covers the same scale and warping of all pairs.
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
#include <utility>
#include <vector>

#include "amdflow.h"
#include "common.hpp"
#include "lockstep.hpp"
#include "tvl1_handle.hpp"
#include "tvl1_launch.hpp"

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
    std::vector<hipEvent_t> wevents;          // one per stream: joins the preparations
    amdflow::LsCtl* mc = nullptr;             // device control of the group
    unsigned long long* hstatus = nullptr;    // host-mapped progress word
    unsigned epoch = 0;
    amdflow::LsPair* h_pairs = nullptr;       // pinned tables, uploaded once per group
    amdflow::LsPair* d_pairs = nullptr;
    size_t pairs_cap = 0;
    unsigned* h_bases = nullptr;
    unsigned* d_bases = nullptr;
    size_t bases_cap = 0;
    int* h_iters = nullptr;
    int* d_iters = nullptr;
    size_t iters_cap = 0;
    std::map<std::pair<int, int>, unsigned> resident;   // (K, R) -> co-resident workgroups
};

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
    for (hipEvent_t e : b->wevents)
        if (e) (void)hipEventDestroy(e);
    b->wevents.clear();
    amdflow::dev_free(b->mc);
    amdflow::dev_free(b->d_pairs);
    amdflow::dev_free(b->d_bases);
    amdflow::dev_free(b->d_iters);
    if (b->hstatus) (void)hipHostFree(b->hstatus);
    if (b->h_pairs) (void)hipHostFree(b->h_pairs);
    if (b->h_bases) (void)hipHostFree(b->h_bases);
    if (b->h_iters) (void)hipHostFree(b->h_iters);
    b->mc = nullptr;
    b->d_pairs = b->h_pairs = nullptr;
    b->d_bases = b->h_bases = nullptr;
    b->d_iters = b->h_iters = nullptr;
    b->hstatus = nullptr;
    b->pairs_cap = b->bases_cap = b->iters_cap = 0;
}

// Pinned host table and its device twin of at least n elements.  Only
// called between groups, when no copy from the old buffers is pending.
template <typename T>
int ls_reserve(T** host, T** dev, size_t* cap, size_t n)
{
    if (n <= *cap) return OF_OK;
    amdflow::dev_free(*dev);
    if (*host) (void)hipHostFree(*host);
    *dev = nullptr;
    *host = nullptr;
    *cap = 0;
    const size_t want = std::max(n, (size_t)64);
    if (hipHostMalloc((void**)host, sizeof(T) * want, hipHostMallocDefault) != hipSuccess) {
        *host = nullptr;
        (void)hipGetLastError();
        return OF_ERR_ALLOC;
    }
    OF_TRY(amdflow::dev_alloc(dev, want));
    *cap = want;
    return OF_OK;
}

int ls_setup(of_tvl1_batch* b)
{
    if (b->mc) return OF_OK;
    OF_TRY(amdflow::dev_alloc(&b->mc, 1));
    OF_HIP(hipMemset(b->mc, 0, sizeof(amdflow::LsCtl)));
    if (hipHostMalloc((void**)&b->hstatus, sizeof(unsigned long long),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        b->hstatus = nullptr;
        (void)hipGetLastError();
        return OF_ERR_ALLOC;
    }
    *b->hstatus = 0ull;
    b->wevents.assign(b->streams, nullptr);
    for (hipEvent_t& e : b->wevents) OF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return OF_OK;
}

// Launch shape of scale s: the batch's rule for the largest participating
// scale size, else (2 iterations, 8 rows) from 200k pixels and (4, 4) below.
void ls_shape(const of_tvl1_batch* b, long long npx, int* K, int* R)
{
    *K = npx >= 200000 ? 2 : 4;
    *R = npx >= 200000 ? 8 : 4;
    for (const auto& r : b->rules)
        if (npx >= std::get<0>(r)) {
            *K = std::get<1>(r);
            *R = std::get<2>(r);
            return;
        }
}

// Iteration launches of one warping of all pairs of `ph`, kept kInFlight
// ahead of the device, until the device reports that every pair stopped.
int ls_run_warping(of_tvl1_batch* b, const amdflow::LsPhase& ph, int K, int R, hipStream_t st)
{
    using namespace amdflow;
    auto key = std::make_pair(K, R);
    auto it = b->resident.find(key);
    if (it == b->resident.end()) {
        unsigned blocks = 0;
        OF_TRY(ls_iter_resident(K, R, &blocks));
        it = b->resident.emplace(key, blocks).first;
    }
    const unsigned grid = std::max(1u, std::min(ph.ntiles, it->second));
    const unsigned epoch = ++b->epoch;
    const long long max_launches = (b->prm.max_iterations + K - 1) / K + 1;
    long long issued = 0, completed = 0;
    bool stopped = false;
    unsigned spins = 0;
    const auto t0 = std::chrono::steady_clock::now();
    while (!stopped && issued < max_launches) {
        while (issued < max_launches && issued - completed < kInFlight) {
            OF_TRY(launch_ls_iter(K, R, ph, b->mc, grid, b->hstatus, epoch, st));
            issued++;
        }
        if (issued >= max_launches) break;
        const unsigned long long word = __atomic_load_n(b->hstatus, __ATOMIC_ACQUIRE);
        if ((unsigned)(word >> 32) == epoch) {
            completed = (long long)(word & 0x7FFFFFFFull);
            stopped = ((word >> 31) & 1ull) != 0;
            if (stopped || issued - completed < kInFlight) continue;
        }
        if (++spins % 4096 == 0) {
            const hipError_t q = hipStreamQuery(st);
            if (q != hipSuccess && q != hipErrorNotReady) {
                record_hip_error(q, "hipStreamQuery", __FILE__, __LINE__);
                return OF_ERR_HIP;
            }
            if (q == hipSuccess) {
                // idle stream: every issued launch has reported
                const unsigned long long w2 = __atomic_load_n(b->hstatus, __ATOMIC_ACQUIRE);
                if ((unsigned)(w2 >> 32) == epoch) {
                    completed = (long long)(w2 & 0x7FFFFFFFull);
                    stopped = ((w2 >> 31) & 1ull) != 0;
                }
                if (!stopped && completed < issued) return OF_ERR_STATE;
            }
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > 60.0)
                return OF_ERR_HIP;
        } else {
            __builtin_ia32_pause();
        }
    }
    return OF_OK;
}

// One lockstep group of n <= kLsMaxPairs pairs.
int ls_group(of_tvl1_batch* b, int n, const int* widths, const int* heights,
             const uint8_t* const* dI0, const uint8_t* const* dI1, float* const* dU,
             int* iterations_total)
{
    using namespace amdflow;
    OF_TRY(ls_setup(b));
    const int S = b->streams;
    hipStream_t main = b->wstreams[0];
    const of_tvl1_params& P = b->prm;

    // a handle per pair, pooled by frame size
    std::vector<of_tvl1*> hs(n, nullptr);
    {
        std::map<std::pair<int, int>, size_t> used;
        for (int i = 0; i < n; i++) {
            const auto key = std::make_pair(widths[i], heights[i]);
            auto& v = b->pool[key];
            const size_t k = used[key]++;
            while (v.size() <= k) {
                of_tvl1* h = nullptr;
                OF_TRY(of_tvl1_create(&P, widths[i], heights[i], &h));
                v.push_back(h);
            }
            hs[i] = v[k];
        }
    }

    // phase 1 of every pair on the batch's streams, joined into `main`
    for (int i = 0; i < n; i++) {
        of_tvl1* h = hs[i];
        hipStream_t st = b->wstreams[i % S];
        OF_TRY(tvl1_prepare(h, dI0[i], dI1[i], true, st));
        const int c = h->ns - 1;   // zero flow at the coarsest scale, tvl1.cpp:199-200
        OF_TRY(launch_fill(h->ubase + 4 * h->off[c], 2 * (size_t)h->nx[c] * h->ny[c], 0.0f, st));
    }
    for (int k = 1; k < S && k < n; k++) {
        OF_HIP(hipEventRecord(b->wevents[k], b->wstreams[k]));
        OF_HIP(hipStreamWaitEvent(main, b->wevents[k], 0));
    }

    // tables: per scale, the pairs that have it
    int maxns = 0;
    for (of_tvl1* h : hs) maxns = std::max(maxns, h->ns);
    std::vector<std::vector<int>> part(maxns);
    size_t nent = 0, nbase = 0;
    for (int s = 0; s < maxns; s++) {
        for (int i = 0; i < n; i++)
            if (hs[i]->ns > s) part[s].push_back(i);
        nent += part[s].size();
        nbase += 3 * (part[s].size() + 1);
    }
    const size_t niters = (size_t)n * kMaxScales * P.warps;
    OF_TRY(ls_reserve(&b->h_pairs, &b->d_pairs, &b->pairs_cap, nent));
    OF_TRY(ls_reserve(&b->h_bases, &b->d_bases, &b->bases_cap, nbase));
    OF_TRY(ls_reserve(&b->h_iters, &b->d_iters, &b->iters_cap, niters));

    std::vector<LsPhase> phases(maxns);
    std::vector<int> shapeK(maxns), shapeR(maxns);
    size_t e = 0, q = 0;
    for (int s = maxns - 1; s >= 0; s--) {
        const std::vector<int>& ps = part[s];
        const int np = (int)ps.size();
        long long maxpx = 0;
        for (int i : ps) maxpx = std::max(maxpx, (long long)hs[i]->nx[s] * hs[i]->ny[s]);
        int K, R;
        ls_shape(b, maxpx, &K, &R);
        shapeK[s] = K;
        shapeR[s] = R;
        LsPhase& ph = phases[s];
        ph.pairs = b->d_pairs + e;
        ph.blk_base = b->d_bases + q;
        ph.fblk_base = b->d_bases + q + (np + 1);
        ph.tile_base = b->d_bases + q + 2 * (np + 1);
        ph.iters = b->d_iters;
        ph.npairs = np;
        ph.level = s;
        ph.warps = P.warps;
        ph.warp = 0;
        unsigned* blk = b->h_bases + q;
        unsigned* fblk = blk + (np + 1);
        unsigned* tiles = fblk + (np + 1);
        blk[0] = fblk[0] = tiles[0] = 0;
        for (int k = 0; k < np; k++) {
            const int i = ps[k];
            of_tvl1* h = hs[i];
            const int nx = h->nx[s], ny = h->ny[s];
            const size_t N = (size_t)nx * ny;
            LsPair& E = b->h_pairs[e + k];
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
            E.dst = dU[i];
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
    OF_HIP(hipMemcpyAsync(b->d_pairs, b->h_pairs, sizeof(LsPair) * nent, hipMemcpyHostToDevice, main));
    OF_HIP(hipMemcpyAsync(b->d_bases, b->h_bases, sizeof(unsigned) * nbase, hipMemcpyHostToDevice, main));
    OF_HIP(hipMemsetAsync(b->d_iters, 0, sizeof(int) * niters, main));

    // phase 2: coarse to fine, all pairs per launch (tvl1.cpp:201-216)
    for (int s = maxns - 1; s >= 0; s--) {
        LsPhase ph = phases[s];
        if (ph.npairs == 0) continue;
        OF_TRY(launch_ls_level_begin(ph, main));
        for (int w = 0; w < P.warps; w++) {
            ph.warp = w;
            OF_TRY(launch_ls_warp_setup(ph, b->mc, main));
            OF_TRY(ls_run_warping(b, ph, shapeK[s], shapeR[s], main));
        }
        if (s > 0) OF_TRY(launch_ls_zoom_in(ph, main));
    }
    OF_TRY(launch_ls_copy_flow(phases[0], main));
    OF_HIP(hipMemcpyAsync(b->h_iters, b->d_iters, sizeof(int) * niters, hipMemcpyDeviceToHost, main));
    OF_HIP(hipStreamSynchronize(main));
    if (iterations_total)
        for (int i = 0; i < n; i++) {
            long long sum = 0;
            for (size_t k = 0; k < (size_t)kMaxScales * P.warps; k++)
                sum += b->h_iters[(size_t)i * kMaxScales * P.warps + k];
            iterations_total[i] = (int)sum;
        }
    return OF_OK;
}

int run_lockstep(of_tvl1_batch* b, int npairs, const int* widths, const int* heights,
                 const uint8_t* const* dI0, const uint8_t* const* dI1, float* const* dU,
                 int* iterations_total)
{
    if (hipSetDevice(b->device) != hipSuccess) return OF_ERR_HIP;
    for (int g = 0; g < npairs; g += amdflow::kLsMaxPairs) {
        const int n = std::min(amdflow::kLsMaxPairs, npairs - g);
        OF_TRY(ls_group(b, n, widths + g, heights + g, dI0 + g, dI1 + g, dU + g,
                        iterations_total ? iterations_total + g : nullptr));
    }
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

int of_tvl1_batch_set_mode(of_tvl1_batch* b, int mode)
{
    if (!b || (mode != OF_BATCH_LOCKSTEP && mode != OF_BATCH_WORKERS)) return OF_ERR_ARG;
    b->mode = mode;
    return OF_OK;
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
