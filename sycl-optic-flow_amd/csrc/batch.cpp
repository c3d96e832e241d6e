// batch.cpp -- many TV-L1 frame pairs of mixed sizes solved concurrently.
//
// No reference counterpart (the reference solves one pair per call).  A
// batch object owns `streams` worker threads; each worker has its own TV-L1
// handle per frame size (hence its own HIP stream) and takes pairs from a
// shared work list, longest first.  The workers' launches interleave on the
// device, so the launch-latency-bound coarse scales of one pair overlap the
// bandwidth-bound fine scales of others.  Handles in a batch use host-paced
// launches only (no cooperative launches run concurrently).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <utility>
#include <vector>

#include "amdflow.h"
#include "common.hpp"

struct of_tvl1_batch {
    of_tvl1_params prm{};
    int streams = 1;
    int device = 0;
    // per worker: its stream and its handles (frame size -> handle)
    std::vector<hipStream_t> wstreams;
    std::vector<std::map<std::pair<int, int>, of_tvl1*>> handles;
    // launch rules applied to every handle: (min_pixels, K, R)
    std::vector<std::tuple<long long, int, int>> rules;
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
        return run_batch(b, npairs, widths, heights, dI0, dI1, dU, iterations_total);
    } catch (...) {
        return OF_ERR_ALLOC;
    }
}

}  // extern "C"
