// common.cpp -- status strings, HIP error recording, Gaussian taps.
#include "common.hpp"

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "amdflow_debug.h"

namespace amdflow {

static std::atomic<long long> g_fail_after{-1};   // < 0: no injected failure
static std::atomic<long long> g_attempts{0};      // since the last of_dbg_fail_alloc_after
static std::atomic<long long> g_live{0};

bool alloc_attempt_fails()
{
    const long long k = g_attempts.fetch_add(1, std::memory_order_relaxed);
    const long long n = g_fail_after.load(std::memory_order_relaxed);
    return n >= 0 && k >= n;
}

void alloc_live_add(long long d) { g_live.fetch_add(d, std::memory_order_relaxed); }

static thread_local char g_last_error[512];

void record_hip_error(hipError_t e, const char* what, const char* file, int line)
{
    std::snprintf(g_last_error, sizeof(g_last_error), "%s failed: %s (%s:%d)", what,
                  hipGetErrorString(e), file, line);
    (void)hipGetLastError();
}

// mask.cpp:212-231: size = (int)(5*sigma)+1 taps exp(-i^2/(2 sigma^2)) /
// (sigma sqrt(2 pi)) in double, normalised by twice the |tap| sum minus the
// centre tap (cblas_dasum, summed left to right), scaled by the reciprocal.
int gaussian_taps(double sigma, GaussTaps* g)
{
    const double den = 2 * sigma * sigma;
    const double sPi = sigma * std::sqrt(M_PI * 2);
    const int size = (int)(5.0 * sigma) + 1;
    if (size < 1 || size > 64) return OF_ERR_SIGMA;
    for (int i = 0; i < size; i++) g->B[i] = 1 / sPi * std::exp(-i * i / den);
    double norm = 0.0;
    for (int i = 0; i < size; i++) norm += std::fabs(g->B[i]);
    norm = norm * 2 - g->B[0];
    const double alpha = 1 / norm;
    for (int i = 0; i < size; i++) g->B[i] = alpha * g->B[i];
    for (int i = size; i < 64; i++) g->B[i] = 0.0;
    g->size = size;
    return OF_OK;
}

}  // namespace amdflow

extern "C" const char* of_status_string(int status)
{
    switch (status) {
    case OF_OK: return "ok";
    case OF_ERR_ARG: return "invalid argument";
    case OF_ERR_SIGMA: return "Gaussian window larger than an image level (sigma too large)";
    case OF_ERR_ALLOC: return "allocation failed";
    case OF_ERR_HIP: return amdflow::g_last_error[0] ? amdflow::g_last_error : "HIP runtime error";
    case OF_ERR_STATE: return "call not valid in this state";
    case OF_ERR_UNSUPPORTED: return "unsupported option";
    case OF_ERR_CAPACITY: return "input larger than the fixed scratch of a no-allocation entry point";
    default: return "unknown status";
    }
}

extern "C" const char* of_build_info(void)
{
    return "amdflow gfx950 hip; -O3 -ffp-contract=off, IEEE fp32 div/sqrt, no packed fp32";
}

extern "C" {

int of_dbg_fail_alloc_after(long long n)
{
    amdflow::g_attempts.store(0);
    amdflow::g_fail_after.store(n < 0 ? -1 : n);
    return OF_OK;
}

long long of_dbg_alloc_attempts(void) { return amdflow::g_attempts.load(); }

long long of_dbg_live_allocs(void) { return amdflow::g_live.load(); }

}  // extern "C"
