// common.hpp -- host-side helpers shared by the TV-L1, HS and LK modules.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "amdflow.h"

namespace amdflow {

// Evaluates a HIP call; on failure records it and returns OF_ERR_HIP from
// the enclosing function.
#define OF_HIP(call)                                                     \
    do {                                                                 \
        hipError_t of_e_ = (call);                                       \
        if (of_e_ != hipSuccess) {                                       \
            ::amdflow::record_hip_error(of_e_, #call, __FILE__, __LINE__); \
            return OF_ERR_HIP;                                           \
        }                                                                \
    } while (0)

#define OF_TRY(expr)                     \
    do {                                 \
        int of_s_ = (expr);              \
        if (of_s_ != OF_OK) return of_s_; \
    } while (0)

void record_hip_error(hipError_t e, const char* what, const char* file, int line);

// Runs an entry point's body; the C ABI never lets an exception out (the
// library throws none itself, but its std containers can report a failed
// allocation that way), so any exception becomes OF_ERR_ALLOC.
template <typename F>
int guarded(F&& body) noexcept
{
    try {
        return body();
    } catch (...) {
        return OF_ERR_ALLOC;
    }
}

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Allocation accounting behind dev_alloc/host_alloc: every attempt is
// counted, and while a failure point is set (of_dbg_fail_alloc_after) the
// attempts from that one on fail as a real allocation failure would; live
// blocks are counted so that tests can see a leak on any error path.
bool alloc_attempt_fails();
void alloc_live_add(long long d);

// Device allocation that reports failure as a status code.
template <typename T>
int dev_alloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    if (alloc_attempt_fails()) return OF_ERR_ALLOC;
    if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return OF_ERR_ALLOC;
    }
    alloc_live_add(1);
    return OF_OK;
}

template <typename T>
void dev_free(T*& p)
{
    if (p) {
        (void)hipFree((void*)p);
        alloc_live_add(-1);
    }
    p = nullptr;
}

// Pinned host allocation (hipHostMalloc flags), accounted as dev_alloc.
template <typename T>
int host_alloc(T** p, size_t count, unsigned flags)
{
    *p = nullptr;
    if (count == 0) count = 1;
    if (alloc_attempt_fails()) return OF_ERR_ALLOC;
    if (hipHostMalloc((void**)p, count * sizeof(T), flags) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return OF_ERR_ALLOC;
    }
    alloc_live_add(1);
    return OF_OK;
}

template <typename T>
void host_free(T*& p)
{
    if (p) {
        (void)hipHostFree((void*)p);
        alloc_live_add(-1);
    }
    p = nullptr;
}

// Gaussian taps of mask.cpp:212-231 computed on the host in double with the
// C library's exp, so CPU and GPU blur with identical coefficients.
struct GaussTaps {
    double B[64];
    int size;
};
int gaussian_taps(double sigma, GaussTaps* g);

// Pyramid size rule of zoom.hpp:14-26.
inline void zoom_size(int nx, int ny, float factor, int* nxx, int* nyy)
{
    *nxx = (int)((double)((float)nx * factor) + 0.5);
    *nyy = (int)((double)((float)ny * factor) + 0.5);
}

}  // namespace amdflow
