// amdflow.hpp -- header-only C++ mirror of the reference's per-algorithm
// classes over the C ABI of libamdflow.so (amdflow.h).  Code written against
// the reference compiles against these with the same names and arguments:
//
//   TV_L1(int width, int height, float tau=0.25, float lambda=0.15,
//         float theta=0.3, int nscales=100, float zfactor=0.5, int warps=5,
//         float epsilon=0.01)                 src/tvl1/sequential/tvl1.hpp:12-17
//     void runDualTVL1Multiscale(uint8_t* I0, uint8_t* I1)      pair solve
//     void runDualTVL1Multiscale(const float* I)                streaming
//     float* getU()                                             planar [u1|u2]
//   void initFlow(int nLevels, int stride, int width, int height)
//   void ComputeFlow(const float* I0, const float* I1, int width, int height,
//                    int stride, float alpha, int nLevels, int nWarpIters,
//                    int nSolverIters, float* u, float* v)
//   void deleteFlow_mem(int nLevels)               src/hs/cuda/flow.cuh:19-33
//     (one process-wide state, as in the reference; initFlowHandle and the
//     of_hs* overloads of the other two are the reentrant form)
//   LucasKanade(int spac, int temp, int window, int nx, int ny)
//     void copy_frames_circularbuffer_GPU_wrapper(unsigned char*, int, int, int)
//     void lucas_kanade(float* Vx, float* Vy, int iframe)  src/lk/cuda/lucaskanade.cuh:6-35
//
// The reference reports failures by throwing or terminating; these wrappers
// throw amdflow::Error carrying the library status.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "amdflow.h"

namespace amdflow {

class Error : public std::runtime_error {
public:
    Error(int status, const char* what)
        : std::runtime_error(std::string(what) + ": " + of_status_string(status)), status(status) {}
    int status;
};

inline void check(int st, const char* what)
{
    if (st != OF_OK) throw Error(st, what);
}

class TV_L1 {
public:
    TV_L1(int width, int height, float tau = 0.25, float lambda = 0.15, float theta = 0.3,
          int nscales = 100, float zfactor = 0.5, int warps = 5, float epsilon = 0.01)
    {
        of_tvl1_params p;
        of_tvl1_default_params(&p);
        p.tau = tau;
        p.lambda = lambda;
        p.theta = theta;
        p.nscales = nscales;
        p.zfactor = zfactor;
        p.warps = warps;
        p.epsilon = epsilon;
        check(of_tvl1_create(&p, width, height, &h_), "TV_L1");
    }
    ~TV_L1() { of_tvl1_destroy(h_); }
    TV_L1(const TV_L1&) = delete;
    TV_L1& operator=(const TV_L1&) = delete;

    // tvl1.hpp:16: the handle owns the planar [u1|u2] buffer
    float* getU() { return const_cast<float*>(of_tvl1_flow(h_)); }
    // tvl1.hpp:17
    void runDualTVL1Multiscale(uint8_t* I0, uint8_t* I1)
    {
        check(of_tvl1_run_u8(h_, I0, I1), "runDualTVL1Multiscale");
    }
    // src/tvl1/cuda/tvl1.cuh:19 (I0 is the previous call's frame)
    void runDualTVL1Multiscale(const float* I)
    {
        check(of_tvl1_push_f32(h_, I), "runDualTVL1Multiscale");
    }
    of_tvl1* handle() { return h_; }

private:
    of_tvl1* h_ = nullptr;
};

// Horn-Schunck with the reference's free-function interface.  The reference
// keeps one process-wide state (flow.cu:26-44) that initFlow allocates,
// ComputeFlow uses and deleteFlow_mem frees; the three functions below with
// the reference's signatures do the same over one library handle, so they are
// not reentrant either.  initFlowHandle returns a handle of its own that the
// of_hs* overloads take: one per thread or per frame size.
inline of_hs*& hs_global_state()
{
    static of_hs* h = nullptr;
    return h;
}

inline of_hs* initFlowHandle(int nLevels, int stride, int width, int height)
{
    of_hs* h = nullptr;
    check(of_hs_create(nLevels, stride, width, height, &h), "initFlow");
    return h;
}

inline void ComputeFlow(of_hs* h, const float* I0, const float* I1, int width, int height,
                        int stride, float alpha, int nLevels, int nWarpIters, int nSolverIters,
                        float* u, float* v)
{
    check(of_hs_compute(h, I0, I1, width, height, stride, alpha, nLevels, nWarpIters,
                        nSolverIters, u, v),
          "ComputeFlow");
}

inline void deleteFlow_mem(of_hs* h) { of_hs_destroy(h); }

// flow.cuh:19: a second initFlow replaces the state of the first
inline void initFlow(int nLevels, int stride, int width, int height)
{
    of_hs* h = initFlowHandle(nLevels, stride, width, height);
    of_hs_destroy(hs_global_state());
    hs_global_state() = h;
}

// flow.cuh:21-31; OF_ERR_STATE before initFlow
inline void ComputeFlow(const float* I0, const float* I1, int width, int height, int stride,
                        float alpha, int nLevels, int nWarpIters, int nSolverIters, float* u,
                        float* v)
{
    if (!hs_global_state()) throw Error(OF_ERR_STATE, "ComputeFlow before initFlow");
    ComputeFlow(hs_global_state(), I0, I1, width, height, stride, alpha, nLevels, nWarpIters,
                nSolverIters, u, v);
}

// flow.cuh:33
inline void deleteFlow_mem(int /*nLevels*/)
{
    of_hs_destroy(hs_global_state());
    hs_global_state() = nullptr;
}

class LucasKanade {
public:
    LucasKanade(int spac_filt_size, int temp_filt_size, int window_size, int nx, int ny)
    {
        check(of_lk_create(spac_filt_size, temp_filt_size, window_size, nx, ny, &h_),
              "LucasKanade");
    }
    ~LucasKanade() { of_lk_destroy(h_); }
    LucasKanade(const LucasKanade&) = delete;
    LucasKanade& operator=(const LucasKanade&) = delete;

    void lucas_kanade(float* Vx, float* Vy, int iframe)
    {
        check(of_lk_compute(h_, Vx, Vy, iframe), "lucas_kanade");
    }
    void copy_frames_circularbuffer_GPU_wrapper(unsigned char* frame, int temp_size, int fr,
                                                int frame_size)
    {
        check(of_lk_push_u8(h_, frame, temp_size, fr, frame_size),
              "copy_frames_circularbuffer_GPU_wrapper");
    }

private:
    of_lk* h_ = nullptr;
};

}  // namespace amdflow
