// debug.cpp -- per-kernel test entry points (include/amdflow_debug.h).
// Each call allocates device buffers, uploads the host inputs, runs one
// stage of the solver on the default stream and downloads the result.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "amdflow_debug.h"
#include "common.hpp"
#include "tvl1_kernels.hpp"
#include "tvl1_launch.hpp"

using namespace amdflow;

namespace {

// Owns a set of device buffers for the duration of one call.
struct DevBufs {
    std::vector<void*> ptrs;
    ~DevBufs()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T* alloc(size_t n)
    {
        void* p = nullptr;
        if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        ptrs.push_back(p);
        return (T*)p;
    }
    template <typename T>
    T* upload(const T* h, size_t n)
    {
        T* d = alloc<T>(n);
        if (d && hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return d;
    }
};

#define OF_NN(p) \
    do {         \
        if (!(p)) return OF_ERR_ALLOC; \
    } while (0)

}  // namespace

extern "C" {

int of_dbg_gaussian(const float* in, float* out, int w, int h, double sigma)
{
    if (!in || !out || w < 1 || h < 1) return OF_ERR_ARG;
    GaussTaps g;
    OF_TRY(gaussian_taps(sigma, &g));
    if (g.size > w || g.size > h) return OF_ERR_SIGMA;
    const size_t n = (size_t)w * h;
    DevBufs B;
    float* d_in = B.upload(in, n);
    float* d_tmp = B.alloc<float>(n);
    float* d_out = B.alloc<float>(n);
    OF_NN(d_in); OF_NN(d_tmp); OF_NN(d_out);
    // the tiled kernel of the solver (the same frame as both of its images),
    // else the two-pass form
    const int st = launch_smooth_pair(d_in, d_in, d_out, d_tmp, w, h, g, nullptr, nullptr);
    if (st == OF_ERR_UNSUPPORTED) OF_TRY(launch_gaussian(d_in, d_tmp, d_out, w, h, g, nullptr));
    else OF_TRY(st);
    OF_HIP(hipMemcpy(out, d_out, n * sizeof(float), hipMemcpyDeviceToHost));
    return OF_OK;
}

int of_dbg_zoom_out(const float* in, float* out, int w, int h, float factor)
{
    if (!in || !out || w < 1 || h < 1 || !(factor > 0 && factor < 1)) return OF_ERR_ARG;
    GaussTaps g;
    const float zsig = (float)((double)0.6f * std::sqrt(1.0 / (double)(factor * factor) - 1.0));
    OF_TRY(gaussian_taps((double)zsig, &g));
    if (g.size > w || g.size > h) return OF_ERR_SIGMA;
    int nxx, nyy;
    zoom_size(w, h, factor, &nxx, &nyy);
    const size_t n = (size_t)w * h;
    DevBufs B;
    float* d_in = B.upload(in, n);
    float* d_t0 = B.alloc<float>(n);
    float* d_t1 = B.alloc<float>(n);
    float* d_out = B.alloc<float>((size_t)nxx * nyy);
    OF_NN(d_in); OF_NN(d_t0); OF_NN(d_t1); OF_NN(d_out);
    // as the solver's pyramid: the tiled kernel at factor 1/2
    int st = OF_ERR_UNSUPPORTED;
    if (factor == 0.5f)
        st = launch_zoom_out_half_pair(d_in, d_in, d_out, d_t0, w, h, nxx, nyy, g, nullptr);
    if (st == OF_ERR_UNSUPPORTED) {
        OF_TRY(launch_gaussian(d_in, d_t0, d_t1, w, h, g, nullptr));
        OF_TRY(launch_zoom_out_sample(d_t1, d_out, w, h, nxx, nyy, factor, nullptr));
    } else {
        OF_TRY(st);
    }
    OF_HIP(hipMemcpy(out, d_out, (size_t)nxx * nyy * sizeof(float), hipMemcpyDeviceToHost));
    return OF_OK;
}

int of_dbg_zoom_in(const float* in, float* out, int w, int h, int nxx, int nyy, float scale)
{
    if (!in || !out || w < 1 || h < 1 || nxx < 1 || nyy < 1) return OF_ERR_ARG;
    const size_t n = (size_t)w * h, nf = (size_t)nxx * nyy;
    DevBufs B;
    // the kernel resamples two planes: pass the same plane twice
    float* d_in = B.alloc<float>(2 * n);
    float* d_out = B.alloc<float>(2 * nf);
    TvCtl* ctl = B.alloc<TvCtl>(1);
    OF_NN(d_in); OF_NN(d_out); OF_NN(ctl);
    OF_HIP(hipMemcpy(d_in, in, n * sizeof(float), hipMemcpyHostToDevice));
    OF_HIP(hipMemcpy(d_in + n, in, n * sizeof(float), hipMemcpyHostToDevice));
    OF_TRY(launch_ctl_init(ctl, nullptr));
    OF_TRY(launch_zoom_in_flow(d_in, d_in, d_out, w, h, nxx, nyy, scale, ctl, 0, nullptr));
    OF_HIP(hipMemcpy(out, d_out, nf * sizeof(float), hipMemcpyDeviceToHost));
    return OF_OK;
}

int of_dbg_centered_gradient(const float* in, float* dx, float* dy, int w, int h)
{
    if (!in || !dx || !dy || w < 2 || h < 2) return OF_ERR_ARG;
    const size_t n = (size_t)w * h;
    DevBufs B;
    float* d_in = B.upload(in, n);
    float* d_x = B.alloc<float>(n);
    float* d_y = B.alloc<float>(n);
    OF_NN(d_in); OF_NN(d_x); OF_NN(d_y);
    OF_TRY(launch_centered_gradient(d_in, d_x, d_y, w, h, nullptr));
    OF_HIP(hipMemcpy(dx, d_x, n * sizeof(float), hipMemcpyDeviceToHost));
    OF_HIP(hipMemcpy(dy, d_y, n * sizeof(float), hipMemcpyDeviceToHost));
    return OF_OK;
}

int of_dbg_normalize(float* I0, float* I1, int w, int h)
{
    if (!I0 || !I1 || w < 1 || h < 1) return OF_ERR_ARG;
    const size_t n = (size_t)w * h;
    DevBufs B;
    float* a = B.upload(I0, n);
    float* b = B.upload(I1, n);
    float* fa = B.alloc<float>(n);
    float* fb = B.alloc<float>(n);
    TvCtl* ctl = B.alloc<TvCtl>(1);
    OF_NN(a); OF_NN(b); OF_NN(fa); OF_NN(fb); OF_NN(ctl);
    OF_TRY(launch_ctl_init(ctl, nullptr));
    OF_TRY(launch_load_minmax_f32(a, b, fa, fb, n, ctl, nullptr));
    OF_TRY(launch_normalize(fa, fb, a, b, n, ctl, nullptr));
    OF_HIP(hipMemcpy(I0, a, n * sizeof(float), hipMemcpyDeviceToHost));
    OF_HIP(hipMemcpy(I1, b, n * sizeof(float), hipMemcpyDeviceToHost));
    return OF_OK;
}

int of_dbg_warp_setup(const float* I0, const float* I1, const float* I1x, const float* I1y,
                      const float* u1, const float* u2, float* I1wx, float* I1wy, float* rho_c,
                      int w, int h)
{
    if (!I0 || !I1 || !I1x || !I1y || !u1 || !u2 || !I1wx || !I1wy || !rho_c || w < 1 || h < 1)
        return OF_ERR_ARG;
    const size_t n = (size_t)w * h;
    DevBufs B;
    float* d0 = B.upload(I0, n);
    float* d1 = B.upload(I1, n);
    float* dx = B.upload(I1x, n);
    float* dy = B.upload(I1y, n);
    float* ub = B.alloc<float>(2 * n);
    float* wx = B.alloc<float>(n);
    float* wy = B.alloc<float>(n);
    float* rc = B.alloc<float>(n);
    TvCtl* ctl = B.alloc<TvCtl>(1);
    float* src = B.alloc<float>(4 * n);
    OF_NN(d0); OF_NN(d1); OF_NN(dx); OF_NN(dy); OF_NN(ub); OF_NN(wx); OF_NN(wy); OF_NN(rc); OF_NN(ctl);
    OF_NN(src);
    OF_HIP(hipMemcpy(ub, u1, n * sizeof(float), hipMemcpyHostToDevice));
    OF_HIP(hipMemcpy(ub + n, u2, n * sizeof(float), hipMemcpyHostToDevice));
    OF_TRY(launch_ctl_init(ctl, nullptr));
    OF_TRY(launch_pack3(d1, dx, dy, src, n, nullptr));
    OF_TRY(launch_warp_setup(d0, src, ub, ub, wx, wy, rc, w, h, ctl, 0, nullptr));
    OF_HIP(hipMemcpy(I1wx, wx, n * sizeof(float), hipMemcpyDeviceToHost));
    OF_HIP(hipMemcpy(I1wy, wy, n * sizeof(float), hipMemcpyDeviceToHost));
    OF_HIP(hipMemcpy(rho_c, rc, n * sizeof(float), hipMemcpyDeviceToHost));
    return OF_OK;
}

int of_dbg_tvl1_iterations(const float* I1wx, const float* I1wy, const float* rho_c, float* u1,
                           float* u2, float* p11, float* p12, float* p21, float* p22, int w, int h,
                           float tau, float lambda, float theta, int iters, int K, int R,
                           double* err_sums)
{
    if (!I1wx || !I1wy || !rho_c || !u1 || !u2 || !p11 || !p12 || !p21 || !p22) return OF_ERR_ARG;
    if (w < 2 || h < 2 || iters < 1 || !(theta > 0)) return OF_ERR_ARG;
    if (!iter_config_ok(K, R)) return OF_ERR_UNSUPPORTED;
    const size_t n = (size_t)w * h;
    unsigned gx, gy;
    iter_grid(K, R, w, h, &gx, &gy);
    DevBufs B;
    float* wx = B.upload(I1wx, n);
    float* wy = B.upload(I1wy, n);
    float* rc = B.upload(rho_c, n);
    float* ub = B.alloc<float>(4 * n);
    float* pb = B.alloc<float>(8 * n);
    TvCtl* ctl = B.alloc<TvCtl>(1);
    double* part = B.alloc<double>((size_t)gx * gy * K);
    double* elog = B.alloc<double>((size_t)iters + K);
    int* slot = B.alloc<int>(1);
    OF_NN(wx); OF_NN(wy); OF_NN(rc); OF_NN(ub); OF_NN(pb); OF_NN(ctl); OF_NN(part); OF_NN(elog); OF_NN(slot);
    const float* st_in[6] = {u1, u2, p11, p12, p21, p22};
    float* st_dev[6] = {ub, ub + n, pb, pb + n, pb + 2 * n, pb + 3 * n};
    for (int i = 0; i < 6; i++)
        OF_HIP(hipMemcpy(st_dev[i], st_in[i], n * sizeof(float), hipMemcpyHostToDevice));
    OF_TRY(launch_ctl_init(ctl, nullptr));
    OF_TRY(launch_ctl_begin(ctl, nullptr));
    IterArgs a{};
    a.wx = wx;
    a.wy = wy;
    a.rhoc = rc;
    a.u[0] = ub;
    a.u[1] = ub + 2 * n;
    a.p[0] = pb;
    a.p[1] = pb + 4 * n;
    a.nx = w;
    a.ny = h;
    a.plane = (int)n;
    a.l_t = lambda * theta;
    a.theta = theta;
    a.taut = tau / theta;
    a.eps2 = -1.0;   // never below: run exactly `iters` iterations
    a.max_iter = iters;
    a.level = 0;
    a.ctl = ctl;
    a.iters_slot = slot;
    a.partials = part;
    a.host_status = nullptr;
    a.epoch = 1;
    a.err_log = elog;
    const int launches = (iters + K - 1) / K;
    for (int l = 0; l < launches; l++) OF_TRY(launch_iter(K, R, a, nullptr));
    OF_HIP(hipDeviceSynchronize());
    TvCtl hc;
    OF_HIP(hipMemcpy(&hc, ctl, sizeof(TvCtl), hipMemcpyDeviceToHost));
    const int cur = hc.cur[0];
    float* st_out[6] = {u1, u2, p11, p12, p21, p22};
    float* src[6] = {ub + 2 * n * cur, ub + 2 * n * cur + n, pb + 4 * n * cur, pb + 4 * n * cur + n,
                     pb + 4 * n * cur + 2 * n, pb + 4 * n * cur + 3 * n};
    for (int i = 0; i < 6; i++)
        OF_HIP(hipMemcpy(st_out[i], src[i], n * sizeof(float), hipMemcpyDeviceToHost));
    if (err_sums) OF_HIP(hipMemcpy(err_sums, elog, sizeof(double) * iters, hipMemcpyDeviceToHost));
    return OF_OK;
}

int of_dbg_persist_force_not_resident(int on)
{
    set_persist_force_not_resident(on != 0);
    return OF_OK;
}

int of_dbg_hypot(const float* a, const float* b, float* out, size_t n)
{
    if (!a || !b || !out) return OF_ERR_ARG;
    DevBufs B;
    float* da = B.upload(a, n);
    float* db = B.upload(b, n);
    float* d_o = B.alloc<float>(n);
    OF_NN(da); OF_NN(db); OF_NN(d_o);
    OF_TRY(launch_hypot(da, db, d_o, n, nullptr));
    OF_HIP(hipMemcpy(out, d_o, n * sizeof(float), hipMemcpyDeviceToHost));
    return OF_OK;
}

}  // extern "C"
