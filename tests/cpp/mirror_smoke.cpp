// mirror_smoke.cpp -- calls the C++ mirror (include/amdflow.hpp) the way the
// reference's apps call their classes (src/tvl1/sequential/main.cpp:151,177-178,
// src/hs/cuda/main.cu:154,180, src/lk/cuda/main.cu:138,157-160) on a smooth
// pattern shifted by a known amount, and checks the recovered motion.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "amdflow.hpp"

static float pattern(double x, double y)
{
    return (float)(128 + 60 * std::sin(2 * M_PI * x / 37.0) * std::cos(2 * M_PI * y / 29.0) +
                   40 * std::sin(2 * M_PI * (x + y) / 53.0));
}

int main()
{
    const int W = 200, H = 160;
    const double du = 1.5, dv = -0.75;
    std::vector<uint8_t> I0(W * H), I1(W * H);
    std::vector<float> F0(W * H), F1(W * H);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            F0[y * W + x] = pattern(x, y);
            F1[y * W + x] = pattern(x - du, y - dv);
            I0[y * W + x] = (uint8_t)std::lround(F0[y * W + x]);
            I1[y * W + x] = (uint8_t)std::lround(F1[y * W + x]);
        }
    try {
        amdflow::TV_L1 tvl1(W, H);
        tvl1.runDualTVL1Multiscale(I0.data(), I1.data());
        const float* u = tvl1.getU();
        double su = 0, sv = 0;
        int n = 0;
        for (int y = 20; y < H - 20; y++)
            for (int x = 20; x < W - 20; x++, n++) {
                su += u[y * W + x];
                sv += u[W * H + y * W + x];
            }
        std::printf("TV-L1 mean flow (%.3f, %.3f), expected (%.2f, %.2f)\n", su / n, sv / n, du, dv);
        if (std::fabs(su / n - du) > 0.05 || std::fabs(sv / n - dv) > 0.05) return 1;

        // the reference's global-state calls, src/hs/cuda/main.cu:154,180
        amdflow::initFlow(1, W, W, H);
        std::vector<float> hu(W * H), hv(W * H);
        amdflow::ComputeFlow(F0.data(), F1.data(), W, H, W, 15.0f, 1, 3, 200, hu.data(), hv.data());
        amdflow::deleteFlow_mem(1);
        std::printf("HS centre flow (%.3f, %.3f)\n", hu[H / 2 * W + W / 2], hv[H / 2 * W + W / 2]);
        // the handle form gives the same flow
        of_hs* hs = amdflow::initFlowHandle(1, W, W, H);
        std::vector<float> hu2(W * H), hv2(W * H);
        amdflow::ComputeFlow(hs, F0.data(), F1.data(), W, H, W, 15.0f, 1, 3, 200, hu2.data(), hv2.data());
        amdflow::deleteFlow_mem(hs);
        if (hu != hu2 || hv != hv2) {
            std::printf("HS handle form differs from the global-state form\n");
            return 1;
        }

        amdflow::LucasKanade lk(3, 2, 5, W, H);
        lk.copy_frames_circularbuffer_GPU_wrapper(I0.data(), 2, 0, W * H);
        lk.copy_frames_circularbuffer_GPU_wrapper(I1.data(), 2, 1, W * H);
        std::vector<float> vx(W * H), vy(W * H);
        lk.lucas_kanade(vx.data(), vy.data(), 1);
        std::printf("LK centre flow (%.3f, %.3f)\n", vx[H / 2 * W + W / 2], vy[H / 2 * W + W / 2]);
    } catch (const amdflow::Error& e) {
        std::printf("error: %s\n", e.what());
        return 2;
    }
    std::printf("mirror_smoke OK\n");
    return 0;
}
