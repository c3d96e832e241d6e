// output.cpp -- output formats of a flow field: the HSV colour rendering of
// the reference application (flowToColor, src/tvl1/sequential/main.cpp:93-126)
// and Middlebury .flo files.
//
// flowToColor: hue = atan2(v, u) in degrees / 2 and value = min(10*|flow|,
// 255), both rounded to 8 bits with saturation (cv::Mat::convertTo to
// CV_8U: round to nearest even, negatives to 0, so flows pointing into the
// upper half plane all get hue 0, as in the reference), saturation 255,
// then HSV -> BGR with 8-bit hue in [0, 180).  The HSV -> RGB step is the
// hexcone model (A. R. Smith, "Color gamut transform pairs", SIGGRAPH 1978)
// evaluated in float.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "amdflow.h"
#include "common.hpp"

namespace amdflow {
namespace {

__host__ __device__ __forceinline__ uint8_t sat_u8(float x)
{
    const float r = rintf(x);
    return (uint8_t)(r <= 0.0f ? 0.0f : (r >= 255.0f ? 255.0f : r));
}

__host__ __device__ __forceinline__ void hsv_to_bgr(uint8_t H, uint8_t S, uint8_t V, uint8_t* bgr)
{
    const float h6 = (float)H * (6.0f / 180.0f);
    const float s = (float)S * (1.0f / 255.0f);
    const float v = (float)V * (1.0f / 255.0f);
    float sector = floorf(h6);
    const float f = h6 - sector;
    int k = (int)sector;
    if (k < 0) k += 6;
    if (k >= 6) k -= 6;
    const float p = v * (1.0f - s);
    const float q = v * (1.0f - s * f);
    const float t = v * (1.0f - s * (1.0f - f));
    float r, g, b;
    switch (k) {
    case 0: r = v; g = t; b = p; break;
    case 1: r = q; g = v; b = p; break;
    case 2: r = p; g = v; b = t; break;
    case 3: r = p; g = q; b = v; break;
    case 4: r = t; g = p; b = v; break;
    default: r = v; g = p; b = q; break;
    }
    bgr[0] = sat_u8(b * 255.0f);
    bgr[1] = sat_u8(g * 255.0f);
    bgr[2] = sat_u8(r * 255.0f);
}

__global__ __launch_bounds__(256) void k_flow_to_color(const float* __restrict__ u,
                                                       const float* __restrict__ v, unsigned n,
                                                       uint8_t* __restrict__ bgr)
{
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float a = u[i], b = v[i];
    const float mag = sqrtf(a * a + b * b);
    const float val = fminf(mag * 10.0f, 255.0f);
    const float hue = atan2f(b, a) * (float)(180.0 / M_PI / 2.0);
    uint8_t px[3];
    hsv_to_bgr(sat_u8(hue), 255, sat_u8(val), px);
    bgr[3 * (size_t)i] = px[0];
    bgr[3 * (size_t)i + 1] = px[1];
    bgr[3 * (size_t)i + 2] = px[2];
}

constexpr float kFloTag = 202021.25f;   // "PIEH" in little-endian bytes

}  // namespace
}  // namespace amdflow

using namespace amdflow;

extern "C" {

int of_flow_to_color_device(const float* dU, int width, int height, uint8_t* dBGR, void* stream)
{
    if (!dU || !dBGR || width < 1 || height < 1 || (long long)width * height >= (1ll << 31))
        return OF_ERR_ARG;
    const unsigned n = (unsigned)width * (unsigned)height;
    hipLaunchKernelGGL(k_flow_to_color, dim3(cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, dU,
                       dU + n, n, dBGR);
    OF_HIP(hipGetLastError());
    return OF_OK;
}

int of_flow_to_color(const float* U, int width, int height, uint8_t* bgr)
{
    if (!U || !bgr || width < 1 || height < 1 || (long long)width * height >= (1ll << 31))
        return OF_ERR_ARG;
    const size_t n = (size_t)width * height;
    float* dU = nullptr;
    uint8_t* dB = nullptr;
    int st = dev_alloc(&dU, 2 * n);
    if (!st) st = dev_alloc(&dB, 3 * n);
    if (!st && hipMemcpy(dU, U, sizeof(float) * 2 * n, hipMemcpyHostToDevice) != hipSuccess) st = OF_ERR_HIP;
    if (!st) st = of_flow_to_color_device(dU, width, height, dB, nullptr);
    if (!st && hipMemcpy(bgr, dB, 3 * n, hipMemcpyDeviceToHost) != hipSuccess) st = OF_ERR_HIP;
    dev_free(dU);
    dev_free(dB);
    return st;
}

int of_flo_write(const char* path, const float* U, int width, int height)
{
    return guarded([&]() -> int {
        if (!path || !U || width < 1 || height < 1) return OF_ERR_ARG;
        std::vector<float> row(2 * (size_t)width);
        FILE* f = std::fopen(path, "wb");
        if (!f) return OF_ERR_ARG;
        const size_t n = (size_t)width * height;
        const int32_t wh[2] = {width, height};
        bool ok = std::fwrite(&kFloTag, 4, 1, f) == 1 && std::fwrite(wh, 4, 2, f) == 2;
        for (int y = 0; y < height && ok; y++) {
            for (int x = 0; x < width; x++) {
                row[2 * x] = U[(size_t)y * width + x];
                row[2 * x + 1] = U[n + (size_t)y * width + x];
            }
            ok = std::fwrite(row.data(), sizeof(float), row.size(), f) == row.size();
        }
        ok = (std::fclose(f) == 0) && ok;
        return ok ? OF_OK : OF_ERR_ARG;
    });
}

int of_flo_read(const char* path, float* U, long long cap, int* width, int* height)
{
    return guarded([&]() -> int {
        if (!path || !width || !height) return OF_ERR_ARG;
        // the handle is closed on every way out, a failed allocation included
        std::unique_ptr<FILE, int (*)(FILE*)> file(std::fopen(path, "rb"), &std::fclose);
        FILE* f = file.get();
        if (!f) return OF_ERR_ARG;
        float tag = 0.0f;
        int32_t wh[2] = {0, 0};
        int st = OF_OK;
        if (std::fread(&tag, 4, 1, f) != 1 || tag != kFloTag || std::fread(wh, 4, 2, f) != 2 || wh[0] < 1 ||
            wh[1] < 1 || (long long)wh[0] * wh[1] > (1ll << 31))
            st = OF_ERR_ARG;
        if (!st) {
            *width = wh[0];
            *height = wh[1];
            const size_t n = (size_t)wh[0] * wh[1];
            if (U) {
                if ((long long)(2 * n) > cap) {
                    st = OF_ERR_ARG;
                } else {
                    std::vector<float> row(2 * (size_t)wh[0]);
                    for (int y = 0; y < wh[1] && !st; y++) {
                        if (std::fread(row.data(), sizeof(float), row.size(), f) != row.size()) {
                            st = OF_ERR_ARG;
                            break;
                        }
                        for (int x = 0; x < wh[0]; x++) {
                            U[(size_t)y * wh[0] + x] = row[2 * x];
                            U[n + (size_t)y * wh[0] + x] = row[2 * x + 1];
                        }
                    }
                }
            }
        }
        return st;
    });
}

}  // extern "C"
