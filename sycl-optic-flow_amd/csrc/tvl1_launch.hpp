// tvl1_launch.hpp -- host-callable launchers of the TV-L1 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "tvl1_kernels.hpp"

namespace amdflow {

// Resets the control block; with `iters`, also clears n_iters counts.
int launch_ctl_init(TvCtl* ctl, hipStream_t s, int* iters = nullptr, int n_iters = 0);
int launch_ctl_begin(TvCtl* ctl, hipStream_t s);
int launch_load_minmax_u8(const uint8_t* a, const uint8_t* b, float* fa, float* fb, size_t n,
                          TvCtl* ctl, hipStream_t s);
int launch_load_minmax_f32(const float* a, const float* b, float* fa, float* fb, size_t n,
                           TvCtl* ctl, hipStream_t s);
int launch_normalize(const float* a, const float* b, float* oa, float* ob, size_t n,
                     const TvCtl* ctl, hipStream_t s);
int launch_gaussian(const float* src, float* tmp, float* dst, int nx, int ny, const GaussTaps& g,
                    hipStream_t s);
int launch_zoom_out_sample(const float* src, float* dst, int nx, int ny, int nxx, int nyy,
                           float factor, hipStream_t s);
// Gaussian of two frames in one tiled launch (a -> oa, b -> ob, not in
// place); with `ctl`, the frames are normalised by the keys in ctl first.
// OF_ERR_UNSUPPORTED when the taps are wider than the tile's halo (16) or
// the image.
int launch_smooth_pair(const float* a, const float* b, float* oa, float* ob, int nx, int ny,
                       const GaussTaps& g, const TvCtl* ctl, hipStream_t s);
// zoom_out by exactly 1/2 of two frames in one launch (blur and sampling at
// even positions); OF_ERR_UNSUPPORTED as launch_smooth_pair.
int launch_zoom_out_half_pair(const float* a, const float* b, float* oa, float* ob, int nx, int ny,
                              int nxx, int nyy, const GaussTaps& g, hipStream_t s);
int launch_zoom_in_flow(const float* uc0, const float* uc1, float* uf, int nx, int ny, int nxx,
                        int nyy, float invz, const TvCtl* ctl, int coarse_level, hipStream_t s);
int launch_zoom_in_planes(const float* a, const float* b, float* oa, float* ob, int nx, int ny,
                          int nxx, int nyy, float scale, hipStream_t s);
int launch_centered_gradient(const float* in, float* dx, float* dy, int nx, int ny,
                             hipStream_t s);
int launch_fill(float* p, size_t n, float v, hipStream_t s);
int launch_scale(float* p, size_t n, float f, hipStream_t s);   // p *= f
// (I1, I1x, I1y, 0) per pixel, 4 floats, 16-byte aligned
// With `pzero` (16-byte aligned), also clears 4 floats per pixel there.
int launch_centered_gradient_packed(const float* in, float* out4, int nx, int ny, hipStream_t s,
                                    float* pzero = nullptr);
int launch_pack3(const float* a, const float* b, const float* c, float* out4, size_t n, hipStream_t s);
int launch_warp_setup(const float* I0, const float* src4, const float* ub0, const float* ub1,
                      float* wx, float* wy, float* rhoc, int nx, int ny, TvCtl* ctl, int level,
                      hipStream_t s);
// With `report` (host-mapped, n_iters + 1 ints), also writes the iteration
// counts `iters` and the control block's error word there.
int launch_copy_flow(const float* ub0, const float* ub1, float* out, size_t n2, const TvCtl* ctl,
                     int level, hipStream_t s, const int* iters = nullptr, int n_iters = 0,
                     int* report = nullptr);
int launch_energy(const float* I0, const float* I1, const float* u1, const float* u2, int nx,
                  int ny, float lambda, double* partials, int max_partials, double* out,
                  hipStream_t s);

int launch_hypot(const float* a, const float* b, float* o, size_t n, hipStream_t s);

}  // namespace amdflow
