// ref_harness.cpp -- C entry points over the reference's own sequential
// bicubic/zoom sources, compiled from /root/reference by oracle/Makefile into
// oracle/_ref/.  TEST INFRASTRUCTURE ONLY: it lets the tests check the
// oracle's restatement of these functions against the reference itself.
// tvl1.cpp and mask/mask.cpp include "mkl.h", which this image lacks, so the
// rest of the reference path is not built here.
#include "bicubic_interpolation/bicubic_interpolation.hpp"
#include "zoom/zoom.hpp"

extern "C" {
// src/tvl1/sequential/bicubic_interpolation/bicubic_interpolation.cpp:74
float ref_bicubic_at(const float* input, float uu, float vv, int nx, int ny, int border_out)
{
    return bicubic_interpolation_at(input, uu, vv, nx, ny, border_out != 0);
}
// src/tvl1/sequential/bicubic_interpolation/bicubic_interpolation.cpp:141
void ref_bicubic_warp(const float* input, const float* u, const float* v, float* output,
                      int nx, int ny, int border_out)
{
    bicubic_interpolation_warp(input, u, v, output, nx, ny, border_out != 0);
}
// src/tvl1/sequential/zoom/zoom.cpp:61
void ref_zoom_in(const float* I, float* Iout, int nx, int ny, int nxx, int nyy)
{
    zoom_in(I, Iout, nx, ny, nxx, nyy);
}
// src/tvl1/sequential/zoom/zoom.hpp:14
void ref_zoom_size(int nx, int ny, float factor, int* nxx, int* nyy)
{
    zoom_size(nx, ny, nxx, nyy, factor);
}
}
