// kt_shade.hpp -- HDR colour of traced rays (kt_shade.hip).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_device.hpp"

namespace ktsh {

constexpr double kLogTMin = 2.5;  // log10 of the table's first temperature (316 K)
constexpr double kLogTMax = 5.3;  // log10 of its last (200,000 K)
// maximum of x^-3 (1 - x^-1/2), reached at x = 49/36: (36/49)^3 / 7
constexpr double kFluxPeak = (36.0 / 49.0) * (36.0 / 49.0) * (36.0 / 49.0) / 7.0;

struct ShadeLaunch {
    long long n;
    const ktd::ViewDev* views;
    int n_views;
    const int8_t* cls;
    const double *theta, *phi, *r_hit, *g;  // FP64 outputs, or null and:
    const float *theta32, *phi32, *r_hit32, *g32;  // the FP32 ones
    const float* table;   // blackbody RGB, table_n entries
    int table_n;
    double disk_temperature;
    float disk_scale;     // disk brightness / luminance of the disk temperature's colour
    int sky;              // 0 black, 1 grid, 2 image
    float sky_brightness, grid_line;
    const float* sky_image;
    int sky_w, sky_h;
    __half* rgba;
};

// CIE XYZ of a blackbody at T (Planck radiance against the colour matching functions)
void planck_xyz(double T, double xyz[3]);
// n linear-sRGB triples over log10 T in [kLogTMin, kLogTMax], luminance 1 at 6500 K
void blackbody_table(int n, float* rgb);
// Rec. 709 luminance of the table's colour at T (linear interpolation, as the kernel reads it)
float table_luminance(const float* tab, int n, double T);

hipError_t shade(const ShadeLaunch& s, hipStream_t st);

}  // namespace ktsh
