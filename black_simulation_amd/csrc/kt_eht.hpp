// kt_eht.hpp -- EHT radiative transfer over the DP5 geodesic engine (kt_eht.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_device.hpp"

namespace kte {

// partial sums the flux reduction needs at most (device scratch of this many doubles)
constexpr int kSumPartials = 1024;

struct EhtLaunch {
    long long n;          // rays: size * size pixels, or the explicit list's length
    long long first;      // this launch traces rays [first, n)
    const double* rays;   // explicit (alpha, beta) pairs, or null for the image's pixels
    int size;
    double half_width, px;
    // the view: spin, inclination (sin, cos and the cosine of the clamped theta), capture
    // radius r+ (1 + 1e-4), M in cm, observing frequency (Hz)
    double a, sinI, cosI, m0, r_cap, len_cm, nu;
    // the flow (eht::Flow): n0, T0, keplerian, sin_theta_b as given; hfac = max(H/R, 0.05);
    // b_coef = 8 pi m_p c^2 / (6 beta), so that B = sqrt(b_coef n_e / r)
    double n0, T0, keplerian, sin_theta_b, hfac, b_coef;
    // its plunge from the prograde ISCO: the ISCO's radius, E and L, the radius where the
    // plunge's density law starts, and start^-1.1 start^2 |u^r(start)| (continuity)
    double isco, E, L, start, start_rate;
    double rtol, atol;
    int max_steps;
    double* intensity;   // erg s^-1 cm^-2 Hz^-1 sr^-1 per pixel
    uint8_t* shadow;     // 1 where the ray fell in
    double* min_radius;  // optional
    int32_t* steps;      // optional
};

// persistent waves over rays [first, n): counter is one device word (zeroed here by a
// kernel), grid the most workgroups of 256 threads to launch
hipError_t render(const EhtLaunch& e, unsigned long long* counter, int grid, hipStream_t st);
// *out = sum of x[0..n) (device), in a fixed order; partial holds kSumPartials doubles
hipError_t sum(const double* x, long long n, double* partial, double* out, hipStream_t st);

}  // namespace kte
