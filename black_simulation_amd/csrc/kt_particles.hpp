// kt_particles.hpp -- massive particles (debris) on Kerr geodesics (kt_particles.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ktp {

enum : int { kMoving = 0, kPlunged = 1, kEscaped = 2, kUnfinished = 3 };

struct ParticleLaunch {
    long long n;
    double a, dt, r_stop, r_escape, rtol, atol;
    int max_steps;
    // state (in and out): r, mu = cos(theta), Kerr-Schild phi, vr = Sigma dr/dtau,
    // vmu = Sigma dmu/dtau; constants E, L, Q per particle; fate (in: 0 = moving)
    double *r, *mu, *phi, *vr, *vmu;
    const double *E, *L, *Q;
    int8_t* fate;
    int32_t* steps;  // optional
    double* h;       // optional: first step size to try; out: the controller's next step
};

// counter: one device word (zeroed here by a kernel); grid: workgroups of 256 threads
hipError_t advance(const ParticleLaunch& p, unsigned long long* counter, int grid, hipStream_t st);

}  // namespace ktp
