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
};

hipError_t advance(const ParticleLaunch& p, hipStream_t st);

}  // namespace ktp
