// kt_particles.hip -- massive particles (mu = 1) on Kerr geodesics, advanced by a span of
// coordinate time: the debris update of shaders/particles_update.comp:207-255 (advance)
// and kerr::stepCoordinateTime (src/kerr.cpp:544-561), with the same right-hand side as
// the photons (kerr::derivatives, src/kerr.cpp:323-351, at mu = 1).
//
// One lane per particle.  The state is (r, mu = cos theta, phi_KS, vr, vmu), vr = dr/dlambda
// and vmu = dmu/dlambda in Mino time, integrated with the ingoing Kerr-Schild time t as the
// independent variable (dy/dt = (dy/dlambda) / (dt/dlambda), regular through the horizon):
//   dr/dlambda = vr,   dvr/dlambda = R'(r)/2 = 2 E r P - (r - 1) K - Delta r
//   dmu/dlambda = vmu, dvmu/dlambda = M'(mu)/2 = -(Q + L^2 - a^2 (E^2 - 1)) mu - 2 a^2 (E^2 - 1) mu^3
//   dphi/dlambda = a F - a E + L / sin^2,  dt/dlambda = (r^2 + a^2) F - a (a E sin^2 - L) - vr
// with P = E (r^2 + a^2) - a L, K = r^2 + (L - a E)^2 + Q, M = sin^2(theta) Theta, and F the
// inflow factor (P + vr) / Delta, taken as K / (P - vr) when P vr < 0 (kerr.cpp:323-329).
// Dormand-Prince 5(4) with the photons' controller; the last step ends exactly at dt.  After
// each accepted step: through the pole (L = 0 and no polar turning point before it) phi
// gains pi, vr and vmu go back onto R and M away from turning points (kerr.cpp:521-542,
// particle_common.glsl enforceMassiveConstraints), and the right-hand side is re-evaluated.
// A particle plunges below r_stop (trajectory.cpp:137: r+ + 0.02 by default) or escapes
// beyond r_escape.  Every operation is a correctly rounded +, -, *, /, sqrt or fma.
#include "kt_particles.hpp"

#include "kt_device.hpp"

using namespace ktd;

namespace ktp {

namespace {

struct Body {
    double a, a2, E, L, Q, aE, aL, K0, C1, A2E2, twoA2E2;
    bool pole;  // L = 0 and mu reaches +-1: phi jumps by pi at each pass over the pole
};

__device__ __forceinline__ void body_init(Body& b, double a, double E, double L, double Q) {
    // a particle that would pass within kAxisSnap radians of the spin axis goes through it
    // (L = 0), as the reference launches photons (src/kerr.cpp:399-405, kerr.h:115): its
    // closest approach sin(theta) ~ |L| / sqrt(Q + a^2 (E^2 - 1)) lies below what steps resolve
    if (fabs(L) < kAxisSnap * sqrt(fmax(Q + a * a * (E * E - 1.0), 0.0))) L = 0.0;
    b.a = a;
    b.a2 = a * a;
    b.E = E;
    b.L = L;
    b.Q = Q;
    b.aE = a * E;
    b.aL = a * L;
    const double x = L - b.aE;
    b.K0 = x * x + Q;
    b.A2E2 = b.a2 * (E * E - 1.0);  // a^2 (E^2 - mu^2)
    b.C1 = (Q + L * L) - b.A2E2;
    b.twoA2E2 = 2.0 * b.A2E2;
    b.pole = L == 0.0 && Q + b.A2E2 > 0.0;
}

// dy/dt from dy/dlambda and dt/dlambda
__device__ __forceinline__ void rhs(const Body& b, const double y[5], double dy[5]) {
    const double r = y[0], m = y[1], w = y[3];
    const double delta = fma(r, r - 2.0, b.a2);
    const double rra = fma(r, r, b.a2);
    const double P = fma(b.E, rra, -b.aL);
    const double K = fma(r, r, b.K0);
    const double s2t = fma(-m, m, 1.0);
    const double s2 = fmax(s2t, 1e-300);
    double F;
    if (P * w < 0.0) {
        F = K / (P - w);
    } else {
        const double d = fabs(delta) > 1e-12 ? delta : copysign(1e-12, delta);
        F = (P + w) / d;
    }
    const double tdot = fma(rra, F, -fma(b.a, fma(b.aE, s2t, -b.L), w));
    const double phidot = fma(b.a, F, -b.aE) + b.L / s2;
    const double accr = fma(2.0 * b.E * r, P, -fma(r - 1.0, K, delta * r));
    const double accm = -(m * fma(b.twoA2E2, m * m, b.C1));
    const double inv = 1.0 / tdot;
    dy[0] = w * inv;
    dy[1] = y[4] * inv;
    dy[2] = phidot * inv;
    dy[3] = accr * inv;
    dy[4] = accm * inv;
}

__device__ __forceinline__ void project(const Body& b, double y[5]) {
    const double r = y[0], m = y[1];
    const double delta = fma(r, r - 2.0, b.a2);
    const double P = fma(b.E, fma(r, r, b.a2), -b.aL);
    const double K = fma(r, r, b.K0);
    const double R = fma(P, P, -(delta * K));
    if (R > 0.05 * fma(P, P, fabs(delta) * K)) y[3] = copysign(sqrt(R), y[3]);
    const double m2 = m * m, s2 = 1.0 - m2;
    const double M = fma(-b.A2E2, m2 * m2, fma(-b.C1, m2, b.Q));
    const double scale = fma(b.L * b.L, m2, fma(fabs(b.Q), s2, fabs(b.A2E2) * (m2 * s2)));
    if (M > 0.05 * scale) y[4] = copysign(sqrt(M), y[4]);
}

__device__ __forceinline__ double dp5(const Body& b, const double y[5], double h, double k[7][5], double yn[5],
                                      double rtol, double atol) {
    double t[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = fma(h, A21 * k[0][i], y[i]);
    rhs(b, t, k[1]);
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = fma(h, sum2(A31, k[0][i], A32, k[1][i]), y[i]);
    rhs(b, t, k[2]);
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = fma(h, sum3(A41, k[0][i], A42, k[1][i], A43, k[2][i]), y[i]);
    rhs(b, t, k[3]);
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = fma(h, sum4(A51, k[0][i], A52, k[1][i], A53, k[2][i], A54, k[3][i]), y[i]);
    rhs(b, t, k[4]);
#pragma unroll
    for (int i = 0; i < 5; ++i)
        t[i] = fma(h, sum5(A61, k[0][i], A62, k[1][i], A63, k[2][i], A64, k[3][i], A65, k[4][i]), y[i]);
    rhs(b, t, k[5]);
#pragma unroll
    for (int i = 0; i < 5; ++i)
        yn[i] = fma(h, sum5(A71, k[0][i], A73, k[2][i], A74, k[3][i], A75, k[4][i], A76, k[5][i]), y[i]);
    rhs(b, yn, k[6]);
    ErrPick<double> ep;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const double e = h * sum6(E1, k[0][i], E3, k[2][i], E4, k[3][i], E5, k[4][i], E6, k[5][i], E7, k[6][i]);
        const double big = amax(fabs(y[i]), fabs(yn[i]));
        double sc;
        if (i == 0) sc = rtol * big;
        else if (i == 3) sc = rtol * amax(big, y[0] * y[0]);
        else sc = fma(rtol, big, atol);
        ep.add(e, sc);
    }
    return ep.value();
}

__global__ __launch_bounds__(128) void k_advance(ParticleLaunch p) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p.n) return;
    if (p.fate[q] != kMoving) {
        if (p.steps) p.steps[q] = 0;
        return;
    }
    Body b;
    body_init(b, p.a, p.E[q], p.L[q], p.Q[q]);
    double y[5] = {p.r[q], p.mu[q], p.phi[q], p.vr[q], p.vmu[q]};
    double k[7][5], yn[5];
    int fate = kMoving, steps = 0;
    bool rejected = false;
    double left = p.dt;
    double h = fmin(p.dt, 0.01 * y[0]);
    rhs(b, y, k[0]);
    while (left > 0.0) {
        if (steps >= p.max_steps) {
            fate = kUnfinished;
            break;
        }
        bool last = false;
        if (h >= left) {
            h = left;
            last = true;
        }
        const double err = dp5(b, y, h, k, yn, p.rtol, p.atol);
        ++steps;
        if (!(err <= 1.0)) {
            h *= step_factor(err, true);
            rejected = true;
            continue;
        }
        const double fac = step_factor(err, rejected);
        rejected = false;
        if (b.pole && y[1] * y[4] > 0.0 && yn[1] * yn[4] < 0.0) yn[2] += kPi;
        project(b, yn);
#pragma unroll
        for (int i = 0; i < 5; ++i) y[i] = yn[i];
        if (!(fabs(y[0]) < 1e300) || !(fabs(y[1]) < 1e300) || y[0] < p.r_stop) {
            fate = kPlunged;
            break;
        }
        if (y[0] > p.r_escape) {
            fate = kEscaped;
            break;
        }
        left = last ? 0.0 : left - h;
        h *= fac;
        rhs(b, y, k[0]);
    }
    p.r[q] = y[0];
    p.mu[q] = y[1];
    p.phi[q] = fabs(y[2]) < 1e300 ? wrap_2pi(y[2]) : y[2];
    p.vr[q] = y[3];
    p.vmu[q] = y[4];
    p.fate[q] = (int8_t)fate;
    if (p.steps) p.steps[q] = steps;
}

}  // namespace

hipError_t advance(const ParticleLaunch& p, hipStream_t st) {
    if (p.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_advance, dim3((unsigned)((p.n + 127) / 128)), dim3(128), 0, st, p);
    return hipGetLastError();
}

}  // namespace ktp
