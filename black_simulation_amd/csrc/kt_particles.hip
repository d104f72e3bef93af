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

// dy/dt = (dy/dlambda) / (dt/dlambda) with one division: the inflow factor F = nf / df
// (kerr.cpp:323-329) and 1 / sin^2 are folded into g = 1 / (sin^2 (rra nf - df X)),
// X = a (a E sin^2 - L) + vr, so dt/dlambda = (rra nf - df X) / df and
// dphi/dlambda = (a nf sin^2 + L df - a E df sin^2) / (df sin^2)
__device__ __forceinline__ void rhs(const Body& b, const double y[5], double dy[5]) {
    const double r = y[0], m = y[1], w = y[3];
    const double delta = fma(r, r - 2.0, b.a2);
    const double rra = fma(r, r, b.a2);
    const double P = fma(b.E, rra, -b.aL);
    const double K = fma(r, r, b.K0);
    const double s2t = fma(-m, m, 1.0);
    const double s2 = fmax(s2t, 1e-200);  // keeps df s2 a normal number at the pole
    double nf, df;
    if (P * w < 0.0) {
        nf = K;
        df = P - w;
    } else {
        nf = P + w;
        df = fabs(delta) > 1e-12 ? delta : copysign(1e-12, delta);
    }
    const double X = fma(b.a, fma(b.aE, s2t, -b.L), w);
    const double den = fma(rra, nf, -(df * X));
    const double g = 1.0 / (s2 * den);
    const double ts = (df * s2) * g;  // 1 / (dt/dlambda)
    const double phin = fma(b.a * nf, s2, fma(b.L, df, -((b.aE * df) * s2)));
    const double accr = fma(2.0 * b.E * r, P, -fma(r - 1.0, K, delta * r));
    const double accm = -(m * fma(b.twoA2E2, m * m, b.C1));
    dy[0] = w * ts;
    dy[1] = y[4] * ts;
    dy[2] = phin * g;
    dy[3] = accr * ts;
    dy[4] = accm * ts;
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

// One particle in one lane: started from its arrays, stepped one DP5 attempt at a time.
struct PLane {
    Body b;
    double y[5], k0[5], h, left, hwant;
    int steps;
    bool rejected;
};

// the next particle of a lane, its columns loaded ahead of the lane's need for them
struct Next {
    long long q;
    double r, mu, phi, vr, vmu, E, L, Q, h;
    int fate;
};

__device__ __forceinline__ void next_load(Next& x, const ParticleLaunch& p, long long q) {
    x.q = q;
    x.fate = p.fate[q];
    x.r = p.r[q];
    x.mu = p.mu[q];
    x.phi = p.phi[q];
    x.vr = p.vr[q];
    x.vmu = p.vmu[q];
    x.E = p.E[q];
    x.L = p.L[q];
    x.Q = p.Q[q];
    x.h = p.h ? p.h[q] : 0.0;
}

__device__ __forceinline__ void lane_start(PLane& s, const ParticleLaunch& p, const Next& x) {
    body_init(s.b, p.a, x.E, x.L, x.Q);
    s.y[0] = x.r;
    s.y[1] = x.mu;
    s.y[2] = x.phi;
    s.y[3] = x.vr;
    s.y[4] = x.vmu;
    s.left = p.dt;
    s.h = x.h > 0.0 ? fmin(p.dt, x.h) : fmin(p.dt, 0.01 * s.y[0]);
    s.hwant = s.h;
    s.steps = 0;
    s.rejected = false;
    rhs(s.b, s.y, s.k0);
}

// one attempt; returns the fate once the particle is done with this call, else -1
__device__ __forceinline__ int lane_step(PLane& s, const ParticleLaunch& p) {
    if (!(s.left > 0.0)) return kMoving;
    if (s.steps >= p.max_steps) return kUnfinished;
    bool last = false;
    double h = s.h;
    s.hwant = h;  // the controller's own choice, before the span's end shortens it
    if (h >= s.left) {
        h = s.left;
        last = true;
    }
    double k[7][5], yn[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) k[0][i] = s.k0[i];
    const double err = dp5(s.b, s.y, h, k, yn, p.rtol, p.atol);
    ++s.steps;
    if (!(err <= 1.0)) {
        s.h = h * step_factor(err, true);
        s.rejected = true;
        return -1;
    }
    const double fac = step_factor(err, s.rejected);
    s.rejected = false;
    if (s.b.pole && s.y[1] * s.y[4] > 0.0 && yn[1] * yn[4] < 0.0) yn[2] += kPi;
    project(s.b, yn);
#pragma unroll
    for (int i = 0; i < 5; ++i) s.y[i] = yn[i];
    if (!(fabs(s.y[0]) < 1e300) || !(fabs(s.y[1]) < 1e300) || s.y[0] < p.r_stop) return kPlunged;
    if (s.y[0] > p.r_escape) return kEscaped;
    s.left = last ? 0.0 : s.left - h;
    if (!(s.left > 0.0)) return kMoving;
    s.h = h * fac;
    rhs(s.b, s.y, s.k0);
    return -1;
}

__device__ __forceinline__ void lane_write(const PLane& s, const ParticleLaunch& p, long long q, int fate) {
    p.r[q] = s.y[0];
    p.mu[q] = s.y[1];
    p.phi[q] = fabs(s.y[2]) < 1e300 ? wrap_2pi(s.y[2]) : s.y[2];
    p.vr[q] = s.y[3];
    p.vmu[q] = s.y[4];
    p.fate[q] = (int8_t)fate;
    if (p.steps) p.steps[q] = s.steps;
    if (p.h) p.h[q] = s.hwant;
}

__global__ void k_zero_counter(unsigned long long* c) {
    if (threadIdx.x == 0) *c = 0ull;
}

// Persistent waves pulling particles from one counter.  Every kRefill attempts the lanes
// whose particle is done take the next one, so the few particles that need hundreds of
// steps near the horizon or the axis do not hold 63 idle lanes each.  Each lane holds its
// next particle's columns, loaded while it steps the current one, and a wave takes indices
// 64 at a time (one atomic per 64 particles), so a refill waits on neither memory nor the
// counter.
constexpr int kRefill = 4;
constexpr int kChunk = 64;

__global__ __launch_bounds__(256) void k_advance(ParticleLaunch p, unsigned long long* counter) {
    const int lane = (int)(threadIdx.x & 63u);
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    PLane s;
    Next nx;
    bool active = false, has_next = false;
    long long my = 0;
    long long cbase = 0;  // the wave's claimed indices not yet handed out: [cbase, cbase + cleft)
    int cleft = 0;
    bool exhausted = false;  // the counter has passed n
    for (;;) {
        // lanes that are done take their prefetched particle
        if (!active && has_next) {
            has_next = false;
            my = nx.q;
            if (nx.fate == kMoving) {
                lane_start(s, p, nx);
                active = true;
            } else if (p.steps) {
                p.steps[my] = 0;
            }
        }
        // lanes without a next particle get one from the wave's chunk, and its loads issue
        const unsigned long long need = __ballot(!has_next);
        const int nneed = __popcll(need);
        if (nneed > 0 && (cleft > 0 || !exhausted)) {
            const int rank = __popcll(need & below);
            long long q = rank < cleft ? cbase + rank : -1;
            if (nneed > cleft && !exhausted) {
                unsigned long long nb = 0;
                if (lane == 0) nb = atomicAdd(counter, (unsigned long long)kChunk);
                nb = __shfl(nb, 0);
                if (nb + kChunk >= (unsigned long long)p.n) exhausted = true;
                const int from_new = nneed - cleft;
                if (rank >= cleft) q = (long long)nb + (rank - cleft);
                cbase = (long long)nb + from_new;
                cleft = kChunk - from_new;
            } else {
                const int used = nneed < cleft ? nneed : cleft;
                cbase += used;
                cleft -= used;
            }
            if (!has_next && q >= 0 && q < p.n) {
                next_load(nx, p, q);
                has_next = true;
            }
        }
        if (__ballot(active) == 0ull) {
            if (__ballot(has_next) == 0ull && exhausted && cleft == 0) break;
            continue;
        }
        for (int k = 0; k < kRefill; ++k) {
            if (active) {
                const int fate = lane_step(s, p);
                if (fate >= 0) {
                    lane_write(s, p, my, fate);
                    active = false;
                }
            }
        }
    }
}

}  // namespace

hipError_t advance(const ParticleLaunch& p, unsigned long long* counter, int grid, hipStream_t st) {
    if (p.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, counter);
    hipLaunchKernelGGL(k_advance, dim3((unsigned)grid), dim3(256), 0, st, p, counter);
    return hipGetLastError();
}

}  // namespace ktp
