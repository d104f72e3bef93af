// kt_device.hpp -- per-lane Kerr photon integrator for gfx950 (one ray per lane).
//
// The integrated system is Carter's separated geodesic equations in Mino time
// (src/kerr.h:17-22, src/kerr.cpp:331-351), traced backward (s = -tau, as
// shaders/trace.frag:28-30), written in (r, mu = cos theta) so that the right-hand
// side is rational -- no sin/cos in the step loop:
//     dr/ds = w_r,   dw_r/ds  = R'(r)/2  = 2 E r P - (r - 1) K
//     dmu/ds = w_mu, dw_mu/ds = M'(mu)/2 = -(Q + L^2 - a^2 E^2) mu - 2 a^2 E^2 mu^3
//     dphi/ds = -(a P / Delta - a E + L / (1 - mu^2))            (Boyer-Lindquist phi)
// with P = E (r^2 + a^2) - a L, Delta = r^2 - 2 r + a^2, K = (L - a E)^2 + Q and
// M(mu) = sin^2(theta) Theta(theta).  Dormand-Prince 5(4) with a max-norm
// controller; events in the order of src/trajectory.cpp:307-392 (disk crossing by a
// root of the dense output, capture, escape); an escaping ray finishes with an
// exact tail in u = 1/r down to u = 0, giving the asymptotic sky angles.
//
// In FP64 every operation in the step loop is a correctly rounded +, -, *, /,
// sqrt (FP contraction is off for this library), so the results do not depend on
// the compiler's choice of instructions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ktd {

constexpr double kPi = 3.14159265358979323846;

// Dormand-Prince 5(4) tableau and Hairer's DOPRI5 dense-output weights.
constexpr double A21 = 1.0 / 5.0;
constexpr double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
constexpr double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
constexpr double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0, A54 = -212.0 / 729.0;
constexpr double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0, A64 = 49.0 / 176.0,
                 A65 = -5103.0 / 18656.0;
constexpr double A71 = 35.0 / 384.0, A73 = 500.0 / 1113.0, A74 = 125.0 / 192.0, A75 = -2187.0 / 6784.0,
                 A76 = 11.0 / 84.0;
constexpr double E1 = 71.0 / 57600.0, E3 = -71.0 / 16695.0, E4 = 71.0 / 1920.0, E5 = -17253.0 / 339200.0,
                 E6 = 22.0 / 525.0, E7 = -1.0 / 40.0;
constexpr double D1 = -12715105075.0 / 11282082432.0, D3 = 87487479700.0 / 32700410799.0,
                 D4 = -10690763975.0 / 1880347072.0, D5 = 701980252875.0 / 199316789632.0,
                 D6 = -1453857185.0 / 822651844.0, D7 = 69997945.0 / 29380423.0;
constexpr double C2 = 1.0 / 5.0, C3 = 3.0 / 10.0, C4 = 4.0 / 5.0, C5 = 8.0 / 9.0;
constexpr double kSafety = 0.9, kFacMin = 0.2, kFacMax = 5.0;
constexpr double kAxisSnap = 3e-4;  // src/kerr.h:115

enum : int { kEscaped = 0, kCaptured = 1, kDisk = 2, kUnfinished = 3 };
enum : int { kFlagClosedForm = 0x08, kFlagRestep = 0x10, kFlagSurrogate = 0x20, kFlagNonfinite = 0x40 };

// Per-view constants, prepared on the host (trig and frames in FP64 there).
struct ViewDev {
    long long offset;  // first ray of this view in the batch
    int kind, width, height, disk, prograde, n_params;
    int sur;  // 1: the view lies in the surrogate's domain (set only for surrogate calls)
    // with sur: the view's network inputs 0-2, 1 / (1.5 R_sh) for inputs 3-4, and its
    // inclination (float, from the host)
    float sur_c[6];
    double a, r_cam, phi_cam;
    double phi0;  // phi at the camera as the integrator carries it: phi_cam - psi(r_cam) (outgoing
                  // Kerr-Schild), or phi_cam itself for observer cameras (ingoing)
    double half_width, half_height, px, py;  // Bardeen plane, pixel pitch
    double sinI, cosI, m0;                   // sin/cos of inclination, cos of the clamped theta
    double r_cap, r_esc, r_in, r_out, h0, r_fp64;  // r_fp64 = 1.2 r+ (mixed precision)
    double escape_radius_;                         // kind 2: the view's escape_radius
    double r_plus;                                 // kind 3: r+ (the far-side rule)
    double lt[4], lp[4], lh[4], er[4];       // pinhole: lowered frame (t, phi, theta) and e^r
    double B[9], tanh_fov, sn, cs, sigma;    // pinhole: camera basis (columns), trig at the camera
    const double* params;                    // explicit rays (alpha, beta) or look vectors, or null
};

static_assert(sizeof(ViewDev) % sizeof(unsigned long long) == 0, "ViewDev is staged into LDS as 8-byte words");

template <class R>
struct Photon {
    R a, a2, E, L, Q, aE, aL, K, C1, A2E2, twoE, twoA2E2;
};

template <class R>
__device__ __forceinline__ R amax(R x, R y) { return x > y ? x : y; }

template <class R>
__device__ __forceinline__ void photon_init(Photon<R>& p, R a, R E, R L, R Q) {
    p.a = a;
    p.a2 = a * a;
    p.E = E;
    p.L = L;
    p.Q = Q;
    p.aE = a * E;
    p.aL = a * L;
    const R x = L - p.aE;
    p.K = x * x + Q;
    p.A2E2 = p.a2 * (E * E);
    p.C1 = (Q + L * L) - p.A2E2;
    p.twoE = R(2) * E;
    p.twoA2E2 = R(2) * p.A2E2;
}

template <class R>
__device__ __forceinline__ R tiny_pos();
template <>
__device__ __forceinline__ double tiny_pos<double>() { return 1e-300; }
template <>
__device__ __forceinline__ float tiny_pos<float>() { return 1e-30f; }

// finite-range limits: 1e300 and 1e-200..1e200 in FP64, 1e30 and 1e-30..1e30 in FP32
template <class R>
__device__ __forceinline__ R huge_val();
template <>
__device__ __forceinline__ double huge_val<double>() { return 1e300; }
template <>
__device__ __forceinline__ float huge_val<float>() { return 1e30f; }
template <class R>
__device__ __forceinline__ R pow_lo();
template <>
__device__ __forceinline__ double pow_lo<double>() { return 1e-200; }
template <>
__device__ __forceinline__ float pow_lo<float>() { return 1e-30f; }
template <class R>
__device__ __forceinline__ R pow_hi();
template <>
__device__ __forceinline__ double pow_hi<double>() { return 1e200; }
template <>
__device__ __forceinline__ float pow_hi<float>() { return 1e30f; }

// Division and square root: correctly rounded in FP64 (the step loop must round as
// written); the hardware reciprocal / square root in FP32, whose bulk steps are judged
// against FP64 anyway.
__device__ __forceinline__ double qdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double qsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float qsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// P(r) = E (r^2 + a^2) - a L,  Delta(r) = r^2 - 2r + a^2
template <class R>
__device__ __forceinline__ R P_of(const Photon<R>& p, R r) { return fma(p.E, fma(r, r, p.a2), -p.aL); }
template <class R>
__device__ __forceinline__ R delta_of(const Photon<R>& p, R r) { return fma(r, r - R(2), p.a2); }

template <class R>
__device__ __forceinline__ R pot_R(const Photon<R>& p, R r) {
    const R P = P_of(p, r);
    return fma(P, P, -(delta_of(p, r) * p.K));
}

template <class R>
__device__ __forceinline__ R pot_M(const Photon<R>& p, R m) {
    const R m2 = m * m;
    return fma(-p.A2E2, m2 * m2, fma(-p.C1, m2, p.Q));
}

// dphi/ds = aE - [a P (1 - mu^2) + L Delta] / [Delta (1 - mu^2)]: one division; the floor
// on 1 - mu^2 keeps L = 0 rays finite at the pole (the ratio tends to aP/Delta).
//
// Which phi a lane carries.  kPhiIn: the ingoing Kerr-Schild phi (dphi_in = dphi_BL +
// a dr / Delta), regular on the future horizon, for cameras that may sit inside it:
// P / Delta becomes (P - w_r) / Delta, which on the regular branch (P w_r > 0, the ray
// rising to the horizon) is K / (P + w_r) by w_r^2 = P^2 - Delta K; on the other branch
// Delta is floored at 1e-12 as src/kerr.cpp:323-329 (inflowFactor) does.  kPhiOut: the
// outgoing one (dphi_out = dphi_BL - a dr / Delta), regular on the past horizon, which a
// ray from a camera outside, followed back into the hole, approaches: P / Delta becomes
// (P + w_r) / Delta = K / (P - w_r) on its regular branch (P w_r < 0).  Its phi stays
// finite down to the capture radius, where the Boyer-Lindquist phi winds without bound
// and would set the step size of rays whose phi is not an output; phi_out - phi_BL =
// -psi(r) vanishes at infinity, so the sky angle needs no conversion.  kPhiBL: the
// Boyer-Lindquist phi (the EHT transfer, whose outputs do not use phi).  One division in
// every form.
enum : int { kPhiBL = 0, kPhiIn = 1, kPhiOut = 2 };

template <class R, int PHI = kPhiBL>
__device__ __forceinline__ void rhs_main(const Photon<R>& p, const R y[5], R dy[5]) {
    const R r = y[0], m = y[1];
    const R delta = delta_of(p, r);
    const R P = P_of(p, r);
    const R s2 = fmax(fma(-m, m, R(1)), tiny_pos<R>());
    R nd = P, dd = delta;
    if (PHI == kPhiIn) {
        const R w = y[3];
        if (P * w > R(0)) {
            nd = p.K;
            dd = P + w;
        } else {
            nd = P - w;
            dd = fabs(delta) > R(1e-12) ? delta : copysign(R(1e-12), delta);
        }
    } else if (PHI == kPhiOut) {
        const R w = y[3];
        if (P * w < R(0)) {
            nd = p.K;
            dd = P - w;
        } else {
            nd = P + w;
            dd = fabs(delta) > R(1e-12) ? delta : copysign(R(1e-12), delta);
        }
    }
    const R num = fma(p.a * nd, s2, p.L * dd);
    dy[0] = y[3];
    dy[1] = y[4];
    dy[2] = p.aE - qdiv(num, dd * s2);
    dy[3] = fma(p.twoE * r, P, -((r - R(1)) * p.K));
    dy[4] = -(m * fma(p.twoA2E2, m * m, p.C1));
}

// tail in u = 1/r on slots (1, 2, 4) = (mu, phi, w_mu); ds/du = -1/sqrt(U),
// U = u^4 R(1/u) = Pu^2 - u^2 Du K.  The tail carries the Boyer-Lindquist phi, or for
// kPhiOut the outgoing Kerr-Schild one: dphi_out/du = dphi_BL/du + a / Du.
template <class R, int PHI = kPhiBL>
__device__ __forceinline__ void rhs_tail(const Photon<R>& p, R u, const R y[5], R dy[5]) {
    const R u2 = u * u;
    const R Pu = fma(p.E, fma(p.a2, u2, R(1)), -(p.aL * u2));
    const R Du = fma(p.a2, u2, fma(R(-2), u, R(1)));
    const R U = fma(Pu, Pu, -(u2 * Du * p.K));
    const R isq = qdiv(R(1), qsqrt(fmax(U, tiny_pos<R>())));
    const R m = y[1];
    const R s2 = fmax(fma(-m, m, R(1)), tiny_pos<R>());
    const R num = fma(p.a * Pu, s2, p.L * Du);
    dy[0] = R(0);
    dy[3] = R(0);
    dy[1] = -(y[4] * isq);
    dy[2] = (qdiv(num, Du * s2) - p.aE) * isq;
    if (PHI == kPhiOut) dy[2] = dy[2] + qdiv(p.a, Du);
    dy[4] = (m * fma(p.twoA2E2, m * m, p.C1)) * isq;
}

template <class R, int PHI = kPhiBL>
__device__ __forceinline__ void eval(const Photon<R>& p, bool tail, R u, const R y[5], R dy[5]) {
    if (tail) rhs_tail<R, PHI>(p, u, y, dy);
    else rhs_main<R, PHI>(p, y, dy);
}

// x^(-1/5): exponent split and three Newton steps on y^-5 = x (+, *, frexp, ldexp only)
template <class R>
__device__ __forceinline__ R pow_m15(R x) {
    const R T0 = R(1.0), T1 = R(0.8705505632961241), T2 = R(0.757858283255199), T3 = R(0.6597539553864471),
            T4 = R(0.5743491774985174);
    int e;
    if (!(x > pow_lo<R>() && x < pow_hi<R>())) return x > R(1) ? R(0) : R(1e20);  // clamped by the caller
    const R m = frexp(x, &e);
    const int q = (e + 1100) / 5 - 220;
    const int s = e - 5 * q;
    const R t = s == 0 ? T0 : s == 1 ? T1 : s == 2 ? T2 : s == 3 ? T3 : T4;
    R y = fma(R(-0.2974), m - R(0.5), R(1.1487));
    y = ldexp(y * t, -q);
#pragma unroll
    for (int it = 0; it < 3; ++it) {
        const R y2 = y * y;
        const R y5 = (y2 * y2) * y;
        y = (y * fma(-x, y5, R(6))) * R(0.2);
    }
    return y;
}

template <class R>
__device__ __forceinline__ R step_factor(R err, bool rejected_before) {
    R fac;
    if (err != err) return R(kFacMin);
    if (err == R(0)) fac = R(kFacMax);
    else fac = R(kSafety) * pow_m15(err);
    if (fac > R(kFacMax)) fac = R(kFacMax);
    if (fac < R(kFacMin)) fac = R(kFacMin);
    if (rejected_before && fac > R(1)) fac = R(1);
    return fac;
}

// Stage sums as FMA chains: y + h (c1 k1 + ... + cn kn).
template <class R>
__device__ __forceinline__ R sum2(R c1, R k1, R c2, R k2) { return fma(c2, k2, c1 * k1); }
template <class R>
__device__ __forceinline__ R sum3(R c1, R k1, R c2, R k2, R c3, R k3) { return fma(c3, k3, sum2(c1, k1, c2, k2)); }
template <class R>
__device__ __forceinline__ R sum4(R c1, R k1, R c2, R k2, R c3, R k3, R c4, R k4) {
    return fma(c4, k4, sum3(c1, k1, c2, k2, c3, k3));
}
template <class R>
__device__ __forceinline__ R sum5(R c1, R k1, R c2, R k2, R c3, R k3, R c4, R k4, R c5, R k5) {
    return fma(c5, k5, sum4(c1, k1, c2, k2, c3, k3, c4, k4));
}
template <class R>
__device__ __forceinline__ R sum6(R c1, R k1, R c2, R k2, R c3, R k3, R c4, R k4, R c5, R k5, R c6, R k6) {
    return fma(c6, k6, sum5(c1, k1, c2, k2, c3, k3, c4, k4, c5, k5));
}

// Max-norm error: the component with the largest |e_i| / sc_i, picked by comparing cross
// products, then one division.  NaN anywhere in e makes the error NaN.
template <class R>
struct ErrPick {
    R be, bs, guard;
    __device__ __forceinline__ ErrPick() : be(R(-1)), bs(R(1)), guard(R(0)) {}
    __device__ __forceinline__ void add(R e, R sc) {
        const R ae = fabs(e);
        guard += ae;
        if (be < R(0) || ae * bs > be * sc) {
            be = ae;
            bs = sc;
        }
    }
    __device__ __forceinline__ R value() const { return guard != guard ? guard : qdiv(be, bs); }
};

// Dense-output coefficients of one component (Hairer's DOPRI5 continuous extension).
template <class R>
struct Dense {
    R y, ydiff, bspl, c4, c5;
    __device__ __forceinline__ R at(R th) const {
        const R th1 = R(1) - th;
        return fma(th, fma(th1, fma(th, fma(th1, c5, c4), bspl), ydiff), y);
    }
};

// One DP5 attempt; k[0] = f(y) on entry; fills k[1..6] and yn; returns the scaled error.
template <class R, int PHI = kPhiBL>
__device__ __forceinline__ R dp5_attempt(const Photon<R>& p, bool tail, R u, const R y[5], R h, R k[7][5],
                                         R yn[5], R rtol, R atol) {
    R t[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = fma(h, R(A21) * k[0][i], y[i]);
    eval<R, PHI>(p, tail, u + R(C2) * h, t, k[1]);
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = fma(h, sum2(R(A31), k[0][i], R(A32), k[1][i]), y[i]);
    eval<R, PHI>(p, tail, u + R(C3) * h, t, k[2]);
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = fma(h, sum3(R(A41), k[0][i], R(A42), k[1][i], R(A43), k[2][i]), y[i]);
    eval<R, PHI>(p, tail, u + R(C4) * h, t, k[3]);
#pragma unroll
    for (int i = 0; i < 5; ++i)
        t[i] = fma(h, sum4(R(A51), k[0][i], R(A52), k[1][i], R(A53), k[2][i], R(A54), k[3][i]), y[i]);
    eval<R, PHI>(p, tail, u + R(C5) * h, t, k[4]);
#pragma unroll
    for (int i = 0; i < 5; ++i)
        t[i] = fma(h, sum5(R(A61), k[0][i], R(A62), k[1][i], R(A63), k[2][i], R(A64), k[3][i], R(A65), k[4][i]),
                   y[i]);
    eval<R, PHI>(p, tail, u + h, t, k[5]);
#pragma unroll
    for (int i = 0; i < 5; ++i)
        yn[i] = fma(h, sum5(R(A71), k[0][i], R(A73), k[2][i], R(A74), k[3][i], R(A75), k[4][i], R(A76), k[5][i]),
                    y[i]);
    eval<R, PHI>(p, tail, u + h, yn, k[6]);
    ErrPick<R> ep;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        if (tail && (i == 0 || i == 3)) continue;
        const R e = h * sum6(R(E1), k[0][i], R(E3), k[2][i], R(E4), k[3][i], R(E5), k[4][i], R(E6), k[5][i],
                             R(E7), k[6][i]);
        const R big = amax(fabs(y[i]), fabs(yn[i]));
        R sc;
        if (!tail && i == 0) sc = rtol * big;
        else if (!tail && i == 3) sc = rtol * amax(big, y[0] * y[0]);
        else sc = fma(rtol, big, atol);
        ep.add(e, sc);
    }
    return ep.value();
}

template <class R>
__device__ __forceinline__ Dense<R> dense_coeffs(const R y[5], const R yn[5], const R k[7][5], R h, int i) {
    Dense<R> d;
    d.y = y[i];
    d.ydiff = yn[i] - y[i];
    d.bspl = fma(h, k[0][i], -d.ydiff);
    d.c4 = fma(-h, k[6][i], d.ydiff - d.bspl);
    d.c5 = h * sum6(R(D1), k[0][i], R(D3), k[2][i], R(D4), k[3][i], R(D5), k[4][i], R(D6), k[5][i], R(D7), k[6][i]);
    return d;
}

// Constraint projection, src/kerr.cpp:521-542 (kTurningZone = 0.05) in (r, mu) form:
// away from a turning point w_r goes back onto w_r^2 = R(r) and w_mu onto
// w_mu^2 = M(mu) (the polar zone test is the reference's times sin^2 theta).  The
// second-order equations conserve w_r^2 - R with its error, and R ~ r^4.
template <class R>
__device__ __forceinline__ void project(const Photon<R>& p, R y[5]) {
    const R r = y[0], m = y[1];
    const R P = P_of(p, r);
    const R delta = delta_of(p, r);
    const R Rr = fma(P, P, -(delta * p.K));
    if (Rr > R(0.05) * fma(P, P, fabs(delta) * p.K)) y[3] = copysign(sqrt(Rr), y[3]);
    const R m2 = m * m, s2 = R(1) - m2;
    const R M = fma(-p.A2E2, m2 * m2, fma(-p.C1, m2, p.Q));
    const R scale = fma(p.L * p.L, m2, fma(fabs(p.Q), s2, p.A2E2 * (m2 * s2)));
    if (M > R(0.05) * scale) y[4] = copysign(sqrt(M), y[4]);
}

__device__ __forceinline__ double wrap_2pi(double x) {
    const double twopi = 2.0 * kPi;
    double w = x - twopi * floor(x / twopi);
    if (w >= twopi) w -= twopi;
    if (w < 0.0) w += twopi;
    return w;
}

// The ray one lane is tracing.
template <class R>
struct Lane {
    Photon<R> p;
    R y[5], k1[5];
    R h, u;
    R r_cap, r_esc, r_in, r_out;
    int steps;
    bool rejected, tail, disk, prograde;
    bool dark;  // observer cameras: the ray came from the far side (ends captured at once)
};

struct Outputs {
    int8_t* cls;
    double* theta;
    double* phi;
    double* r_hit;
    double* g;
    int32_t* steps;
    float *theta32, *phi32, *r_hit32, *g32;  // optional FP32 copies (rounded once)
    double *turns, *r_min, *delay, *b, *q;   // optional single-ray report (FP64 kernels only)
};

struct Result {
    int cls, flags;
    double theta, phi, r_hit, g;
};

// A launched ray: its photon, initial state and the radii its events test.
struct Launch {
    Photon<double> p;
    double y0[5];
    double r_cap, r_esc, r_in, r_out, h0, r_fp64;
    int disk, prograde;
    int dark;  // kind 3 inside the horizon: P(r+) = 2 r+ E - a L < 0
};

// phi_KS - phi_BL = psi(r) = integral from infinity to r of a / Delta, so that the two
// agree at infinity: (a / (r+ - r-)) log((r - r+) / (r - r-)) outside the horizon, the
// limit -a / (r - 1) at |a| = 1, and past it (no horizon, Delta = (r - 1)^2 + b^2,
// b^2 = a^2 - 1) (a / b)(atan((r - 1) / b) - pi / 2).
__device__ __forceinline__ double ks_shift(double a, double r) {
    if (a == 0.0) return 0.0;
    if (fabs(a) > 1.0) {
        const double b = sqrt(a * a - 1.0);
        return a / b * (atan((r - 1.0) / b) - 0.5 * kPi);
    }
    const double d = 2.0 * sqrt(fmax(0.0, 1.0 - a * a));
    const double rm = 1.0 - 0.5 * d;
    if (d == 0.0) return -a / (r - 1.0);
    return a / d * log1p(-d / (r - rm));
}

// Bardeen launch (src/eht.cpp:262-271): E = 1, L = -alpha sin i,
// Q = beta^2 + (alpha^2 - a^2) cos^2 i, inward, toward smaller theta when beta > 0.
__device__ __forceinline__ void bardeen_photon(double a, double r_cam, double phi_cam, double sinI, double cosI,
                                               double m0, double alpha, double beta, Photon<double>& p, double y0[5]) {
    const double L = -alpha * sinI;
    const double Q = beta * beta + (alpha * alpha - a * a) * cosI * cosI;
    photon_init(p, a, 1.0, L, Q);
    y0[0] = r_cam;
    y0[1] = m0;
    y0[2] = phi_cam;
    y0[3] = -sqrt(fmax(pot_R(p, r_cam), 0.0));
    y0[4] = (beta > 0.0 ? 1.0 : -1.0) * sqrt(fmax(pot_M(p, m0), 0.0));
}

// BPT ISCO radius (src/kerr.cpp:179-190)
__device__ __forceinline__ double isco_dev(double a, int prograde) {
    const double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
    const double z2 = sqrt(3.0 * a * a + z1 * z1);
    const double root = sqrt(fmax(0.0, (3.0 - z1) * (3.0 + z1 + 2.0 * z2)));
    return prograde ? 3.0 + z2 - root : 3.0 + z2 + root;
}

// Row and column of pixel `local` of a view `width` pixels wide, row 0 at the top: a 32-bit
// division where the index fits (a few instructions; a 64-bit one is a long sequence).
__device__ __forceinline__ void pixel_rc(long long local, int width, int& row, int& col) {
    if (local <= 0xffffffffLL) {
        const unsigned l = (unsigned)local, w = (unsigned)width, r = l / w;
        row = (int)r;
        col = (int)(l - r * w);
    } else {
        row = (int)(local / width);
        col = (int)(local - (long long)row * width);
    }
}

// Launch: the photon of ray `local` of a view -- a pixel of a Bardeen plane (kind 0) or of
// the pinhole camera (kind 1), an explicit (alpha, beta) or look, or (kind 2) a ray with
// its own camera: params hold (a, r_cam, inclination, alpha, beta) per ray and the
// per-ray constants are derived here with device libm.
__device__ __forceinline__ void launch_photon(const ViewDev& v, long long local, Launch& L) {
    L.r_cap = v.r_cap;
    L.r_esc = v.r_esc;
    L.r_in = v.r_in;
    L.r_out = v.r_out;
    L.h0 = v.h0;
    L.r_fp64 = v.r_fp64;
    L.disk = v.disk;
    L.prograde = v.prograde;
    L.dark = 0;
    Photon<double>& p = L.p;
    double* y0 = L.y0;
    const double a = v.a;
    if (v.kind == 2) {
        const double* q = v.params + 5 * local;
        const double ar = q[0], rc = q[1], inc = q[2];
        const double rp = 1.0 + sqrt(fmax(0.0, 1.0 - ar * ar));
        double th = inc;
        if (th < 1e-4) th = 1e-4;
        if (th > kPi - 1e-4) th = kPi - 1e-4;
        L.r_cap = v.r_cap * rp;  // kind 2 stores the capture factor in r_cap
        L.r_esc = v.escape_radius_ > 1.5 * rc ? v.escape_radius_ : 1.5 * rc;
        L.r_in = isco_dev(ar, v.prograde);
        L.h0 = 0.01 / rc;
        L.r_fp64 = 1.2 * rp;
        bardeen_photon(ar, rc, -ks_shift(ar, rc), sin(inc), cos(inc), cos(th), q[3], q[4], p, y0);  // phi_out
        return;
    }
    if (v.kind == 0) {  // Bardeen image plane, src/eht.cpp:262-271, 332-334
        double alpha, beta;
        if (v.params) {
            alpha = v.params[2 * local];
            beta = v.params[2 * local + 1];
        } else {
            int row, col;
            pixel_rc(local, v.width, row, col);
            alpha = -v.half_width + ((double)col + 0.5) * v.px;
            beta = v.half_height - ((double)row + 0.5) * v.py;
        }
        bardeen_photon(a, v.r_cam, v.phi0, v.sinI, v.cosI, v.m0, alpha, beta, p, y0);
        return;
    }
    // pinhole camera (kind 1: the static ZAMO; kind 3: any observer, frame from the host):
    // trace.frag:1494-1496 + launchPhoton (trace.frag:1470-1488)
    double n0, n1, n2;
    if (v.params) {
        n0 = v.params[3 * local];
        n1 = v.params[3 * local + 1];
        n2 = v.params[3 * local + 2];
    } else {
        int row, col;
        pixel_rc(local, v.width, row, col);
        const double W = (double)v.width, H = (double)v.height;
        const double nx = (((double)col + 0.5) - 0.5 * W) / (0.5 * H);
        const double ny = (((double)(v.height - row) - 0.5) - 0.5 * H) / (0.5 * H);
        const double x = nx * v.tanh_fov, yv = ny * v.tanh_fov;
        const double d0 = (x * v.B[0] + yv * v.B[3]) + v.B[6];
        const double d1 = (x * v.B[1] + yv * v.B[4]) + v.B[7];
        const double d2 = (x * v.B[2] + yv * v.B[5]) + v.B[8];
        const double inv = 1.0 / sqrt((d0 * d0 + d1 * d1) + d2 * d2);
        n0 = d0 * inv;
        n1 = d1 * inv;
        n2 = d2 * inv;
    }
    const double qt = ((v.lt[0] - n0 * v.lt[1]) - n1 * v.lt[2]) - n2 * v.lt[3];
    const double qp = ((v.lp[0] - n0 * v.lp[1]) - n1 * v.lp[2]) - n2 * v.lp[3];
    const double qh = ((v.lh[0] - n0 * v.lh[1]) - n1 * v.lh[2]) - n2 * v.lh[3];
    const double qr = ((v.er[0] - n0 * v.er[1]) - n1 * v.er[2]) - n2 * v.er[3];
    const double E = -qt;
    double Lz = qp;
    const double cs2 = v.cs * v.cs, sn2 = fmax(v.sn * v.sn, 1e-18), a2 = a * a;
    double Q = qh * qh + cs2 * (Lz * Lz / sn2 - a2 * E * E);
    if (fabs(Lz) < kAxisSnap * sqrt(fmax(Q + a2 * E * E, 0.0))) {
        Lz = 0.0;
        Q = qh * qh - cs2 * a2 * E * E;
    }
    photon_init(p, a, E, Lz, Q);
    y0[0] = v.r_cam;
    y0[1] = v.cs;
    y0[2] = v.phi0;
    y0[3] = -(v.sigma * qr);
    y0[4] = v.sn * qh;
    // inside the horizon every ray, followed back, rises to it; one that arrives with
    // P(r+) < 0 entered from the far side of the eternal spacetime: dark
    // (src/trajectory.cpp:234-241, shaders/trace.frag:1538-1541)
    if (v.kind == 3 && v.r_cam < v.r_plus && 2.0 * v.r_plus * E - a * Lz < 0.0) L.dark = 1;
}

template <class R, int PHI>
__device__ __forceinline__ void lane_start(Lane<R>& s, const Launch& la) {
    const Photon<double>& p = la.p;
    const double* y0 = la.y0;
    s.p.a = R(p.a);
    s.p.a2 = R(p.a2);
    s.p.E = R(p.E);
    s.p.L = R(p.L);
    s.p.Q = R(p.Q);
    s.p.aE = R(p.aE);
    s.p.aL = R(p.aL);
    s.p.K = R(p.K);
    s.p.C1 = R(p.C1);
    s.p.A2E2 = R(p.A2E2);
    s.p.twoE = R(p.twoE);
    s.p.twoA2E2 = R(p.twoA2E2);
#pragma unroll
    for (int i = 0; i < 5; ++i) s.y[i] = R(y0[i]);
    s.h = R(la.h0);
    s.u = R(0);
    s.r_cap = R(la.r_cap);
    s.r_esc = R(la.r_esc);
    s.r_in = R(la.r_in);
    s.r_out = R(la.r_out);
    s.steps = 0;
    s.rejected = false;
    s.tail = false;
    s.disk = la.disk != 0;
    s.prograde = la.prograde != 0;
    s.dark = PHI == kPhiIn && la.dark != 0;
    rhs_main<R, PHI>(s.p, s.y, s.k1);
}

// After an attempt: the controller, the pole, the projection and the events.  k7 is the
// attempt's f(yn); dr, dm its dense output of r and mu when the step crossed the plane.
// Returns true when the ray has ended (res filled).
template <class R, int PHI>
__device__ __forceinline__ bool lane_finish(Lane<R>& s, R err, R h, bool last, R yn[5], const R k7[5], bool crossing,
                                            const Dense<R>& dr, const Dense<R>& dm, Result& res) {
    if (!(err <= R(1))) {
        s.h = h * step_factor(err, true);
        s.rejected = true;
        return false;
    }
    const R fac = step_factor(err, s.rejected);
    s.rejected = false;
    if (s.p.L == R(0) && s.y[1] * s.y[4] > R(0) && yn[1] * yn[4] < R(0)) yn[2] += R(kPi);

    if (s.tail) {
        s.y[1] = yn[1];
        s.y[2] = yn[2];
        s.y[4] = yn[4];
#pragma unroll
        for (int i = 0; i < 5; ++i) s.k1[i] = k7[i];
        if (last) {
            R m = s.y[1];
            if (m > R(1)) m = R(1);
            if (m < R(-1)) m = R(-1);
            res.cls = kEscaped;
            res.theta = acos((double)m);
            res.phi = wrap_2pi((double)s.y[2]);
            return true;
        }
        s.u = s.u + h;
        s.h = h * fac;
        return false;
    }

    project(s.p, yn);
    if (!(fabs(yn[0]) < huge_val<R>()) || !(fabs(yn[1]) < huge_val<R>())) {  // non-finite: captured (eht.cpp:285)
        res.cls = kCaptured;
        res.flags |= kFlagNonfinite;
        return true;
    }
    if (crossing) {
        R lo = R(0), hi = R(1);
        const bool pos = s.y[1] > R(0);
        for (int it = 0; it < 48; ++it) {
            const R mid = R(0.5) * (lo + hi);
            const R f = dm.at(mid);
            if ((f > R(0)) == pos) lo = mid;
            else hi = mid;
        }
        const R rc = dr.at(R(0.5) * (lo + hi));
        if (rc >= s.r_in && rc <= s.r_out) {
            // trajectory.cpp:311-321, in FP64
            const double a = (double)s.p.a, r = (double)rc;
            const double sg = s.prograde ? 1.0 : -1.0;
            const double omega = sg / (r * sqrt(r) + sg * a);
            const double gtt = -(1.0 - 2.0 / r), gtp = -2.0 * a / r;
            const double gpp = r * r + a * a + 2.0 * a * a / r;
            const double norm = -(gtt + 2.0 * gtp * omega + gpp * omega * omega);
            if (norm > 0.0) {
                res.cls = kDisk;
                res.r_hit = r;
                res.g = sqrt(norm) / ((double)s.p.E - (double)s.p.L * omega);
                return true;
            }
        }
    }
    // KS cameras may start inside the capture radius: there a ray is captured only when
    // it heads inward (src/trajectory.cpp:377)
    if (yn[0] < s.r_cap && (PHI != kPhiIn || yn[3] < R(0))) {
        res.cls = kCaptured;
        return true;
    }
    if (yn[0] > s.r_esc && yn[3] > R(0)) {
        if (PHI == kPhiIn) yn[2] -= R(ks_shift((double)s.p.a, (double)yn[0]));  // its tail runs in BL phi
#pragma unroll
        for (int i = 0; i < 5; ++i) s.y[i] = yn[i];
        s.tail = true;
        s.u = R(1) / yn[0];
        s.h = -s.u;
        rhs_tail<R, PHI>(s.p, s.u, s.y, s.k1);
        return false;
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) s.y[i] = yn[i];
#pragma unroll
    for (int i = 0; i < 5; ++i) s.k1[i] = k7[i];
    s.k1[0] = s.y[3];  // FSAL after the projection: only dr/ds = w_r, dmu/ds = w_mu move
    s.k1[1] = s.y[4];
    s.h = h * fac;
    return false;
}

// One DP5 attempt plus events, all in precision R.
template <class R, int PHI>
__device__ __forceinline__ bool lane_step(Lane<R>& s, R rtol, R atol, int max_steps, Result& res) {
    res.flags = 0;
    if (PHI == kPhiIn && s.dark) {
        res.cls = kCaptured;
        return true;
    }
    if (s.steps >= max_steps) {
        res.cls = kUnfinished;
        return true;
    }
    R k[7][5], yn[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) k[0][i] = s.k1[i];
    bool last = false;
    if (s.tail && s.h <= -s.u) {
        s.h = -s.u;
        last = true;
    }
    const R h = s.h;
    const R err = dp5_attempt<R, PHI>(s.p, s.tail, s.u, s.y, h, k, yn, rtol, atol);
    ++s.steps;
    // dense output for the disk crossing, taken now so the stage vectors die here
    const bool crossing = !s.tail && s.disk && s.y[1] * yn[1] < R(0);
    Dense<R> dr, dm;
    if (crossing) {
        dr = dense_coeffs(s.y, yn, k, h, 0);
        dm = dense_coeffs(s.y, yn, k, h, 1);
    }
    return lane_finish<R, PHI>(s, err, h, last, yn, k[6], crossing, dr, dm, res);
}

// ---- mixed precision --------------------------------------------------------------
//
// The state and the controller live in FP64; a step is attempted in FP32 unless the ray
// is in an FP64 zone: r < 1.2 r+, inside the turning-point zones of kerr.cpp:521-542
// (R or M below 0.05 of their scale), within sin^2(theta) < 0.05 of the axis (where
// 1 - mu^2 loses its digits in FP32), or in the tail.  An FP32 step that fails the
// error test, or passes it but moves H = 0 by more than `drift` (|w_r^2 - R| or
// |w_mu^2 - M| relative to their scales; E, L and Q are exact constants here), is thrown
// away and re-taken in FP64 with the same h, so only FP64 attempts steer the step size:
// an FP32 error estimate held above tolerance by rounding noise cannot shrink h.  Both
// precisions use the mixed tolerances.
struct MixedLane {
    Lane<double> d;
    Photon<float> pf;
    float r_fp64;     // 1.2 r+
    int fp64_steps;
};

constexpr float kFp32AxisZone = 0.05f;  // FP64 while 1 - mu^2 is below this

__device__ __forceinline__ bool fp64_zone(const MixedLane& s) {
    if (s.d.tail) return true;
    const float r = (float)s.d.y[0], m = (float)s.d.y[1];
    if (r < s.r_fp64) return true;
    const Photon<float>& p = s.pf;
    const float P = P_of(p, r), delta = delta_of(p, r);
    const float Rr = fmaf(P, P, -(delta * p.K));
    if (!(Rr > 0.05f * fmaf(P, P, fabsf(delta) * p.K))) return true;
    const float m2 = m * m, s2 = 1.0f - m2;
    if (!(s2 >= kFp32AxisZone)) return true;
    const float M = fmaf(-p.A2E2, m2 * m2, fmaf(-p.C1, m2, p.Q));
    const float scale = fmaf(p.L * p.L, m2, fmaf(fabsf(p.Q), s2, p.A2E2 * (m2 * s2)));
    return !(M > 0.05f * scale);
}

// relative violation of H = 0 by an FP32 state
__device__ __forceinline__ float h_drift(const Photon<float>& p, const float y[5]) {
    const float r = y[0], m = y[1];
    const float P = P_of(p, r), delta = delta_of(p, r);
    const float dr = qdiv(fabsf(fmaf(y[3], y[3], -fmaf(P, P, -(delta * p.K)))), fmaf(P, P, fabsf(delta) * p.K));
    const float m2 = m * m, s2 = 1.0f - m2;
    const float M = fmaf(-p.A2E2, m2 * m2, fmaf(-p.C1, m2, p.Q));
    const float scale = fmaf(p.L * p.L, m2, fmaf(fabsf(p.Q), s2, p.A2E2 * (m2 * s2))) + 1e-30f;
    const float dm = qdiv(fabsf(fmaf(y[4], y[4], -M)), scale);
    return dr > dm ? dr : dm;
}

__device__ __forceinline__ Dense<double> widen(const Dense<float>& f) {
    Dense<double> d;
    d.y = f.y;
    d.ydiff = f.ydiff;
    d.bspl = f.bspl;
    d.c4 = f.c4;
    d.c5 = f.c5;
    return d;
}

__device__ __forceinline__ bool lane_step_mixed(MixedLane& s, double rtol, double atol, float drift, int max_steps,
                                                Result& res) {
    Lane<double>& d = s.d;
    res.flags = 0;
    if (d.steps >= max_steps) {
        res.cls = kUnfinished;
        return true;
    }
    bool last = false;
    if (d.tail && d.h <= -d.u) {
        d.h = -d.u;
        last = true;
    }
    const double h = d.h;
    double yn[5], k7[5], err = 0.0;
    bool crossing = false;
    Dense<double> dr, dm;
    bool use64 = fp64_zone(s);
    if (!use64) {
        float y[5], k[7][5], ynf[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            y[i] = (float)d.y[i];
            k[0][i] = (float)d.k1[i];
        }
        const float hf = (float)h;
        const float errf = dp5_attempt<float, kPhiOut>(s.pf, false, 0.0f, y, hf, k, ynf, (float)rtol, (float)atol);
        if (!(errf <= 1.0f) || h_drift(s.pf, ynf) > drift) {
            use64 = true;  // the re-step
        } else {
            err = errf;
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                yn[i] = ynf[i];
                k7[i] = k[6][i];
            }
            crossing = d.disk && y[1] * ynf[1] < 0.0f;
            if (crossing) {
                dr = widen(dense_coeffs(y, ynf, k, hf, 0));
                dm = widen(dense_coeffs(y, ynf, k, hf, 1));
            }
        }
    }
    if (use64) {
        double k[7][5];
#pragma unroll
        for (int i = 0; i < 5; ++i) k[0][i] = d.k1[i];
        err = dp5_attempt<double, kPhiOut>(d.p, d.tail, d.u, d.y, h, k, yn, rtol, atol);
#pragma unroll
        for (int i = 0; i < 5; ++i) k7[i] = k[6][i];
        crossing = !d.tail && d.disk && d.y[1] * yn[1] < 0.0;
        if (crossing) {
            dr = dense_coeffs(d.y, yn, k, h, 0);
            dm = dense_coeffs(d.y, yn, k, h, 1);
        }
        ++s.fp64_steps;
    }
    ++d.steps;
    return lane_finish<double, kPhiOut>(d, err, h, last, yn, k7, crossing, dr, dm, res);
}

// ---- the single-ray report: RayReport's turns, rMin and delay -------------------------
// (src/trajectory.h:97-113, accumulated along the march of src/trajectory.cpp:259-396)
//
// Over each accepted step, up to the step's event when it has one (the disk crossing, or
// r = r_cap, or r = r_esc, each located on the step's dense output):
//   turns  the angle swept by the direction of the position in Kerr-Schild Cartesian form
//          (kerr::toCartesian, src/kerr.cpp:638-641), over 2 pi: great-circle angles
//          between the directions at equal sub-steps of the dense output (report_parts),
//          summed;
//   r_min  the least r: the camera, the step ends, the event radius, and where w_r turns
//          from in- to outgoing inside a step, the root of R(r) = 0 refined by Newton from
//          the dense output's estimate;
//   delay  ingoing Kerr-Schild coordinate time between the event and the camera (the
//          reference's State::t, src/kerr.cpp:350: dt/dtau = (r^2 + a^2) I - a (a E
//          sin^2 - L) - v_r with I its inflowFactor), by 5-point Gauss-Legendre on equal
//          parts of each step's dense output, half as many as the swept angle's (at least
//          16: the integrand grows as 1/(r - r+) toward the capture radius).
struct ReportAcc {
    double swept, r_min, delay;
    bool done;  // the event has been reached: nothing more accumulates
};

// parts of a step for the swept angle (and half as many for the delay): at least 32, and
// one per 0.02 rad of the ingoing phi the step turns through, which near the capture
// radius the outgoing phi carried (phi_in = phi_out + 2 psi(r)) does not show: bounded by
// |dr| |a| / Delta at the step's lower end.  Plain arithmetic, so host and device agree.
constexpr int kReportSub = 32, kReportMaxSub = 65536;
__device__ __forceinline__ int report_parts(const Photon<double>& p, bool ingoing, double r0, double r1, double dphi) {
    double turn = fabs(dphi);
    if (!ingoing) {
        const double lo = fmin(r0, r1);
        const double dlo = fmax(delta_of(p, lo), 1e-12);
        turn += 2.0 * fabs(r1 - r0) * fabs(p.a) / dlo;
    }
    const double n = (double)kReportSub + ceil(turn / 0.02);
    return n < (double)kReportMaxSub ? (int)n : kReportMaxSub;
}
// Gauss-Legendre on [0, 1], 5 nodes (x = (1 + t) / 2, w = W / 2 of the [-1, 1] rule)
constexpr double kGLx[5] = {0.04691007703066800, 0.23076534494715845, 0.5, 0.76923465505284155,
                            0.95308992296933200};
constexpr double kGLw[5] = {0.11846344252809454, 0.23931433524968324, 0.28444444444444444,
                            0.23931433524968324, 0.11846344252809454};

// unit direction of kerr::toCartesian(a, r, theta, phi_ks)
__device__ __forceinline__ void ks_direction(double a, double r, double mu, double phi_ks, double d[3]) {
    const double sn = sqrt(fmax(0.0, 1.0 - mu * mu));
    const double cp = cos(phi_ks), sp = sin(phi_ks);
    const double x = (r * cp - a * sp) * sn, y = (r * sp + a * cp) * sn, z = r * mu;
    const double n = sqrt(x * x + y * y + z * z);
    d[0] = x / n;
    d[1] = y / n;
    d[2] = z / n;
}

// the angle between unit vectors, accurate for small ones too
__device__ __forceinline__ double angle_between(const double u[3], const double v[3]) {
    const double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2], cz = u[0] * v[1] - u[1] * v[0];
    return atan2(sqrt(cx * cx + cy * cy + cz * cz), u[0] * v[0] + u[1] * v[1] + u[2] * v[2]);
}

// ingoing Kerr-Schild dt/dtau of the photon at (r, mu) with w = dr/ds = -v_r
__device__ __forceinline__ double dt_dtau(const Photon<double>& p, double r, double mu, double w) {
    const double delta = delta_of(p, r), P = P_of(p, r);
    double inflow;
    if (P * w > 0.0) {
        inflow = p.K / (P + w);
    } else {
        const double dd = fabs(delta) > 1e-12 ? delta : copysign(1e-12, delta);
        inflow = (P - w) / dd;
    }
    return fma(r, r, p.a2) * inflow + p.a * (p.L - p.aE * (1.0 - mu * mu)) + w;
}

// the fraction of a step where the dense output d first reaches `target` from `from_below`
__device__ __forceinline__ double dense_root(const Dense<double>& d, double target, bool from_below) {
    double lo = 0.0, hi = 1.0;
    for (int it = 0; it < 48; ++it) {
        const double mid = 0.5 * (lo + hi);
        if ((d.at(mid) < target) == from_below) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

template <int PHI>
__device__ __forceinline__ void report_start(const Lane<double>& s, ReportAcc& acc) {
    acc.swept = 0.0;
    acc.r_min = s.y[0];
    acc.delay = 0.0;
    acc.done = false;
}

// lane_step in FP64 with the report accumulated over the accepted steps (the integration
// itself is the same operation for operation, so every other output equals lane_step's)
template <int PHI>
__device__ __forceinline__ bool report_step(Lane<double>& s, ReportAcc& acc, double rtol, double atol, int max_steps,
                                            Result& res) {
    res.flags = 0;
    if (PHI == kPhiIn && s.dark) {
        res.cls = kCaptured;
        return true;
    }
    if (s.steps >= max_steps) {
        res.cls = kUnfinished;
        return true;
    }
    double k[7][5], yn[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) k[0][i] = s.k1[i];
    bool last = false;
    if (s.tail && s.h <= -s.u) {
        s.h = -s.u;
        last = true;
    }
    const double h = s.h;
    const double err = dp5_attempt<double, PHI>(s.p, s.tail, s.u, s.y, h, k, yn, rtol, atol);
    ++s.steps;
    const bool main = !s.tail;
    const bool crossing = main && s.disk && s.y[1] * yn[1] < 0.0;
    Dense<double> d[5];
    double y0[5];
    if (main) {
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            d[i] = dense_coeffs(s.y, yn, k, h, i);
            y0[i] = s.y[i];
        }
    }
    const bool flip = s.p.L == 0.0 && s.y[1] * s.y[4] > 0.0 && yn[1] * yn[4] < 0.0;
    const bool ended = lane_finish<double, PHI>(s, err, h, last, yn, k[6], crossing, d[0], d[1], res);
    if (!main || acc.done || (!ended && s.rejected)) return ended;
    if (ended && (res.flags & kFlagNonfinite)) {
        acc.done = true;
        return ended;
    }
    // the part of this step before its event
    double th_end = 1.0;
    if (ended) {
        if (res.cls == kDisk) th_end = dense_root(d[1], 0.0, !(y0[1] > 0.0));
        else if (res.cls == kCaptured && y0[0] >= s.r_cap) th_end = dense_root(d[0], s.r_cap, false);
        acc.done = true;
    } else if (s.tail) {
        th_end = dense_root(d[0], s.r_esc, true);
        acc.done = true;
    }
    const double a = (double)s.p.a;
    // r_min
    double rm = fmin(acc.r_min, d[0].at(th_end));
    if (y0[3] < 0.0 && d[3].at(th_end) > 0.0) {
        const double tt = dense_root(d[3], 0.0, true);
        const double r0 = d[0].at(tt);
        double r = r0;
        for (int it = 0; it < 4; ++it) {
            const double P = P_of(s.p, r);
            const double dR = 4.0 * s.p.E * r * P - 2.0 * (r - 1.0) * s.p.K;
            if (dR == 0.0) break;
            r -= pot_R(s.p, r) / dR;
        }
        rm = fmin(rm, fabs(r - r0) <= 1e-6 * r0 ? r : r0);
    }
    acc.r_min = rm;
    // turns: sub-step directions (through the pole for L = 0 the dense phi lacks the
    // half-turn the step adds, so it is added where mu w_mu has changed sign)
    double dir0[3], dir1[3];
    const int nsub = report_parts(s.p, PHI == kPhiIn, y0[0], d[0].at(th_end), d[2].at(th_end) - y0[2]);
    // toCartesian takes the ingoing phi: phi_in = phi_out + 2 psi(r)
    ks_direction(a, y0[0], y0[1], PHI == kPhiIn ? y0[2] : y0[2] + 2.0 * ks_shift(a, y0[0]), dir0);
    for (int j = 1; j <= nsub; ++j) {
        const double th = th_end * (double)j / (double)nsub;
        const double r = d[0].at(th), mu = d[1].at(th);
        double ph = d[2].at(th);
        if (flip && mu * d[4].at(th) < 0.0) ph += kPi;
        ks_direction(a, r, mu, PHI == kPhiIn ? ph : ph + 2.0 * ks_shift(a, r), dir1);
        acc.swept += angle_between(dir0, dir1);
        dir0[0] = dir1[0];
        dir0[1] = dir1[1];
        dir0[2] = dir1[2];
    }
    // delay
    const int nquad = nsub / 2;
    double q = 0.0;
    for (int m = 0; m < nquad; ++m)
#pragma unroll
        for (int g = 0; g < 5; ++g) {
            const double th = th_end * ((double)m + kGLx[g]) / (double)nquad;
            q = fma(kGLw[g], dt_dtau(s.p, d[0].at(th), d[1].at(th), d[3].at(th)), q);
        }
    acc.delay = fma(h * th_end / (double)nquad, q, acc.delay);
    return ended;
}

__device__ __forceinline__ void write_report(const Outputs& o, long long idx, const Photon<double>& p,
                                             const ReportAcc& acc) {
    const double invE = fabs(p.E) > 1e-12 ? 1.0 / p.E : 0.0;  // src/trajectory.cpp:221
    if (o.turns) o.turns[idx] = acc.swept / (2.0 * kPi);
    if (o.r_min) o.r_min[idx] = acc.r_min;
    if (o.delay) o.delay[idx] = acc.delay;
    if (o.b) o.b[idx] = p.L * invE;
    if (o.q) o.q[idx] = p.Q * invE * invE;
}

__device__ __forceinline__ void write_result(const Outputs& o, long long idx, const Result& r, int steps, int flags) {
    const double nan = __builtin_nan("");
    if (o.cls) o.cls[idx] = (int8_t)(r.cls | r.flags | flags);
    if (o.theta) o.theta[idx] = r.cls == kEscaped ? r.theta : nan;
    if (o.phi) o.phi[idx] = r.cls == kEscaped ? r.phi : nan;
    if (o.r_hit) o.r_hit[idx] = r.cls == kDisk ? r.r_hit : nan;
    if (o.g) o.g[idx] = r.cls == kDisk ? r.g : nan;
    if (o.steps) o.steps[idx] = steps;
    if (o.theta32) o.theta32[idx] = (float)(r.cls == kEscaped ? r.theta : nan);
    if (o.phi32) o.phi32[idx] = (float)(r.cls == kEscaped ? r.phi : nan);
    if (o.r_hit32) o.r_hit32[idx] = (float)(r.cls == kDisk ? r.r_hit : nan);
    if (o.g32) o.g32[idx] = (float)(r.cls == kDisk ? r.g : nan);
}

}  // namespace ktd
