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

// Per-view constants, prepared on the host (trig and frames in FP64 there).
struct ViewDev {
    long long offset;  // first ray of this view in the batch
    int kind, width, height, disk, prograde, n_params;
    double a, r_cam, phi_cam;
    double half_width, half_height, px, py;  // Bardeen plane, pixel pitch
    double sinI, cosI, m0;                   // sin/cos of inclination, cos of the clamped theta
    double r_cap, r_esc, r_in, r_out, h0;
    double lt[4], lp[4], lh[4], er[4];       // pinhole: lowered frame (t, phi, theta) and e^r
    double B[9], tanh_fov, sn, cs, sigma;    // pinhole: camera basis (columns), trig at the camera
    const double* params;                    // explicit rays (alpha, beta) or look vectors, or null
};

static_assert(sizeof(ViewDev) % sizeof(unsigned long long) == 0, "ViewDev is staged into LDS as 8-byte words");

template <class R>
struct Photon {
    R a, a2, E, L, Q, aE, aL, K, C1, A2E2;
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
}

template <class R>
__device__ __forceinline__ R pot_R(const Photon<R>& p, R r) {
    const R P = p.E * (r * r + p.a2) - p.aL;
    const R delta = (r * r - R(2) * r) + p.a2;
    return P * P - delta * p.K;
}

template <class R>
__device__ __forceinline__ R pot_M(const Photon<R>& p, R m) {
    const R m2 = m * m;
    return (p.Q - p.C1 * m2) - p.A2E2 * (m2 * m2);
}

template <class R>
__device__ __forceinline__ void rhs_main(const Photon<R>& p, const R y[5], R dy[5]) {
    const R r = y[0], m = y[1];
    const R r2 = r * r;
    const R delta = (r2 - R(2) * r) + p.a2;
    const R P = p.E * (r2 + p.a2) - p.aL;
    dy[0] = y[3];
    dy[1] = y[4];
    dy[2] = -(((p.a * P) / delta - p.aE) + p.L / (R(1) - m * m));
    dy[3] = R(2) * p.E * r * P - (r - R(1)) * p.K;
    dy[4] = -(p.C1 * m) - R(2) * p.A2E2 * (m * m * m);
}

template <class R>
__device__ __forceinline__ R tiny_pos();
template <>
__device__ __forceinline__ double tiny_pos<double>() { return 1e-300; }
template <>
__device__ __forceinline__ float tiny_pos<float>() { return 1e-30f; }

// tail in u = 1/r on slots (1, 2, 4) = (mu, phi, w_mu); ds/du = -1/sqrt(U),
// U = u^4 R(1/u) = Pu^2 - u^2 Du K
template <class R>
__device__ __forceinline__ void rhs_tail(const Photon<R>& p, R u, const R y[5], R dy[5]) {
    const R u2 = u * u;
    const R Pu = p.E * (R(1) + p.a2 * u2) - p.aL * u2;
    const R Du = (R(1) - R(2) * u) + p.a2 * u2;
    const R U = Pu * Pu - u2 * Du * p.K;
    const R isq = R(1) / sqrt(fmax(U, tiny_pos<R>()));
    const R m = y[1];
    dy[0] = R(0);
    dy[3] = R(0);
    dy[1] = -y[4] * isq;
    dy[2] = (((p.a * Pu) / Du - p.aE) + p.L / (R(1) - m * m)) * isq;
    dy[4] = (p.C1 * m + R(2) * p.A2E2 * (m * m * m)) * isq;
}

template <class R>
__device__ __forceinline__ void eval(const Photon<R>& p, bool tail, R u, const R y[5], R dy[5]) {
    if (tail) rhs_tail(p, u, y, dy);
    else rhs_main(p, y, dy);
}

// x^(-1/5): exponent split and three Newton steps on y^-5 = x (+, *, frexp, ldexp only)
template <class R>
__device__ __forceinline__ R pow_m15(R x) {
    const R T0 = R(1.0), T1 = R(0.8705505632961241), T2 = R(0.757858283255199), T3 = R(0.6597539553864471),
            T4 = R(0.5743491774985174);
    int e;
    if (!(x > R(1e-30) && x < R(1e30))) return x > R(1) ? R(0) : R(1e20);  // inf/NaN/0: clamped by the caller
    const R m = frexp(x, &e);
    const int q = (e + 1100) / 5 - 220;
    const int s = e - 5 * q;
    const R t = s == 0 ? T0 : s == 1 ? T1 : s == 2 ? T2 : s == 3 ? T3 : T4;
    R y = R(1.1487) - R(0.2974) * (m - R(0.5));
    y = ldexp(y * t, -q);
#pragma unroll
    for (int it = 0; it < 3; ++it) {
        const R y5 = y * y * y * y * y;
        y = y * (R(6) - x * y5) * R(0.2);
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

// One DP5 attempt; k[0] = f(y) on entry; fills k[1..6] and yn; returns the scaled error.
template <class R>
__device__ __forceinline__ R dp5_attempt(const Photon<R>& p, bool tail, R u, const R y[5], R h, R k[7][5],
                                         R yn[5], R rtol, R atol) {
    R t[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = y[i] + h * (R(A21) * k[0][i]);
    eval(p, tail, u + R(C2) * h, t, k[1]);
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = y[i] + h * (R(A31) * k[0][i] + R(A32) * k[1][i]);
    eval(p, tail, u + R(C3) * h, t, k[2]);
#pragma unroll
    for (int i = 0; i < 5; ++i) t[i] = y[i] + h * (R(A41) * k[0][i] + R(A42) * k[1][i] + R(A43) * k[2][i]);
    eval(p, tail, u + R(C4) * h, t, k[3]);
#pragma unroll
    for (int i = 0; i < 5; ++i)
        t[i] = y[i] + h * (R(A51) * k[0][i] + R(A52) * k[1][i] + R(A53) * k[2][i] + R(A54) * k[3][i]);
    eval(p, tail, u + R(C5) * h, t, k[4]);
#pragma unroll
    for (int i = 0; i < 5; ++i)
        t[i] = y[i] + h * (R(A61) * k[0][i] + R(A62) * k[1][i] + R(A63) * k[2][i] + R(A64) * k[3][i] +
                           R(A65) * k[4][i]);
    eval(p, tail, u + h, t, k[5]);
#pragma unroll
    for (int i = 0; i < 5; ++i)
        yn[i] = y[i] + h * (R(A71) * k[0][i] + R(A73) * k[2][i] + R(A74) * k[3][i] + R(A75) * k[4][i] +
                            R(A76) * k[5][i]);
    eval(p, tail, u + h, yn, k[6]);
    R err = R(0);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        if (tail && (i == 0 || i == 3)) continue;
        const R e = h * (R(E1) * k[0][i] + R(E3) * k[2][i] + R(E4) * k[3][i] + R(E5) * k[4][i] +
                         R(E6) * k[5][i] + R(E7) * k[6][i]);
        const R big = amax(fabs(y[i]), fabs(yn[i]));
        R sc;
        if (!tail && i == 0) sc = rtol * big;
        else if (!tail && i == 3) sc = rtol * amax(big, y[0] * y[0]);
        else sc = atol + rtol * big;
        const R q = fabs(e) / sc;
        if (!(q <= err)) err = q;
    }
    return err;
}

template <class R>
__device__ __forceinline__ R dense(const R y[5], const R yn[5], const R k[7][5], R h, int i, R th) {
    const R ydiff = yn[i] - y[i];
    const R bspl = h * k[0][i] - ydiff;
    const R c4 = ydiff - h * k[6][i] - bspl;
    const R c5 = h * (R(D1) * k[0][i] + R(D3) * k[2][i] + R(D4) * k[3][i] + R(D5) * k[4][i] + R(D6) * k[5][i] +
                      R(D7) * k[6][i]);
    const R th1 = R(1) - th;
    return y[i] + th * (ydiff + th1 * (bspl + th * (c4 + th1 * c5)));
}

// Constraint projection, src/kerr.cpp:521-542 (kTurningZone = 0.05) in (r, mu) form:
// away from a turning point w_r goes back onto w_r^2 = R(r) and w_mu onto
// w_mu^2 = M(mu) (the polar zone test is the reference's times sin^2 theta).  The
// second-order equations conserve w_r^2 - R with its error, and R ~ r^4.
template <class R>
__device__ __forceinline__ void project(const Photon<R>& p, R y[5]) {
    const R r = y[0], m = y[1];
    const R r2 = r * r;
    const R P = p.E * (r2 + p.a2) - p.aL;
    const R delta = (r2 - R(2) * r) + p.a2;
    const R PP = P * P;
    const R Rr = PP - delta * p.K;
    if (Rr > R(0.05) * (PP + fabs(delta) * p.K)) y[3] = copysign(sqrt(Rr), y[3]);
    const R m2 = m * m, s2 = R(1) - m2;
    const R M = (p.Q - p.C1 * m2) - p.A2E2 * (m2 * m2);
    const R scale = (fabs(p.Q) * s2 + p.A2E2 * (m2 * s2)) + p.L * p.L * m2;
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
};

struct Outputs {
    int8_t* cls;
    double* theta;
    double* phi;
    double* r_hit;
    double* g;
    int32_t* steps;
};

struct Result {
    int cls;
    double theta, phi, r_hit, g;
};

// Launch: per-pixel photon of a view (Bardeen plane or pinhole camera).
__device__ __forceinline__ void launch_photon(const ViewDev& v, long long local, Photon<double>& p, double y0[5]) {
    const double a = v.a;
    if (v.kind == 0) {  // Bardeen image plane, src/eht.cpp:262-271, 332-334
        double alpha, beta;
        if (v.params) {
            alpha = v.params[2 * local];
            beta = v.params[2 * local + 1];
        } else {
            const int row = (int)(local / v.width), col = (int)(local - (long long)row * v.width);
            alpha = -v.half_width + ((double)col + 0.5) * v.px;
            beta = v.half_height - ((double)row + 0.5) * v.py;
        }
        const double L = -alpha * v.sinI;
        const double Q = beta * beta + (alpha * alpha - a * a) * v.cosI * v.cosI;
        photon_init(p, a, 1.0, L, Q);
        y0[0] = v.r_cam;
        y0[1] = v.m0;
        y0[2] = v.phi_cam;
        y0[3] = -sqrt(fmax(pot_R(p, v.r_cam), 0.0));
        y0[4] = (beta > 0.0 ? 1.0 : -1.0) * sqrt(fmax(pot_M(p, v.m0), 0.0));
        return;
    }
    // pinhole camera: trace.frag:1494-1496 + launchPhoton (trace.frag:1470-1488)
    double n0, n1, n2;
    if (v.params) {
        n0 = v.params[3 * local];
        n1 = v.params[3 * local + 1];
        n2 = v.params[3 * local + 2];
    } else {
        const int row = (int)(local / v.width), col = (int)(local - (long long)row * v.width);
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
    double L = qp;
    const double cs2 = v.cs * v.cs, sn2 = fmax(v.sn * v.sn, 1e-18), a2 = a * a;
    double Q = qh * qh + cs2 * (L * L / sn2 - a2 * E * E);
    if (fabs(L) < kAxisSnap * sqrt(fmax(Q + a2 * E * E, 0.0))) {
        L = 0.0;
        Q = qh * qh - cs2 * a2 * E * E;
    }
    photon_init(p, a, E, L, Q);
    y0[0] = v.r_cam;
    y0[1] = v.cs;
    y0[2] = v.phi_cam;
    y0[3] = -(v.sigma * qr);
    y0[4] = v.sn * qh;
}

template <class R>
__device__ __forceinline__ void lane_start(Lane<R>& s, const ViewDev& v, const Photon<double>& p, const double y0[5]) {
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
#pragma unroll
    for (int i = 0; i < 5; ++i) s.y[i] = R(y0[i]);
    s.h = R(v.h0);
    s.u = R(0);
    s.r_cap = R(v.r_cap);
    s.r_esc = R(v.r_esc);
    s.r_in = R(v.r_in);
    s.r_out = R(v.r_out);
    s.steps = 0;
    s.rejected = false;
    s.tail = false;
    s.disk = v.disk != 0;
    s.prograde = v.prograde != 0;
    rhs_main(s.p, s.y, s.k1);
}

// One DP5 attempt plus events.  Returns true when the ray has ended (res filled).
template <class R>
__device__ __forceinline__ bool lane_step(Lane<R>& s, R rtol, R atol, int max_steps, Result& res) {
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
    const R err = dp5_attempt(s.p, s.tail, s.u, s.y, h, k, yn, rtol, atol);
    ++s.steps;
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
        for (int i = 0; i < 5; ++i) s.k1[i] = k[6][i];
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
    if (s.disk && s.y[1] * yn[1] < R(0)) {
        R lo = R(0), hi = R(1);
        const bool pos = s.y[1] > R(0);
        for (int it = 0; it < 48; ++it) {
            const R mid = R(0.5) * (lo + hi);
            const R f = dense(s.y, yn, k, h, 1, mid);
            if ((f > R(0)) == pos) lo = mid;
            else hi = mid;
        }
        const R rc = dense(s.y, yn, k, h, 0, R(0.5) * (lo + hi));
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
    if (yn[0] < s.r_cap) {
        res.cls = kCaptured;
        return true;
    }
    if (yn[0] > s.r_esc && yn[3] > R(0)) {
#pragma unroll
        for (int i = 0; i < 5; ++i) s.y[i] = yn[i];
        s.tail = true;
        s.u = R(1) / yn[0];
        s.h = -s.u;
        rhs_tail(s.p, s.u, s.y, s.k1);
        return false;
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) s.y[i] = yn[i];
#pragma unroll
    for (int i = 0; i < 5; ++i) s.k1[i] = k[6][i];
    s.k1[0] = s.y[3];  // FSAL after the projection: only dr/ds = w_r, dmu/ds = w_mu move
    s.k1[1] = s.y[4];
    s.h = h * fac;
    return false;
}

__device__ __forceinline__ void write_result(const Outputs& o, long long idx, const Result& r, int steps, int flags) {
    const double nan = __builtin_nan("");
    if (o.cls) o.cls[idx] = (int8_t)(r.cls | flags);
    if (o.theta) o.theta[idx] = r.cls == kEscaped ? r.theta : nan;
    if (o.phi) o.phi[idx] = r.cls == kEscaped ? r.phi : nan;
    if (o.r_hit) o.r_hit[idx] = r.cls == kDisk ? r.r_hit : nan;
    if (o.g) o.g[idx] = r.cls == kDisk ? r.g : nan;
    if (o.steps) o.steps[idx] = steps;
}

}  // namespace ktd
