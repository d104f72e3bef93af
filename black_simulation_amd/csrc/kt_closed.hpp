// kt_closed.hpp -- the photon's outcome from the closed-form solution of Carter's equations.
//
// Kerr null geodesics separate in Mino time s: with lambda = L/E and eta = Q/E^2 the radial
// and polar motions are independent quadratures,
//     ds = dr / sqrt(R(r)),   R(r) = r^4 + A r^2 + B r + C,
//          A = a^2 - eta - lambda^2,  B = 2 (eta + (lambda - a)^2),  C = -a^2 eta
//     ds = dmu / sqrt(M(mu)), M(mu) = a^2 (u+ - mu^2)(mu^2 - u-),  u+ u- = -eta / a^2,
// (the integrator's right-hand side, kt_device.hpp, divided by E^2), and both reduce to
// Legendre's elliptic integrals once the roots of R and of M are known (Gralla & Lupsasca,
// Phys. Rev. D 101, 044032 (2020), sections IV-V; Dexter & Agol, ApJ 696, 1616 (2009)).
// So the answers the integrator reaches after hundreds of DP5 steps follow from a few
// integrals:
//   * polar: with mu = sqrt(u+) sin(psi), F(psi | m) (m = u+/u- <= 0) grows by
//     omega = sqrt(-a^2 u-) per unit of s, so the equator (psi = j pi) is crossed at the
//     Mino times s_1, s_1 + 2K/omega, ... (K = K(m));
//   * radial: four real roots r1 <= r2 <= r3 <= r4 (with r4 > r_cap the ray turns at r4
//     and escapes) or two real and a complex pair (it falls in).  The Mino time from the
//     largest real root up to r is c F(x(r) | k) in closed form, and r at a given Mino time
//     is a rational function of sn or cn of it;
//   * the disk: the first crossing whose radius lies in [r_in, r_out] is the hit, with the
//     redshift of the Keplerian emitter there (src/trajectory.cpp:311-321);
//   * an escaping ray's sky: theta_inf from sn at the total Mino time, and phi_inf from the
//     integrals of dphi/ds = -(a (2r - a lambda) / Delta + lambda / (1 - mu^2)): elliptic
//     integrals of the third kind, one per horizon radius, and one in psi.
// Special functions: Carlson's symmetric integrals R_F, R_J, R_C by duplication (B. C.
// Carlson, Numer. Algorithms 10, 13 (1995)), K(m) by the arithmetic-geometric mean, and
// Jacobi's sn, cn, dn by descending Landen transformations (Abramowitz & Stegun 16.12).
// In FP64 (solve) the answers agree with the FP64 DP5 integrator at rtol 1e-10 to ~1e-9
// (its own error).  solve_t<float> runs the integrals and Jacobi's functions in FP32 and
// keeps the roots and every class decision in FP64: disk radius and g to ~1e-6, and the
// rays whose answer FP32 cannot settle (an escape's sky, a crossing near a disk edge) are
// marked for the FP64 form.  Rays the forms here do not cover (polar vortical orbits that
// escape, L = 0 escapes whose phi jumps at the pole, a turning point at the capture
// radius, a crossing within 1e-9 of the disk's edges, outgoing rays whose R has no real
// root, or a complex pair that the escape's quadrature cannot resolve) are declined and left to the integrator.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#define KTC_FN __host__ __device__ __forceinline__
// the iterated special functions are called, not inlined at each of their call sites: one
// copy each keeps the kernel's code and register count down
#define KTC_CALL __host__ __device__ __noinline__
#else
#define KTC_FN inline
#define KTC_CALL inline
#endif

// Inside this solver a * b + c may fuse into one FMA (one rounding instead of two): its
// answers are held to tolerances, not to bit patterns, and the fused form is the more
// accurate one.  The library's own setting, no fusing (the integrators' FP64 results are
// reproduced bit for bit on the CPU only without it), is set again at the end of the file.
#pragma clang fp contract(fast)

namespace ktc {

constexpr double kPiC = 3.14159265358979323846;

// ---- arithmetic ------------------------------------------------------------------------
// On the device: square root, reciprocal square root and reciprocal from the hardware
// estimates (relative error ~2^-23) refined by Newton-Raphson steps to about one ulp --
// a third of the instructions of the correctly rounded forms, which nothing here needs.
// The host uses the correctly rounded ones.
KTC_FN double qrsqrt(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);  // 1 - x y^2
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
#else
    return 1.0 / sqrt(x);
#endif
}
// x >= 0 (0 gives 1e-150, which no use here tells from 0)
KTC_FN double qsqrt(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    x = fmax(x, 1e-300);
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, h = 0.5 * y;
    const double r = fma(-s, h, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    const double d = fma(-s, s, x);
    return fma(d, h, s);
#else
    return sqrt(x);
#endif
}
KTC_FN double qrcp(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
#else
    return 1.0 / x;
#endif
}

// FP32: the hardware's square root, reciprocal square root and reciprocal (about one ulp)
KTC_FN float qrsqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}
KTC_FN float qsqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
KTC_FN float qrcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
// FP32 sin and cos: u - j pi/2 by two fused steps, pi/2 in two parts (its FP32 rounding and
// the remainder; each step rounds once, so x is good to 6e-8 for |j| < 2^20), then the
// Taylor series to the 11th and 12th order (truncation below 1e-11)
KTC_FN void qsincos(float u, float& s, float& c) {
    const float j = rintf(u * 0.636619772f);
    float x = fmaf(-j, 1.57079637f, u);
    x = fmaf(-j, -4.37113883e-8f, x);
    const float x2 = x * x;
    float ps = -1.0f / 39916800.0f;
    ps = fmaf(ps, x2, 1.0f / 362880.0f);
    ps = fmaf(ps, x2, -1.0f / 5040.0f);
    ps = fmaf(ps, x2, 1.0f / 120.0f);
    ps = fmaf(ps, x2, -1.0f / 6.0f);
    const float sx = fmaf(ps * x2, x, x);
    float pc = -1.0f / 479001600.0f;
    pc = fmaf(pc, x2, 1.0f / 3628800.0f);
    pc = fmaf(pc, x2, -1.0f / 40320.0f);
    pc = fmaf(pc, x2, 1.0f / 720.0f);
    pc = fmaf(pc, x2, -1.0f / 24.0f);
    pc = fmaf(pc, x2, 0.5f);
    const float cx = fmaf(-pc, x2, 1.0f);
    const int q = (int)j & 3;
    const float a = (q & 1) ? cx : sx, b = (q & 1) ? sx : cx;
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}

// sin and cos of |u| < 1e6: u - j pi/2 with pi/2 in three parts (Cody-Waite; the product
// of j and each part is exact for j < 2^20), then the Taylor series of both on
// [-pi/4, pi/4] to the 19th and 20th order (truncation below 1e-18)
KTC_FN void qsincos(double u, double& s, double& c) {
    constexpr double kTwoOverPi = 0.63661977236758134308;
    constexpr double kP1 = 1.5707963267341256;      // pi/2 to 33 bits
    constexpr double kP2 = 6.0771005065061922e-11;  // the next 33 bits
    constexpr double kP3 = 2.0222662487959506e-21;  // the rest
    const double j = rint(u * kTwoOverPi);
    const double x = ((u - j * kP1) - j * kP2) - j * kP3;
    const double x2 = x * x;
    // sin x = x (1 - x^2/3! + x^4/5! - ...), cos x = 1 - x^2/2! + x^4/4! - ...
    double ps = -1.0 / 121645100408832000.0;  // -1/19!
    ps = fma(ps, x2, 1.0 / 355687428096000.0);
    ps = fma(ps, x2, -1.0 / 1307674368000.0);
    ps = fma(ps, x2, 1.0 / 6227020800.0);
    ps = fma(ps, x2, -1.0 / 39916800.0);
    ps = fma(ps, x2, 1.0 / 362880.0);
    ps = fma(ps, x2, -1.0 / 5040.0);
    ps = fma(ps, x2, 1.0 / 120.0);
    ps = fma(ps, x2, -1.0 / 6.0);
    const double sx = fma(ps * x2, x, x);
    double pc = 1.0 / 2432902008176640000.0;  // 1/20!
    pc = fma(pc, x2, -1.0 / 6402373705728000.0);
    pc = fma(pc, x2, 1.0 / 20922789888000.0);
    pc = fma(pc, x2, -1.0 / 87178291200.0);
    pc = fma(pc, x2, 1.0 / 479001600.0);
    pc = fma(pc, x2, -1.0 / 3628800.0);
    pc = fma(pc, x2, 1.0 / 40320.0);
    pc = fma(pc, x2, -1.0 / 720.0);
    pc = fma(pc, x2, 1.0 / 24.0);
    pc = fma(pc, x2, -0.5);
    const double cx = fma(pc, x2, 1.0);
    // quadrant j mod 4: (s, c) = (sx, cx), (cx, -sx), (-sx, -cx), (-cx, sx)
    const long long q = (long long)j & 3;
    const double a = (q & 1) ? cx : sx, b = (q & 1) ? sx : cx;
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}

// ---- special functions -------------------------------------------------------------------

// R_F(x, y, z), x, y, z >= 0, at most one zero: duplication until the arguments agree to
// 3e-2 of their mean in FP64 (1e-1 in FP32), then the seventh-order series (DLMF 19.36.1;
// truncation ~1e-14, ~1e-8 in FP32)
template <class R>
struct Tol;
template <>
struct Tol<double> {
    static constexpr double dup = 3e-2, agm = 1e-9;
};
template <>
struct Tol<float> {
    static constexpr float dup = 1e-1f, agm = 1e-4f;
};

template <class R>
KTC_CALL R rf(R x, R y, R z) {
    R A = (x + y + z) * R(1.0 / 3.0);
    for (int it = 0; it < 24; ++it) {
        if (fmax(fabs(A - x), fmax(fabs(A - y), fabs(A - z))) < Tol<R>::dup * A) break;
        const R sx = qsqrt(x), sy = qsqrt(y), sz = qsqrt(z);
        const R lam = sx * (sy + sz) + sy * sz;
        x = R(0.25) * (x + lam);
        y = R(0.25) * (y + lam);
        z = R(0.25) * (z + lam);
        A = (x + y + z) * R(1.0 / 3.0);
    }
    const R inv = qrcp(A);
    const R X = (A - x) * inv, Y = (A - y) * inv, Z = -X - Y;
    const R E2 = X * Y - Z * Z, E3 = X * Y * Z;
    return (R(1) - E2 * R(1.0 / 10.0) + E3 * R(1.0 / 14.0) + E2 * E2 * R(1.0 / 24.0) - E2 * E3 * R(3.0 / 44.0) -
            E2 * E2 * E2 * R(5.0 / 208.0) + E3 * E3 * R(3.0 / 104.0) + E2 * E2 * E3 * R(1.0 / 16.0)) *
           qrsqrt(A);
}

// R_C(x, y), y > 0: the elementary closed forms, written so that x ~ y loses nothing
KTC_FN double rc(double x, double y) {
    if (x < y) {
        const double d = y - x;
        return x > 0.0 ? atan(sqrt(d / x)) / sqrt(d) : 0.5 * kPiC / sqrt(y);
    }
    if (x > y) {  // atanh(t) / sqrt(d), t = sqrt(d / x): near t = 1 as the logarithm
        const double d = x - y, t = sqrt(d / x);
        if (t < 0.5) return atanh(t) / sqrt(d);
        const double sd = sqrt(d);
        return log((sqrt(x) + sd) / sqrt(y)) / sd;
    }
    return 1.0 / sqrt(x);
}

// R_J(x, y, z, p), p > 0: duplication with the R_C terms, then the fifth-order series
constexpr double kDupTolJ = 6e-3;

KTC_CALL double rj(double x, double y, double z, double p) {
    double s = 0.0, fac = 1.0;
    double A = (x + y + z + 2.0 * p) * 0.2;
    for (int it = 0; it < 24; ++it) {
        if (fmax(fmax(fabs(A - x), fabs(A - y)), fmax(fabs(A - z), fabs(A - p))) < kDupTolJ * A) break;
        const double sx = qsqrt(x), sy = qsqrt(y), sz = qsqrt(z);
        const double lam = sx * (sy + sz) + sy * sz;
        double al = p * (sx + sy + sz) + sx * sy * sz;
        al *= al;
        const double pl = p + lam;
        const double be = p * pl * pl;
        s += fac * rc(al, be);
        fac *= 0.25;
        x = 0.25 * (x + lam);
        y = 0.25 * (y + lam);
        z = 0.25 * (z + lam);
        p = 0.25 * (p + lam);
        A = (x + y + z + 2.0 * p) * 0.2;
    }
    const double inv = qrcp(A);
    const double X = (A - x) * inv, Y = (A - y) * inv, Z = (A - z) * inv;
    const double P = -0.5 * (X + Y + Z);
    const double P2 = P * P, XYZ = X * Y * Z;
    const double E2 = X * Y + X * Z + Y * Z - 3.0 * P2;
    const double E3 = XYZ + 2.0 * E2 * P + 4.0 * P2 * P;
    const double E4 = (2.0 * XYZ + E2 * P + 3.0 * P2 * P) * P;
    const double E5 = XYZ * P2;
    const double ser = 1.0 - E2 * (3.0 / 14.0) + E3 * (1.0 / 6.0) + E2 * E2 * (9.0 / 88.0) - E4 * (3.0 / 22.0) -
                       E2 * E3 * (9.0 / 52.0) + E5 * (3.0 / 26.0);
    return 3.0 * s + fac * ser * inv * qrsqrt(A);
}

// Legendre's F(asin s | m) and Pi(n; asin s | m), |s| <= 1, m < 1, n s^2 < 1
template <class R>
KTC_FN R ell_f(R s, R m) {
    return s * rf<R>((R(1) - s) * (R(1) + s), R(1) - m * s * s, R(1));
}
// the same from s and its cofactors c2 = 1 - s^2, d2 = 1 - m s^2, which the caller forms
// from FP64 differences: near s = 1 (a turning point) the amplitude is ill-conditioned in s,
// not in c2
template <class R>
KTC_FN R ell_f2(R s, R c2, R d2) {
    return s * rf<R>(c2, d2, R(1));
}
KTC_FN double ell_pi(double n, double s, double m) {
    const double c2 = (1.0 - s) * (1.0 + s), d2 = 1.0 - m * s * s, s2 = s * s;
    return s * rf<double>(c2, d2, 1.0) + n * (1.0 / 3.0) * s * s2 * rj(c2, d2, 1.0, 1.0 - n * s2);
}
KTC_FN double ell_pi_complete(double n, double m) {
    return rf<double>(0.0, 1.0 - m, 1.0) + n * (1.0 / 3.0) * rj(0.0, 1.0 - m, 1.0, 1.0 - n);
}

// The arithmetic-geometric mean of (1, k'), k'^2 = 1 - m (m < 1), kept level by level:
//   a_{j+1} = (a_j + b_j) / 2,  b_{j+1} = sqrt(a_j b_j),  c_{j+1} = (a_j - b_j) / 2,
// until |a - b| / a < 1e-9 (at most seven levels from m = 1 - 1e-12; m < 0 is K's only).  K(m) = pi / (a + b)
// at the end.  Jacobi's functions at any argument u follow from it by descending Landen
// transformations (Abramowitz & Stegun 16.12): with k_{j+1} = c_{j+1} / a_{j+1}, the
// argument u a_N at the bottom level, whose parameter (1 - b/a)(1 + b/a) is below 2e-9,
// takes the first-order small-parameter form there, and each level up is
//   sn_j = (a + c) sn / (a + c sn^2),  cn_j = a cn dn / (a + c sn^2),
//   dn_j = (a - c sn^2) / (a + c sn^2)          (a, c of level j + 1).
// One modulus serves all the equatorial crossings of a ray.
template <class R>
struct Agm {
    static constexpr int kLevels = 7;
    R a[kLevels], c[kLevels];
    int n;
    R aN, bN;
};

// kp = sqrt(1 - m), given so that m near 1 keeps its digits
template <class R>
KTC_FN void agm_init(Agm<R>& G, R kp) {
    R a = R(1), b = kp;
    G.n = 0;
#pragma unroll
    for (int j = 0; j < Agm<R>::kLevels; ++j) {
        G.a[j] = R(1);
        G.c[j] = R(0);
        if (fabs(a - b) > Tol<R>::agm * a) {
            const R an = R(0.5) * (a + b), cn = R(0.5) * (a - b);
            b = qsqrt(a * b);
            a = an;
            G.a[j] = a;
            G.c[j] = cn;
            G.n = j + 1;
        }
    }
    G.aN = a;
    G.bN = b;
}

template <class R>
KTC_FN R agm_k(const Agm<R>& G) {
    return R(kPiC) * qrcp(G.aN + G.bN);
}

template <class R>
KTC_FN void agm_sncndn(const Agm<R>& G, R u, R& sn, R& cn, R& dn) {
    u *= G.aN;
    const R kp = G.bN * qrcp(G.aN);
    const R mn = (R(1) - kp) * (R(1) + kp);
    R s, c;
    qsincos(u, s, c);
    const R d = R(0.25) * mn * (u - s * c);  // sn(u|m) = sin u - (m/4)(u - sin u cos u) cos u + O(m^2)
    const R s1 = s - d * c, c1 = c + d * s;
    s = s1;
    c = c1;
    R dd = qsqrt(R(1) - mn * s * s);
#pragma unroll
    for (int j = Agm<R>::kLevels - 1; j >= 0; --j) {
        if (j < G.n) {
            const R a = G.a[j], cj = G.c[j];
            const R cs2 = cj * s * s;
            const R inv = qrcp(a + cs2);
            const R ns = (a + cj) * s * inv, nc = a * c * dd * inv, nd = (a - cs2) * inv;
            s = ns;
            c = nc;
            dd = nd;
        }
    }
    sn = s;
    cn = c;
    dn = dd;
}

// K(m), m < 1
template <class R>
KTC_CALL R ell_k(R m) {
    Agm<R> G;
    agm_init<R>(G, qsqrt(R(1) - m));
    return agm_k(G);
}

// Jacobi sn, cn, dn (u | m), 0 <= m < 1, for a single argument
KTC_CALL void sncndn(double u, double m, double& sn, double& cn, double& dn) {
    Agm<double> G;
    agm_init<double>(G, qsqrt(1.0 - m));
    agm_sncndn(G, u, sn, cn, dn);
}

// the largest real root of the resolvent cubic xi^3 + A xi^2 + (A^2/4 - C) xi - B^2/8 of
// the depressed quartic r^4 + A r^2 + B r + C (positive when B != 0): Cardano's or the
// trigonometric form, then Newton steps.  Their cube root, arccosine and cosine run
// in FP32 (an estimate good to ~1e-6, which the Newton steps make exact) unless the
// result is not positive or leaves a residual, as where the cubic's roots crowd together
// (B ~ 0); then in FP64.
KTC_FN double resolvent_from(double c2, double c1, double c0, bool fp64) {
    const double p = c1 - c2 * c2 * (1.0 / 3.0);
    const double q = 2.0 * c2 * c2 * c2 * (1.0 / 27.0) - c2 * c1 * (1.0 / 3.0) + c0;
    const double hq = 0.5 * q, tp = p * (1.0 / 3.0);
    const double disc = hq * hq + tp * tp * tp;
    double t;
    if (fp64 || !(fabs(disc) < 1e30 && fabs(tp) < 1e12)) {
        if (disc > 0.0) {
            const double sd = qsqrt(disc);
            const double w = hq < 0.0 ? -hq + sd : -hq - sd;
            const double u = cbrt(w);
            t = u != 0.0 ? u - tp * qrcp(u) : 0.0;
        } else {
            const double r = qsqrt(-tp);
            const double ca = fmin(1.0, fmax(-1.0, -hq * qrcp(r * r * r)));
            t = 2.0 * r * cos(acos(ca) * (1.0 / 3.0));
        }
    } else {  // the estimate in FP32 once the branch is chosen (w is formed without cancellation)
        if (disc > 0.0) {
            const double sd = (double)qsqrt((float)disc);
            const float u = cbrtf((float)(hq < 0.0 ? -hq + sd : -hq - sd));
            t = u != 0.0f ? (double)(u - (float)tp * qrcp(u)) : 0.0;
        } else {
            const float r = qsqrt((float)-tp);
            const float ca = fminf(1.0f, fmaxf(-1.0f, (float)-hq * qrcp(r * r * r)));
            t = (double)(2.0f * r * cosf(acosf(ca) * (1.0f / 3.0f)));
        }
    }
    double xi = t - c2 * (1.0 / 3.0);
    // The step's correction needs only a few digits: an FP32 reciprocal of the derivative
    // (relative error ~1e-7) keeps the convergence quadratic down to 1e-7 of the error.  The
    // FP32 estimate is good to ~1e-7 (1e-6 at p99.9), so two steps end at FP64 rounding; a
    // worse one (where the roots crowd) leaves the residual that sends the root to the FP64
    // attempt, which divides in FP64 and takes three steps.  On the orbit frames, the a =
    // 0.998 single frame and the 64 batch views, 5 rays in 433,000 take it.
    for (int it = 0; it < (fp64 ? 3 : 2); ++it) {
        const double f = ((xi + c2) * xi + c1) * xi + c0;
        const double df = (3.0 * xi + 2.0 * c2) * xi + c1;
        if (fp64) {
            if (df != 0.0) xi -= f * qrcp(df);
        } else {
            const float d = (float)df;
            if (d != 0.0f) xi -= f * (double)qrcp(d);
        }
    }
    return xi;
}

KTC_CALL double resolvent_root(double A, double B, double C) {
    const double c2 = A, c1 = 0.25 * A * A - C, c0 = -0.125 * B * B;
    double xi = resolvent_from(c2, c1, c0, false);
    const double f = ((xi + c2) * xi + c1) * xi + c0;
    const double scale = fabs(xi * xi * xi) + fabs(c2 * xi * xi) + fabs(c1 * xi) + fabs(c0);
    if (!(xi > 0.0) || !(fabs(f) <= 1e-12 * scale)) xi = resolvent_from(c2, c1, c0, true);
    return xi;
}

// ---- one ray --------------------------------------------------------------------------------

// one ray in E = 1 units: spin a, camera (r_o, mu_o, BL phi_o), lambda = L/E, eta = Q/E^2,
// the signs of dr/ds and dmu/ds at the camera (s the backward Mino time), the disk and the
// capture radius; E only scales g
struct Ray {
    double a, r_o, mu_o, phi_o, lam, eta, E;
    int s_r, s_mu;
    double r_in, r_out, r_cap, r_esc;
    int disk, prograde;
};

enum : int { kDeclined = -1 };

struct Outcome {
    int cls;  // ktd::kEscaped / kCaptured / kDisk, or kDeclined
    double theta, phi, r_hit, g;
};

// the Keplerian emitter's redshift at r (src/trajectory.cpp:311-321): 1 with g, 0 where no
// timelike circular orbit exists, -1 (FP32 only) where the FP32 pass cannot settle g (the ray
// is then solved again in FP64).  FP64 evaluates the metric form as the integrator does.  Its
// sum cancels near the ISCO of a fast spin (a hundredfold at a = 0.998) and at the photon
// orbit, so FP32 uses the same quantity in Bardeen, Press and Teukolsky's form,
// -g(u/u^t, u/u^t) = x^3 d / (x^3 + s a)^2, x = sqrt(r), with its one cancelling factor
// d = x^3 - 3 x + 2 s a (zero at the photon orbit) formed in FP64.  Near that orbit g
// varies steeply with r: a relative error e in the crossing radius becomes kappa e in g,
// kappa = r |d'(r)| / (2 d).  The FP32 radius is good to ~1e-5, so where kappa > 10 -- or
// where the sign of d, whether the orbit exists at all, could flip -- the ray goes to FP64.
// (kappa stays below 6 down to the ISCO of a = 0.998; it grows without bound where an inner
// edge set below the ISCO brings the disk to the photon orbit.)
template <class R>
KTC_FN int disk_g(double a_, double r_, int prograde, double E, double L, double& g) {
    if constexpr (sizeof(R) == 4) {
        const double sa = prograde ? a_ : -a_, x64 = qsqrt(r_);
        const double d = x64 * (r_ - 3.0) + 2.0 * sa;
        if (!(fabs(d) > 0.075 * x64 * fabs(r_ - 1.0))) return -1;  // kappa > 10: r d'(r) = 1.5 x (r - 1)
        if (!(d > 0.0)) return 0;
        const float x = (float)x64, x3 = x * x * x, q = x3 + (float)sa;
        const float iq = qrcp(q);
        const float norm = (float)d * x3 * iq * iq;
        const float om = (prograde ? 1.0f : -1.0f) * iq;
        const float Lom = (float)L * om, den = (float)E - Lom;
        if (!(fabsf(den) > 1e-2f * (fabsf((float)E) + fabsf(Lom)))) return -1;
        g = (double)(qsqrt(norm) * qrcp(den));
        return 1;
    } else {
        const R a = R(a_), r = R(r_);
        const R sg = prograde ? R(1) : R(-1);
        const R om = sg * qrcp(r * qsqrt(r) + sg * a);
        const R ir = qrcp(r);
        const R gtt = -(R(1) - R(2) * ir), gtp = R(-2) * a * ir;
        const R gpp = r * r + a * a + R(2) * a * a * ir;
        const R norm = -(gtt + R(2) * gtp * om + gpp * om * om);
        if (!(norm > R(0))) return 0;
        g = (double)(qsqrt(norm) * qrcp(R(E) - R(L) * om));
        return 1;
    }
}

KTC_FN double wrap2pi(double x) {
    const double t = 2.0 * kPiC;
    double y = fmod(x, t);
    if (y < 0.0) y += t;
    return y;
}

// a crossing radius within this relative distance of a disk edge is left to the
// integrator (FP64, whose crossing radius is good to ~1e-10) or, from the FP32 pass, to the
// FP64 one
constexpr double kEdgeBand = 1e-9, kEdgeBand32 = 1e-5;
constexpr int kMaxCrossings = 24;
enum : int { kUncertain = -2 };  // the FP32 pass leaves the ray to the FP64 one

// The polar motion: omega, u+, m, x = mu_o / sqrt(u+) (FP64), F(asin x | m), K(m), and the
// Mino time of the first equatorial crossing (eta > 0), the last three in precision R.
template <class R>
struct Polar {
    double om, up, m, x;
    R Fo, K, s1;
};

// u+ and the cofactor 1 - x^2 = (u+ - mu_o^2) / u+ need FP64 (mu_o^2 may lie close to u+);
// omega, m and x feed only precision-R integrals and are formed in R.
template <class R>
KTC_FN Polar<R> polar_motion(const Ray& ry, double A) {
    Polar<R> P;
    const double a2 = ry.a * ry.a, eta = ry.eta;
    const double disc = qsqrt(A * A + 4.0 * a2 * eta);
    const double om2 = A <= 0.0 ? 0.5 * (disc - A) : 2.0 * a2 * eta * qrcp(A + disc);  // a^2 (-u-)
    const double iom2 = qrcp(om2);
    P.up = eta * iom2;
    P.m = -a2 * eta * iom2 * iom2;
    const R up = R(P.up), iup = qrcp(up);
    P.om = double(qsqrt(R(om2)));
    P.x = fmin(1.0, fmax(-1.0, double(R(ry.mu_o) * qrsqrt(up))));
    const R c2 = R(fmax(0.0, P.up - ry.mu_o * ry.mu_o)) * iup;  // 1 - x^2
    const R xr = R(P.x), mr = R(P.m);
    P.Fo = ell_f2<R>(xr, c2, R(1) - mr * xr * xr);
    P.K = ell_k<R>(mr);
    // F at the next psi = j pi in the direction of motion; starting on the plane is no crossing
    const double x = P.x;
    R F1;
    if (ry.s_mu > 0) F1 = (x > 0.0 ? R(2) * P.K : R(0)) - P.Fo;
    else F1 = P.Fo - (x > 0.0 ? R(0) : R(-2) * P.K);
    if (x == 0.0) F1 = R(2) * P.K;
    P.s1 = F1 * qrcp(R(P.om));
    return P;
}

// Bernstein parameter rho of the point t = re + i im for quadrature over t in [0, 1]: the
// ellipse with foci 0 and 1 through t has semi-axis sum rho / 2 (after [0, 1] -> [-1, 1]).
KTC_CALL double bernstein_rho(double re, double im) {
    const double x = 2.0 * re - 1.0, y = 2.0 * im;
    const double s = 0.5 * (sqrt((x - 1.0) * (x - 1.0) + y * y) + sqrt((x + 1.0) * (x + 1.0) + y * y));
    return s + sqrt(fmax(0.0, s * s - 1.0));
}

// An outward start past a complex pair of roots (r1 < r2 below the camera, z +- i b): the
// radial part of the escape's phi, the integral of a (2r - a lambda) / (Delta sqrt R) dr from
// the camera to infinity, by 32-point Gauss-Legendre quadrature in t = r_o / r.  With u = 1 / r
// the integrand is a u (2 - a lambda u) / (r_o D S), D = 1 - 2u + a^2 u^2 and
// S = sqrt((1 - r1 u)(1 - r2 u)((1 - z u)^2 + b^2 u^2)): analytic on [0, 1] but for t = r_o / r
// at r = r+, r-, r1, r2 and z +- i b.  Its error falls as rho^-64 with rho that of the nearest
// of these; below rho = 1.6 it returns false and the ray goes to the integrator.
KTC_CALL bool radial_phi_quad(const Ray& ry, double r1, double r2, double z, double b, double& Ir) {
    constexpr double kX[16] = {0.04830766568773831, 0.1444719615827965, 0.23928736225213706, 0.33186860228212767, 0.42135127613063533, 0.5068999089322294, 0.5877157572407623, 0.6630442669302152, 0.7321821187402897, 0.7944837959679424, 0.84936761373257, 0.8963211557660522, 0.9349060759377397, 0.9647622555875064, 0.9856115115452684, 0.9972638618494816};
    constexpr double kW[16] = {0.09654008851472781, 0.09563872007927483, 0.09384439908080457, 0.09117387869576386, 0.08765209300440391, 0.08331192422694685, 0.07819389578707031, 0.07234579410884845, 0.06582222277636175, 0.058684093478535704, 0.050998059262376244, 0.042835898022226426, 0.034273862913021626, 0.025392065309262427, 0.016274394730905965, 0.007018610009469298};
    const double a = ry.a, lam = ry.lam, ro = ry.r_o;
    const double sq1 = sqrt(fmax(0.0, 1.0 - a * a));
    const double pts[4] = {1.0 + sq1, 1.0 - sq1, r1, r2};
    double rho = bernstein_rho(ro * z / (z * z + b * b), ro * b / (z * z + b * b));
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (pts[i] != 0.0) rho = fmin(rho, bernstein_rho(ro / pts[i], 0.0));
    if (!(rho >= 1.6)) return false;
    const double iro = 1.0 / ro;
    double sum = 0.0;
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        const double t = 0.5 + 0.5 * (i < 16 ? kX[i] : -kX[i - 16]);
        const double u = t * iro;
        const double D = 1.0 - 2.0 * u + a * a * u * u;
        const double zu = 1.0 - z * u;
        const double S = sqrt((1.0 - r1 * u) * (1.0 - r2 * u) * (zu * zu + b * b * u * u));
        sum += kW[i & 15] * a * u * (2.0 - a * lam * u) / (D * S);
    }
    Ir = 0.5 * iro * sum;
    return Ir == Ir;
}

// An escaping ray that the polar forms cover (eta > 0, lambda != 0): the sky angles at
// infinity from the total Mino time S_tot.  With four real roots (turning at r4, or outward
// from beyond it) x_o is x(r_o) of the radial substitution and x_inf its limit; otherwise
// (real4 false: outward past the complex pair zc +- i bc, r1 < r3 the real roots) the radial
// part is the quadrature above.  Returns false if a value is not finite or not resolved.
KTC_CALL bool escape_sky(const Ray& ry, const Polar<double>& P, double S_tot, bool real4, double r1, double r3,
                         double r4, double zc, double bc, double k, double pref, double x_o, double x_inf, double& theta, double& phi) {
    const double a = ry.a, a2 = a * a, lam = ry.lam;
    const double r31 = r3 - r1, r41 = r4 - r1;
    // theta_inf from sn at the total Mino time (m <= 0:
    // sn(u | -M) = sd(u sqrt(1 + M) | M / (1 + M)) / sqrt(1 + M))
    const double Phi = P.Fo + (double)ry.s_mu * P.om * S_tot;  // F(psi_inf | m), psi unwrapped
    const double M = -P.m, sq = sqrt(1.0 + M);
    double sn, cn, dn;
    sncndn(Phi * sq, M / (1.0 + M), sn, cn, dn);
    const double sn_m = sn / (dn * sq);  // sn(Phi | m)
    const double mu_inf = sqrt(P.up) * sn_m;
    theta = acos(fmin(1.0, fmax(-1.0, mu_inf)));
    // the unwrapped amplitude: psi_inf in [j pi - pi/2, j pi + pi/2]
    const double jj = floor((Phi + P.K) / (2.0 * P.K));
    const double rem = fmod(jj, 2.0) == 0.0 ? sn_m : -sn_m;
    const double P_inf = 2.0 * jj * ell_pi_complete(P.up, P.m) + ell_pi(P.up, rem, P.m);
    const double P_o = ell_pi(P.up, P.x, P.m);
    const double G_phi = (double)ry.s_mu * (P_inf - P_o) / P.om;  // integral of ds / (1 - mu^2)
    // the radial part, integral of a (2r - a lambda) / Delta ds, by partial fractions over
    // the horizons: a (2 rp - a lambda) / (rp - rq) / (r - rp) for rp = r+, r-; over the
    // path, Mino time from r4 to r of 1 / (r - p) is
    //   pref [F / (r3 - p) + (1 / (r4 - p) - 1 / (r3 - p)) Pi(n_p; x | k)],
    //   n_p = (r3 - p) r41 / ((r4 - p) r31)
    double Ir = 0.0;
    if (a != 0.0 && !real4) {
        if (!radial_phi_quad(ry, r1, r3, zc, bc, Ir)) return false;
    } else if (a != 0.0) {
        const double sq1 = sqrt(fmax(0.0, 1.0 - a2));
        if (sq1 == 0.0) return false;  // |a| = 1: the partial fractions degenerate
        const double Fo = ell_f<double>(x_o, k), Fi = ell_f<double>(x_inf, k);
        for (int h = 0; h < 2; ++h) {
            const double p = h == 0 ? 1.0 + sq1 : 1.0 - sq1, q = h == 0 ? 1.0 - sq1 : 1.0 + sq1;
            const double Ap = a * (2.0 * p - a * lam) / (p - q);
            const double np = (r3 - p) * r41 / ((r4 - p) * r31);
            const double ca = 1.0 / (r3 - p), cb = 1.0 / (r4 - p) - ca;
            const double Jo = ca * Fo + cb * ell_pi(np, x_o, k);
            const double Ji = ca * Fi + cb * ell_pi(np, x_inf, k);
            // inward start: camera to r4 and r4 to infinity; outward: camera to infinity
            Ir += Ap * pref * (ry.s_r < 0 ? Jo + Ji : Ji - Jo);
        }
    }
    phi = wrap2pi(ry.phi_o - (Ir + lam * G_phi));
    return theta == theta && phi == phi;
}

// The outcome of one ray with its elliptic integrals and Jacobi functions in precision R
// (the roots of R(r) and the class decisions they make are FP64 either way).  R = float
// answers disk hits and captures to ~1e-6 and returns kUncertain where FP32 could place a
// crossing on the wrong side of a disk edge or where an escape's sky is needed; those rays
// take R = double.
template <class R>
KTC_FN Outcome solve_t(const Ray& ry) {
    constexpr bool fp32 = sizeof(R) == 4;
    const int unsure = fp32 ? (int)kUncertain : (int)kDeclined;
    const double band = fp32 ? kEdgeBand32 : kEdgeBand;
    Outcome out{kDeclined, 0.0, 0.0, 0.0, 0.0};
    const double a = ry.a, a2 = a * a, lam = ry.lam, eta = ry.eta;
    const double A = a2 - eta - lam * lam;
    const double lma = lam - a;
    const double B = 2.0 * (eta + lma * lma);
    const double C = -a2 * eta;
    if (!(B > 1e-12) || !(ry.E > 0.0) || !(ry.r_out < ry.r_esc)) return out;
    const double xi = resolvent_root(A, B, C);
    if (!(xi > 0.0)) return out;
    const double h = 0.5 * xi, rs = qrsqrt(h);  // z = sqrt(xi / 2) and 1 / (4 z) from one estimate
    const double z = h * rs;
    const double Bz = 0.25 * B * rs;
    const double D12 = -0.5 * A - z * z + Bz;
    const double D34 = -0.5 * A - z * z - Bz;
    const bool cross = eta > 0.0 && ry.disk;  // eta <= 0: the ray never reaches the equator

    // ---- radial motion.  J, the Mino time from the largest real root rt of R (r4 of four
    // real roots, r2 of two real and a complex pair) up to r, is pref F(phi(r) | k), and r at
    // J is the Moebius form (P0 + P1 w) / (Q0 + Q1 w) in w = sn^2(J / pref | k) (four real
    // roots: sin^2 phi = r31 (r - r4) / (r41 (r - r3))) or w = cn(J / pref | k) (a complex
    // pair: cos phi = (Aq (r - r1) - Bq (r - r2)) / (Aq (r - r1) + Bq (r - r2)),
    // Aq = |r3 - r2|, Bq = |r3 - r1|).  S_t: J at the camera, signed so that the ray's own
    // Mino time s reaches rt at s = S_t (S_t < 0: rt lies behind an outward start).  A camera
    // between r2 and r3 measures J from r2 instead (below).
    // The roots, the Moebius coefficients and every comparison with a radius are FP64; the
    // modulus k (k' = sqrt(1 - k)), the prefactor and the amplitudes with their cofactors
    // feed only precision-R integrals and are formed in R from FP64 root differences.
    bool real4, turning = false;
    double P0, P1, Q0, Q1, r1 = 0.0, r3 = 0.0, r4 = 0.0, zc = 0.0, bc = 0.0;
    R pref, k, kp, x_o = R(0), x_inf = R(0), c2_o = R(1), c2_inf = R(1);
    R S_t;
    R I_inf = R(-1);  // Mino time from the radial origin to infinity, once needed
    Agm<R> G;  // the radial modulus's mean (for the crossings)
    if (D34 >= 0.0) {
        real4 = true;
        const double sD12 = qsqrt(fmax(D12, 0.0)), sD34 = qsqrt(D34);
        r1 = -z - sD12;
        const double r2 = -z + sD12;
        r3 = z - sD34;
        r4 = z + sD34;
        const double r31 = r3 - r1, r41 = r4 - r1, r42 = r4 - r2, r32 = r3 - r2, r43 = r4 - r3;
        const R i3142 = qrcp(R(r31 * r42));
        k = R(r32 * r41) * i3142;
        kp = qsqrt(R((r2 - r1) * r43) * i3142);  // sqrt(1 - k)
        pref = R(2) * qrsqrt(R(r31 * r42));
        // a camera at the turning point r4 itself (a tangential ray, R(r_cam) ~ 0) may find r4
        // a rounding above r_cam: it starts at r4
        const double ro = ry.r_o < r4 && ry.r_o > r4 * (1.0 - 1e-12) ? r4 : ry.r_o;
        if (!(ro >= r4)) {
            // The camera between r2 and r3 (a close camera's wide rays; R < 0 between r3 and
            // r4, where no ray starts): with the same k and prefactor, the Mino time from r2
            // up to r is pref F(phi | k), sin^2 phi = r31 (r - r2) / (r32 (r - r1)), and r at
            // it is (r2 r31 - r1 r32 w) / (r31 - r32 w), w = sn^2.  The ray falls to r2, below
            // the capture radius (no photon is bound between r2 and r3), after turning at r3
            // when it starts outward: it is captured unless a crossing lands on the disk.
            const double r21 = r2 - r1;
            if (!(ry.r_o < r3) || !(r2 < ry.r_cap)) return out;
            if (!cross) {
                out.cls = 1;
                return out;
            }
            P0 = r2 * r31;
            P1 = -r1 * r32;
            Q0 = r31;
            Q1 = -r32;
            const R io = qrcp(R(r32 * (ry.r_o - r1)));
            x_o = qsqrt(R(r31 * (ry.r_o - r2)) * io);
            c2_o = R(r21 * (r3 - ry.r_o)) * io;  // 1 - x_o^2
            const R I_o = pref * ell_f2<R>(x_o, c2_o, kp * kp + k * c2_o);
            agm_init<R>(G, kp);
            S_t = ry.s_r < 0 ? I_o : R(2) * pref * agm_k(G) - I_o;  // the Mino time at which r = r2
        } else {
            // beyond every root: inward, the ray turns at r4 and escapes if r4 lies above the
            // capture radius and falls in otherwise; outward, it escapes
            const bool outward = ry.s_r > 0;
            turning = outward || r4 > ry.r_cap;
            if (!outward && (r4 == ry.r_cap || (r4 > ry.r_cap && r4 < ry.r_cap * (1.0 + 1e-7)))) return out;
            P0 = r31 * r4;
            P1 = -r41 * r3;
            Q0 = r31;
            Q1 = -r41;
            const R i41 = qrcp(R(r41)), io3 = qrcp(R(r41 * (ro - r3)));
            x_o = qsqrt(R(r31 * (ro - r4)) * io3);
            x_inf = qsqrt(R(r31) * i41);
            c2_o = R(r43 * (ro - r1)) * io3;  // 1 - x_o^2
            c2_inf = R(r43) * i41;                // 1 - x_inf^2
            const R I_o = pref * ell_f2<R>(x_o, c2_o, kp * kp + k * c2_o);
            S_t = outward ? -I_o : I_o;
            if (cross) agm_init<R>(G, kp);
        }
    } else {
        const bool outward = ry.s_r > 0;
        if (!(D12 >= 0.0)) {
            // all four roots complex: no turning point, and (eta < 0) no equatorial crossing
            if (cross || outward) return out;
            out.cls = 1;
            return out;
        }
        real4 = false;
        const double sD12 = qsqrt(D12);
        r1 = -z - sD12;
        const double r2 = -z + sD12;
        // inward the ray falls to r2, which must lie below the capture radius; outward (the
        // camera above r2, where R > 0) it runs out to infinity without turning
        if (!outward && !(r2 < ry.r_cap)) return out;
        if (!cross && !outward) {
            out.cls = 1;
            return out;
        }
        r3 = r2;
        zc = z;
        bc = qsqrt(-D34);
        const double d2 = z - r2, d1 = z - r1;
        const double Aq = qsqrt(d2 * d2 - D34), Bq = qsqrt(d1 * d1 - D34);  // |r3 - r2|, |r3 - r1|
        const double r21 = r2 - r1;
        const double i4AB = qrcp(4.0 * Aq * Bq);
        k = R(((Aq + Bq) * (Aq + Bq) - r21 * r21) * i4AB);
        kp = R(qsqrt(fmax(0.0, (r21 * r21 - (Aq - Bq) * (Aq - Bq)) * i4AB)));  // sqrt(1 - k)
        pref = qrsqrt(R(Aq * Bq));
        P0 = Aq * r1 - Bq * r2;
        P1 = -(Aq * r1 + Bq * r2);
        Q0 = Aq - Bq;
        Q1 = -(Aq + Bq);
        const double ro = ry.r_o;
        const double c = (Aq * (ro - r1) - Bq * (ro - r2)) * qrcp(Aq * (ro - r1) + Bq * (ro - r2));
        const R cc = R(c * c);
        const R f = ell_f2<R>(qsqrt(R(fmax(0.0, (1.0 - c) * (1.0 + c)))), cc, kp * kp + k * cc);
        agm_init<R>(G, kp);
        S_t = pref * (c >= 0.0 ? f : R(2) * agm_k(G) - f);  // F(acos c) past pi/2
        if (outward) {
            // J runs from I_o at the camera to I_inf, where c reaches (Aq - Bq) / (Aq + Bq)
            const double ci = (Aq - Bq) * qrcp(Aq + Bq);
            const R ci2 = R(ci * ci);
            const R fi = ell_f2<R>(qsqrt(R(fmax(0.0, (1.0 - ci) * (1.0 + ci)))), ci2, kp * kp + k * ci2);
            I_inf = pref * (ci >= 0.0 ? fi : R(2) * agm_k(G) - fi);
            S_t = -S_t;
            turning = true;
        }
    }

    // ---- polar motion, and the disk at its equatorial crossings: inward of rt first; for a
    // turning ray past rt until the escape, whose time is needed only then or for the sky
    Polar<R> P{};
    if (eta > 0.0) P = polar_motion<R>(ry, A);
    if (cross) {
        const R ds = R(2) * P.K * qrcp(R(P.om));
        const R ipref = qrcp(pref);
        R s = P.s1, s_end = I_inf >= R(0) ? S_t + I_inf : S_t;
        int j = 0;
        for (; j < kMaxCrossings; ++j) {
            if (!(s < s_end)) {
                if (!turning || I_inf >= R(0)) break;
                I_inf = pref * ell_f2<R>(x_inf, c2_inf, kp * kp + k * c2_inf);
                s_end = S_t + I_inf;
                if (!(s < s_end)) break;
            }
            const R J = s < S_t ? S_t - s : s - S_t;
            R sn, cn, dn;
            agm_sncndn(G, J * ipref, sn, cn, dn);
            const double w = real4 ? (double)(sn * sn) : (double)cn;
            const double rc = (P0 + P1 * w) * qrcp(Q0 + Q1 * w);
            if (!(rc == rc)) return out;
            if (rc < ry.r_cap) break;  // inward past the capture radius: later crossings lie deeper
            if (fabs(rc - ry.r_in) < band * ry.r_in || fabs(rc - ry.r_out) < band * ry.r_out) {
                out.cls = unsure;
                return out;
            }
            double g;
            const int dg = rc >= ry.r_in && rc <= ry.r_out ? disk_g<R>(a, rc, ry.prograde, ry.E, lam * ry.E, g) : 0;
            if (dg < 0) {
                out.cls = unsure;
                return out;
            }
            if (dg > 0) {
                out.cls = 2;
                out.r_hit = rc;
                out.g = g;
                return out;
            }
            s += ds;
        }
        if (j == kMaxCrossings) return out;
    }
    if (!turning) {
        out.cls = 1;
        return out;
    }
    if (!(eta > 0.0) || lam == 0.0) return out;  // vortical or through the pole: the integrator
    if constexpr (fp32) {
        out.cls = kUncertain;  // the sky angles in FP64
        return out;
    } else {
        double th, ph;
        if (I_inf < 0.0) I_inf = pref * ell_f2<double>(x_inf, c2_inf, kp * kp + k * c2_inf);
        if (!escape_sky(ry, P, S_t + I_inf, real4, r1, r3, r4, zc, bc, k, pref, x_o, x_inf, th, ph)) return out;
        out.cls = 0;
        out.theta = th;
        out.phi = ph;
        return out;
    }
}

KTC_FN Outcome solve(const Ray& ry) { return solve_t<double>(ry); }

}  // namespace ktc

#pragma clang fp contract(off)
