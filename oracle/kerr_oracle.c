/*
 * kerr_oracle.c -- FP64 CPU oracle for the Kerr per-pixel photon trace.
 *
 * TEST INFRASTRUCTURE ONLY (see kerr_oracle.h): the checker for the HIP
 * library, never part of it.  Build with -ffp-contract=off (oracle/Makefile).
 */
#include "kerr_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* radii (src/kerr.cpp:157-223) and circular orbits (src/kerr.cpp:296-299)   */
/* ------------------------------------------------------------------------- */

static int naked(double a) { return fabs(a) > 1.0; }

/* sin and cos each behind its own call: compilers otherwise fuse sin(x), cos(x) into
   sincos(x), whose glibc result can differ from sin(x) in the last bit. */
static __attribute__((noinline)) double ko_sin(double x) { return sin(x); }
static __attribute__((noinline)) double ko_cos(double x) { return cos(x); }

double ko_horizon_radius(double a) {
    if (naked(a)) return 0.0;
    return 1.0 + sqrt(fmax(0.0, 1.0 - a * a));
}

double ko_inner_horizon_radius(double a) {
    if (naked(a)) return 0.0;
    return 1.0 - sqrt(fmax(0.0, 1.0 - a * a));
}

/* Past a = 1 (naked, kerr.cpp:179-206): roots by bisection to adjacent doubles.  The ISCO is
   the root of the marginal-stability condition x^4 - 6 x^2 + 8 s a x - 3 a^2 = 0, x = sqrt(r)
   (dE/dr = 0 of the circular orbits, which the reference finds by a golden-section search
   on E): in (0, 1) prograde, where f rises through zero once, and in (sqrt 3, 6) retrograde,
   where f falls to its least value and then rises through zero once.  The retrograde
   photon orbit solves r (r - 3)^2 = 4 a^2 in (3, 3 + 2|a|); the prograde one is 0. */
static double isco_f(double x, double sa, double a2) { return ((x * x - 6.0) * x + 8.0 * sa) * x - 3.0 * a2; }
static double orbit_f(double r, double sa, double a2) { (void)sa; return r * (r - 3.0) * (r - 3.0) - 4.0 * a2; }

static double bisect_root(double (*f)(double, double, double), double sa, double a2, double lo, double hi) {
    for (int k = 0; k < 200; ++k) {
        const double mid = 0.5 * (lo + hi);
        if (!(mid > lo && mid < hi)) break;
        if (f(mid, sa, a2) < 0.0) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

/* Bardeen, Press & Teukolsky (1972) closed form, kerr.cpp:179-190 */
double ko_isco_radius(double a, int prograde) {
    if (naked(a)) {
        const double x = prograde ? bisect_root(isco_f, a, a * a, 0.0, 1.0)
                                  : bisect_root(isco_f, -a, a * a, sqrt(3.0), 6.0);
        return x * x;
    }
    double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
    double z2 = sqrt(3.0 * a * a + z1 * z1);
    double root = sqrt(fmax(0.0, (3.0 - z1) * (3.0 + z1 + 2.0 * z2)));
    return prograde ? 3.0 + z2 - root : 3.0 + z2 + root;
}

/* kerr.cpp:192-206 */
double ko_photon_orbit_radius(double a, int prograde) {
    if (naked(a)) return prograde ? 0.0 : bisect_root(orbit_f, 0.0, a * a, 3.0, 3.0 + 2.0 * fabs(a));
    double s = prograde ? -1.0 : 1.0;
    return 2.0 * (1.0 + ko_cos(2.0 / 3.0 * acos(s * a)));
}

/* kerr.cpp:208-223 (|a| <= 1) */
double ko_marginally_bound_radius(double a, int prograde) {
    double s = prograde ? 1.0 : -1.0;
    return 2.0 - s * a + 2.0 * sqrt(1.0 - s * a);
}

/* kerr.cpp:296-299 */
double ko_circular_omega(double a, double r, int prograde) {
    double s = prograde ? 1.0 : -1.0;
    return s / (r * sqrt(r) + s * a);
}

/* ------------------------------------------------------------------------- */
/* reference-form equations (src/kerr.cpp:323-363, 521-542)                   */
/* ------------------------------------------------------------------------- */

/* kerr.cpp:323-329 */
double ko_inflow_factor(double P, double vr, double delta, double K) {
    if (P * vr < 0.0) return K / (P - vr);
    double d = fabs(delta) > 1e-12 ? delta : copysign(1e-12, delta);
    return (P + vr) / d;
}

static double safe_sin_cubed(double s) {
    double m = fmax(fabs(s), 1e-9);
    return (s < 0.0 ? -1.0 : 1.0) * m * m * m;
}

/* kerr.cpp:331-351: Kerr-Schild t, phi; forward Mino time */
ko_derivs ko_derivatives(double a, ko_consts c, ko_state s) {
    const double r = s.r;
    const double sn = ko_sin(s.theta), cs = ko_cos(s.theta);
    const double sn2 = fmax(sn * sn, 1e-18);
    const double a2 = a * a, mu2 = c.mu * c.mu;
    const double delta = r * r - 2.0 * r + a2;
    const double rra = r * r + a2;
    const double P = c.E * rra - a * c.L;
    const double x = c.L - a * c.E;
    const double K = mu2 * r * r + x * x + c.Q;
    const double inflow = ko_inflow_factor(P, s.vr, delta, K);
    ko_derivs d;
    d.dr = s.vr;
    d.dtheta = s.vtheta;
    d.dvr = 2.0 * c.E * r * P - (r - 1.0) * K - delta * mu2 * r;
    d.dvtheta = sn * cs * a2 * (mu2 - c.E * c.E) + c.L * c.L * cs / safe_sin_cubed(sn);
    d.dphi = a * inflow - a * c.E + c.L / sn2;
    d.dt = rra * inflow - a * (a * c.E * sn * sn - c.L) - s.vr;
    return d;
}

/* kerr.cpp:353-358 */
double ko_radial_potential(double a, ko_consts c, double r) {
    const double delta = r * r - 2.0 * r + a * a;
    const double P = c.E * (r * r + a * a) - a * c.L;
    const double x = c.L - a * c.E;
    return P * P - delta * (c.mu * c.mu * r * r + x * x + c.Q);
}

/* kerr.cpp:360-363 */
double ko_polar_potential(double a, ko_consts c, double theta) {
    const double sn = ko_sin(theta), cs = ko_cos(theta);
    return c.Q - cs * cs * (a * a * (c.mu * c.mu - c.E * c.E) + c.L * c.L / fmax(sn * sn, 1e-18));
}

/* kerr.cpp:521-542 (kTurningZone = 0.05) */
void ko_enforce_constraints(double a, ko_consts c, ko_state* s) {
    const double zone = 0.05;
    const double r = s->r;
    const double delta = r * r - 2.0 * r + a * a;
    const double P = c.E * (r * r + a * a) - a * c.L;
    const double x = c.L - a * c.E;
    const double K = c.mu * c.mu * r * r + x * x + c.Q;
    const double R = P * P - delta * K;
    if (R > zone * (P * P + fabs(delta) * K)) s->vr = copysign(sqrt(R), s->vr);
    const double sn = ko_sin(s->theta), cs = ko_cos(s->theta);
    const double cot2 = cs * cs / fmax(sn * sn, 1e-18);
    const double polarMass = a * a * (c.mu * c.mu - c.E * c.E) * cs * cs;
    const double Theta = c.Q - polarMass - c.L * c.L * cot2;
    if (Theta > zone * (fabs(c.Q) + fabs(polarMass) + c.L * c.L * cot2))
        s->vtheta = copysign(sqrt(Theta), s->vtheta);
}

/* ------------------------------------------------------------------------- */
/* metric, frames and photon launch (src/kerr.cpp:238-270, 424-483)          */
/* ------------------------------------------------------------------------- */

/* covariant Kerr-Schild metric, kerr.cpp:238-254 */
void ko_metric_tensor(double a, double r, double theta, double g[4][4]) {
    const double s = ko_sin(theta), c = ko_cos(theta);
    const double s2 = s * s;
    const double sigma = r * r + a * a * c * c;
    const double rra = r * r + a * a;
    const double A = rra * rra - a * a * (r * r - 2.0 * r + a * a) * s2;
    const double f = 2.0 * r / sigma;
    memset(g, 0, 16 * sizeof(double));
    g[0][0] = -(1.0 - f);
    g[0][1] = g[1][0] = f;
    g[0][3] = g[3][0] = -f * a * s2;
    g[1][1] = 1.0 + f;
    g[1][3] = g[3][1] = -a * s2 * (1.0 + f);
    g[2][2] = sigma;
    g[3][3] = A * s2 / sigma;
}

static double dot4(const double g[4][4], const double x[4], const double y[4]) {
    double sum = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) sum += g[i][j] * x[i] * y[j];
    return sum;
}

static void lower4(const double g[4][4], const double x[4], double out[4]) {
    for (int i = 0; i < 4; ++i) {
        out[i] = 0.0;
        for (int j = 0; j < 4; ++j) out[i] += g[i][j] * x[j];
    }
}

/* kerr.cpp:424-430 */
void ko_zamo_four_velocity(double a, double r, double theta, double u[4]) {
    const double s = ko_sin(theta), c = ko_cos(theta);
    const double sigma = r * r + a * a * c * c;
    const double delta = r * r - 2.0 * r + a * a;
    const double rra = r * r + a * a;
    const double A = rra * rra - a * a * delta * s * s;
    const double omega = 2.0 * a * r / A;
    const double ut = sqrt(A / (sigma * delta));
    u[0] = ut;
    u[1] = 0.0;
    u[2] = 0.0;
    u[3] = omega * ut;
}

/* Gram-Schmidt frame, kerr.cpp:432-456 */
void ko_observer_frame(double a, double r, double theta, const double u[4], double e[4][4]) {
    double g[4][4];
    ko_metric_tensor(a, r, theta, g);
    static const double cand[3][4] = {{0, 0, 1, 0}, {0, 0, 0, 1}, {0, 1, 0, 0}};
    static const int slot[3] = {2, 3, 1};
    for (int i = 0; i < 4; ++i) e[0][i] = u[i];
    for (int k = 0; k < 3; ++k) {
        double w[4];
        for (int i = 0; i < 4; ++i) w[i] = cand[k][i];
        const double along = dot4(g, w, u);
        for (int i = 0; i < 4; ++i) w[i] += along * u[i];
        for (int j = 0; j < k; ++j) {
            const double* ej = e[slot[j]];
            const double p = dot4(g, w, ej);
            for (int i = 0; i < 4; ++i) w[i] -= p * ej[i];
        }
        const double n = sqrt(fmax(dot4(g, w, w), 1e-300));
        for (int i = 0; i < 4; ++i) w[i] /= n;
        for (int i = 0; i < 4; ++i) e[slot[k]][i] = w[i];
    }
}

#define KO_AXIS_SNAP 3e-4 /* kerr.h:115 */

/* kerr.cpp:458-483 (forward photon, Kerr-Schild state) */
void ko_photon_from_frame(double a, double r, double theta, double phi, const double e[4][4],
                          const double look[3], ko_consts* c, ko_state* s) {
    double p[4], low[4], g[4][4];
    for (int i = 0; i < 4; ++i) p[i] = e[0][i] - look[0] * e[1][i] - look[1] * e[2][i] - look[2] * e[3][i];
    ko_metric_tensor(a, r, theta, g);
    lower4(g, p, low);
    const double sn = ko_sin(theta), cs = ko_cos(theta);
    const double sigma = r * r + a * a * cs * cs;
    c->mu = 0.0;
    c->E = -low[0];
    c->L = low[3];
    s->t = 0.0;
    s->r = r;
    s->theta = theta;
    s->phi = phi;
    s->vr = sigma * p[1];
    s->vtheta = low[2];
    c->Q = s->vtheta * s->vtheta + cs * cs * (c->L * c->L / fmax(sn * sn, 1e-18) - a * a * c->E * c->E);
    if (fabs(c->L) < KO_AXIS_SNAP * sqrt(fmax(c->Q + a * a * c->E * c->E, 0.0))) {
        c->L = 0.0;
        c->Q = s->vtheta * s->vtheta - cs * cs * a * a * c->E * c->E;
    }
}

/* camera.cpp:6-21 in FP64 */
void ko_camera_basis(double yaw, double pitch, double roll, double B[9]) {
    const double f0[3] = {-1.0, 0.0, 0.0}, u0[3] = {0.0, -1.0, 0.0}, r0[3] = {0.0, 0.0, 1.0};
    const double cy = ko_cos(yaw), sy = ko_sin(yaw), cp = ko_cos(pitch), sp = ko_sin(pitch);
    const double cr = ko_cos(roll), sr = ko_sin(roll);
    double yawed[3], right[3], fwd[3], up[3];
    for (int i = 0; i < 3; ++i) {
        yawed[i] = cy * f0[i] + sy * r0[i];
        right[i] = cy * r0[i] - sy * f0[i];
    }
    for (int i = 0; i < 3; ++i) fwd[i] = cp * yawed[i] + sp * u0[i];
    up[0] = right[1] * fwd[2] - right[2] * fwd[1];
    up[1] = right[2] * fwd[0] - right[0] * fwd[2];
    up[2] = right[0] * fwd[1] - right[1] * fwd[0];
    for (int i = 0; i < 3; ++i) {
        B[0 + i] = cr * right[i] + sr * up[i];
        B[3 + i] = cr * up[i] - sr * right[i];
        B[6 + i] = fwd[i];
    }
}

/* ------------------------------------------------------------------------- */
/* the integrated system                                                      */
/* ------------------------------------------------------------------------- */

void ko_photon_init(ko_photon* p, double a, double E, double L, double Q) {
    p->a = a;
    p->a2 = a * a;
    p->E = E;
    p->L = L;
    p->Q = Q;
    p->aE = a * E;
    p->aL = a * L;
    const double x = L - p->aE;
    p->K = x * x + Q;
    p->A2E2 = p->a2 * (E * E);
    p->C1 = (Q + L * L) - p->A2E2;
    p->twoE = 2.0 * E;
    p->twoA2E2 = 2.0 * p->A2E2;
    p->ks = 0;
}

/* P(r) = E (r^2 + a^2) - a L and Delta(r) = r^2 - 2r + a^2 */
static double P_of(const ko_photon* p, double r) { return fma(p->E, fma(r, r, p->a2), -p->aL); }
static double delta_of(const ko_photon* p, double r) { return fma(r, r - 2.0, p->a2); }

double ko_R(const ko_photon* p, double r) {
    const double P = P_of(p, r);
    return fma(P, P, -(delta_of(p, r) * p->K));
}

double ko_M(const ko_photon* p, double m) {
    const double m2 = m * m;
    return fma(-p->A2E2, m2 * m2, fma(-p->C1, m2, p->Q));
}

/* Carter's equations in Mino time, backward (s = -tau):
 *   dr/ds = w_r,  dw_r/ds = R'(r)/2 = 2 E r P - (r - 1) K          (kerr.cpp:346, mu = 0)
 *   dmu/ds = w_mu, dw_mu/ds = M'(mu)/2 = -(Q + L^2 - a^2 E^2) mu - 2 a^2 E^2 mu^3
 *   dphi/ds = -(a P / Delta - a E + L / (1 - mu^2))                  (BL; kerr.cpp:348 less a vr/Delta)
 * M(mu) = sin^2(theta) Theta(theta); on the constraint surface w_mu = -ko_sin(theta) w_theta. */
#define KO_TINY 1e-300

/* dphi/ds = aE - [a P (1 - mu^2) + L Delta] / [Delta (1 - mu^2)]: one division.  The
   floor on 1 - mu^2 keeps L = 0 rays finite at the pole (the ratio -> aP/Delta). */
void ko_rhs(const ko_photon* p, const double y[5], double dy[5]) {
    const double r = y[0], m = y[1];
    const double delta = delta_of(p, r);
    const double P = P_of(p, r);
    const double s2 = fmax(fma(-m, m, 1.0), KO_TINY);
    /* Kerr-Schild phi (ks): dphi_KS = dphi_BL + a dr / Delta, so P / Delta becomes
       (P - w_r) / Delta = inflowFactor of kerr.cpp:323-329 in backward time: K / (P + w_r)
       when P w_r > 0 (by w_r^2 = P^2 - Delta K), else with Delta floored at 1e-12 */
    double nd = P, dd = delta;
    const double w = y[3];
    if (p->ks == 1) {
        if (P * w > 0.0) {
            nd = p->K;
            dd = P + w;
        } else {
            nd = P - w;
            dd = fabs(delta) > 1e-12 ? delta : copysign(1e-12, delta);
        }
    } else if (p->ks == 2) {
        /* outgoing Kerr-Schild phi (dphi_out = dphi_BL - a dr / Delta), regular on the past
           horizon the ray approaches when followed back into the hole from outside:
           (P + w_r) / Delta = K / (P - w_r) when P w_r < 0 */
        if (P * w < 0.0) {
            nd = p->K;
            dd = P - w;
        } else {
            nd = P + w;
            dd = fabs(delta) > 1e-12 ? delta : copysign(1e-12, delta);
        }
    }
    const double num = fma(p->a * nd, s2, p->L * dd);
    dy[0] = y[3];
    dy[1] = y[4];
    dy[2] = p->aE - num / (dd * s2);
    dy[3] = fma(p->twoE * r, P, -((r - 1.0) * p->K));
    dy[4] = -(m * fma(p->twoA2E2, m * m, p->C1));
}

/* tail system in u = 1/r for z = (mu, phi, w_mu), outgoing (w_r = +sqrt(R)):
 *   ds/du = -1/sqrt(U), U(u) = u^4 R(1/u) = Pu^2 - u^2 Du K,
 *   Pu = E (1 + a^2 u^2) - a L u^2, Du = 1 - 2u + a^2 u^2 */
static void tail_rhs(const ko_photon* p, double u, const double z[3], double dz[3]) {
    const double u2 = u * u;
    const double Pu = fma(p->E, fma(p->a2, u2, 1.0), -(p->aL * u2));
    const double Du = fma(p->a2, u2, fma(-2.0, u, 1.0));
    const double U = fma(Pu, Pu, -(u2 * Du * p->K));
    const double isq = 1.0 / sqrt(fmax(U, KO_TINY));
    const double m = z[0];
    const double s2 = fmax(fma(-m, m, 1.0), KO_TINY);
    const double num = fma(p->a * Pu, s2, p->L * Du);
    dz[0] = -(z[2] * isq);
    dz[1] = (num / (Du * s2) - p->aE) * isq;
    if (p->ks == 2) dz[1] = dz[1] + p->a / Du; /* dphi_out/du = dphi_BL/du + a / Du */
    dz[2] = (m * fma(p->twoA2E2, m * m, p->C1)) * isq;
}

/* ------------------------------------------------------------------------- */
/* Dormand-Prince 5(4) (Hairer, Norsett & Wanner I, Table II.5.2)             */
/* ------------------------------------------------------------------------- */

#define A21 (1.0 / 5.0)
#define A31 (3.0 / 40.0)
#define A32 (9.0 / 40.0)
#define A41 (44.0 / 45.0)
#define A42 (-56.0 / 15.0)
#define A43 (32.0 / 9.0)
#define A51 (19372.0 / 6561.0)
#define A52 (-25360.0 / 2187.0)
#define A53 (64448.0 / 6561.0)
#define A54 (-212.0 / 729.0)
#define A61 (9017.0 / 3168.0)
#define A62 (-355.0 / 33.0)
#define A63 (46732.0 / 5247.0)
#define A64 (49.0 / 176.0)
#define A65 (-5103.0 / 18656.0)
#define A71 (35.0 / 384.0)
#define A73 (500.0 / 1113.0)
#define A74 (125.0 / 192.0)
#define A75 (-2187.0 / 6784.0)
#define A76 (11.0 / 84.0)
#define E1 (71.0 / 57600.0)
#define E3 (-71.0 / 16695.0)
#define E4 (71.0 / 1920.0)
#define E5 (-17253.0 / 339200.0)
#define E6 (22.0 / 525.0)
#define E7 (-1.0 / 40.0)
/* dense output (Hairer's DOPRI5 continuous extension) */
#define D1 (-12715105075.0 / 11282082432.0)
#define D3 (87487479700.0 / 32700410799.0)
#define D4 (-10690763975.0 / 1880347072.0)
#define D5 (701980252875.0 / 199316789632.0)
#define D6 (-1453857185.0 / 822651844.0)
#define D7 (69997945.0 / 29380423.0)
#define C2 (1.0 / 5.0)
#define C3 (3.0 / 10.0)
#define C4 (4.0 / 5.0)
#define C5 (8.0 / 9.0)

#define SAFETY 0.9
#define FACMIN 0.2
#define FACMAX 5.0

/* x^(-1/5) for x > 0 by exponent split + three Newton steps on y^-5 = x.
   Only +, *, frexp, ldexp: the same bits on any IEEE machine. */
double ko_pow_m15(double x) {
    static const double T[5] = {1.0, 0.8705505632961241, 0.757858283255199, 0.6597539553864471,
                                0.5743491774985174};
    int e;
    if (!(x > 1e-200 && x < 1e200)) return x > 1.0 ? 0.0 : 1e20; /* inf/NaN/0: clamped by the caller */
    const double m = frexp(x, &e); /* x = m 2^e, m in [0.5, 1) */
    const int q = (e + 1100) / 5 - 220;
    const int s = e - 5 * q;
    double y = fma(-0.2974, m - 0.5, 1.1487);
    y = ldexp(y * T[s], -q);
    for (int it = 0; it < 3; ++it) {
        const double y2 = y * y;
        const double y5 = (y2 * y2) * y;
        y = (y * fma(-x, y5, 6.0)) * 0.2;
    }
    return y;
}

static double step_factor(double err, int rejected_before) {
    double fac;
    if (err != err) return FACMIN;
    if (err == 0.0) fac = FACMAX;
    else fac = SAFETY * ko_pow_m15(err);
    if (fac > FACMAX) fac = FACMAX;
    if (fac < FACMIN) fac = FACMIN;
    if (rejected_before && fac > 1.0) fac = 1.0;
    return fac;
}

static double amax(double x, double y) { return x > y ? x : y; }

/* Stage sums as FMA chains: t = y + h (a_1 k_1 + ... + a_n k_n). */
#define SUM2(c1, k1, c2, k2) fma(c2, k2, (c1) * (k1))
#define SUM3(c1, k1, c2, k2, c3, k3) fma(c3, k3, SUM2(c1, k1, c2, k2))
#define SUM4(c1, k1, c2, k2, c3, k3, c4, k4) fma(c4, k4, SUM3(c1, k1, c2, k2, c3, k3))
#define SUM5(c1, k1, c2, k2, c3, k3, c4, k4, c5, k5) fma(c5, k5, SUM4(c1, k1, c2, k2, c3, k3, c4, k4))
#define SUM6(c1, k1, c2, k2, c3, k3, c4, k4, c5, k5, c6, k6) fma(c6, k6, SUM5(c1, k1, c2, k2, c3, k3, c4, k4, c5, k5))

/* The max-norm error picks the component with the largest |e_i| / sc_i by comparing
   cross products, then divides once.  NaN anywhere in e makes the error NaN. */
typedef struct { double be, bs, guard; } errpick;
static void pick(errpick* ep, double e, double sc) {
    const double ae = fabs(e);
    ep->guard += ae;
    if (ep->be < 0.0 || ae * ep->bs > ep->be * sc) {
        ep->be = ae;
        ep->bs = sc;
    }
}
static double picked(const errpick* ep) { return ep->guard != ep->guard ? NAN : ep->be / ep->bs; }

/* the five stage inputs and the solution; k[0] = f(y) on entry, k[1..6] filled */
#define DP5_STAGES(N, EVAL)                                                                                  \
    for (i = 0; i < N; ++i) t[i] = fma(h, A21 * k[0][i], y[i]);                                              \
    EVAL(C2, t, k[1]);                                                                                      \
    for (i = 0; i < N; ++i) t[i] = fma(h, SUM2(A31, k[0][i], A32, k[1][i]), y[i]);                          \
    EVAL(C3, t, k[2]);                                                                                      \
    for (i = 0; i < N; ++i) t[i] = fma(h, SUM3(A41, k[0][i], A42, k[1][i], A43, k[2][i]), y[i]);            \
    EVAL(C4, t, k[3]);                                                                                      \
    for (i = 0; i < N; ++i)                                                                                 \
        t[i] = fma(h, SUM4(A51, k[0][i], A52, k[1][i], A53, k[2][i], A54, k[3][i]), y[i]);                  \
    EVAL(C5, t, k[4]);                                                                                      \
    for (i = 0; i < N; ++i)                                                                                 \
        t[i] = fma(h, SUM5(A61, k[0][i], A62, k[1][i], A63, k[2][i], A64, k[3][i], A65, k[4][i]), y[i]);    \
    EVAL(1.0, t, k[5]);                                                                                     \
    for (i = 0; i < N; ++i)                                                                                 \
        ynew[i] = fma(h, SUM5(A71, k[0][i], A73, k[2][i], A74, k[3][i], A75, k[4][i], A76, k[5][i]), y[i]); \
    EVAL(1.0, ynew, k[6]);

#define ERR_OF(i) (h * SUM6(E1, k[0][i], E3, k[2][i], E4, k[3][i], E5, k[4][i], E6, k[5][i], E7, k[6][i]))

static double dp5_main(const ko_photon* p, const ko_opts* o, const double y[5], double h, double k[7][5],
                       double ynew[5]) {
    double t[5];
    int i;
#define EVAL_MAIN(c, yy, kk) ko_rhs(p, yy, kk)
    DP5_STAGES(5, EVAL_MAIN)
#undef EVAL_MAIN
    errpick ep = {-1.0, 1.0, 0.0};
    for (i = 0; i < 5; ++i) {
        const double big = amax(fabs(y[i]), fabs(ynew[i]));
        double sc;
        if (i == 0) sc = o->rtol * big;
        else if (i == 3) sc = o->rtol * amax(big, y[0] * y[0]);
        else sc = fma(o->rtol, big, o->atol);
        pick(&ep, ERR_OF(i), sc);
    }
    return picked(&ep);
}

static double dp5_tail(const ko_photon* p, const ko_opts* o, double u, const double y[3], double h,
                       double k[7][3], double ynew[3]) {
    double t[3];
    int i;
#define EVAL_TAIL(c, yy, kk) tail_rhs(p, u + (c) * h, yy, kk)
    DP5_STAGES(3, EVAL_TAIL)
#undef EVAL_TAIL
    errpick ep = {-1.0, 1.0, 0.0};
    for (i = 0; i < 3; ++i) pick(&ep, ERR_OF(i), fma(o->rtol, amax(fabs(y[i]), fabs(ynew[i])), o->atol));
    return picked(&ep);
}

/* dense output of component i at fraction th of the step */
typedef struct { double y, ydiff, bspl, c4, c5; } ko_dense;

static ko_dense dense_coeffs(const double* y, const double* ynew, double (*k)[5], double h, int i) {
    ko_dense d;
    d.y = y[i];
    d.ydiff = ynew[i] - y[i];
    d.bspl = fma(h, k[0][i], -d.ydiff);
    d.c4 = fma(-h, k[6][i], d.ydiff - d.bspl);
    d.c5 = h * SUM6(D1, k[0][i], D3, k[2][i], D4, k[3][i], D5, k[4][i], D6, k[5][i], D7, k[6][i]);
    return d;
}

static double dense_at(const ko_dense* d, double th) {
    const double th1 = 1.0 - th;
    return fma(th, fma(th1, fma(th, fma(th1, d->c5, d->c4), d->bspl), d->ydiff), d->y);
}

/* Constraint projection, kerr.cpp:521-542 (kTurningZone = 0.05) in (r, mu) form: away
   from a turning point put w_r back on w_r^2 = R(r) and w_mu on w_mu^2 = M(mu), where
   M = sin^2(theta) Theta and the polar zone test is the reference's multiplied by
   sin^2(theta).  The second-order equations conserve w_r^2 - R including its error,
   and R ~ r^4: without this, a rounding-level error at r = 1000 M is O(1) near the hole. */
static void project(const ko_photon* p, double y[5]) {
    const double r = y[0], m = y[1];
    const double P = P_of(p, r);
    const double delta = delta_of(p, r);
    const double R = fma(P, P, -(delta * p->K));
    if (R > 0.05 * fma(P, P, fabs(delta) * p->K)) y[3] = copysign(sqrt(R), y[3]);
    const double m2 = m * m, s2 = 1.0 - m2;
    const double M = fma(-p->A2E2, m2 * m2, fma(-p->C1, m2, p->Q));
    const double scale = fma(p->L * p->L, m2, fma(fabs(p->Q), s2, p->A2E2 * (m2 * s2)));
    if (M > 0.05 * scale) y[4] = copysign(sqrt(M), y[4]);
}

static double wrap_2pi(double x) {
    const double twopi = 2.0 * KO_PI;
    double w = x - twopi * floor(x / twopi);
    if (w >= twopi) w -= twopi;
    if (w < 0.0) w += twopi;
    return w;
}

/* ------------------------------------------------------------------------- */
/* massive particles (particles_update.comp:207-255; kerr.cpp:323-351, 365-406, 544-561) */
/* ------------------------------------------------------------------------- */

/* a particle's constants; mu = 1 everywhere: A2E2 = a^2 (E^2 - 1) */
typedef struct {
    double a, a2, E, L, Q, aE, aL, K0, C1, A2E2, twoA2E2;
    int pole; /* L = 0 with no polar turning point short of the pole: phi + pi per pass */
} ko_body;

static void body_init(ko_body* b, double a, double E, double L, double Q) {
    /* within KO_AXIS_SNAP radians of the axis: through it (L = 0), as photons (kerr.cpp:399-405) */
    if (fabs(L) < KO_AXIS_SNAP * sqrt(fmax(Q + a * a * (E * E - 1.0), 0.0))) L = 0.0;
    b->a = a;
    b->a2 = a * a;
    b->E = E;
    b->L = L;
    b->Q = Q;
    b->aE = a * E;
    b->aL = a * L;
    const double x = L - b->aE;
    b->K0 = x * x + Q;
    b->A2E2 = b->a2 * (E * E - 1.0);
    b->C1 = (Q + L * L) - b->A2E2;
    b->twoA2E2 = 2.0 * b->A2E2;
    b->pole = L == 0.0 && Q + b->A2E2 > 0.0;
}

/* dy/dt for y = (r, mu, phi_KS, vr, vmu), t the ingoing Kerr-Schild time: the reference's
   Kerr-Schild derivatives (kerr.cpp:331-351, mu = 1) in mu form over dt/dlambda, with one
   division: F = nf / df (kerr.cpp:323-329) and 1 / sin^2 folded into
   g = 1 / (sin^2 (rra nf - df X)), X = a (a E sin^2 - L) + vr */
static void particle_rhs(const ko_body* b, const double y[5], double dy[5]) {
    const double r = y[0], m = y[1], w = y[3];
    const double delta = fma(r, r - 2.0, b->a2);
    const double rra = fma(r, r, b->a2);
    const double P = fma(b->E, rra, -b->aL);
    const double K = fma(r, r, b->K0);
    const double s2t = fma(-m, m, 1.0);
    const double s2 = fmax(s2t, 1e-200); /* keeps df s2 a normal number at the pole */
    double nf, df;
    if (P * w < 0.0) {
        nf = K;
        df = P - w;
    } else {
        nf = P + w;
        df = fabs(delta) > 1e-12 ? delta : copysign(1e-12, delta);
    }
    const double X = fma(b->a, fma(b->aE, s2t, -b->L), w);
    const double den = fma(rra, nf, -(df * X));
    const double g = 1.0 / (s2 * den);
    const double ts = (df * s2) * g;  /* 1 / (dt/dlambda) */
    const double phin = fma(b->a * nf, s2, fma(b->L, df, -((b->aE * df) * s2)));
    const double accr = fma(2.0 * b->E * r, P, -fma(r - 1.0, K, delta * r));
    const double accm = -(m * fma(b->twoA2E2, m * m, b->C1));
    dy[0] = w * ts;
    dy[1] = y[4] * ts;
    dy[2] = phin * g;
    dy[3] = accr * ts;
    dy[4] = accm * ts;
}

/* particle_common.glsl enforceMassiveConstraints / kerr.cpp:521-542 in mu form */
static void particle_project(const ko_body* b, double y[5]) {
    const double r = y[0], m = y[1];
    const double delta = fma(r, r - 2.0, b->a2);
    const double P = fma(b->E, fma(r, r, b->a2), -b->aL);
    const double K = fma(r, r, b->K0);
    const double R = fma(P, P, -(delta * K));
    if (R > 0.05 * fma(P, P, fabs(delta) * K)) y[3] = copysign(sqrt(R), y[3]);
    const double m2 = m * m, s2 = 1.0 - m2;
    const double M = fma(-b->A2E2, m2 * m2, fma(-b->C1, m2, b->Q));
    const double scale = fma(b->L * b->L, m2, fma(fabs(b->Q), s2, fabs(b->A2E2) * (m2 * s2)));
    if (M > 0.05 * scale) y[4] = copysign(sqrt(M), y[4]);
}

static double dp5_particle(const ko_body* b, const ko_opts* o, const double y[5], double h, double k[7][5],
                           double ynew[5]) {
    double t[5];
    int i;
#define EVAL_PARTICLE(c, yy, kk) particle_rhs(b, yy, kk)
    DP5_STAGES(5, EVAL_PARTICLE)
#undef EVAL_PARTICLE
    errpick ep = {-1.0, 1.0, 0.0};
    for (i = 0; i < 5; ++i) {
        const double big = amax(fabs(y[i]), fabs(ynew[i]));
        double sc;
        if (i == 0) sc = o->rtol * big;
        else if (i == 3) sc = o->rtol * amax(big, y[0] * y[0]);
        else sc = fma(o->rtol, big, o->atol);
        pick(&ep, ERR_OF(i), sc);
    }
    return picked(&ep);
}

int ko_particle_advance(double a, double E, double L, double Q, double st[5], double dt, double r_stop,
                        double r_escape, const ko_opts* o, int* steps_out, double* h_io) {
    ko_body b;
    body_init(&b, a, E, L, Q);
    double y[5], k[7][5], yn[5];
    for (int i = 0; i < 5; ++i) y[i] = st[i];
    int fate = KO_PARTICLE_MOVING, steps = 0, rejected = 0;
    double left = dt;
    double h = h_io && *h_io > 0.0 ? fmin(dt, *h_io) : fmin(dt, 0.01 * y[0]);
    double hwant = h; /* the controller's own step before the span's end shortens it */
    particle_rhs(&b, y, k[0]);
    while (left > 0.0) {
        if (steps >= o->max_steps) {
            fate = KO_PARTICLE_UNFINISHED;
            break;
        }
        int last = 0;
        hwant = h;
        if (h >= left) {
            h = left;
            last = 1;
        }
        const double err = dp5_particle(&b, o, y, h, k, yn);
        ++steps;
        if (!(err <= 1.0)) {
            h *= step_factor(err, 1);
            rejected = 1;
            continue;
        }
        const double fac = step_factor(err, rejected);
        rejected = 0;
        if (b.pole && y[1] * y[4] > 0.0 && yn[1] * yn[4] < 0.0) yn[2] += KO_PI;
        particle_project(&b, yn);
        for (int i = 0; i < 5; ++i) y[i] = yn[i];
        if (!(fabs(y[0]) < 1e300) || !(fabs(y[1]) < 1e300) || y[0] < r_stop) {
            fate = KO_PARTICLE_PLUNGED;
            break;
        }
        if (y[0] > r_escape) {
            fate = KO_PARTICLE_ESCAPED;
            break;
        }
        left = last ? 0.0 : left - h;
        h *= fac;
        particle_rhs(&b, y, k[0]);
    }
    for (int i = 0; i < 5; ++i) st[i] = y[i];
    if (fabs(st[2]) < 1e300) st[2] = wrap_2pi(st[2]);
    if (steps_out) *steps_out = steps;
    if (h_io) *h_io = hwant;
    return fate;
}

/* kerr::launchFromZamo with mu = 1 (kerr.cpp:365-406); out = {E, L, Q, vr, vmu} */
void ko_particle_from_zamo(double a, double r, double theta, const double v[3], double out[5]) {
    const double sn = ko_sin(theta), cs = ko_cos(theta), a2 = a * a;
    const double sigma = r * r + a2 * cs * cs, delta = r * r - 2.0 * r + a2;
    const double A = (r * r + a2) * (r * r + a2) - a2 * delta * sn * sn;
    const double omega = 2.0 * a * r / A;
    double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (v2 > 1.0 - 1e-12) v2 = 1.0 - 1e-12;
    const double gamma = 1.0 / sqrt(1.0 - v2);
    const double L = gamma * v[2] * sqrt(A / sigma) * sn;
    const double E = gamma * sqrt(sigma * delta / A) + omega * L;
    const double vr = gamma * v[0] * sqrt(sigma * delta);
    const double vth = gamma * v[1] * sqrt(sigma);
    const double Q = vth * vth + cs * cs * (a2 * (1.0 - E * E) + L * L / fmax(sn * sn, 1e-18));
    out[0] = E;
    out[1] = L;
    out[2] = Q;
    out[3] = vr;
    out[4] = -sn * vth;
}

/* ------------------------------------------------------------------------- */
/* launch                                                                     */
/* ------------------------------------------------------------------------- */

/* eht.cpp:262-271: E = 1, L = -alpha sin i, Q = beta^2 + (alpha^2 - a^2) cos^2 i,
   r = r_cam, theta = clamp(i, 1e-4, pi - 1e-4), inward, heading for smaller theta
   when beta > 0.  The view's theta_cam is the inclination. */
void ko_launch_bardeen(const ko_view* v, double alpha, double beta, ko_photon* p, double y0[5]) {
    const double a = v->spin;
    const double sinI = ko_sin(v->theta_cam), cosI = ko_cos(v->theta_cam);
    double th = v->theta_cam;
    if (th < 1e-4) th = 1e-4;
    if (th > KO_PI - 1e-4) th = KO_PI - 1e-4;
    const double m0 = ko_cos(th);
    const double L = -alpha * sinI;
    const double Q = beta * beta + (alpha * alpha - a * a) * cosI * cosI;
    ko_photon_init(p, a, 1.0, L, Q);
    y0[0] = v->r_cam;
    y0[1] = m0;
    y0[2] = v->phi_cam;
    y0[3] = -sqrt(fmax(ko_R(p, v->r_cam), 0.0));
    y0[4] = (beta > 0.0 ? 1.0 : -1.0) * sqrt(fmax(ko_M(p, m0), 0.0));
}

/* Pinhole camera in the ZAMO frame (camera.cpp:6-40 with Hover, no drift), one
   pixel as trace.frag:1494-1496 + launchPhoton (trace.frag:1470-1488), with the
   frame vectors lowered once per view (renderer.cpp:345-357). */
static void pinhole_photon(const ko_view* v, double n0, double n1, double n2, ko_photon* p, double y0[5]);

static void launch_pinhole(const ko_view* v, int col, int row, ko_photon* p, double y0[5]) {
    double B[9];
    ko_camera_basis(v->yaw, v->pitch, v->roll, B);
    const double tanh_fov = tan(0.5 * v->fov_deg * KO_PI / 180.0);
    const double W = (double)v->width, H = (double)v->height;
    const double nx = (((double)col + 0.5) - 0.5 * W) / (0.5 * H);
    const double ny = (((double)(v->height - row) - 0.5) - 0.5 * H) / (0.5 * H);
    const double x = nx * tanh_fov, yv = ny * tanh_fov;
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = (x * B[i] + yv * B[3 + i]) + B[6 + i];
    const double inv = 1.0 / sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    pinhole_photon(v, d[0] * inv, d[1] * inv, d[2] * inv, p, y0);
}

void ko_launch_look(const ko_view* v, const double look[3], ko_photon* p, double y0[5]) {
    pinhole_photon(v, look[0], look[1], look[2], p, y0);
}

/* the observer of an observer view (camera.cpp:23-36): u_obs normalised to u.u = -1, or
   the ZAMO boosted by drift, u = gamma (e0 + v^i e_i) */
void ko_observer_u(const ko_view* v, double u[4]) {
    const double a = v->spin, r = v->r_cam, th = v->theta_cam;
    if (v->u_obs[0] != 0.0 || v->u_obs[1] != 0.0 || v->u_obs[2] != 0.0 || v->u_obs[3] != 0.0) {
        double g[4][4], n = 0.0;
        ko_metric_tensor(a, r, th, g);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) n += g[i][j] * v->u_obs[i] * v->u_obs[j];
        const double s = 1.0 / sqrt(-n);
        for (int i = 0; i < 4; ++i) u[i] = v->u_obs[i] * s;
        return;
    }
    ko_zamo_four_velocity(a, r, th, u);
    const double v2 = v->drift[0] * v->drift[0] + v->drift[1] * v->drift[1] + v->drift[2] * v->drift[2];
    if (v2 > 0.0 && v2 < 1.0) {
        double e[4][4];
        ko_observer_frame(a, r, th, u, e);
        const double gamma = 1.0 / sqrt(1.0 - v2);
        for (int i = 0; i < 4; ++i)
            u[i] = gamma * (e[0][i] + v->drift[0] * e[1][i] + v->drift[1] * e[2][i] + v->drift[2] * e[3][i]);
    }
}

/* the photon arriving from unit direction n of the camera's frame (the ZAMO's, or the
   observer's for observer views) */
static void pinhole_photon(const ko_view* v, double n0, double n1, double n2, ko_photon* p, double y0[5]) {
    const double a = v->spin, r = v->r_cam, th = v->theta_cam;
    double u[4], e[4][4], g[4][4], low[4][4];
    int given = 0;
    if (v->camera_kind == KO_CAMERA_OBSERVER)
        for (int k = 0; k < 16; ++k) given |= v->frame[k] != 0.0;
    if (given) { /* the caller's frame, as uCamPhoton carries Camera::frame (renderer.cpp:345-357) */
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) e[j][i] = v->frame[4 * j + i];
    } else {
        if (v->camera_kind == KO_CAMERA_OBSERVER) ko_observer_u(v, u);
        else ko_zamo_four_velocity(a, r, th, u);
        ko_observer_frame(a, r, th, u, e);
    }
    ko_metric_tensor(a, r, th, g);
    for (int j = 0; j < 4; ++j) lower4(g, e[j], low[j]);
    const double sn = ko_sin(th), cs = ko_cos(th);
    const double sigma = r * r + a * a * cs * cs;
    /* q = frame^T (1, -n) on the lowered (t, phi, theta) components and e^r */
    const double qt = ((low[0][0] - n0 * low[1][0]) - n1 * low[2][0]) - n2 * low[3][0];
    const double qp = ((low[0][3] - n0 * low[1][3]) - n1 * low[2][3]) - n2 * low[3][3];
    const double qh = ((low[0][2] - n0 * low[1][2]) - n1 * low[2][2]) - n2 * low[3][2];
    const double qr = ((e[0][1] - n0 * e[1][1]) - n1 * e[2][1]) - n2 * e[3][1];
    const double E = -qt;
    double L = qp;
    const double cs2 = cs * cs, sn2 = fmax(sn * sn, 1e-18), a2 = a * a;
    double Q = qh * qh + cs2 * (L * L / sn2 - a2 * E * E);
    if (fabs(L) < KO_AXIS_SNAP * sqrt(fmax(Q + a2 * E * E, 0.0))) {
        L = 0.0;
        Q = qh * qh - cs2 * a2 * E * E;
    }
    ko_photon_init(p, a, E, L, Q);
    p->ks = v->camera_kind == KO_CAMERA_OBSERVER;
    y0[0] = r;
    y0[1] = cs;
    y0[2] = v->phi_cam;
    y0[3] = -(sigma * qr);
    y0[4] = sn * qh;
}

void ko_launch_pixel(const ko_view* v, int col, int row, ko_photon* p, double y0[5]) {
    if (v->camera_kind == KO_CAMERA_PINHOLE || v->camera_kind == KO_CAMERA_OBSERVER) {
        launch_pinhole(v, col, row, p, y0);
        return;
    }
    const double px = 2.0 * v->half_width / (double)v->width;
    const double py = 2.0 * v->half_height / (double)v->height;
    const double alpha = -v->half_width + ((double)col + 0.5) * px;
    const double beta = v->half_height - ((double)row + 0.5) * py;
    ko_launch_bardeen(v, alpha, beta, p, y0);
}

/* ------------------------------------------------------------------------- */
/* the tracer                                                                 */
/* ------------------------------------------------------------------------- */

/* phi_KS - phi_BL = psi(r), the integral of a / Delta from infinity: zero at infinity.  Past
   a = 1, Delta = (r - 1)^2 + b^2 and psi = (a / b)(atan((r - 1) / b) - pi / 2); kerr.cpp:284-291
   (azimuthShift) is the same function without the constant. */
double ko_ks_shift(double a, double r) {
    if (a == 0.0) return 0.0;
    if (naked(a)) {
        const double b = sqrt(a * a - 1.0);
        return a / b * (atan((r - 1.0) / b) - 0.5 * 3.14159265358979323846);
    }
    const double d = 2.0 * sqrt(fmax(0.0, 1.0 - a * a));
    const double rm = 1.0 - 0.5 * d;
    if (d == 0.0) return -a / (r - 1.0);
    return a / d * log1p(-d / (r - rm));
}

static void finish_escape(const ko_photon* p, const ko_opts* o, const double y[5], int* steps, ko_result* out) {
    double z[3] = {y[1], y[2], y[4]}, zn[3], k[7][3];
    double u = 1.0 / y[0];
    double h = -u;
    int rejected = 0;
    tail_rhs(p, u, z, k[0]);
    while (*steps < o->max_steps) {
        int last = 0;
        if (h <= -u) {
            h = -u;
            last = 1;
        }
        const double err = dp5_tail(p, o, u, z, h, k, zn);
        ++*steps;
        if (!(err <= 1.0)) {
            h *= step_factor(err, 1);
            rejected = 1;
            continue;
        }
        const double fac = step_factor(err, rejected);
        rejected = 0;
        if (p->L == 0.0 && z[0] * z[2] > 0.0 && zn[0] * zn[2] < 0.0) zn[1] += KO_PI;
        for (int i = 0; i < 3; ++i) z[i] = zn[i];
        for (int i = 0; i < 3; ++i) k[0][i] = k[6][i];
        if (last) {
            u = 0.0;
            double m = z[0];
            if (m > 1.0) m = 1.0;
            if (m < -1.0) m = -1.0;
            out->cls = KO_ESCAPED;
            out->theta_inf = acos(m);
            out->phi_inf = wrap_2pi(z[1]);
            out->y_end[0] = INFINITY;
            out->y_end[1] = z[0];
            out->y_end[2] = z[1];
            out->y_end[3] = INFINITY;
            out->y_end[4] = z[2];
            return;
        }
        u = u + h;
        h *= fac;
    }
    out->cls = KO_UNFINISHED;
}

/* ---- the single-ray report: RayReport's turns, rMin and delay (trajectory.h:97-113,
   accumulated along the march of trajectory.cpp:259-396) ----
   Over each accepted step, up to the step's event when it has one (the disk crossing,
   r = r_cap or r = r_esc, located on the step's dense output):
     turns  the angle swept by the direction of kerr::toCartesian(a, r, theta, phi_KS)
            (kerr.cpp:638-641), over 2 pi, as great-circle angles between the directions at
            equal sub-steps of the dense output (report_parts);
     r_min  the least r: the camera, the step ends, the event radius, and where w_r turns
            from in- to outgoing inside a step the root of R(r) = 0, Newton-refined from
            the dense output's estimate;
     delay  ingoing Kerr-Schild coordinate time from the event to the camera, the integral
            of dt/dtau (kerr.cpp:350, with inflowFactor kerr.cpp:322-329 and v_r = -w_r)
            over s by 5-point Gauss-Legendre on equal parts of each step's dense output, half
            as many as the swept angle's (the integrand grows as 1/(r - r+) toward the capture
            radius). */
static const double GL_X[5] = {0.04691007703066800, 0.23076534494715845, 0.5, 0.76923465505284155,
                               0.95308992296933200};
static const double GL_W[5] = {0.11846344252809454, 0.23931433524968324, 0.28444444444444444,
                               0.23931433524968324, 0.11846344252809454};

static void ks_direction(double a, double r, double mu, double phi_ks, double d[3]) {
    const double sn = sqrt(fmax(0.0, 1.0 - mu * mu));
    const double cp = cos(phi_ks), sp = sin(phi_ks);
    const double x = (r * cp - a * sp) * sn, y = (r * sp + a * cp) * sn, z = r * mu;
    const double n = sqrt(x * x + y * y + z * z);
    d[0] = x / n;
    d[1] = y / n;
    d[2] = z / n;
}

static double angle_between(const double u[3], const double v[3]) {
    const double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2], cz = u[0] * v[1] - u[1] * v[0];
    return atan2(sqrt(cx * cx + cy * cy + cz * cz), u[0] * v[0] + u[1] * v[1] + u[2] * v[2]);
}

/* ingoing Kerr-Schild dt/dtau at (r, mu) with w = dr/ds = -v_r */
static double dt_dtau(const ko_photon* p, double r, double mu, double w) {
    const double delta = delta_of(p, r), P = P_of(p, r);
    double inflow;
    if (P * w > 0.0) {
        inflow = p->K / (P + w);
    } else {
        const double dd = fabs(delta) > 1e-12 ? delta : copysign(1e-12, delta);
        inflow = (P - w) / dd;
    }
    return fma(r, r, p->a2) * inflow + p->a * (p->L - p->aE * (1.0 - mu * mu)) + w;
}

/* the fraction of a step where the dense output first reaches `target` */
static double dense_root(const ko_dense* d, double target, int from_below) {
    double lo = 0.0, hi = 1.0;
    for (int it = 0; it < 48; ++it) {
        const double mid = 0.5 * (lo + hi);
        if ((dense_at(d, mid) < target) == (from_below != 0)) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

typedef struct { double swept, r_min, delay; } report_acc;

/* parts of a step for the swept angle (half as many for the delay): at least 32, and one per
   0.02 rad of the ingoing phi the step turns through, bounded for an outgoing phi by
   |dr| |a| / Delta at the step's lower end (phi_in = phi_out + 2 psi(r)) */
static int report_parts(const ko_photon* p, double r0, double r1, double dphi) {
    double turn = fabs(dphi);
    if (p->ks != 1) {
        const double lo = fmin(r0, r1);
        const double dlo = fmax(delta_of(p, lo), 1e-12);
        turn += 2.0 * fabs(r1 - r0) * fabs(p->a) / dlo;
    }
    const double n = 32.0 + ceil(turn / 0.02);
    return n < 65536.0 ? (int)n : 65536;
}

static void report_accumulate(const ko_photon* p, const double y0[5], const ko_dense d[5], double h, double th_end,
                              int flip, report_acc* acc) {
    const double a = p->a;
    double rm = fmin(acc->r_min, dense_at(&d[0], th_end));
    if (y0[3] < 0.0 && dense_at(&d[3], th_end) > 0.0) {
        const double tt = dense_root(&d[3], 0.0, 1);
        const double r0 = dense_at(&d[0], tt);
        double r = r0;
        for (int it = 0; it < 4; ++it) {
            const double P = P_of(p, r);
            const double dR = 4.0 * p->E * r * P - 2.0 * (r - 1.0) * p->K;
            if (dR == 0.0) break;
            r -= ko_R(p, r) / dR;
        }
        rm = fmin(rm, fabs(r - r0) <= 1e-6 * r0 ? r : r0);
    }
    acc->r_min = rm;
    double dir0[3], dir1[3];
    const int nsub = report_parts(p, y0[0], dense_at(&d[0], th_end), dense_at(&d[2], th_end) - y0[2]);
    /* toCartesian takes the ingoing phi: phi_in = phi_out + 2 psi(r) */
    ks_direction(a, y0[0], y0[1], p->ks == 1 ? y0[2] : y0[2] + 2.0 * ko_ks_shift(a, y0[0]), dir0);
    for (int j = 1; j <= nsub; ++j) {
        const double th = th_end * (double)j / (double)nsub;
        const double r = dense_at(&d[0], th), mu = dense_at(&d[1], th);
        double ph = dense_at(&d[2], th);
        if (flip && mu * dense_at(&d[4], th) < 0.0) ph += KO_PI;
        ks_direction(a, r, mu, p->ks == 1 ? ph : ph + 2.0 * ko_ks_shift(a, r), dir1);
        acc->swept += angle_between(dir0, dir1);
        dir0[0] = dir1[0];
        dir0[1] = dir1[1];
        dir0[2] = dir1[2];
    }
    const int nquad = nsub / 2;
    double q = 0.0;
    for (int m = 0; m < nquad; ++m)
        for (int g = 0; g < 5; ++g) {
            const double th = th_end * ((double)m + GL_X[g]) / (double)nquad;
            q = fma(GL_W[g], dt_dtau(p, dense_at(&d[0], th), dense_at(&d[1], th), dense_at(&d[3], th)), q);
        }
    acc->delay = fma(h * th_end / (double)nquad, q, acc->delay);
}

static void trace_photon(const ko_view* v, const ko_opts* o, const ko_photon* p, const double y0[5], ko_result* out,
                         ko_report* rep);

void ko_trace_photon(const ko_view* v, const ko_opts* o, const ko_photon* p, const double y0[5],
                     ko_result* out) {
    trace_photon(v, o, p, y0, out, NULL);
}

void ko_trace_photon_report(const ko_view* v, const ko_opts* o, const ko_photon* p, const double y0[5],
                            ko_result* out, ko_report* rep) {
    trace_photon(v, o, p, y0, out, rep);
}

static void trace_photon(const ko_view* v, const ko_opts* o, const ko_photon* p_launch, const double y0_launch[5],
                         ko_result* out, ko_report* rep) {
    /* a launch in Boyer-Lindquist phi (Bardeen and pinhole views) is traced in the outgoing
       Kerr-Schild phi, phi_out = phi_BL - psi(r), which equals phi_BL at infinity */
    ko_photon pp = *p_launch;
    double y0[5];
    for (int i = 0; i < 5; ++i) y0[i] = y0_launch[i];
    if (pp.ks == 0) {
        pp.ks = 2;
        y0[2] = y0_launch[2] - ko_ks_shift(pp.a, y0_launch[0]);
    }
    const ko_photon* p = &pp;
    const double a = v->spin;
    const double r_cap = v->capture_factor * ko_horizon_radius(a);
    const double r_esc = amax(v->escape_radius, 1.5 * v->r_cam);
    const int disk = v->disk_enabled;
    const double r_in = v->disk_inner > 0.0 ? v->disk_inner : ko_isco_radius(a, v->disk_prograde);
    const double r_out = v->disk_outer;

    double y[5], yn[5], k[7][5];
    for (int i = 0; i < 5; ++i) y[i] = y0[i];
    out->cls = KO_UNFINISHED;
    out->flags = 0;
    out->theta_inf = out->phi_inf = out->r_hit = out->g = NAN;
    int steps = 0, rejected = 0;
    double h = 0.01 / v->r_cam;
    report_acc acc = {0.0, y0[0], 0.0};
    if (rep) {
        const double invE = fabs(p->E) > 1e-12 ? 1.0 / p->E : 0.0; /* trajectory.cpp:221 */
        rep->b = p->L * invE;
        rep->q = p->Q * invE * invE;
    }
#define KO_REPORT_OUT()                              \
    do {                                             \
        if (rep) {                                   \
            rep->turns = acc.swept / (2.0 * KO_PI);  \
            rep->r_min = acc.r_min;                  \
            rep->delay = acc.delay;                  \
        }                                            \
    } while (0)
    /* inside the horizon every ray, followed back, rises to it; one arriving with
       P(r+) = 2 r+ E - a L < 0 came from the far side: dark (trajectory.cpp:234-241) */
    const double rh = ko_horizon_radius(a);
    if (p->ks == 1 && v->r_cam < rh && 2.0 * rh * p->E - a * p->L < 0.0) {
        out->cls = KO_CAPTURED;
        for (int i = 0; i < 5; ++i) out->y_end[i] = y[i];
        out->steps = 0;
        KO_REPORT_OUT();
        return;
    }
    ko_rhs(p, y, k[0]);
    while (steps < o->max_steps) {
        const double err = dp5_main(p, o, y, h, k, yn);
        ++steps;
        if (!(err <= 1.0)) {
            h *= step_factor(err, 1);
            rejected = 1;
            continue;
        }
        const double fac = step_factor(err, rejected);
        rejected = 0;
        /* the report's dense output, from the step as taken (before the half-turn and the projection) */
        ko_dense dd[5];
        if (rep)
            for (int i = 0; i < 5; ++i) dd[i] = dense_coeffs(y, yn, k, h, i);
        /* through the pole (L = 0 only): same point, phi + pi */
        const int flip = p->L == 0.0 && y[1] * y[4] > 0.0 && yn[1] * yn[4] < 0.0;
        if (flip) yn[2] += KO_PI;
        project(p, yn);

        /* a non-finite state counts as captured (eht.cpp:285) */
        if (!(fabs(yn[0]) < 1e300) || !(fabs(yn[1]) < 1e300)) {
            out->cls = KO_CAPTURED;
            out->flags = KO_FLAG_NONFINITE;
            for (int i = 0; i < 5; ++i) out->y_end[i] = yn[i];
            out->steps = steps;
            KO_REPORT_OUT();
            return;
        }
        /* events in the reference's order: disk, horizon, escape (trajectory.cpp:307-392) */
        if (disk && y[1] * yn[1] < 0.0) {
            const ko_dense dr = dense_coeffs(y, yn, k, h, 0), dm = dense_coeffs(y, yn, k, h, 1);
            double lo = 0.0, hi = 1.0;
            const int pos = y[1] > 0.0;
            for (int it = 0; it < 48; ++it) {
                const double mid = 0.5 * (lo + hi);
                const double f = dense_at(&dm, mid);
                if ((f > 0.0) == pos) lo = mid;
                else hi = mid;
            }
            const double rc = dense_at(&dr, 0.5 * (lo + hi));
            if (rc >= r_in && rc <= r_out) {
                /* trajectory.cpp:311-321 */
                const double omega = ko_circular_omega(a, rc, v->disk_prograde);
                const double gtt = -(1.0 - 2.0 / rc), gtp = -2.0 * a / rc;
                const double gpp = rc * rc + a * a + 2.0 * a * a / rc;
                const double norm = -(gtt + 2.0 * gtp * omega + gpp * omega * omega);
                if (norm > 0.0) {
                    out->cls = KO_DISK;
                    out->r_hit = rc;
                    out->g = sqrt(norm) / (p->E - p->L * omega);
                    for (int i = 0; i < 5; ++i) out->y_end[i] = yn[i];
                    out->steps = steps;
                    if (rep) report_accumulate(p, y, dd, h, dense_root(&dd[1], 0.0, !(y[1] > 0.0)), flip, &acc);
                    KO_REPORT_OUT();
                    return;
                }
            }
        }
        /* observer cameras may start below r_cap: capture only heading inward (trajectory.cpp:377) */
        if (yn[0] < r_cap && (p->ks != 1 || yn[3] < 0.0)) {
            out->cls = KO_CAPTURED;
            for (int i = 0; i < 5; ++i) out->y_end[i] = yn[i];
            out->steps = steps;
            if (rep) report_accumulate(p, y, dd, h, y[0] >= r_cap ? dense_root(&dd[0], r_cap, 0) : 1.0, flip, &acc);
            KO_REPORT_OUT();
            return;
        }
        if (yn[0] > r_esc && yn[3] > 0.0) {
            if (rep) report_accumulate(p, y, dd, h, dense_root(&dd[0], r_esc, 1), flip, &acc);
            KO_REPORT_OUT();
            if (p->ks == 1) yn[2] -= ko_ks_shift(a, yn[0]); /* its tail runs in Boyer-Lindquist phi */
            finish_escape(p, o, yn, &steps, out);
            out->steps = steps;
            return;
        }
        if (rep) report_accumulate(p, y, dd, h, 1.0, flip, &acc);
        for (int i = 0; i < 5; ++i) y[i] = yn[i];
        for (int i = 0; i < 5; ++i) k[0][i] = k[6][i];
        k[0][0] = y[3]; /* FSAL after the projection: only dr/ds = w_r, dmu/ds = w_mu move */
        k[0][1] = y[4];
        h *= fac;
    }
    for (int i = 0; i < 5; ++i) out->y_end[i] = y[i];
    out->steps = steps;
    KO_REPORT_OUT();
#undef KO_REPORT_OUT
}

void ko_trace_pixel(const ko_view* v, const ko_opts* o, int col, int row, ko_result* out) {
    ko_photon p;
    double y0[5];
    ko_launch_pixel(v, col, row, &p, y0);
    ko_trace_photon(v, o, &p, y0, out);
}

/* ------------------------------------------------------------------------- */
/* threaded image driver (eht.cpp:321-347)                                    */
/* ------------------------------------------------------------------------- */

typedef struct {
    const ko_view* v;
    const ko_opts* o;
    const volatile int* cancel;
    int stride;
    int next_row;
    int cancelled;
    int8_t* cls;
    double *theta, *phi, *rhit, *g;
    int32_t* steps;
} render_job;

static void* render_worker(void* arg) {
    render_job* j = (render_job*)arg;
    const int W = j->v->width, H = j->v->height;
    for (;;) {
        const int row = __atomic_fetch_add(&j->next_row, 1, __ATOMIC_RELAXED);
        if (row >= H) break;
        if (j->cancel && *j->cancel) {
            __atomic_store_n(&j->cancelled, 1, __ATOMIC_RELAXED);
            break;
        }
        for (int col = 0; col < W; ++col) {
            const long idx = (long)row * W + col;
            if (j->stride > 1 && idx % j->stride != 0) continue;
            const long o = j->stride > 1 ? idx / j->stride : idx;
            ko_result res;
            ko_trace_pixel(j->v, j->o, col, row, &res);
            if (j->cls) j->cls[o] = (int8_t)(res.cls | res.flags);
            if (j->theta) j->theta[o] = res.theta_inf;
            if (j->phi) j->phi[o] = res.phi_inf;
            if (j->rhit) j->rhit[o] = res.r_hit;
            if (j->g) j->g[o] = res.g;
            if (j->steps) j->steps[o] = res.steps;
        }
    }
    return NULL;
}

long ko_render(const ko_view* v, const ko_opts* o, int threads, const volatile int* cancel, int stride,
               int8_t* cls, double* theta_inf, double* phi_inf, double* r_hit, double* g, int32_t* steps) {
    render_job j = {v, o, cancel, stride < 1 ? 1 : stride, 0, 0, cls, theta_inf, phi_inf, r_hit, g, steps};
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t pool[256];
    int started = 0;
    for (int t = 1; t < threads; ++t)
        if (pthread_create(&pool[started], NULL, render_worker, &j) == 0) ++started;
    render_worker(&j);
    for (int t = 0; t < started; ++t) pthread_join(pool[t], NULL);
    if (j.cancelled) return -1;
    const long n = (long)v->width * v->height;
    return j.stride > 1 ? (n + j.stride - 1) / j.stride : n;
}

/* ------------------------------------------------------------------------- */
/* critical curve (eht.cpp:52-97)                                             */
/* ------------------------------------------------------------------------- */

int ko_critical_curve(double a, double inclination, int samples, double* alpha, double* beta, int cap) {
    const double sinI = ko_sin(inclination), cosI = ko_cos(inclination);
    int n_out = 0;
    if (fabs(a) < 1e-6) {
        for (int k = 0; k < samples && n_out < cap; ++k) {
            const double psi = 2.0 * KO_PI * k / samples;
            alpha[n_out] = sqrt(27.0) * ko_cos(psi);
            beta[n_out] = sqrt(27.0) * ko_sin(psi);
            ++n_out;
        }
        return n_out;
    }
    const double r1 = ko_photon_orbit_radius(a, 1), r2 = ko_photon_orbit_radius(a, 0);
    const int n = (samples > 16 ? samples : 16) * 8;
    int nu = 0;
    /* upper half first, into the front of the buffers */
    for (int k = 0; k <= n; ++k) {
        const double x = 0.5 - 0.5 * ko_cos(KO_PI * k / n);
        const double r = r1 + (r2 - r1) * x;
        const double xi = (r * r * (3.0 - r) - a * a * (r + 1.0)) / (a * (r - 1.0));
        const double eta = r * r * r * (4.0 * a * a - r * (r - 3.0) * (r - 3.0)) / (a * a * (r - 1.0) * (r - 1.0));
        const double b2 = eta + a * a * cosI * cosI - xi * xi * cosI * cosI / (sinI * sinI);
        if (b2 < 0.0) continue;
        if (nu >= cap) break;
        alpha[nu] = -xi / sinI;
        beta[nu] = sqrt(b2);
        ++nu;
    }
    n_out = nu;
    for (int k = nu; k-- > 0 && n_out < cap;) {
        alpha[n_out] = alpha[k];
        beta[n_out] = -beta[k];
        ++n_out;
    }
    return n_out;
}

double ko_critical_radius(const double* alpha, const double* beta, int n, double psi) {
    const double dx = ko_cos(psi), dy = ko_sin(psi);
    double best = 0.0;
    for (int k = 0; k < n; ++k) {
        const int q = (k + 1) % n;
        const double ex = alpha[q] - alpha[k], ey = beta[q] - beta[k];
        const double det = ex * dy - ey * dx;
        if (fabs(det) < 1e-15) continue;
        const double t = (beta[k] * dx - alpha[k] * dy) / det;
        const double s = (ex * beta[k] - ey * alpha[k]) / det;
        if (t >= 0.0 && t <= 1.0 && s > 0.0 && s > best) best = s;
    }
    return best;
}

/* ---- HDR colour of a traced ray ------------------------------------------------------ */

/* Wyman, Sloan and Shirley (2013) multi-lobe fit of the CIE 1931 2-degree observer, as
   blackbody.cpp:16-26 uses it (lambda in nm) */
static double ko_lobe(double x, double mu, double s1, double s2) {
    double t = (x - mu) / (x < mu ? s1 : s2);
    return exp(-0.5 * t * t);
}

void ko_planck_xyz(double T, double xyz[3]) {
    const double h = 6.62607015e-34, c = 2.99792458e8, k = 1.380649e-23;
    xyz[0] = xyz[1] = xyz[2] = 0.0;
    for (int nm = 360; nm <= 830; nm += 2) {
        double lam = nm * 1e-9, x = (double)nm;
        double B = 2.0 * h * c * c / (lam * lam * lam * lam * lam) / expm1(h * c / (lam * k * T));
        xyz[0] += B * (1.056 * ko_lobe(x, 599.8, 37.9, 31.0) + 0.362 * ko_lobe(x, 442.0, 16.0, 26.7) -
                       0.065 * ko_lobe(x, 501.1, 20.4, 26.2));
        xyz[1] += B * (0.821 * ko_lobe(x, 568.8, 46.9, 40.5) + 0.286 * ko_lobe(x, 530.9, 16.3, 31.1));
        xyz[2] += B * (1.217 * ko_lobe(x, 437.0, 11.8, 36.0) + 0.681 * ko_lobe(x, 459.0, 26.0, 13.8));
    }
}

void ko_blackbody_table(int n, float* rgb) {
    double ref[3];
    ko_planck_xyz(6500.0, ref);
    for (int i = 0; i < n; ++i) {
        double logT = KO_LOGT_MIN + (KO_LOGT_MAX - KO_LOGT_MIN) * i / (n - 1), v[3];
        ko_planck_xyz(pow(10.0, logT), v);
        /* XYZ -> linear sRGB (IEC 61966-2-1), negatives clipped (blackbody.cpp:36-41) */
        double r = 3.2406 * v[0] - 1.5372 * v[1] - 0.4986 * v[2];
        double g = -0.9689 * v[0] + 1.8758 * v[1] + 0.0415 * v[2];
        double b = 0.0557 * v[0] - 0.2040 * v[1] + 1.0570 * v[2];
        rgb[3 * i + 0] = (float)((r > 0.0 ? r : 0.0) / ref[1]);
        rgb[3 * i + 1] = (float)((g > 0.0 ? g : 0.0) / ref[1]);
        rgb[3 * i + 2] = (float)((b > 0.0 ? b : 0.0) / ref[1]);
    }
}

/* the table at T, linearly interpolated in log10 T and clamped (trace.frag:308-313) */
static void ko_table_rgb(const ko_shade* s, double T, double rgb[3]) {
    double u = (log10(T > 1.0 ? T : 1.0) - KO_LOGT_MIN) / (KO_LOGT_MAX - KO_LOGT_MIN);
    u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
    double x = u * (s->table_n - 1);
    int i = (int)x;
    if (i > s->table_n - 2) i = s->table_n - 2;
    double f = x - i;
    const float* a = s->table + 3 * i;
    for (int k = 0; k < 3; ++k) rgb[k] = a[k] + f * ((double)a[k + 3] - a[k]);
}

static double ko_smoothstep(double e0, double e1, double x) {
    double t = (x - e0) / (e1 - e0);
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return t * t * (3.0 - 2.0 * t);
}

void ko_shade_ray(const ko_shade* s, int cls, double theta, double phi, double r_hit, double g, double r_in,
                  double r_out, double rgb[3]) {
    rgb[0] = rgb[1] = rgb[2] = 0.0;
    if (cls == KO_ESCAPED) {
        double d[3] = {sin(theta) * cos(phi), sin(theta) * sin(phi), cos(theta)};
        double lon = atan2(d[1], d[0]), lat = asin(d[2] > 1.0 ? 1.0 : (d[2] < -1.0 ? -1.0 : d[2]));
        if (s->sky == 1) { /* gridSky, trace.frag:1419-1436 */
            static const double hue[4][3] = {{0.95, 0.3, 0.25}, {0.25, 0.75, 0.35}, {0.3, 0.45, 0.95},
                                             {0.95, 0.8, 0.25}};
            double u = lon / KO_PI * 12.0, v = lat / KO_PI * 12.0, cu = floor(u), cv = floor(v);
            int quadrant = ((int)floor((lon + KO_PI) / (0.5 * KO_PI))) & 3;
            double k = lat < 0.0 ? 0.55 : 1.0;
            double parity = (cu + cv) - 2.0 * floor((cu + cv) * 0.5);
            k *= 0.45 + 0.55 * parity;
            double eu = 0.5 - fabs(u - cu - 0.5), ev = 0.5 - fabs(v - cv - 0.5);
            double lu = ko_smoothstep(0.0, s->grid_line, eu), lv = ko_smoothstep(0.0, s->grid_line, ev);
            k *= (0.15 + 0.85 * (lu < lv ? lu : lv)) * 0.6;
            for (int c = 0; c < 3; ++c) rgb[c] = hue[quadrant][c] * k;
        } else if (s->sky == 2) { /* imageSky, trace.frag:1438-1445, bilinear */
            int W = s->sky_width, H = s->sky_height;
            double x = (lon / (2.0 * KO_PI) + 0.5) * W - 0.5, y = (lat / KO_PI + 0.5) * H - 0.5;
            y = y < 0.0 ? 0.0 : (y > H - 1 ? H - 1 : y);
            double fx = floor(x), fy = floor(y), tx = x - fx, ty = y - fy;
            int x0 = ((int)fx % W + W) % W, x1 = (x0 + 1) % W, y0 = (int)fy, y1 = y0 + 1 < H ? y0 + 1 : H - 1;
            for (int c = 0; c < 3; ++c) {
                double p00 = s->sky_image[3 * ((long)y0 * W + x0) + c], p01 = s->sky_image[3 * ((long)y0 * W + x1) + c];
                double p10 = s->sky_image[3 * ((long)y1 * W + x0) + c], p11 = s->sky_image[3 * ((long)y1 * W + x1) + c];
                rgb[c] = (1 - ty) * ((1 - tx) * p00 + tx * p01) + ty * ((1 - tx) * p10 + tx * p11);
            }
        }
        for (int c = 0; c < 3; ++c) rgb[c] *= s->sky_brightness;
    } else if (cls == KO_DISK) {
        /* diskEmission (trace.frag:373-403): T^4 ~ x^-3 (1 - x^-1/2), peak (x = 49/36) at the
           disk temperature; seen as a blackbody at g T; normalised by the luminance of the
           disk temperature's colour; edges as trace.frag:396-398 */
        double x = r_hit / r_in, peak = pow(36.0 / 49.0, 3.0) / 7.0;
        double flux = pow(x, -3.0) * (1.0 - 1.0 / sqrt(x));
        double T = s->disk_temperature * pow((flux > 0.0 ? flux : 0.0) / peak, 0.25), bb[3], ref[3];
        ko_table_rgb(s, T * g, bb);
        ko_table_rgb(s, s->disk_temperature, ref);
        double lum = 0.2126 * ref[0] + 0.7152 * ref[1] + 0.0722 * ref[2];
        double edges = (1.0 - ko_smoothstep(0.7 * r_out, r_out, r_hit)) * ko_smoothstep(r_in, 1.03 * r_in, r_hit);
        for (int c = 0; c < 3; ++c) rgb[c] = bb[c] * s->disk_brightness / (lum > 1e-30 ? lum : 1e-30) * edges;
    }
    /* trace.frag:1905-1906: non-finite -> black, capped at 30000 */
    if (!(fabs(rgb[0]) <= 1e300 && fabs(rgb[1]) <= 1e300 && fabs(rgb[2]) <= 1e300)) rgb[0] = rgb[1] = rgb[2] = 0.0;
    for (int c = 0; c < 3; ++c) rgb[c] = rgb[c] < 30000.0 ? rgb[c] : 30000.0;
}

/* ---- EHT radiative transfer (eht.h:22-122; eht.cpp:98-316, 321-354, 432-457) ------------
   Physical constants in cgs as eht.h:22-33.  The polynomial approximations of I0, I1, K0
   and K1 are Abramowitz and Stegun 9.8.1-9.8.8, as eht.cpp:98-152 uses them. */
#define KO_QE 4.80320425e-10
#define KO_ME 9.1093837015e-28
#define KO_MP 1.67262192369e-24
#define KO_CL 2.99792458e10
#define KO_KB 1.380649e-16
#define KO_HP 6.62607015e-27
#define KO_MSUN_CM 1.476625e5
#define KO_JY 1e-23
#define KO_UAS (KO_PI / 180.0 / 3600.0 / 1e6)
#define KO_EHT_START 1000.0 /* eht.cpp:14 */
#define KO_EHT_LIGHT 60.0   /* eht.cpp:15 */
#define KO_EHT_DEPTH 30.0   /* eht.cpp:16 */

static double ko_i0(double x) {
    double ax = fabs(x);
    if (ax < 3.75) {
        double y = (x / 3.75) * (x / 3.75);
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))));
    }
    double y = 3.75 / ax;
    return (exp(ax) / sqrt(ax)) * (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 +
           y * (-0.2057706e-1 + y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))));
}

static double ko_i1(double x) {
    double ax = fabs(x), res;
    if (ax < 3.75) {
        double y = (x / 3.75) * (x / 3.75);
        res = ax * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 + y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
    } else {
        double y = 3.75 / ax;
        double t = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
        t = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 + y * (-0.1031555e-1 + y * t))));
        res = t * exp(ax) / sqrt(ax);
    }
    return x < 0.0 ? -res : res;
}

/* K0, K1 times exp(x) */
static double ko_k0s(double x) {
    if (x <= 2.0) {
        double y = x * x / 4.0;
        return exp(x) * ((-log(x / 2.0) * ko_i0(x)) + (-0.57721566 + y * (0.42278420 + y * (0.23069756 + y * (0.3488590e-1 +
               y * (0.262698e-2 + y * (0.10750e-3 + y * 0.74e-5)))))));
    }
    double y = 2.0 / x;
    return (1.0 / sqrt(x)) * (1.25331414 + y * (-0.7832358e-1 + y * (0.2189568e-1 + y * (-0.1062446e-1 + y * (0.587872e-2 +
           y * (-0.251540e-2 + y * 0.53208e-3))))));
}

static double ko_k1s(double x) {
    if (x <= 2.0) {
        double y = x * x / 4.0;
        return exp(x) * ((log(x / 2.0) * ko_i1(x)) + (1.0 / x) * (1.0 + y * (0.15443144 + y * (-0.67278579 + y * (-0.18156897 +
               y * (-0.1919402e-1 + y * (-0.110404e-2 + y * (-0.4686e-4))))))));
    }
    double y = 2.0 / x;
    return (1.0 / sqrt(x)) * (1.25331414 + y * (0.23498619 + y * (-0.3655620e-1 + y * (0.1504268e-1 + y * (-0.780353e-2 +
           y * (0.325614e-2 + y * (-0.68245e-3)))))));
}

static double ko_k2s(double x) { return ko_k0s(x) + 2.0 / x * ko_k1s(x); } /* K2 = K0 + (2/x) K1 */

double ko_bessel_k(int which, double x) {
    double s = which == 0 ? ko_k0s(x) : which == 1 ? ko_k1s(x) : ko_k2s(x);
    return s * exp(-x);
}

/* Leung, Gammie and Noble (2011) eq. 72, eht.cpp:163-177 */
double ko_synchrotron_emissivity(double nu, double ne, double thetaE, double B, double sinThetaB) {
    if (ne <= 0.0 || thetaE < 0.05 || B <= 0.0 || sinThetaB <= 0.0) return 0.0;
    double nuC = KO_QE * B / (2.0 * KO_PI * KO_ME * KO_CL);
    double nuS = (2.0 / 9.0) * nuC * thetaE * thetaE * sinThetaB;
    double X = nu / nuS;
    double shape = sqrt(X) + pow(2.0, 11.0 / 12.0) * pow(X, 1.0 / 6.0);
    double inv = 1.0 / thetaE;
    double expo = inv - cbrt(X); /* 1/K2(1/thetaE) = exp(1/thetaE)/scaledK2, with exp(-X^1/3) */
    if (expo < -700.0) return 0.0;
    return ne * sqrt(2.0) * KO_PI * KO_QE * KO_QE * nuS / (3.0 * ko_k2s(inv) * KO_CL) * shape * shape * exp(expo);
}

double ko_planck_nu(double nu, double T) {
    double x = KO_HP * nu / (KO_KB * T);
    return 2.0 * KO_HP * nu * nu * nu / (KO_CL * KO_CL) / expm1(x);
}

/* the flow's plunge from the prograde ISCO, per spin (eht.cpp:183-192, 222-233) */
typedef struct ko_plunge {
    double a, isco, horizon, E, L, start, start_speed;
} ko_plunge;

static double plunge_ur(const ko_plunge* q, double r) {
    double delta = r * r - 2.0 * r + q->a * q->a, x = q->L - q->a * q->E;
    double P = q->E * (r * r + q->a * q->a) - q->a * q->L;
    return sqrt(fmax(P * P - delta * (r * r + x * x), 0.0)) / (r * r);
}

static void plunge_init(ko_plunge* q, double a) {
    q->a = a;
    q->isco = ko_isco_radius(a, 1);
    q->horizon = ko_horizon_radius(a);
    /* circular equatorial orbit at the ISCO (kerr.cpp:308-319, prograde) */
    double r = q->isco, sr = sqrt(r), r32 = r * sr;
    double den = pow(r, 0.75) * sqrt(r32 - 3.0 * sr + 2.0 * a);
    q->E = (r32 - 2.0 * sr + a) / den;
    q->L = (r * r - 2.0 * a * sr + a * a) / den;
    q->start = fmax(q->isco - 0.25, 0.5 * (q->isco + q->horizon));
    q->start_speed = plunge_ur(q, q->start);
}

/* n_e, T_e, B at (r, mu = cos theta) (eht.cpp:196-215) */
static void plasma_mu(const ko_flow* f, const ko_plunge* q, double r, double mu, double out[3]) {
    double z = r * mu, rho = fmax(r * sqrt(fmax(1.0 - mu * mu, 0.0)), 1e-6);
    double h = fmax(f->height, 0.05) * rho;
    double radial = pow(r, -1.1);
    if (r < q->isco && r < q->start) {
        double here = plunge_ur(q, r);
        if (here > 0.0) radial = fmin(radial, pow(q->start, -1.1) * (q->start * q->start * q->start_speed) / (r * r * here));
    }
    out[0] = f->n0 * radial * exp(-z * z / (2.0 * h * h));
    out[1] = f->T0 * pow(r, -0.84);
    out[2] = sqrt(8.0 * KO_PI * out[0] * KO_MP * KO_CL * KO_CL / (6.0 * f->beta * r));
}

/* the gas's u^mu (BL, contravariant) at (r, mu) (eht.cpp:217-253) */
static void flow_mu(const ko_flow* f, const ko_plunge* q, double r, double mu, double u[4]) {
    double a = q->a, s2 = 1.0 - mu * mu, c2 = mu * mu;
    double sigma = r * r + a * a * c2, delta = r * r - 2.0 * r + a * a;
    double A = (r * r + a * a) * (r * r + a * a) - a * a * delta * s2;
    double gtt = -(1.0 - 2.0 * r / sigma), gtp = -2.0 * a * r * s2 / sigma, gpp = A * s2 / sigma, grr = sigma / delta;
    double omega, ur = 0.0;
    if (r >= q->isco) {
        omega = f->keplerian / (pow(r, 1.5) + a);
    } else {
        double x = q->L - a * q->E, P = q->E * (r * r + a * a) - a * q->L;
        double R = fmax(P * P - delta * (r * r + x * x), 0.0);
        double ut = ((r * r + a * a) * P / delta + a * x) / (r * r);
        double up = (a * P / delta + x) / (r * r);
        omega = up / ut;
        ur = -sqrt(R) / (r * r);
    }
    double den = gtt + 2.0 * omega * gtp + omega * omega * gpp;
    double num = -1.0 - grr * ur * ur;
    if (!(den < 0.0) || !(num < 0.0)) { /* the local ZAMO */
        omega = -gtp / gpp;
        ur = 0.0;
        den = gtt + 2.0 * omega * gtp + omega * omega * gpp;
        num = -1.0;
    }
    double ut = sqrt(num / den);
    u[0] = ut;
    u[1] = ur;
    u[2] = 0.0;
    u[3] = omega * ut;
}

void ko_plasma_at(const ko_flow* f, double r, double theta, double a, double out[3]) {
    ko_plunge q;
    plunge_init(&q, a);
    plasma_mu(f, &q, r, cos(theta), out);
}

void ko_flow_velocity(const ko_flow* f, double a, double r, double theta, double u[4]) {
    ko_plunge q;
    plunge_init(&q, a);
    flow_mu(f, &q, r, cos(theta), u);
}

/* One pixel: the Bardeen ray from r = 1000 M (eht.cpp:262-271) stepped by the DP5 engine;
   after every accepted step inside r = 60 M the transfer of eht.cpp:290-316 over that step
   (emission and absorption at the step's end, affine length from the mean Sigma). */
void ko_eht_trace(const ko_eht_view* v, const ko_flow* f, const ko_opts* o, double alpha, double beta,
                  ko_eht_ray* out) {
    ko_view kv;
    memset(&kv, 0, sizeof kv);
    kv.camera_kind = KO_CAMERA_BARDEEN;
    kv.spin = v->a;
    kv.r_cam = KO_EHT_START;
    kv.theta_cam = v->inclination;
    ko_photon p;
    double y[5], yn[5], k[7][5];
    ko_launch_bardeen(&kv, alpha, beta, &p, y);
    ko_plunge q;
    plunge_init(&q, v->a);
    const double a = v->a, r_cap = q.horizon * (1.0 + 1e-4), len_cm = KO_MSUN_CM * v->mass_solar;
    out->intensity = 0.0;
    out->captured = 0;
    out->min_radius = y[0];
    double depth = 0.0;
    int steps = 0, rejected = 0;
    double h = 0.01 / KO_EHT_START;
    ko_rhs(&p, y, k[0]);
    while (steps < o->max_steps) {
        const double err = dp5_main(&p, o, y, h, k, yn);
        ++steps;
        if (!(err <= 1.0)) {
            h *= step_factor(err, 1);
            rejected = 1;
            continue;
        }
        const double fac = step_factor(err, rejected);
        rejected = 0;
        if (p.L == 0.0 && y[1] * y[4] > 0.0 && yn[1] * yn[4] < 0.0) yn[2] += KO_PI;
        project(&p, yn);
        if (!(fabs(yn[0]) < 1e300) || !(fabs(yn[1]) < 1e300) || yn[0] < r_cap) {
            out->captured = 1;
            break;
        }
        if (yn[0] < out->min_radius) out->min_radius = yn[0];
        if (yn[0] > KO_EHT_START * 1.01 && yn[3] > 0.0) break;
        if (yn[0] <= KO_EHT_LIGHT) {
            const double r = yn[0], mu = yn[1];
            const double sig_b = y[0] * y[0] + a * a * y[1] * y[1], sig_a = r * r + a * a * mu * mu;
            const double dl = 0.5 * (sig_a + sig_b) * h * len_cm; /* affine length of the step, cm */
            double pl[3], u[4];
            plasma_mu(f, &q, r, mu, pl);
            flow_mu(f, &q, r, mu, u);
            const double delta = r * r - 2.0 * r + a * a;
            /* -p.u with p_t = -1, p_phi = L, p_r = -w_r / Delta (towards the camera); u^theta = 0 */
            const double mpu = u[0] - p.L * u[3] + yn[3] / delta * u[1];
            if (mpu > 0.0) {
                const double g = 1.0 / mpu, nu_f = v->frequency / g;
                const double thetaE = KO_KB * pl[1] / (KO_ME * KO_CL * KO_CL);
                const double j = ko_synchrotron_emissivity(nu_f, pl[0], thetaE, pl[2], f->sin_theta_b);
                if (j > 0.0) {
                    const double absorb = j / ko_planck_nu(nu_f, pl[1]);
                    const double dtau = absorb * dl / g;
                    const double add = dtau < 1e-6 ? g * g * j * dl : g * g * g * (j / absorb) * -expm1(-dtau);
                    out->intensity += exp(-depth) * add;
                    depth += dtau;
                    if (depth > KO_EHT_DEPTH) break;
                }
            }
        }
        for (int i = 0; i < 5; ++i) y[i] = yn[i];
        for (int i = 0; i < 5; ++i) k[0][i] = k[6][i];
        k[0][0] = y[3];
        k[0][1] = y[4];
        h *= fac;
    }
    out->steps = steps;
}

/* eht.cpp:321-347: pixel centres, rows from the top */
void ko_eht_render(const ko_eht_view* v, const ko_flow* f, const ko_opts* o, double* intensity, uint8_t* shadow) {
    const double px = 2.0 * v->half_width / v->size;
    for (int row = 0; row < v->size; ++row)
        for (int col = 0; col < v->size; ++col) {
            ko_eht_ray ray;
            ko_eht_trace(v, f, o, -v->half_width + (col + 0.5) * px, v->half_width - (row + 0.5) * px, &ray);
            intensity[(long)row * v->size + col] = ray.intensity;
            if (shadow) shadow[(long)row * v->size + col] = (uint8_t)ray.captured;
        }
}

/* eht.cpp:349-354: Jy, the pixel's solid angle from theta_g in microarcseconds */
double ko_eht_flux(const double* intensity, int size, double half_width, double theta_guas) {
    double pix = 2.0 * half_width / size, sa = pow(pix * theta_guas * KO_UAS, 2.0), sum = 0.0;
    for (long i = 0; i < (long)size * size; ++i) sum += intensity[i];
    return sum * sa / KO_JY;
}

/* eht.cpp:432-457: the density n0 that gives target_jy, by a secant in log n0 against log
   flux on 48 x 48 renders of the view (a dark start is first brightened tenfold at a time) */
double ko_eht_calibrate(const ko_eht_view* v, const ko_flow* flow, const ko_opts* o, double target_jy,
                        double theta_guas) {
    ko_eht_view small = *v;
    small.size = 48;
    ko_flow f = *flow;
    const long n = (long)small.size * small.size;
    double* img = (double*)malloc(sizeof(double) * n);
    if (!img) return NAN;
#define FLUX_AT(x) (f.n0 = (x), ko_eht_render(&small, &f, o, img, NULL), ko_eht_flux(img, small.size, small.half_width, theta_guas))
    double n0 = f.n0, f0 = FLUX_AT(n0);
    for (int k = 0; k < 8 && f0 <= 0.0; ++k) {
        n0 *= 10.0;
        f0 = FLUX_AT(n0);
    }
    if (f0 <= 0.0) {
        free(img);
        return n0;
    }
    double n1 = n0 * target_jy / f0, f1 = FLUX_AT(n1);
    for (int k = 0; k < 8; ++k) {
        if (f1 <= 0.0 || fabs(f1 / target_jy - 1.0) < 0.01) break;
        const double slope = (log(f1) - log(f0)) / (log(n1) - log(n0));
        const double next = exp(log(n1) + (log(target_jy) - log(f1)) / (fabs(slope) > 0.05 ? slope : 1.0));
        n0 = n1;
        f0 = f1;
        n1 = next < n0 / 30.0 ? n0 / 30.0 : next > n0 * 30.0 ? n0 * 30.0 : next;
        f1 = FLUX_AT(n1);
    }
#undef FLUX_AT
    free(img);
    return n1;
}
