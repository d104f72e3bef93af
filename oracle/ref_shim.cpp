// ref_shim.cpp -- C entry points onto the reference's own Kerr code, for golden vectors.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with the
// reference's src/kerr.cpp, src/eht.cpp, src/trajectory.cpp and src/central.cpp, read
// where they lie under /root/reference, into oracle/_ref/libkerr_ref.so.  Nothing
// from the reference is copied into this repository; tests/golden/make_golden.py
// calls these wrappers through ctypes to write the fixtures under tests/golden/.

#include "central.h"
#include "eht.h"
#include "kerr.h"
#include "trajectory.h"

#include <cstring>

extern "C" {

double ref_horizon_radius(double a) { return kerr::horizonRadius(a); }
double ref_isco_radius(double a, int prograde) { return kerr::iscoRadius(a, prograde != 0); }
double ref_photon_orbit_radius(double a, int prograde) { return kerr::photonOrbitRadius(a, prograde != 0); }
double ref_marginally_bound_radius(double a, int prograde) {
    return kerr::marginallyBoundRadius(a, prograde != 0);
}
double ref_circular_omega(double a, double r, int prograde) { return kerr::circularOmega(a, r, prograde != 0); }
double ref_azimuth_shift(double a, double r) { return kerr::azimuthShift(a, r); }

// consts = {E, L, Q, mu}; state = {t, r, theta, phi, vr, vtheta}; out = {dt, dr, dtheta, dphi, dvr, dvtheta}
void ref_derivatives(double a, const double* consts, const double* state, double* out) {
    kerr::Constants c{consts[0], consts[1], consts[2], consts[3]};
    kerr::State s{state[0], state[1], state[2], state[3], state[4], state[5]};
    kerr::Derivatives d = kerr::derivatives(a, c, s);
    out[0] = d.dt;
    out[1] = d.dr;
    out[2] = d.dtheta;
    out[3] = d.dphi;
    out[4] = d.dvr;
    out[5] = d.dvtheta;
}

double ref_radial_potential(double a, const double* consts, double r) {
    return kerr::radialPotential(a, {consts[0], consts[1], consts[2], consts[3]}, r);
}

double ref_polar_potential(double a, const double* consts, double theta) {
    return kerr::polarPotential(a, {consts[0], consts[1], consts[2], consts[3]}, theta);
}

void ref_enforce_constraints(double a, const double* consts, double* state) {
    kerr::State s{state[0], state[1], state[2], state[3], state[4], state[5]};
    kerr::enforceConstraints(a, {consts[0], consts[1], consts[2], consts[3]}, s);
    state[4] = s.vr;
    state[5] = s.vtheta;
}

// ZAMO frame at (r, theta): e[16] row-major e[k][mu]
void ref_zamo_frame(double a, double r, double theta, double* e) {
    kerr::Frame f = kerr::observerFrame(a, r, theta, kerr::zamoFourVelocity(a, r, theta));
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 4; ++i) e[4 * k + i] = f.e[k][i];
}

void ref_metric_tensor(double a, double r, double theta, double* g) {
    kerr::MetricTensor m = kerr::metricTensor(a, r, theta);
    std::memcpy(g, m.g, sizeof(m.g));
}

// photon seen by a ZAMO at (r, theta, phi) from direction look; consts {E,L,Q,mu}, state {t,r,th,ph,vr,vth}
void ref_photon_from_zamo(double a, double r, double theta, double phi, const double* look, double* consts,
                          double* state) {
    kerr::Frame f = kerr::observerFrame(a, r, theta, kerr::zamoFourVelocity(a, r, theta));
    kerr::Constants c;
    kerr::State s;
    kerr::photonFromFrame(a, r, theta, phi, f, look, c, s);
    consts[0] = c.E;
    consts[1] = c.L;
    consts[2] = c.Q;
    consts[3] = c.mu;
    state[0] = s.t;
    state[1] = s.r;
    state[2] = s.theta;
    state[3] = s.phi;
    state[4] = s.vr;
    state[5] = s.vtheta;
}

// eht::traceRay without the light: out = {captured, minRadius}
void ref_eht_trace(double a, double inclination, double alpha, double beta, double* out) {
    eht::View v;
    v.a = a;
    v.inclination = inclination;
    eht::Ray ray = eht::traceRay(v, eht::Flow{}, alpha, beta, false);
    out[0] = ray.captured ? 1.0 : 0.0;
    out[1] = ray.minRadius;
}

int ref_critical_curve(double a, double inclination, int samples, double* alpha, double* beta, int cap) {
    std::vector<eht::Point> c = eht::criticalCurve(a, inclination, samples);
    int n = 0;
    for (const auto& p : c) {
        if (n >= cap) break;
        alpha[n] = p.alpha;
        beta[n] = p.beta;
        ++n;
    }
    return n;
}

double ref_critical_radius(const double* alpha, const double* beta, int n, double psi) {
    std::vector<eht::Point> c(n);
    for (int k = 0; k < n; ++k) c[k] = {alpha[k], beta[k]};
    return eht::criticalRadius(c, psi);
}

// trajectory::traceRay from a ZAMO camera; out = {end, rEnd, g, turns, rMin, sky x, y, z, energy, b, q}
void ref_trajectory_trace(double a, double r, double theta, double phi, const double* look, int disk,
                          double disk_inner, double disk_outer, double escape, double* out) {
    trajectory::RayScene scene;
    scene.disk = disk != 0;
    scene.diskInner = disk_inner;
    scene.diskOuter = disk_outer;
    scene.escapeRadius = escape;
    kerr::Frame f = kerr::observerFrame(a, r, theta, kerr::zamoFourVelocity(a, r, theta));
    trajectory::RayReport rep = trajectory::traceRay(a, r, theta, phi, f, look, scene);
    out[0] = static_cast<double>(static_cast<int>(rep.end));
    out[1] = rep.rEnd;
    out[2] = rep.g;
    out[3] = rep.turns;
    out[4] = rep.rMin;
    out[5] = rep.skyDirection[0];
    out[6] = rep.skyDirection[1];
    out[7] = rep.skyDirection[2];
    out[8] = rep.energy;
    out[9] = rep.b;
    out[10] = rep.q;
}

// trajectory::traceRay from a ZAMO camera with the whole RayReport the inspector prints
// (src/app.cpp:776-777); out = {end, rEnd, g, turns, rMin, delay, energy, b, q}
void ref_trajectory_report(double a, double r, double theta, double phi, const double* look, int disk,
                           double disk_inner, double disk_outer, double escape, double* out) {
    trajectory::RayScene scene;
    scene.disk = disk != 0;
    scene.diskInner = disk_inner;
    scene.diskOuter = disk_outer;
    scene.escapeRadius = escape;
    kerr::Frame f = kerr::observerFrame(a, r, theta, kerr::zamoFourVelocity(a, r, theta));
    trajectory::RayReport rep = trajectory::traceRay(a, r, theta, phi, f, look, scene);
    out[0] = static_cast<double>(static_cast<int>(rep.end));
    out[1] = rep.rEnd;
    out[2] = rep.g;
    out[3] = rep.turns;
    out[4] = rep.rMin;
    out[5] = rep.delay;
    out[6] = rep.energy;
    out[7] = rep.b;
    out[8] = rep.q;
}

double ref_bending_angle(double R, double alpha) { return central::bendingAngle(R, alpha); }

}  // extern "C"

// ---- EHT radiative transfer: the physics functions and traceRay with the light ----------

static eht::Flow make_flow(const double* f) {
    eht::Flow flow;
    flow.n0 = f[0];
    flow.T0 = f[1];
    flow.beta = f[2];
    flow.keplerian = f[3];
    flow.height = f[4];
    flow.sinThetaB = f[5];
    return flow;
}

extern "C" {

double ref_bessel_k(int which, double x) {
    return which == 0 ? eht::besselK0(x) : which == 1 ? eht::besselK1(x) : eht::besselK2(x);
}

double ref_synchrotron_emissivity(double nu, double ne, double thetaE, double B, double s) {
    return eht::synchrotronEmissivity(nu, ne, thetaE, B, s);
}

double ref_planck_nu(double nu, double T) { return eht::planck(nu, T); }

void ref_plasma_at(const double* flow, double r, double theta, double a, double* out) {
    eht::Plasma p = eht::plasmaAt(make_flow(flow), r, theta, a);
    out[0] = p.ne;
    out[1] = p.Te;
    out[2] = p.B;
}

void ref_flow_velocity(const double* flow, double a, double r, double theta, double* u) {
    eht::flowVelocity(make_flow(flow), a, r, theta, u);
}

// out = {intensity, captured, minRadius}
void ref_eht_trace_light(double a, double inclination, double frequency, double mass_solar, const double* flow,
                         double alpha, double beta, double* out) {
    eht::View v;
    v.a = a;
    v.inclination = inclination;
    v.frequency = frequency;
    v.massSolar = mass_solar;
    eht::Ray ray = eht::traceRay(v, make_flow(flow), alpha, beta, true);
    out[0] = ray.intensity;
    out[1] = ray.captured ? 1.0 : 0.0;
    out[2] = ray.minRadius;
}

double ref_eht_calibrate(double a, double inclination, double half_width, double frequency, double mass_solar,
                         const double* flow, double target_jy, double theta_guas) {
    eht::View v;
    v.a = a;
    v.inclination = inclination;
    v.halfWidth = half_width;
    v.frequency = frequency;
    v.massSolar = mass_solar;
    return eht::calibrate(v, make_flow(flow), target_jy, theta_guas, 8, nullptr);
}

}  // extern "C"

// ---- observers other than the static ZAMO: boosted, falling, inside the horizon ----------

extern "C" {

// kerr::fourVelocity; consts = {E, L, Q, mu}, state = {t, r, theta, phi, vr, vtheta}
void ref_four_velocity(double a, const double* consts, const double* state, double* u) {
    kerr::Vec4 v = kerr::fourVelocity(a, {consts[0], consts[1], consts[2], consts[3]},
                                      {state[0], state[1], state[2], state[3], state[4], state[5]});
    for (int i = 0; i < 4; ++i) u[i] = v[i];
}

// kerr::observerFrame of the observer of 4-velocity u (Kerr-Schild): e[4 k + i] = e_k^i
void ref_observer_frame(double a, double r, double theta, const double* u, double* e) {
    kerr::Frame f = kerr::observerFrame(a, r, theta, {u[0], u[1], u[2], u[3]});
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 4; ++i) e[4 * k + i] = f.e[k][i];
}

// trajectory::traceRay from the frame of the observer of 4-velocity u (Kerr-Schild);
// out as ref_trajectory_trace
void ref_trajectory_trace_u(double a, double r, double theta, double phi, const double* u, const double* look,
                            int disk, double disk_inner, double disk_outer, double escape, double* out) {
    trajectory::RayScene scene;
    scene.disk = disk != 0;
    scene.diskInner = disk_inner;
    scene.diskOuter = disk_outer;
    scene.escapeRadius = escape;
    kerr::Frame f = kerr::observerFrame(a, r, theta, {u[0], u[1], u[2], u[3]});
    trajectory::RayReport rep = trajectory::traceRay(a, r, theta, phi, f, look, scene);
    out[0] = static_cast<double>(static_cast<int>(rep.end));
    out[1] = rep.rEnd;
    out[2] = rep.g;
    out[3] = rep.turns;
    out[4] = rep.rMin;
    out[5] = rep.skyDirection[0];
    out[6] = rep.skyDirection[1];
    out[7] = rep.skyDirection[2];
    out[8] = rep.energy;
    out[9] = rep.b;
    out[10] = rep.q;
}

}  // extern "C"

// ---- massive particles --------------------------------------------------------------------

extern "C" {

// kerr::launchFromZamo with mu = 1: out = {E, L, Q, vr, vtheta}
void ref_particle_from_zamo(double a, double r, double theta, const double* v, double* out) {
    kerr::Constants c;
    kerr::State s;
    kerr::launchFromZamo(a, r, theta, 0.0, v, 1.0, c, s);
    out[0] = c.E;
    out[1] = c.L;
    out[2] = c.Q;
    out[3] = s.vr;
    out[4] = s.vtheta;
}

// trajectory::predict up to max_time of coordinate time (escape beyond `escape`):
// out = {fate, x, y, z of the last point, rMin, rMax, periapses, precessionDeg, duration}
void ref_predict(double a, const double* consts, const double* state, double max_time, double escape, double* out) {
    kerr::Constants c{consts[0], consts[1], consts[2], consts[3]};
    kerr::State s{state[0], state[1], state[2], state[3], state[4], state[5]};
    trajectory::Path p = trajectory::predict(a, c, s, max_time, 1e30, 4000, 0.0, escape);
    const auto& x = p.points.back();
    out[0] = static_cast<double>(static_cast<int>(p.fate));
    out[1] = x[0];
    out[2] = x[1];
    out[3] = x[2];
    out[4] = p.rMin;
    out[5] = p.rMax;
    out[6] = p.periapses;
    out[7] = p.precessionDeg;
    out[8] = p.duration;
}

}  // extern "C"
