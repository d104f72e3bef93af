/*
 * kerr_oracle.h -- FP64 CPU oracle for the Kerr per-pixel photon trace.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only the
 * test suite (tests/), the smoke check in __graft_entry__.py and the
 * cpu_baseline leg of bench.py may load it.  The HIP library under
 * black_simulation_amd/ never links or calls anything in oracle/.
 *
 * What it restates (all citations are files of Dhruv-Deku/Black_simulation):
 *   - characteristic radii                  src/kerr.cpp:157-223
 *   - circular-orbit angular velocity       src/kerr.cpp:296-299
 *   - Mino-time geodesic right-hand side    src/kerr.cpp:323-351 (ko_derivatives, reference form)
 *   - radial / polar potentials             src/kerr.cpp:353-363
 *   - metric tensor, lower, observer frame  src/kerr.cpp:238-270, 424-456
 *   - photon launch from a frame            src/kerr.cpp:458-483 (axis snap kerr.h:115)
 *   - constraint projection                 src/kerr.cpp:521-542
 *   - Bardeen image-plane launch            src/eht.cpp:262-271, pixel centres eht.cpp:332-334
 *   - threaded row-claiming image driver    src/eht.cpp:321-347
 *   - critical curve / critical radius      src/eht.cpp:52-97
 *   - disk crossing + Keplerian redshift    src/trajectory.cpp:307-324
 *   - capture / escape tests                src/trajectory.cpp:377-392, src/eht.cpp:285-289
 *   - camera basis (FP64 here, FP32 glm in the reference)  src/camera.cpp:6-21
 *
 * New (no reference counterpart, pinned by closed forms and self-convergence):
 *   - the integrated system: Carter's equations in Mino time in (r, mu = cos theta),
 *     traced backward (s = -tau), with Boyer-Lindquist phi;
 *   - Dormand-Prince 5(4) with dense output (Hairer, Norsett & Wanner, Solving
 *     ODEs I, sec. II.5 and II.6), a max-norm error controller, a dense-output root
 *     for the disk crossing and an exact tail integration in u = 1/r to u = 0
 *     for the asymptotic sky angles.
 *
 * Arithmetic is plain IEEE binary64 with FP contraction disabled; the step loop
 * uses only +, -, *, /, sqrt and explicit fma, each correctly rounded, so an
 * implementation that performs the same operations in the same order gets the
 * same bits.
 */
#ifndef KERR_ORACLE_H
#define KERR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KO_PI 3.14159265358979323846

/* ray classes (kt_class in include/kerr_trace.h uses the same codes) */
enum { KO_ESCAPED = 0, KO_CAPTURED = 1, KO_DISK = 2, KO_UNFINISHED = 3 };
enum { KO_FLAG_NONFINITE = 0x40 };

enum { KO_CAMERA_BARDEEN = 0, KO_CAMERA_PINHOLE = 1, KO_CAMERA_OBSERVER = 3 };

/* ---- radii and circular orbits (kerr.cpp:157-223, 296-299) ---- */
double ko_horizon_radius(double a);
double ko_inner_horizon_radius(double a);
double ko_isco_radius(double a, int prograde);
double ko_photon_orbit_radius(double a, int prograde);
double ko_marginally_bound_radius(double a, int prograde);
double ko_circular_omega(double a, double r, int prograde);

/* ---- reference-form geodesic equations (kerr.cpp:323-363) ---- */
typedef struct { double E, L, Q, mu; } ko_consts;
typedef struct { double t, r, theta, phi, vr, vtheta; } ko_state;
typedef struct { double dt, dr, dtheta, dphi, dvr, dvtheta; } ko_derivs;

double ko_inflow_factor(double P, double vr, double delta, double K);
ko_derivs ko_derivatives(double a, ko_consts c, ko_state s);
double ko_radial_potential(double a, ko_consts c, double r);
double ko_polar_potential(double a, ko_consts c, double theta);
void ko_enforce_constraints(double a, ko_consts c, ko_state* s);

/* ---- metric, frames, launch (kerr.cpp:227-270, 424-483) ---- */
void ko_metric_tensor(double a, double r, double theta, double g[4][4]);
void ko_zamo_four_velocity(double a, double r, double theta, double u[4]);
void ko_observer_frame(double a, double r, double theta, const double u[4], double e[4][4]);
void ko_photon_from_frame(double a, double r, double theta, double phi, const double e[4][4],
                          const double look[3], ko_consts* c, ko_state* s);
/* camera basis of src/camera.cpp:6-21 as a column-major 3x3 (right, up, forward),
   components along the frame's (r, theta, phi) axes */
void ko_camera_basis(double yaw, double pitch, double roll, double basis[9]);

/* ---- the integrated system: backward Mino time, (r, mu, phi, w_r, w_mu) ---- */
typedef struct {
    double a, a2, E, L, Q;
    double aE, aL, K, C1, A2E2, twoE, twoA2E2;
    int ks; /* phi carried: 0 Boyer-Lindquist, 1 ingoing Kerr-Schild (observer cameras),
               2 outgoing Kerr-Schild (how ko_trace_photon traces a Boyer-Lindquist launch) */
} ko_photon;

void ko_photon_init(ko_photon* p, double a, double E, double L, double Q);
/* dy/ds for y = (r, mu, phi_BL, w_r = dr/ds, w_mu = dmu/ds), s = -tau */
void ko_rhs(const ko_photon* p, const double y[5], double dy[5]);
/* R(r) and M(mu) = sin^2(theta) Theta(theta) for the photon */
double ko_R(const ko_photon* p, double r);
double ko_M(const ko_photon* p, double m);

/* ---- views and the tracer ---- */
typedef struct {
    int camera_kind;           /* KO_CAMERA_BARDEEN or KO_CAMERA_PINHOLE */
    int width, height;
    double spin;
    double r_cam, theta_cam, phi_cam;
    double half_width, half_height;          /* Bardeen image plane (M)           */
    double yaw, pitch, roll, fov_deg;        /* pinhole camera (ZAMO frame)        */
    double disk_inner, disk_outer;           /* disk_inner <= 0 means r_ISCO       */
    int disk_prograde, disk_enabled;
    double escape_radius;                    /* effective: max(this, 1.5 r_cam)    */
    double capture_factor;                   /* captured when r < this * r_+       */
    double u_obs[4];                         /* observer: KS 4-velocity, or zero   */
    double drift[3];                         /* observer, u_obs zero: ZAMO boost   */
    double frame[16];                        /* observer: e_a^mu given, or zero    */
} ko_view;

typedef struct {
    double rtol, atol;
    int max_steps;
} ko_opts;

typedef struct {
    int cls, flags, steps;
    double theta_inf, phi_inf, r_hit, g;
    double y_end[5];            /* state at termination (tail: u = 0 -> y_end[0] = inf) */
} ko_result;

/* launch helpers: initial photon + state for pixel (col, row) of a view */
void ko_launch_pixel(const ko_view* v, int col, int row, ko_photon* p, double y0[5]);
void ko_launch_bardeen(const ko_view* v, double alpha, double beta, ko_photon* p, double y0[5]);
/* pinhole view, explicit unit look vector on the ZAMO frame's (r, theta, phi) axes */
void ko_launch_look(const ko_view* v, const double look[3], ko_photon* p, double y0[5]);
/* ---- massive particles (particles_update.comp:207-255, kerr.cpp:544-561, 365-406) ---- */
enum { KO_PARTICLE_MOVING = 0, KO_PARTICLE_PLUNGED = 1, KO_PARTICLE_ESCAPED = 2, KO_PARTICLE_UNFINISHED = 3 };
/* advances st = (r, mu, phi_KS, vr, vmu) of a particle with constants E, L, Q by dt of
   ingoing Kerr-Schild time; returns its fate */
int ko_particle_advance(double a, double E, double L, double Q, double st[5], double dt, double r_stop,
                        double r_escape, const ko_opts* o, int* steps, double* h); /* h: NULL or in/out */
void ko_particle_from_zamo(double a, double r, double theta, const double v[3], double out[5]);

/* observer views: the camera's 4-velocity (camera.cpp:23-36) */
void ko_observer_u(const ko_view* v, double u[4]);
/* phi_KS - phi_BL at r (zero at infinity) */
double ko_ks_shift(double a, double r);

void ko_trace_photon(const ko_view* v, const ko_opts* o, const ko_photon* p, const double y0[5],
                     ko_result* out);
void ko_trace_pixel(const ko_view* v, const ko_opts* o, int col, int row, ko_result* out);

/* the single-ray report (trajectory::RayReport, trajectory.h:97-113): the angle the
   position's direction sweeps over 2 pi, the least r and the ingoing Kerr-Schild time from
   the event to the camera, each up to the event (disk crossing, r_cap, r_esc); b = L/E and
   q = Q/E^2 (0 for |E| <= 1e-12).  ko_trace_photon with the report beside it. */
typedef struct { double turns, r_min, delay, b, q; } ko_report;
void ko_trace_photon_report(const ko_view* v, const ko_opts* o, const ko_photon* p, const double y0[5],
                            ko_result* out, ko_report* rep);

/* whole image, `threads` workers claiming rows from an atomic counter (eht.cpp:321-347).
   Outputs are row-major from the top row; any pointer may be NULL. `stride` > 1 traces
   only pixels with (row * width + col) % stride == 0 (a sparse timing sample) and writes
   them densely.  Returns the number of rays traced, or -1 if cancelled. */
long ko_render(const ko_view* v, const ko_opts* o, int threads, const volatile int* cancel, int stride,
               int8_t* cls, double* theta_inf, double* phi_inf, double* r_hit, double* g, int32_t* steps);

/* ---- critical curve (eht.cpp:52-97) ---- */
int ko_critical_curve(double a, double inclination, int samples, double* alpha, double* beta, int cap);
double ko_critical_radius(const double* alpha, const double* beta, int n, double psi);

/* the controller's deterministic x^(-1/5) */
double ko_pow_m15(double x);

/* ---- HDR colour of a traced ray (trace.frag:330-335, 373-403, 1419-1458, 1893-1906;
   blackbody.cpp:16-60) ---- */
#define KO_LOGT_MIN 2.5
#define KO_LOGT_MAX 5.3
/* CIE XYZ of Planck radiance at T, 360-830 nm in 2 nm steps (blackbody.cpp:16-41) */
void ko_planck_xyz(double T, double xyz[3]);
/* n linear-sRGB triples over log10 T in [2.5, 5.3], luminance 1 at 6500 K (blackbody.cpp:48-60) */
void ko_blackbody_table(int n, float* rgb);
typedef struct ko_shade {
    double disk_temperature, disk_brightness;
    int sky;                      /* 0 black, 1 grid, 2 image */
    double sky_brightness, grid_line;
    const float* sky_image;       /* equirectangular RGB, row 0 at latitude -90 deg */
    int sky_width, sky_height;
    const float* table;           /* ko_blackbody_table output */
    int table_n;
} ko_shade;
/* ---- EHT radiative transfer (eht.h:88-140; eht.cpp:98-354) ---- */
typedef struct ko_flow {
    double n0, T0, beta, keplerian, height, sin_theta_b; /* eht::Flow (eht.h:97-104) */
} ko_flow;
typedef struct ko_eht_view {
    double a, inclination, half_width; /* eht::View (eht.h:113-120) */
    int size;
    double frequency, mass_solar;
} ko_eht_view;
typedef struct ko_eht_ray {
    double intensity, min_radius; /* erg s^-1 cm^-2 Hz^-1 sr^-1; M */
    int captured, steps;
} ko_eht_ray;
double ko_bessel_k(int which, double x); /* K0, K1, K2 (A&S 9.8.5-9.8.8) */
double ko_synchrotron_emissivity(double nu, double ne, double thetaE, double B, double sinThetaB);
double ko_planck_nu(double nu, double T);
void ko_plasma_at(const ko_flow* f, double r, double theta, double a, double out[3]); /* ne, Te, B */
void ko_flow_velocity(const ko_flow* f, double a, double r, double theta, double u[4]);
void ko_eht_trace(const ko_eht_view* v, const ko_flow* f, const ko_opts* o, double alpha, double beta,
                  ko_eht_ray* out);
void ko_eht_render(const ko_eht_view* v, const ko_flow* f, const ko_opts* o, double* intensity, uint8_t* shadow);
double ko_eht_flux(const double* intensity, int size, double half_width, double theta_guas);
/* eht.cpp:432-457: secant in log n0 on 48 x 48 renders */
double ko_eht_calibrate(const ko_eht_view* v, const ko_flow* flow, const ko_opts* o, double target_jy,
                        double theta_guas);

/* colour of one ray from its traced class, sky angles, disk radius and g */
void ko_shade_ray(const ko_shade* s, int cls, double theta, double phi, double r_hit, double g, double r_in,
                  double r_out, double rgb[3]);

#ifdef __cplusplus
}
#endif
#endif
