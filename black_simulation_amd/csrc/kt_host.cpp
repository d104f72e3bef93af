// kt_host.cpp -- host side of libkerr_trace.so: the C ABI of include/kerr_trace.h.
//
// Prepares per-view constants in FP64 (trig, ZAMO frame, camera basis), owns the
// device scratch (work queue head, counters, view table, staging for host
// outputs), launches the trace kernels on the caller's stream and times them
// with HIP events on that same stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/kerr_trace.h"
#include "kt_internal.hpp"
#include "kt_surrogate.hpp"
#include "kt_nn.hpp"
#include "kt_shade.hpp"
#include "kt_eht.hpp"
#include "kt_particles.hpp"

using ktd::ViewDev;

namespace {

constexpr double kPi = 3.14159265358979323846;

// sin and cos each behind its own call: compilers otherwise fuse sin(x), cos(x) into
// sincos(x), whose glibc result can differ from sin(x) in the last bit.
__attribute__((noinline)) double hsin(double x) { return std::sin(x); }
__attribute__((noinline)) double hcos(double x) { return std::cos(x); }

// phi_in - phi_BL = psi(r) = integral from infinity to r of a / Delta (zero at infinity);
// phi_out - phi_BL = -psi(r).  The same operations as the device's ks_shift.  Past |a| = 1
// Delta = (r - 1)^2 + b^2 (b^2 = a^2 - 1) has no zero and psi = (a / b)(atan((r - 1) / b) -
// pi / 2), finite for every r (src/kerr.cpp:284-291 writes the same function without the
// constant, which only moves the Kerr-Schild phi's origin).
double ks_shift_host(double a, double r) {
    if (a == 0.0) return 0.0;
    if (std::fabs(a) > 1.0) {
        const double b = std::sqrt(a * a - 1.0);
        return a / b * (std::atan((r - 1.0) / b) - 0.5 * kPi);
    }
    const double d = 2.0 * std::sqrt(std::fmax(0.0, 1.0 - a * a));
    const double rm = 1.0 - 0.5 * d;
    if (d == 0.0) return -a / (r - 1.0);
    return a / d * std::log1p(-d / (r - rm));
}

// ---- closed forms (src/kerr.cpp:157-206, 296-299) ----

double horizon(double a) {
    if (std::fabs(a) > 1.0) return 0.0;
    return 1.0 + std::sqrt(std::fmax(0.0, 1.0 - a * a));
}

// ---- naked singularities, 1 < a <= 1.5 (src/kerr.cpp:117-153, 157-206): no horizon ----
// The ISCO is the root of Bardeen, Press and Teukolsky's marginal-stability condition
// r^2 - 6 r + 8 s a sqrt(r) - 3 a^2 = 0 (dE/dr = 0 of the circular orbits; s = 1 prograde),
// which holds past a = 1 too: in x = sqrt(r), f(x) = x^4 - 6 x^2 + 8 s a x - 3 a^2.  Prograde,
// f increases from f(0) < 0 to f(1) > 0 (a < 5/3): the root lies inside r = 1.  Retrograde,
// f falls from f(sqrt 3) < 0 to its least value and then rises through zero once before
// x = 6.  The reference finds the same radius as the least circular-orbit energy by a scan
// and a golden-section search, good to ~1e-8; this bisection ends at adjacent doubles.
// Circular photon orbits: r (r - 3)^2 = 4 a^2, only the retrograde one past a = 1 (the
// prograde radius is reported as 0, as the reference does).
constexpr double kMaxNakedSpin = 1.5;

bool naked(double a) { return std::fabs(a) > 1.0; }

template <class F>
double bisect_root(F f, double lo, double hi) {  // f(lo) < 0 <= f(hi), one sign change
    for (int k = 0; k < 200; ++k) {
        const double mid = 0.5 * (lo + hi);
        if (!(mid > lo && mid < hi)) break;
        if (f(mid) < 0.0) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

double isco_naked(double a, int prograde) {
    const double sa = prograde ? a : -a, a2 = a * a;
    const auto f = [=](double x) { return ((x * x - 6.0) * x + 8.0 * sa) * x - 3.0 * a2; };
    const double x = prograde ? bisect_root(f, 0.0, 1.0) : bisect_root(f, std::sqrt(3.0), 6.0);
    return x * x;
}

double photon_orbit_naked(double a, int prograde) {
    if (prograde) return 0.0;
    const double a2 = a * a;
    return bisect_root([=](double r) { return r * (r - 3.0) * (r - 3.0) - 4.0 * a2; }, 3.0, 3.0 + 2.0 * std::fabs(a));
}

double isco(double a, int prograde) {
    if (naked(a)) return isco_naked(a, prograde);
    double z1 = 1.0 + std::cbrt(1.0 - a * a) * (std::cbrt(1.0 + a) + std::cbrt(1.0 - a));
    double z2 = std::sqrt(3.0 * a * a + z1 * z1);
    double root = std::sqrt(std::fmax(0.0, (3.0 - z1) * (3.0 + z1 + 2.0 * z2)));
    return prograde ? 3.0 + z2 - root : 3.0 + z2 + root;
}

double photon_orbit(double a, int prograde) {
    if (naked(a)) return photon_orbit_naked(a, prograde);
    double s = prograde ? -1.0 : 1.0;
    return 2.0 * (1.0 + hcos(2.0 / 3.0 * std::acos(s * a)));
}

// ---- Kerr-Schild metric and the ZAMO frame (src/kerr.cpp:238-270, 424-456) ----

void metric_tensor(double a, double r, double theta, double g[4][4]) {
    const double s = hsin(theta), c = hcos(theta);
    const double s2 = s * s;
    const double sigma = r * r + a * a * c * c;
    const double rra = r * r + a * a;
    const double A = rra * rra - a * a * (r * r - 2.0 * r + a * a) * s2;
    const double f = 2.0 * r / sigma;
    std::memset(g, 0, 16 * sizeof(double));
    g[0][0] = -(1.0 - f);
    g[0][1] = g[1][0] = f;
    g[0][3] = g[3][0] = -f * a * s2;
    g[1][1] = 1.0 + f;
    g[1][3] = g[3][1] = -a * s2 * (1.0 + f);
    g[2][2] = sigma;
    g[3][3] = A * s2 / sigma;
}

double dot4(const double g[4][4], const double* x, const double* y) {
    double sum = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) sum += g[i][j] * x[i] * y[j];
    return sum;
}

void lower4(const double g[4][4], const double* x, double* out) {
    for (int i = 0; i < 4; ++i) {
        out[i] = 0.0;
        for (int j = 0; j < 4; ++j) out[i] += g[i][j] * x[j];
    }
}

void zamo_u(double a, double r, double theta, double u[4]) {
    const double s = hsin(theta), c = hcos(theta);
    const double sigma = r * r + a * a * c * c;
    const double delta = r * r - 2.0 * r + a * a;
    const double rra = r * r + a * a;
    const double A = rra * rra - a * a * delta * s * s;
    const double omega = 2.0 * a * r / A;
    const double ut = std::sqrt(A / (sigma * delta));
    u[0] = ut;
    u[1] = 0.0;
    u[2] = 0.0;
    u[3] = omega * ut;
}

void observer_frame(double a, double r, double theta, const double u[4], double e[4][4]) {
    double g[4][4];
    metric_tensor(a, r, theta, g);
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
        const double n = std::sqrt(std::fmax(dot4(g, w, w), 1e-300));
        for (int i = 0; i < 4; ++i) w[i] /= n;
        for (int i = 0; i < 4; ++i) e[slot[k]][i] = w[i];
    }
}

// camera basis, src/camera.cpp:6-21 (FP64): columns right, up, forward on (r, theta, phi)
void camera_basis(double yaw, double pitch, double roll, double B[9]) {
    const double f0[3] = {-1.0, 0.0, 0.0}, u0[3] = {0.0, -1.0, 0.0}, r0[3] = {0.0, 0.0, 1.0};
    const double cy = hcos(yaw), sy = hsin(yaw), cp = hcos(pitch), sp = hsin(pitch);
    const double cr = hcos(roll), sr = hsin(roll);
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

bool has_u_obs(const kt_view& v) { return v.u_obs[0] != 0.0 || v.u_obs[1] != 0.0 || v.u_obs[2] != 0.0 || v.u_obs[3] != 0.0; }

bool has_frame(const kt_view& v) {
    for (double x : v.frame)
        if (x != 0.0) return true;
    return false;
}

// a caller's observer frame: orthonormal (g(e_a, e_b) = diag(-1, 1, 1, 1) to 1e-8) and
// future-directed (e_0^t > 0)
bool frame_ok(const kt_view& v) {
    double g[4][4];
    metric_tensor(v.spin, v.r_cam, v.theta_cam, g);
    const double* e = v.frame;
    if (!(e[0] > 0.0)) return false;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const double d = dot4(g, e + 4 * i, e + 4 * j) - (i != j ? 0.0 : i == 0 ? -1.0 : 1.0);
            if (!(std::fabs(d) <= 1e-8)) return false;
        }
    return true;
}

// the observer of an observer view: u_obs normalised, or the ZAMO boosted by drift
// (src/camera.cpp:23-36); false if it is not a future-directed timelike vector
bool observer_u(const kt_view& v, double u[4]) {
    const double a = v.spin, r = v.r_cam, t = v.theta_cam;
    double g[4][4];
    metric_tensor(a, r, t, g);
    if (has_u_obs(v)) {
        for (int i = 0; i < 4; ++i) u[i] = v.u_obs[i];
        const double n = dot4(g, u, u);
        if (!(n < 0.0) || !(u[0] > 0.0)) return false;
        const double s = 1.0 / std::sqrt(-n);
        for (int i = 0; i < 4; ++i) u[i] *= s;
        return true;
    }
    zamo_u(a, r, t, u);
    const double v2 = v.drift[0] * v.drift[0] + v.drift[1] * v.drift[1] + v.drift[2] * v.drift[2];
    if (v2 > 0.0 && v2 < 1.0) {  // u = gamma (e0 + v^i e_i) of the ZAMO frame
        double e[4][4];
        observer_frame(a, r, t, u, e);
        const double gamma = 1.0 / std::sqrt(1.0 - v2);
        for (int i = 0; i < 4; ++i)
            u[i] = gamma * (e[0][i] + v.drift[0] * e[1][i] + v.drift[1] * e[2][i] + v.drift[2] * e[3][i]);
    }
    return true;
}

const char* check_view(const kt_view& v) {
    if (v.camera_kind != KT_CAMERA_BARDEEN && v.camera_kind != KT_CAMERA_PINHOLE &&
        v.camera_kind != KT_CAMERA_OBSERVER)
        return "unknown camera_kind";
    // every real field finite: a NaN or infinity would reach the rays unnoticed (NaN
    // fails no range test written as !(x in range), but phi, the angles and the frame have
    // no range to test)
    const double scalars[] = {v.spin, v.r_cam, v.theta_cam, v.phi_cam, v.half_width, v.half_height,
                              v.yaw, v.pitch, v.roll, v.fov_deg, v.disk_inner, v.disk_outer,
                              v.escape_radius, v.capture_factor};
    for (double x : scalars)
        if (!std::isfinite(x)) return "view fields must be finite";
    for (double x : v.u_obs)
        if (!std::isfinite(x)) return "u_obs must be finite";
    for (double x : v.drift)
        if (!std::isfinite(x)) return "drift must be finite";
    for (double x : v.frame)
        if (!std::isfinite(x)) return "frame must be finite";
    if (v.width <= 0 || v.height <= 0) return "image size must be positive";
    if ((long long)v.width * v.height > (1ll << 31)) return "view larger than 2^31 pixels";
    if (!(v.spin >= -1.0 && v.spin <= kMaxNakedSpin))
        return "spin must lie in [-1, 1.5] (past 1: a naked singularity)";
    if (v.camera_kind == KT_CAMERA_OBSERVER) {
        if (naked(v.spin)) return "observer cameras need a horizon: naked singularities take Bardeen and pinhole views";
        if (!(v.fov_deg > 0.0 && v.fov_deg < 180.0)) return "fov out of range";
        if (!(v.theta_cam > 0.0 && v.theta_cam < kPi)) return "an observer camera must be off the axis";
        if (has_frame(v)) {
            if (!(v.r_cam > 1.0 - std::sqrt(std::fmax(0.0, 1.0 - v.spin * v.spin)) + 1e-9))
                return "an observer camera must be outside the inner horizon";
            if (!frame_ok(v)) return "frame must be orthonormal (to 1e-8) with a future-directed e_0";
            if (!(v.capture_factor >= 1.0)) return "capture_factor must be >= 1";
            if (v.disk_enabled && !(v.disk_outer > 0.0)) return "disk_outer must be positive";
            return nullptr;
        }
        if (has_u_obs(v)) {
            if (!(v.r_cam > 1.0 - std::sqrt(std::fmax(0.0, 1.0 - v.spin * v.spin)) + 1e-9))
                return "an observer camera must be outside the inner horizon";
        } else {
            if (!(v.r_cam > horizon(v.spin))) return "no observer hovers inside the horizon: give u_obs";
            const double v2 = v.drift[0] * v.drift[0] + v.drift[1] * v.drift[1] + v.drift[2] * v.drift[2];
            if (!(v2 < 1.0)) return "drift must be slower than light";
        }
        double u[4];
        if (!observer_u(v, u)) return "u_obs must be a future-directed timelike vector (u.u < 0, u^t > 0)";
        if (!(v.capture_factor >= 1.0)) return "capture_factor must be >= 1";
        if (v.disk_enabled && !(v.disk_outer > 0.0)) return "disk_outer must be positive";
        return nullptr;
    }
    if (!(v.r_cam > horizon(v.spin))) return "camera must be outside the horizon";
    if (!(v.theta_cam >= 0.0 && v.theta_cam <= kPi)) return "theta_cam must lie in [0, pi]";
    if (!(v.capture_factor >= 1.0)) return "capture_factor must be >= 1";
    if (v.camera_kind == KT_CAMERA_BARDEEN && !(v.half_width > 0.0 && v.half_height > 0.0))
        return "Bardeen half widths must be positive";
    if (v.camera_kind == KT_CAMERA_PINHOLE && !(v.fov_deg > 0.0 && v.fov_deg < 180.0)) return "fov out of range";
    if (v.disk_enabled && !(v.disk_outer > 0.0)) return "disk_outer must be positive";
    return nullptr;
}

ViewDev make_view_dev(const kt_view& v, long long offset) {
    ViewDev d;
    std::memset(&d, 0, sizeof d);
    const double a = v.spin;
    d.offset = offset;
    d.kind = v.camera_kind;
    d.width = v.width;
    d.height = v.height;
    d.disk = v.disk_enabled ? 1 : 0;
    d.prograde = v.disk_prograde ? 1 : 0;
    d.a = a;
    d.r_cam = v.r_cam;
    d.phi_cam = v.phi_cam;
    // the integrator carries the outgoing Kerr-Schild phi (ingoing for observer cameras)
    d.phi0 = v.camera_kind == KT_CAMERA_OBSERVER ? v.phi_cam : v.phi_cam - ks_shift_host(a, v.r_cam);
    d.half_width = v.half_width;
    d.half_height = v.half_height;
    d.px = 2.0 * v.half_width / (double)v.width;
    d.py = 2.0 * v.half_height / (double)v.height;
    d.sinI = hsin(v.theta_cam);
    d.cosI = hcos(v.theta_cam);
    double th = v.theta_cam;
    if (th < 1e-4) th = 1e-4;
    if (th > kPi - 1e-4) th = kPi - 1e-4;
    d.m0 = hcos(th);
    d.r_cap = v.capture_factor * horizon(a);
    d.r_esc = v.escape_radius > 1.5 * v.r_cam ? v.escape_radius : 1.5 * v.r_cam;
    d.r_in = v.disk_inner > 0.0 ? v.disk_inner : isco(a, v.disk_prograde);
    d.r_out = v.disk_outer;
    d.h0 = 0.01 / v.r_cam;
    d.r_fp64 = 1.2 * horizon(a);
    d.r_plus = horizon(a);
    if (v.camera_kind == KT_CAMERA_PINHOLE || v.camera_kind == KT_CAMERA_OBSERVER) {
        const double r = v.r_cam, t = v.theta_cam;
        double u[4], e[4][4], g[4][4], low[4][4];
        if (v.camera_kind == KT_CAMERA_OBSERVER && has_frame(v)) {  // as given (uCamPhoton)
            for (int j = 0; j < 4; ++j)
                for (int i = 0; i < 4; ++i) e[j][i] = v.frame[4 * j + i];
        } else {
            if (v.camera_kind == KT_CAMERA_OBSERVER) observer_u(v, u);
            else zamo_u(a, r, t, u);
            observer_frame(a, r, t, u, e);
        }
        metric_tensor(a, r, t, g);
        for (int j = 0; j < 4; ++j) lower4(g, e[j], low[j]);
        for (int j = 0; j < 4; ++j) {
            d.lt[j] = low[j][0];
            d.lp[j] = low[j][3];
            d.lh[j] = low[j][2];
            d.er[j] = e[j][1];
        }
        camera_basis(v.yaw, v.pitch, v.roll, d.B);
        d.tanh_fov = std::tan(0.5 * v.fov_deg * kPi / 180.0);
        d.sn = hsin(t);
        d.cs = hcos(t);
        d.sigma = r * r + a * a * d.cs * d.cs;
    }
    d.params = nullptr;
    return d;
}

}  // namespace

struct kt_ctx {
    int device = 0;
    int cu_count = 0;
    std::string err;
    kt_stats stats{};
    unsigned long long* d_counter = nullptr;  // [0] queue head, [1..] stats
    ViewDev* d_views = nullptr;
    int views_cap = 0;
    void* d_stage = nullptr;  // staging for host outputs
    size_t stage_cap = 0;
    void* d_params = nullptr;  // staging for host explicit-ray parameters
    size_t params_cap = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool timed = false;
    kts::Surrogate* surrogate = nullptr;
    kts::Queue cf_queue;  // the closed form's queue of declined rays (kt_trace_views)
    kti::Scratch scratch{};  // the no-allocation entry point's fixed device arrays (per device)
    float* d_blackbody = nullptr;  // shading: the blackbody colour table
    std::vector<float> h_blackbody;
    double* d_eht = nullptr;  // EHT: flux partials, the sum, calibrate's 48 x 48 image
    // autotuned launch shape per precision (0 = not tuned)
    unsigned long long* stats_src = nullptr;  // counters of the last launch (a plan's or d_counter)
    int tuned_mode[3] = {0, 0, 0};  // 1 + KT_MODE_* once tuned
    int tuned_block[3] = {0, 0, 0}, tuned_wpc[3] = {0, 0, 0}, tuned_spd[3] = {0, 0, 0}, tuned_wps[3] = {0, 0, 0};
};

struct kt_plan {
    kt_ctx* ctx = nullptr;
    ViewDev* d_views = nullptr;
    unsigned long long* d_counter = nullptr;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    kti::TraceLaunch L{};
    kts::Queue queue;           // the surrogate's fallback queue, owned by the plan
    hipEvent_t ev[2] = {};      // around the last replay
    bool timed = false;
    int surrogate = 0;
};

namespace {

int fail(kt_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(kt_ctx* c, hipError_t e, const char* where) {
    return fail(c, e == hipErrorOutOfMemory ? KT_ENOMEM : KT_EHIP,
                std::string(where) + ": " + hipGetErrorString(e));
}

#define KT_HIP(call)                                    \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

void fill_defaults(const kt_trace_opts* in, kt_trace_opts& o) {
    kt_trace_opts_default(&o);
    if (!in) return;
    o = *in;
    kt_trace_opts d;
    kt_trace_opts_default(&d);
    if (o.max_steps <= 0) o.max_steps = d.max_steps;
    if (!(o.rtol > 0.0)) o.rtol = d.rtol;
    if (!(o.atol > 0.0)) o.atol = d.atol;
    if (!(o.rtol32 > 0.0)) o.rtol32 = d.rtol32;
    if (!(o.atol32 > 0.0)) o.atol32 = d.atol32;
    if (!(o.restep_drift > 0.0)) o.restep_drift = d.restep_drift;
    if (!(o.gate_prob > 0.0)) o.gate_prob = d.gate_prob;
    if (!(o.gate_ring > 0.0)) o.gate_ring = d.gate_ring;
}

// 1 when a view lies in the surrogate's domain: a Bardeen plane of the trained scene (disk
// on, prograde, from the ISCO to disk_outer; escape radius and capture factor as trained)
// with spin, inclination and camera radius inside the trained ranges
int surrogate_covers(const kts::Domain& d, const ViewDev& v) {
    if (v.kind != KT_CAMERA_BARDEEN || !v.disk || !v.prograde) return 0;
    const auto near = [](double x, double y) { return std::fabs(x - y) <= 1e-12 * std::fabs(y); };
    const double inc = std::atan2(v.sinI, v.cosI);
    if (!(v.a >= d.spin_min && v.a <= d.spin_max)) return 0;
    if (!(inc >= d.inc_min && inc <= d.inc_max)) return 0;
    if (!(v.r_cam >= d.r_cam_min && v.r_cam <= d.r_cam_max)) return 0;
    if (!near(v.r_in, isco(v.a, 1)) || !near(v.r_out, d.disk_outer)) return 0;
    if (!near(v.r_cap, d.capture_factor * horizon(v.a))) return 0;
    if (!near(v.r_esc, std::max(d.escape_radius, 1.5 * v.r_cam))) return 0;
    return 1;
}

void mark_coverage(const kt_ctx* ctx, std::vector<ViewDev>& vd) {
    const kts::Domain d = kts::get_domain(ctx->surrogate);
    for (ViewDev& v : vd) {
        v.sur = surrogate_covers(d, v);
        if (v.sur) kts::view_constants(ctx->surrogate, v.a, v.sinI, v.cosI, v.r_cam, v.sur_c);
    }
}

// the launch-shape and precision fields a caller may set; null when valid
const char* check_opts(const kt_trace_opts& o) {
    if (o.precision < KT_PREC_FP64 || o.precision > KT_PREC_FP32) return "unknown precision";
    if (o.mode < KT_MODE_PERSISTENT || o.mode > KT_MODE_AUTO) return "unknown mode";
    // the kernels are built for at most 256 threads per block and count in whole waves
    if (o.block != 0 && (o.block < 64 || o.block > 256 || o.block % 64 != 0))
        return "block must be 0 (tuned) or 64, 128, 192 or 256";
    if (o.waves_per_cu < 0 || o.steps_per_dequeue < 0) return "negative launch shape";
    if (o.waves_per_simd < 0 || o.waves_per_simd > 4) return "waves_per_simd must be 0 (tuned) to 4";
    if (o.closed_form < 0 || o.closed_form > 2) return "closed_form must be 0, 1 or 2";
    return nullptr;
}

int ensure_stage(kt_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->stage_cap) return KT_OK;
    if (ctx->d_stage) (void)hipFree(ctx->d_stage);
    ctx->d_stage = nullptr;
    ctx->stage_cap = 0;
    size_t cap = std::max(bytes, (size_t)1 << 20);
    KT_HIP(hipMalloc(&ctx->d_stage, cap));
    ctx->stage_cap = cap;
    return KT_OK;
}

int ensure_views(kt_ctx* ctx, int n) {
    if (n <= ctx->views_cap) return KT_OK;
    if (ctx->d_views) (void)hipFree(ctx->d_views);
    ctx->d_views = nullptr;
    ctx->views_cap = 0;
    int cap = std::max(n, 64);
    KT_HIP(hipMalloc(&ctx->d_views, sizeof(ViewDev) * (size_t)cap));
    ctx->views_cap = cap;
    return KT_OK;
}

// tuned defaults for one MI355X (see the autotuner in black_simulation_amd/autotune.py)
void launch_shape(const kt_ctx* ctx, const kt_trace_opts& o, int& block, int& grid, int& spd, int& wps) {
    const int pr = std::clamp(o.precision, 0, 2);
    block = o.block > 0 ? o.block : (ctx->tuned_block[pr] > 0 ? ctx->tuned_block[pr] : 256);
    const int wpc = o.waves_per_cu > 0 ? o.waves_per_cu
                                       : (ctx->tuned_wpc[pr] > 0 ? ctx->tuned_wpc[pr] : (pr == KT_PREC_FP64 ? 8 : 16));
    spd = o.steps_per_dequeue > 0 ? o.steps_per_dequeue : (ctx->tuned_spd[pr] > 0 ? ctx->tuned_spd[pr] : 4);
    wps = o.waves_per_simd > 0 ? std::min(o.waves_per_simd, 4) : (ctx->tuned_wps[pr] > 0 ? ctx->tuned_wps[pr] : 1);
    const int waves = std::max(1, ctx->cu_count) * wpc;
    grid = std::max(1, (waves * 64 + block - 1) / block);
}

// validates views and builds their device table; returns KT_OK or a status
int prepare_views(kt_ctx* ctx, const kt_view* views, int n_views, std::vector<ViewDev>& vd, long long& total) {
    if (!views || n_views <= 0) return fail(ctx, KT_EINVAL, "no views");
    vd.clear();
    vd.reserve(n_views);
    total = 0;
    for (int i = 0; i < n_views; ++i) {
        if (const char* why = check_view(views[i]))
            return fail(ctx, KT_EINVAL, "view " + std::to_string(i) + ": " + why);
        vd.push_back(make_view_dev(views[i], total));
        total += (long long)views[i].width * views[i].height;
    }
    return KT_OK;
}

// observer views (Kerr-Schild phi, FP64 only) cannot share a launch with other kinds;
// sets ks = 1 when the batch is all observer views
int observer_batch(kt_ctx* ctx, const std::vector<ViewDev>& vd, const kt_trace_opts& o, int& ks) {
    int n = 0;
    for (const ViewDev& d : vd) n += d.kind == KT_CAMERA_OBSERVER ? 1 : 0;
    ks = n > 0 ? 1 : 0;
    if (n == 0) return KT_OK;
    if (n != (int)vd.size()) return fail(ctx, KT_EINVAL, "observer cameras cannot share a call with other kinds");
    if (o.precision != KT_PREC_FP64) return fail(ctx, KT_EINVAL, "observer cameras trace in FP64 (KT_PREC_FP64)");
    if (o.use_surrogate) return fail(ctx, KT_EINVAL, "the surrogate does not cover observer cameras");
    return KT_OK;
}

kti::TraceLaunch make_launch(kt_ctx* ctx, const ViewDev* d_views, int n_views, long long total,
                             unsigned long long* counter, const ktd::Outputs& dout, const kt_trace_opts& o) {
    kti::TraceLaunch L{};
    L.views = d_views;
    L.n_views = n_views;
    L.total = total;
    L.counter = counter;
    L.stats = counter + 1;
    L.out = dout;
    L.ray_list = nullptr;
    L.total_dev = nullptr;
    L.bucket_count = nullptr;
    L.bucket_stride = 0;
    L.rtol = o.rtol;
    L.atol = o.atol;
    L.rtol32 = o.rtol32;
    L.atol32 = o.atol32;
    L.restep_drift = o.restep_drift;
    L.max_steps = o.max_steps;
    L.precision = o.precision;
    const int pr = std::clamp(o.precision, 0, 2);
    L.mode = o.mode == KT_MODE_AUTO ? (ctx->tuned_mode[pr] ? ctx->tuned_mode[pr] - 1 : KT_MODE_PERSISTENT) : o.mode;
    launch_shape(ctx, o, L.block, L.grid, L.steps_per_dequeue, L.waves_per_simd);
    return L;
}

int trace_common(kt_ctx* ctx, std::vector<ViewDev>& vd, long long total, const kt_trace_opts& o,
                 const kt_rays_out* out, hipStream_t stream) {
    if (const char* why = check_opts(o)) return fail(ctx, KT_EINVAL, why);
    if (o.use_surrogate && ctx->surrogate) mark_coverage(ctx, vd);
    KT_HIP(hipSetDevice(ctx->device));
    if (ensure_views(ctx, (int)vd.size()) != KT_OK) return KT_ENOMEM;

    // outputs: device pointers as given, or staging for host pointers
    const bool report = out->turns || out->r_min || out->delay || out->b || out->q;
    if (report && o.use_surrogate) return fail(ctx, KT_EINVAL, "the ray report is not given with the surrogate");
    if (report && o.closed_form) return fail(ctx, KT_EINVAL, "the ray report is not given with the closed form");
    if (report && o.precision != KT_PREC_FP64) return fail(ctx, KT_EINVAL, "the ray report is traced in FP64 (KT_PREC_FP64)");
    ktd::Outputs dout{};
    size_t off[15] = {};
    if (out->on_device) {
        dout.turns = out->turns;
        dout.r_min = out->r_min;
        dout.delay = out->delay;
        dout.b = out->b;
        dout.q = out->q;
        dout.cls = out->cls;
        dout.theta = out->theta_inf;
        dout.phi = out->phi_inf;
        dout.r_hit = out->r_hit;
        dout.g = out->g;
        dout.steps = out->steps;
        dout.theta32 = out->theta_inf_f32;
        dout.phi32 = out->phi_inf_f32;
        dout.r_hit32 = out->r_hit_f32;
        dout.g32 = out->g_f32;
    } else {
        size_t bytes = 0;
        const size_t n = (size_t)total;
        auto take = [&](size_t each) {
            size_t at = (bytes + 255) & ~(size_t)255;
            bytes = at + each * n;
            return at;
        };
        off[0] = out->cls ? take(1) : 0;
        off[1] = out->theta_inf ? take(8) : 0;
        off[2] = out->phi_inf ? take(8) : 0;
        off[3] = out->r_hit ? take(8) : 0;
        off[4] = out->g ? take(8) : 0;
        off[5] = out->steps ? take(4) : 0;
        off[6] = out->theta_inf_f32 ? take(4) : 0;
        off[7] = out->phi_inf_f32 ? take(4) : 0;
        off[8] = out->r_hit_f32 ? take(4) : 0;
        off[9] = out->g_f32 ? take(4) : 0;
        off[10] = out->turns ? take(8) : 0;
        off[11] = out->r_min ? take(8) : 0;
        off[12] = out->delay ? take(8) : 0;
        off[13] = out->b ? take(8) : 0;
        off[14] = out->q ? take(8) : 0;
        if (ensure_stage(ctx, bytes + 256) != KT_OK) return KT_ENOMEM;
        char* s = static_cast<char*>(ctx->d_stage);
        if (out->cls) dout.cls = reinterpret_cast<int8_t*>(s + off[0]);
        if (out->theta_inf) dout.theta = reinterpret_cast<double*>(s + off[1]);
        if (out->phi_inf) dout.phi = reinterpret_cast<double*>(s + off[2]);
        if (out->r_hit) dout.r_hit = reinterpret_cast<double*>(s + off[3]);
        if (out->g) dout.g = reinterpret_cast<double*>(s + off[4]);
        if (out->steps) dout.steps = reinterpret_cast<int32_t*>(s + off[5]);
        if (out->theta_inf_f32) dout.theta32 = reinterpret_cast<float*>(s + off[6]);
        if (out->phi_inf_f32) dout.phi32 = reinterpret_cast<float*>(s + off[7]);
        if (out->r_hit_f32) dout.r_hit32 = reinterpret_cast<float*>(s + off[8]);
        if (out->g_f32) dout.g32 = reinterpret_cast<float*>(s + off[9]);
        if (out->turns) dout.turns = reinterpret_cast<double*>(s + off[10]);
        if (out->r_min) dout.r_min = reinterpret_cast<double*>(s + off[11]);
        if (out->delay) dout.delay = reinterpret_cast<double*>(s + off[12]);
        if (out->b) dout.b = reinterpret_cast<double*>(s + off[13]);
        if (out->q) dout.q = reinterpret_cast<double*>(s + off[14]);
    }

    KT_HIP(hipMemcpyAsync(ctx->d_views, vd.data(), sizeof(ViewDev) * vd.size(), hipMemcpyHostToDevice, stream));

    kti::TraceLaunch L = make_launch(ctx, ctx->d_views, (int)vd.size(), total, ctx->d_counter, dout, o);
    if (int st = observer_batch(ctx, vd, o, L.ks)) return st;

    ctx->stats = kt_stats{};
    ctx->stats.rays = total;
    ctx->stats.block = L.block;
    ctx->stats.grid = L.grid;
    ctx->stats.steps_per_dequeue = L.steps_per_dequeue;
    ctx->stats.waves_per_simd = L.waves_per_simd;
    ctx->stats.mode = L.mode;
    ctx->stats.waves_per_cu = (L.grid * L.block / 64) / std::max(1, ctx->cu_count);

    if (o.use_surrogate) {
        if (!ctx->surrogate) return fail(ctx, KT_EUNSUPPORTED, "use_surrogate set but no surrogate is loaded");
        KT_HIP(hipEventRecord(ctx->ev[0], stream));
        hipError_t e = kts::trace_with_surrogate(ctx->surrogate, L, o.gate_prob, o.gate_ring, o.closed_form, nullptr,
                                                 stream, ctx->ev);
        if (e != hipSuccess) return hip_fail(ctx, e, "surrogate trace");
        KT_HIP(hipEventRecord(ctx->ev[3], stream));
    } else if (o.closed_form) {
        KT_HIP(kts::queue_reserve(ctx->cf_queue, total));
        const kti::RayQueue rq{ctx->cf_queue.list, ctx->cf_queue.count, ctx->cf_queue.total, ctx->cf_queue.cap};
        KT_HIP(hipEventRecord(ctx->ev[0], stream));
        hipError_t e = kti::trace_closed_form(L, rq, o.closed_form == 2, stream, ctx->ev);
        if (e != hipSuccess) return hip_fail(ctx, e, "closed-form trace");
        KT_HIP(hipEventRecord(ctx->ev[3], stream));
    } else {
        KT_HIP(hipEventRecord(ctx->ev[0], stream));
        KT_HIP(hipEventRecord(ctx->ev[1], stream));
        hipError_t e = kti::launch_trace(L, stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "trace kernel launch");
        KT_HIP(hipEventRecord(ctx->ev[2], stream));
        KT_HIP(hipEventRecord(ctx->ev[3], stream));
    }
    ctx->timed = true;
    ctx->stats_src = ctx->d_counter;

    if (!out->on_device) {
        const size_t n = (size_t)total;
        if (out->cls) KT_HIP(hipMemcpyAsync(out->cls, dout.cls, n, hipMemcpyDeviceToHost, stream));
        if (out->theta_inf) KT_HIP(hipMemcpyAsync(out->theta_inf, dout.theta, 8 * n, hipMemcpyDeviceToHost, stream));
        if (out->phi_inf) KT_HIP(hipMemcpyAsync(out->phi_inf, dout.phi, 8 * n, hipMemcpyDeviceToHost, stream));
        if (out->r_hit) KT_HIP(hipMemcpyAsync(out->r_hit, dout.r_hit, 8 * n, hipMemcpyDeviceToHost, stream));
        if (out->g) KT_HIP(hipMemcpyAsync(out->g, dout.g, 8 * n, hipMemcpyDeviceToHost, stream));
        if (out->steps) KT_HIP(hipMemcpyAsync(out->steps, dout.steps, 4 * n, hipMemcpyDeviceToHost, stream));
        if (out->theta_inf_f32)
            KT_HIP(hipMemcpyAsync(out->theta_inf_f32, dout.theta32, 4 * n, hipMemcpyDeviceToHost, stream));
        if (out->phi_inf_f32) KT_HIP(hipMemcpyAsync(out->phi_inf_f32, dout.phi32, 4 * n, hipMemcpyDeviceToHost, stream));
        if (out->r_hit_f32) KT_HIP(hipMemcpyAsync(out->r_hit_f32, dout.r_hit32, 4 * n, hipMemcpyDeviceToHost, stream));
        if (out->g_f32) KT_HIP(hipMemcpyAsync(out->g_f32, dout.g32, 4 * n, hipMemcpyDeviceToHost, stream));
        if (out->turns) KT_HIP(hipMemcpyAsync(out->turns, dout.turns, 8 * n, hipMemcpyDeviceToHost, stream));
        if (out->r_min) KT_HIP(hipMemcpyAsync(out->r_min, dout.r_min, 8 * n, hipMemcpyDeviceToHost, stream));
        if (out->delay) KT_HIP(hipMemcpyAsync(out->delay, dout.delay, 8 * n, hipMemcpyDeviceToHost, stream));
        if (out->b) KT_HIP(hipMemcpyAsync(out->b, dout.b, 8 * n, hipMemcpyDeviceToHost, stream));
        if (out->q) KT_HIP(hipMemcpyAsync(out->q, dout.q, 8 * n, hipMemcpyDeviceToHost, stream));
        KT_HIP(hipStreamSynchronize(stream));
    }
    return KT_OK;
}

}  // namespace

extern "C" {

void kt_trace_opts_default(kt_trace_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->precision = KT_PREC_FP64;
    o->mode = KT_MODE_AUTO;
    o->use_surrogate = 0;
    o->max_steps = 200000;
    o->rtol = 1e-10;
    o->atol = 1e-12;
    o->rtol32 = 1e-6;
    o->atol32 = 1e-7;
    o->restep_drift = 1e-6;
    o->gate_prob = 0.999;
    o->gate_ring = 0.01;
}

void kt_view_default(kt_view* v) {
    if (!v) return;
    std::memset(v, 0, sizeof *v);
    v->camera_kind = KT_CAMERA_BARDEEN;
    v->width = v->height = 256;
    v->spin = 0.9;
    v->r_cam = 30.0;
    v->theta_cam = 80.0 * kPi / 180.0;
    v->half_width = v->half_height = 12.0;
    v->fov_deg = 60.0;
    v->disk_inner = 0.0;
    v->disk_outer = 20.0;
    v->disk_prograde = 1;
    v->disk_enabled = 1;
    v->escape_radius = 1000.0;
    v->capture_factor = 1.01;
}

int kt_create(const kt_device_opts* opts, kt_ctx** out) {
    if (!out) return KT_EINVAL;
    *out = nullptr;
    kt_ctx* ctx = new (std::nothrow) kt_ctx();
    if (!ctx) return KT_ENOMEM;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        delete ctx;
        return KT_ENODEVICE;
    }
    ctx->device = opts ? opts->device : 0;
    if (ctx->device < 0 || ctx->device >= n) {
        delete ctx;
        return KT_EINVAL;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        delete ctx;
        return KT_EHIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
    if (hipMalloc(&ctx->d_counter, sizeof(unsigned long long) * (1 + kti::kStatCount)) != hipSuccess) {
        delete ctx;
        return KT_ENOMEM;
    }
    for (auto& e : ctx->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            kt_destroy(ctx);
            return KT_EHIP;
        }
    }
    // the device address of this GPU's fixed scratch (kt_trace_views_noalloc); the device
    // is current here
    if (kti::scratch_setup(ctx->device, &ctx->scratch) != hipSuccess) {
        kt_destroy(ctx);
        return KT_EHIP;
    }
    *out = ctx;
    return KT_OK;
}

int kt_destroy(kt_ctx* ctx) {
    if (!ctx) return KT_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    if (ctx->surrogate) kts::destroy(ctx->surrogate);
    kts::queue_release(ctx->cf_queue);
    if (ctx->d_blackbody) (void)hipFree(ctx->d_blackbody);
    if (ctx->d_eht) (void)hipFree(ctx->d_eht);
    if (ctx->d_counter) (void)hipFree(ctx->d_counter);
    if (ctx->d_views) (void)hipFree(ctx->d_views);
    if (ctx->d_stage) (void)hipFree(ctx->d_stage);
    if (ctx->d_params) (void)hipFree(ctx->d_params);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    delete ctx;
    return KT_OK;
}

const char* kt_last_error_string(const kt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int kt_last_stats(const kt_ctx* ctx_c, kt_stats* s) {
    kt_ctx* ctx = const_cast<kt_ctx*>(ctx_c);
    if (!ctx || !s) return KT_EINVAL;
    if (ctx->timed) {
        KT_HIP(hipEventSynchronize(ctx->ev[3]));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]) == hipSuccess) ctx->stats.kernel_ms = ms;
        if (hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]) == hipSuccess) ctx->stats.integrator_ms = ms;
        // ev[0] .. ev[1] bracket the network or closed-form kernels (nothing on the integrator path)
        if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) == hipSuccess) ctx->stats.surrogate_ms = ms;
        unsigned long long h[1 + kti::kStatCount];
        KT_HIP(hipMemcpy(h, ctx->stats_src ? ctx->stats_src : ctx->d_counter, sizeof h, hipMemcpyDeviceToHost));
        ctx->stats.steps = (long long)h[1 + kti::kStatSteps];
        ctx->stats.restep_rays = (long long)h[1 + kti::kStatRestepRays];
        ctx->stats.fallback_rays = (long long)h[1 + kti::kStatFallback];
        ctx->stats.fp64_steps = (long long)h[1 + kti::kStatFp64Steps];
        ctx->timed = false;
    }
    *s = ctx->stats;
    return KT_OK;
}

int kt_trace_views(kt_ctx* ctx, const kt_view* views, int n_views, const kt_trace_opts* opts,
                   const kt_rays_out* out, void* hip_stream) {
    if (!ctx) return KT_EINVAL;
    try {
        if (!views || n_views <= 0 || !out) return fail(ctx, KT_EINVAL, "views/out missing");
        kt_trace_opts o;
        fill_defaults(opts, o);
        std::vector<ViewDev> vd;
        long long total = 0;
        if (int st = prepare_views(ctx, views, n_views, vd, total)) return st;
        return trace_common(ctx, vd, total, o, out, static_cast<hipStream_t>(hip_stream));
    } catch (const std::exception& e) {
        return fail(ctx, KT_EINVAL, e.what());
    } catch (...) {
        return fail(ctx, KT_EINVAL, "unknown error");
    }
}

static_assert((int)KT_NET_WIDE == (int)ktn::kNetWide && (int)KT_NET_FAST == (int)ktn::kNetFast, "network ids");
static_assert(KT_NOALLOC_MAX_VIEWS == kti::kMaxLdsViews, "the scratch holds the LDS-staged view table");

int kt_trace_views_noalloc(kt_ctx* ctx, const kt_view* views, int n_views, const kt_trace_opts* opts,
                           const kt_rays_out* out, void* hip_stream) {
    if (!ctx) return KT_EINVAL;
    try {
        if (!views || n_views <= 0 || !out) return fail(ctx, KT_EINVAL, "views/out missing");
        if (!ctx->scratch.views) return fail(ctx, KT_EHIP, "no device scratch on this context's GPU");
        if (n_views > ctx->scratch.max_views)
            return fail(ctx, KT_EINVAL, "kt_trace_views_noalloc takes at most KT_NOALLOC_MAX_VIEWS (64) views");
        if (!out->on_device)
            return fail(ctx, KT_EINVAL, "kt_trace_views_noalloc writes device outputs only (on_device = 1)");
        if (out->turns || out->r_min || out->delay || out->b || out->q)
            return fail(ctx, KT_EINVAL, "kt_trace_views_noalloc does not give the ray report (use kt_trace_views)");
        kt_trace_opts o;
        fill_defaults(opts, o);
        if (const char* why = check_opts(o)) return fail(ctx, KT_EINVAL, why);
        if (o.use_surrogate)
            return fail(ctx, KT_EINVAL, "kt_trace_views_noalloc does not run the surrogate (its queue is allocated)");
        if (o.closed_form)
            return fail(ctx, KT_EINVAL, "kt_trace_views_noalloc does not run the closed form (its queue is allocated)");
        // the view table on the stack: no allocation on the way to the kernels
        ViewDev vd[kti::kMaxLdsViews];
        long long total = 0;
        int n_obs = 0;
        for (int i = 0; i < n_views; ++i) {
            if (const char* why = check_view(views[i])) return fail(ctx, KT_EINVAL, why);
            vd[i] = make_view_dev(views[i], total);
            total += (long long)views[i].width * views[i].height;
            n_obs += views[i].camera_kind == KT_CAMERA_OBSERVER ? 1 : 0;
        }
        if (n_obs && n_obs != n_views) return fail(ctx, KT_EINVAL, "observer cameras cannot share a call with other kinds");
        if (n_obs && o.precision != KT_PREC_FP64) return fail(ctx, KT_EINVAL, "observer cameras trace in FP64 (KT_PREC_FP64)");
        hipStream_t stream = static_cast<hipStream_t>(hip_stream);
        KT_HIP(hipSetDevice(ctx->device));
        // a pageable source: the copy has read vd before the call returns
        KT_HIP(hipMemcpyAsync(ctx->scratch.views, vd, sizeof(ViewDev) * (size_t)n_views, hipMemcpyHostToDevice,
                              stream));
        ktd::Outputs dout{out->cls, out->theta_inf, out->phi_inf, out->r_hit, out->g, out->steps,
                          out->theta_inf_f32, out->phi_inf_f32, out->r_hit_f32, out->g_f32};
        kti::TraceLaunch L = make_launch(ctx, ctx->scratch.views, n_views, total, ctx->scratch.counter, dout, o);
        L.ks = n_obs ? 1 : 0;
        ctx->stats = kt_stats{};
        ctx->stats.rays = total;
        ctx->stats.block = L.block;
        ctx->stats.grid = L.grid;
        ctx->stats.steps_per_dequeue = L.steps_per_dequeue;
        ctx->stats.waves_per_simd = L.waves_per_simd;
        ctx->stats.mode = L.mode;
        ctx->stats.waves_per_cu = (L.grid * L.block / 64) / std::max(1, ctx->cu_count);
        KT_HIP(hipEventRecord(ctx->ev[0], stream));
        KT_HIP(hipEventRecord(ctx->ev[1], stream));
        hipError_t e = kti::launch_trace(L, stream);  // the counters are cleared by a kernel
        if (e != hipSuccess) return hip_fail(ctx, e, "trace kernel launch");
        KT_HIP(hipEventRecord(ctx->ev[2], stream));
        KT_HIP(hipEventRecord(ctx->ev[3], stream));
        ctx->timed = true;
        ctx->stats_src = ctx->scratch.counter;
        return KT_OK;
    } catch (const std::exception& e) {
        return fail(ctx, KT_EINVAL, e.what());
    } catch (...) {
        return fail(ctx, KT_EINVAL, "unknown error");
    }
}

int kt_trace_rays(kt_ctx* ctx, const kt_view* view, const double* params, long long n, const kt_trace_opts* opts,
                  const kt_rays_out* out, void* hip_stream) {
    if (!ctx) return KT_EINVAL;
    try {
        if (!view || !params || n <= 0 || !out) return fail(ctx, KT_EINVAL, "view/params/out missing");
        kt_view v = *view;
        // explicit rays: the image size only has to be valid
        v.width = 1;
        v.height = 1;
        if (const char* why = check_view(v)) return fail(ctx, KT_EINVAL, why);
        kt_trace_opts o;
        fill_defaults(opts, o);
        hipStream_t stream = static_cast<hipStream_t>(hip_stream);
        std::vector<ViewDev> vd(1, make_view_dev(v, 0));
        const int per = view->camera_kind == KT_CAMERA_BARDEEN ? 2 : 3;
        if (out->on_device) {
            vd[0].params = params;
        } else {
            KT_HIP(hipSetDevice(ctx->device));
            const size_t bytes = (size_t)n * per * 8;
            if (bytes > ctx->params_cap) {
                if (ctx->d_params) (void)hipFree(ctx->d_params);
                ctx->d_params = nullptr;
                ctx->params_cap = 0;
                KT_HIP(hipMalloc(&ctx->d_params, bytes));
                ctx->params_cap = bytes;
            }
            KT_HIP(hipMemcpyAsync(ctx->d_params, params, bytes, hipMemcpyHostToDevice, stream));
            vd[0].params = static_cast<const double*>(ctx->d_params);
        }
        vd[0].n_params = (int)std::min<long long>(n, 0x7fffffff);
        vd[0].width = (int)std::min<long long>(n, 0x7fffffff);
        vd[0].height = 1;
        return trace_common(ctx, vd, n, o, out, stream);
    } catch (const std::exception& e) {
        return fail(ctx, KT_EINVAL, e.what());
    } catch (...) {
        return fail(ctx, KT_EINVAL, "unknown error");
    }
}

int kt_autotune(kt_ctx* ctx, const kt_view* views, int n_views, kt_trace_opts* opts) {
    if (!ctx || !opts) return KT_EINVAL;
    try {
        kt_trace_opts o;
        fill_defaults(opts, o);
        if (const char* why = check_opts(o)) return fail(ctx, KT_EINVAL, why);
        o.mode = KT_MODE_PERSISTENT;
        o.use_surrogate = 0;
        o.closed_form = 0;
        std::vector<ViewDev> vd;
        long long total = 0;
        if (int st = prepare_views(ctx, views, n_views, vd, total)) return st;
        KT_HIP(hipSetDevice(ctx->device));
        if (ensure_views(ctx, (int)vd.size()) != KT_OK) return KT_ENOMEM;
        if (ensure_stage(ctx, (size_t)total + 256) != KT_OK) return KT_ENOMEM;
        KT_HIP(hipMemcpy(ctx->d_views, vd.data(), sizeof(ViewDev) * vd.size(), hipMemcpyHostToDevice));
        int ks = 0;
        if (int st = observer_batch(ctx, vd, o, ks)) return st;
        ktd::Outputs dout{};
        dout.cls = static_cast<int8_t*>(ctx->d_stage);  // classes only: the timing is the same
        auto time_of = [&](const kt_trace_opts& c) -> double {
            kti::TraceLaunch L = make_launch(ctx, ctx->d_views, (int)vd.size(), total, ctx->d_counter, dout, c);
            L.ks = ks;
            double best = 1e30;
            for (int rep = 0; rep < 2; ++rep) {
                if (hipEventRecord(ctx->ev[1], nullptr) != hipSuccess) return 1e30;
                if (kti::launch_trace(L, nullptr) != hipSuccess) return 1e30;
                if (hipEventRecord(ctx->ev[2], nullptr) != hipSuccess) return 1e30;
                if (hipEventSynchronize(ctx->ev[2]) != hipSuccess) return 1e30;
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]) != hipSuccess) return 1e30;
                best = std::min(best, (double)ms);
            }
            return best;
        };
        kt_trace_opts cur = o;
        int b, g, sp, w;
        launch_shape(ctx, cur, b, g, sp, w);
        cur.block = b;
        cur.waves_per_cu = std::max(1, (g * b / 64) / std::max(1, ctx->cu_count));
        cur.steps_per_dequeue = sp;
        cur.waves_per_simd = w;
        // the persistent work queue, tuned by coordinate descent over its four knobs
        double best = time_of(cur);
        const std::vector<int> wps_c = {1, 2, 3, 4}, wpc_c = {4, 8, 12, 16, 24, 32}, spd_c = {1, 2, 4, 8, 16},
                               blk_c = {64, 128, 256};
        for (int pass = 0; pass < 2; ++pass) {
            for (int knob = 0; knob < 4; ++knob) {
                const std::vector<int>& vals = knob == 0 ? wps_c : knob == 1 ? wpc_c : knob == 2 ? spd_c : blk_c;
                for (int v : vals) {
                    kt_trace_opts c = cur;
                    int& f = knob == 0 ? c.waves_per_simd : knob == 1 ? c.waves_per_cu
                                                          : knob == 2 ? c.steps_per_dequeue : c.block;
                    if (f == v) continue;
                    f = v;
                    const double t = time_of(c);
                    if (t < best) {
                        best = t;
                        cur = c;
                    }
                }
            }
        }
        // against one lane per ray without refill (the hardware's block scheduler balances)
        for (int blk : blk_c) {
            kt_trace_opts c = cur;
            c.mode = KT_MODE_ONESHOT;
            c.block = blk;
            const double t = time_of(c);
            if (t < best) {
                best = t;
                cur = c;
            }
        }
        const int pr = o.precision;
        ctx->tuned_mode[pr] = cur.mode + 1;
        ctx->tuned_block[pr] = cur.block;
        ctx->tuned_wpc[pr] = cur.waves_per_cu;
        ctx->tuned_spd[pr] = cur.steps_per_dequeue;
        ctx->tuned_wps[pr] = cur.waves_per_simd;
        opts->mode = cur.mode;
        opts->block = cur.block;
        opts->waves_per_cu = cur.waves_per_cu;
        opts->steps_per_dequeue = cur.steps_per_dequeue;
        opts->waves_per_simd = cur.waves_per_simd;
        ctx->stats = kt_stats{};
        ctx->stats.integrator_ms = ctx->stats.kernel_ms = best;
        ctx->stats.rays = total;
        ctx->timed = false;
        return KT_OK;
    } catch (const std::exception& e) {
        return fail(ctx, KT_EINVAL, e.what());
    }
}

int kt_plan_create(kt_ctx* ctx, const kt_view* views, int n_views, const kt_trace_opts* opts, const kt_rays_out* out,
                   kt_plan** plan_out) {
    if (!ctx || !plan_out) return KT_EINVAL;
    *plan_out = nullptr;
    try {
        if (!out || !out->on_device) return fail(ctx, KT_EINVAL, "a plan needs device output pointers");
        if (out->turns || out->r_min || out->delay || out->b || out->q)
            return fail(ctx, KT_EINVAL, "a plan does not give the ray report (use kt_trace_views)");
        kt_trace_opts o;
        fill_defaults(opts, o);
        if (const char* why = check_opts(o)) return fail(ctx, KT_EINVAL, why);
        if (o.use_surrogate && !kts::ready(ctx->surrogate))
            return fail(ctx, KT_EUNSUPPORTED, "use_surrogate set but no surrogate is loaded");
        std::vector<ViewDev> vd;
        long long total = 0;
        if (int st = prepare_views(ctx, views, n_views, vd, total)) return st;
        if (o.use_surrogate) mark_coverage(ctx, vd);
        KT_HIP(hipSetDevice(ctx->device));
        kt_plan* p = new kt_plan();
        p->ctx = ctx;
        p->surrogate = o.use_surrogate || o.closed_form;
        auto cleanup = [&](int code, const char* where, hipError_t e) {
            hip_fail(ctx, e, where);
            kt_plan_destroy(p);
            return code;
        };
        hipError_t e;
        if ((e = hipMalloc(&p->d_views, sizeof(ViewDev) * vd.size())) != hipSuccess)
            return cleanup(KT_ENOMEM, "plan views", e);
        if ((e = hipMalloc(&p->d_counter, sizeof(unsigned long long) * (1 + kti::kStatCount))) != hipSuccess)
            return cleanup(KT_ENOMEM, "plan counters", e);
        if ((e = hipMemcpy(p->d_views, vd.data(), sizeof(ViewDev) * vd.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return cleanup(KT_EHIP, "plan upload", e);
        for (auto& ev : p->ev)
            if ((e = hipEventCreate(&ev)) != hipSuccess) return cleanup(KT_EHIP, "plan events", e);
        if ((o.use_surrogate || o.closed_form) && (e = kts::queue_reserve(p->queue, total)) != hipSuccess)
            return cleanup(KT_ENOMEM, "plan fallback queue", e);
        ktd::Outputs dout{out->cls, out->theta_inf, out->phi_inf, out->r_hit, out->g, out->steps,
                          out->theta_inf_f32, out->phi_inf_f32, out->r_hit_f32, out->g_f32};
        p->L = make_launch(ctx, p->d_views, (int)vd.size(), total, p->d_counter, dout, o);
        if (int st = observer_batch(ctx, vd, o, p->L.ks)) {
            kt_plan_destroy(p);
            return st;
        }
        if ((e = hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking)) != hipSuccess)
            return cleanup(KT_EHIP, "plan stream", e);
        if ((e = hipStreamBeginCapture(p->capture, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return cleanup(KT_EHIP, "begin capture", e);
        const kti::RayQueue rq{p->queue.list, p->queue.count, p->queue.total, p->queue.cap};
        hipError_t le = o.use_surrogate ? kts::trace_with_surrogate(ctx->surrogate, p->L, o.gate_prob, o.gate_ring,
                                                                    o.closed_form, &p->queue, p->capture, nullptr)
                        : o.closed_form ? kti::trace_closed_form(p->L, rq, o.closed_form == 2, p->capture, nullptr)
                                        : kti::launch_trace(p->L, p->capture);
        if ((e = hipStreamEndCapture(p->capture, &p->graph)) != hipSuccess) return cleanup(KT_EHIP, "end capture", e);
        if (le != hipSuccess) return cleanup(KT_EHIP, "captured launch", le);
        if ((e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0)) != hipSuccess)
            return cleanup(KT_EHIP, "graph instantiate", e);
        *plan_out = p;
        return KT_OK;
    } catch (const std::exception& ex) {
        return fail(ctx, KT_EINVAL, ex.what());
    }
}

int kt_plan_launch(kt_plan* p, void* hip_stream) {
    if (!p) return KT_EINVAL;
    kt_ctx* ctx = p->ctx;
    hipStream_t stream = static_cast<hipStream_t>(hip_stream);
    KT_HIP(hipSetDevice(ctx->device));
    KT_HIP(hipEventRecord(p->ev[0], stream));
    KT_HIP(hipEventRecord(ctx->ev[0], stream));
    KT_HIP(hipEventRecord(ctx->ev[1], stream));
    KT_HIP(hipGraphLaunch(p->exec, stream));
    KT_HIP(hipEventRecord(ctx->ev[2], stream));
    KT_HIP(hipEventRecord(ctx->ev[3], stream));
    KT_HIP(hipEventRecord(p->ev[1], stream));
    p->timed = true;
    ctx->stats = kt_stats{};
    ctx->stats.rays = p->L.total;
    ctx->stats.block = p->L.block;
    ctx->stats.grid = p->L.grid;
    ctx->stats.steps_per_dequeue = p->L.steps_per_dequeue;
    ctx->stats.waves_per_simd = p->L.waves_per_simd;
    ctx->stats.waves_per_cu = (p->L.grid * p->L.block / 64) / std::max(1, ctx->cu_count);
    ctx->timed = true;
    ctx->stats_src = p->d_counter;
    return KT_OK;
}

int kt_plan_destroy(kt_plan* p) {
    if (!p) return KT_OK;
    (void)hipSetDevice(p->ctx->device);
    if (p->capture) (void)hipStreamSynchronize(p->capture);
    (void)hipDeviceSynchronize();
    if (p->ctx->stats_src == p->d_counter) p->ctx->stats_src = nullptr;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->capture) (void)hipStreamDestroy(p->capture);
    for (auto& ev : p->ev)
        if (ev) (void)hipEventDestroy(ev);
    kts::queue_release(p->queue);
    if (p->d_views) (void)hipFree(p->d_views);
    if (p->d_counter) (void)hipFree(p->d_counter);
    delete p;
    return KT_OK;
}

int kt_plan_stats(kt_plan* p, kt_stats* s) {
    if (!p || !s) return KT_EINVAL;
    kt_ctx* ctx = p->ctx;
    KT_HIP(hipSetDevice(ctx->device));
    kt_stats r{};
    r.rays = p->L.total;
    r.block = p->L.block;
    r.grid = p->L.grid;
    r.steps_per_dequeue = p->L.steps_per_dequeue;
    r.waves_per_simd = p->L.waves_per_simd;
    r.mode = p->L.mode;
    r.waves_per_cu = (p->L.grid * p->L.block / 64) / std::max(1, ctx->cu_count);
    if (p->timed) {
        KT_HIP(hipEventSynchronize(p->ev[1]));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p->ev[0], p->ev[1]) == hipSuccess) r.kernel_ms = ms;
        if (!p->surrogate) r.integrator_ms = r.kernel_ms;
        unsigned long long h[1 + kti::kStatCount];
        KT_HIP(hipMemcpy(h, p->d_counter, sizeof h, hipMemcpyDeviceToHost));
        r.steps = (long long)h[1 + kti::kStatSteps];
        r.restep_rays = (long long)h[1 + kti::kStatRestepRays];
        r.fallback_rays = (long long)h[1 + kti::kStatFallback];
        r.fp64_steps = (long long)h[1 + kti::kStatFp64Steps];
    }
    *s = r;
    return KT_OK;
}

// ---- shading ---------------------------------------------------------------------------

static constexpr int kBlackbodyEntries = 1024;

void kt_shade_opts_default(kt_shade_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->disk_temperature = 10000.0;
    o->disk_brightness = 1.0;
    o->sky = KT_SKY_GRID;
    o->sky_brightness = 1.0;
    o->grid_line = 0.02;
}

int kt_blackbody_table(int n, float* rgb) {
    if (n < 2 || !rgb) return KT_EINVAL;
    ktsh::blackbody_table(n, rgb);
    return KT_OK;
}

int kt_shade_views(kt_ctx* ctx, const kt_view* views, int n_views, const kt_rays_out* rays, const kt_shade_opts* opts,
                   void* rgba, void* hip_stream) {
    if (!ctx) return KT_EINVAL;
    try {
        if (!rays || !rgba) return fail(ctx, KT_EINVAL, "rays/rgba missing");
        if (!rays->on_device) return fail(ctx, KT_EINVAL, "shading reads device outputs (on_device = 1)");
        if (!rays->cls || !(rays->theta_inf || rays->theta_inf_f32) || !(rays->phi_inf || rays->phi_inf_f32) ||
            !(rays->r_hit || rays->r_hit_f32) || !(rays->g || rays->g_f32))
            return fail(ctx, KT_EINVAL, "shading needs cls, theta_inf, phi_inf, r_hit and g (FP64 or FP32)");
        kt_shade_opts o;
        kt_shade_opts_default(&o);
        if (opts) o = *opts;
        if (o.sky < KT_SKY_BLACK || o.sky > KT_SKY_IMAGE) return fail(ctx, KT_EINVAL, "unknown sky");
        if (o.sky == KT_SKY_IMAGE && (!o.sky_image || o.sky_width < 1 || o.sky_height < 1))
            return fail(ctx, KT_EINVAL, "image sky needs sky_image and its size");
        if (!(o.disk_temperature > 0.0)) return fail(ctx, KT_EINVAL, "disk_temperature must be positive");
        std::vector<ViewDev> vd;
        long long total = 0;
        if (int st = prepare_views(ctx, views, n_views, vd, total)) return st;
        KT_HIP(hipSetDevice(ctx->device));
        hipStream_t stream = static_cast<hipStream_t>(hip_stream);
        if (!ctx->d_blackbody) {
            ctx->h_blackbody.resize(3 * kBlackbodyEntries);
            ktsh::blackbody_table(kBlackbodyEntries, ctx->h_blackbody.data());
            KT_HIP(hipMalloc(&ctx->d_blackbody, sizeof(float) * ctx->h_blackbody.size()));
            KT_HIP(hipMemcpy(ctx->d_blackbody, ctx->h_blackbody.data(), sizeof(float) * ctx->h_blackbody.size(),
                             hipMemcpyHostToDevice));
        }
        if (ensure_views(ctx, (int)vd.size()) != KT_OK) return KT_ENOMEM;
        KT_HIP(hipMemcpyAsync(ctx->d_views, vd.data(), sizeof(ViewDev) * vd.size(), hipMemcpyHostToDevice, stream));
        ktsh::ShadeLaunch s{};
        s.n = total;
        s.views = ctx->d_views;
        s.n_views = (int)vd.size();
        s.cls = rays->cls;
        s.theta = rays->theta_inf;
        s.phi = rays->phi_inf;
        s.r_hit = rays->r_hit;
        s.g = rays->g;
        s.theta32 = rays->theta_inf_f32;
        s.phi32 = rays->phi_inf_f32;
        s.r_hit32 = rays->r_hit_f32;
        s.g32 = rays->g_f32;
        s.table = ctx->d_blackbody;
        s.table_n = kBlackbodyEntries;
        s.disk_temperature = o.disk_temperature;
        const float lum = ktsh::table_luminance(ctx->h_blackbody.data(), kBlackbodyEntries, o.disk_temperature);
        s.disk_scale = (float)(o.disk_brightness / std::max((double)lum, 1e-30));
        s.sky = o.sky;
        s.sky_brightness = (float)o.sky_brightness;
        s.grid_line = (float)o.grid_line;
        s.sky_image = o.sky_image;
        s.sky_w = o.sky_width;
        s.sky_h = o.sky_height;
        s.rgba = static_cast<__half*>(rgba);
        hipError_t e = ktsh::shade(s, stream);
        return e == hipSuccess ? KT_OK : hip_fail(ctx, e, "shade kernel launch");
    } catch (const std::exception& ex) {
        return fail(ctx, KT_EINVAL, ex.what());
    }
}

}  // extern "C"

// ---- EHT radiative transfer -------------------------------------------------------------

namespace {

constexpr double kSolarLengthCm = 1.476625e5;  // G M_sun / c^2 (src/eht.h:30)
constexpr double kProtonMass = 1.67262192369e-24;  // g
constexpr double kLightCm = 2.99792458e10;         // cm/s
constexpr double kParsecCm = 3.0856775814913673e18;
constexpr double kJansky = 1e-23;
constexpr double kMicroarcsecond = kPi / 180.0 / 3600.0 / 1e6;
constexpr int kCalibrateSize = 48;                 // eht.cpp:435
constexpr size_t kEhtSumSlot = kte::kSumPartials;  // d_eht: partials, then the sum
constexpr size_t kEhtImage = kte::kSumPartials + 8;
constexpr size_t kEhtDoubles = kEhtImage + (size_t)kCalibrateSize * kCalibrateSize;

const char* check_eht(const kt_eht_view& v, const kt_eht_flow& f) {
    const double reals[] = {v.a, v.inclination, v.half_width, v.frequency, v.mass_solar,
                            f.n0, f.T0, f.beta, f.keplerian, f.height, f.sin_theta_b};
    for (double x : reals)
        if (!std::isfinite(x)) return "view and flow fields must be finite";
    if (!(std::fabs(v.a) < 1.0)) return "spin must be in (-1, 1)";
    if (!(v.inclination >= 0.0 && v.inclination <= kPi)) return "inclination must be in [0, pi]";
    if (!(v.half_width > 0.0) || !std::isfinite(v.half_width)) return "half_width must be positive";
    if (v.size < 1 || v.size > 32768) return "size must be in [1, 32768]";
    if (!(v.frequency > 0.0) || !(v.mass_solar > 0.0)) return "frequency and mass_solar must be positive";
    if (!(f.n0 >= 0.0) || !(f.T0 > 0.0) || !(f.beta > 0.0)) return "flow needs n0 >= 0, T0 > 0, beta > 0";
    if (!std::isfinite(f.keplerian) || !std::isfinite(f.height) || !std::isfinite(f.sin_theta_b))
        return "flow parameters must be finite";
    return nullptr;
}

// the ISCO's plunging geodesic (src/eht.cpp:183-192, 222-233): |u^r| from its E, L
double plunge_speed(double a, double E, double L, double r) {
    const double delta = r * r - 2.0 * r + a * a, x = L - a * E;
    const double P = E * (r * r + a * a) - a * L;
    return std::sqrt(std::fmax(P * P - delta * (r * r + x * x), 0.0)) / (r * r);
}

kte::EhtLaunch make_eht(const kt_eht_view& v, const kt_eht_flow& f, const kt_trace_opts& o) {
    kte::EhtLaunch e{};
    e.size = v.size;
    e.n = (long long)v.size * v.size;
    e.half_width = v.half_width;
    e.px = 2.0 * v.half_width / v.size;
    e.a = v.a;
    e.sinI = hsin(v.inclination);
    e.cosI = hcos(v.inclination);
    e.m0 = hcos(std::clamp(v.inclination, 1e-4, kPi - 1e-4));
    const double rh = horizon(v.a);
    e.r_cap = rh * (1.0 + 1e-4);
    e.len_cm = kSolarLengthCm * v.mass_solar;
    e.nu = v.frequency;
    e.n0 = f.n0;
    e.T0 = f.T0;
    e.keplerian = f.keplerian;
    e.sin_theta_b = f.sin_theta_b;
    e.hfac = std::fmax(f.height, 0.05);
    e.b_coef = 8.0 * kPi * kProtonMass * kLightCm * kLightCm / (6.0 * f.beta);
    // circular equatorial orbit at the prograde ISCO (src/kerr.cpp:308-319)
    const double r = isco(v.a, 1), sr = std::sqrt(r), r32 = r * sr;
    const double den = std::pow(r, 0.75) * std::sqrt(r32 - 3.0 * sr + 2.0 * v.a);
    e.isco = r;
    e.E = (r32 - 2.0 * sr + v.a) / den;
    e.L = (r * r - 2.0 * v.a * sr + v.a * v.a) / den;
    e.start = std::fmax(r - 0.25, 0.5 * (r + rh));
    e.start_rate = std::pow(e.start, -1.1) * (e.start * e.start * plunge_speed(v.a, e.E, e.L, e.start));
    e.rtol = o.rtol;
    e.atol = o.atol;
    e.max_steps = o.max_steps;
    return e;
}

int ensure_eht(kt_ctx* ctx) {
    if (ctx->d_eht) return KT_OK;
    KT_HIP(hipMalloc(&ctx->d_eht, sizeof(double) * kEhtDoubles));
    return KT_OK;
}

// two workgroups of 256 per CU: the kernel's 2 waves per SIMD
int eht_grid(const kt_ctx* ctx) { return std::max(1, ctx->cu_count) * 2; }

// the launch itself: one kernel, or slice after slice while a cancel flag stays clear
// (eht::render polls its flag per row, src/eht.cpp:330-331).  Rays of slices not traced
// are zeroed.  A slice is an eighth of the rays, at least 2^17 (two waves per SIMD).
int eht_launch(kt_ctx* ctx, kte::EhtLaunch e, const volatile int* cancel, hipStream_t stream) {
    if (!cancel) {
        hipError_t err = kte::render(e, ctx->d_counter, eht_grid(ctx), stream);
        return err == hipSuccess ? KT_OK : hip_fail(ctx, err, "EHT kernel launch");
    }
    const long long n = e.n, slice = std::max<long long>((n + 7) / 8, 1LL << 17);
    for (long long first = 0; first < n; first += slice) {
        if (*cancel) {
            const size_t k = (size_t)first, rest = (size_t)(n - first);
            KT_HIP(hipMemsetAsync(e.intensity + k, 0, 8 * rest, stream));
            if (e.shadow) KT_HIP(hipMemsetAsync(e.shadow + k, 0, rest, stream));
            if (e.min_radius) KT_HIP(hipMemsetAsync(e.min_radius + k, 0, 8 * rest, stream));
            if (e.steps) KT_HIP(hipMemsetAsync(e.steps + k, 0, 4 * rest, stream));
            return KT_ECANCELLED;
        }
        e.first = first;
        e.n = std::min(n, first + slice);
        hipError_t err = kte::render(e, ctx->d_counter, eht_grid(ctx), stream);
        if (err != hipSuccess) return hip_fail(ctx, err, "EHT kernel launch");
        KT_HIP(hipStreamSynchronize(stream));
    }
    return KT_OK;
}

// ray count n: explicit (ab != null) or the view's pixels; outputs staged when on the host
int eht_run(kt_ctx* ctx, const kt_eht_view* view, const kt_eht_flow* flow, const double* ab, long long n,
            const kt_trace_opts* opts, const kt_eht_out* out, hipStream_t stream) {
    if (!view || !out || !out->intensity) return fail(ctx, KT_EINVAL, "view/out/intensity missing");
    kt_eht_flow fl;
    kt_eht_flow_default(&fl);
    if (flow) fl = *flow;
    if (const char* why = check_eht(*view, fl)) return fail(ctx, KT_EINVAL, why);
    kt_trace_opts o;
    fill_defaults(opts, o);
    kte::EhtLaunch e = make_eht(*view, fl, o);
    if (ab) e.n = n;
    if (e.n <= 0) return KT_OK;
    KT_HIP(hipSetDevice(ctx->device));
    if (out->on_device) {
        e.rays = ab;
        e.intensity = out->intensity;
        e.shadow = out->shadow;
        e.min_radius = out->min_radius;
        e.steps = out->steps;
        const int st = eht_launch(ctx, e, o.cancel, stream);
        if (st == KT_ECANCELLED) return fail(ctx, KT_ECANCELLED, "cancelled");
        return st;
    }
    // host buffers: explicit rays, then intensity, min radius, steps, shadow on the device
    const size_t nr = (size_t)e.n, in_bytes = ab ? 16 * nr : 0;
    if (int st = ensure_stage(ctx, in_bytes + nr * (8 + 8 + 4 + 1) + 64)) return st;
    char* base = static_cast<char*>(ctx->d_stage);
    if (ab) {
        KT_HIP(hipMemcpyAsync(base, ab, in_bytes, hipMemcpyHostToDevice, stream));
        e.rays = reinterpret_cast<const double*>(base);
    }
    e.intensity = reinterpret_cast<double*>(base + in_bytes);
    e.min_radius = out->min_radius ? e.intensity + nr : nullptr;
    e.steps = out->steps ? reinterpret_cast<int32_t*>(e.intensity + 2 * nr) : nullptr;
    e.shadow = out->shadow ? reinterpret_cast<uint8_t*>(e.intensity + 2 * nr) + 4 * nr : nullptr;
    const int st = eht_launch(ctx, e, o.cancel, stream);
    if (st != KT_OK && st != KT_ECANCELLED) return st;
    KT_HIP(hipMemcpyAsync(out->intensity, e.intensity, 8 * nr, hipMemcpyDeviceToHost, stream));
    if (out->min_radius) KT_HIP(hipMemcpyAsync(out->min_radius, e.min_radius, 8 * nr, hipMemcpyDeviceToHost, stream));
    if (out->steps) KT_HIP(hipMemcpyAsync(out->steps, e.steps, 4 * nr, hipMemcpyDeviceToHost, stream));
    if (out->shadow) KT_HIP(hipMemcpyAsync(out->shadow, e.shadow, nr, hipMemcpyDeviceToHost, stream));
    KT_HIP(hipStreamSynchronize(stream));
    return st == KT_OK ? KT_OK : fail(ctx, KT_ECANCELLED, "cancelled");
}

int eht_flux(kt_ctx* ctx, const double* intensity, int size, double half_width, double theta_guas, double* jy,
             hipStream_t stream) {
    if (int st = ensure_eht(ctx)) return st;
    hipError_t err = kte::sum(intensity, (long long)size * size, ctx->d_eht, ctx->d_eht + kEhtSumSlot, stream);
    if (err != hipSuccess) return hip_fail(ctx, err, "EHT flux reduction");
    double sum = 0.0;
    KT_HIP(hipMemcpyAsync(&sum, ctx->d_eht + kEhtSumSlot, sizeof sum, hipMemcpyDeviceToHost, stream));
    KT_HIP(hipStreamSynchronize(stream));
    const double pixel = 2.0 * half_width / size;
    const double solid_angle = std::pow(pixel * theta_guas * kMicroarcsecond, 2.0);
    *jy = sum * solid_angle / kJansky;
    return KT_OK;
}

}  // namespace

extern "C" {

void kt_eht_flow_default(kt_eht_flow* f) {
    if (!f) return;
    f->n0 = 1e5;
    f->T0 = 4e11;
    f->beta = 10.0;
    f->keplerian = 0.9;
    f->height = 1.0;
    f->sin_theta_b = 0.8165;
}

void kt_eht_view_default(kt_eht_view* v) {
    if (!v) return;
    v->a = 0.9;
    v->inclination = 0.3;
    v->half_width = 16.0;
    v->size = 128;
    v->frequency = 230e9;
    v->mass_solar = 6.5e9;
}

int kt_eht_render(kt_ctx* ctx, const kt_eht_view* view, const kt_eht_flow* flow, const kt_trace_opts* opts,
                  const kt_eht_out* out, void* hip_stream) {
    if (!ctx) return KT_EINVAL;
    try {
        return eht_run(ctx, view, flow, nullptr, 0, opts, out, static_cast<hipStream_t>(hip_stream));
    } catch (const std::exception& ex) {
        return fail(ctx, KT_EINVAL, ex.what());
    }
}

int kt_eht_trace_rays(kt_ctx* ctx, const kt_eht_view* view, const kt_eht_flow* flow, const double* ab, long long n,
                      const kt_trace_opts* opts, const kt_eht_out* out, void* hip_stream) {
    if (!ctx) return KT_EINVAL;
    if (!ab || n <= 0) return fail(ctx, KT_EINVAL, "explicit rays: ab and n > 0 needed");
    try {
        return eht_run(ctx, view, flow, ab, n, opts, out, static_cast<hipStream_t>(hip_stream));
    } catch (const std::exception& ex) {
        return fail(ctx, KT_EINVAL, ex.what());
    }
}

int kt_eht_flux(kt_ctx* ctx, const double* intensity, int size, double half_width, double theta_guas, double* jy,
                void* hip_stream) {
    if (!ctx) return KT_EINVAL;
    if (!intensity || !jy || size < 1 || !(half_width > 0.0) || !(theta_guas > 0.0))
        return fail(ctx, KT_EINVAL, "flux: intensity, jy, size >= 1, half_width > 0, theta_guas > 0 needed");
    KT_HIP(hipSetDevice(ctx->device));
    return eht_flux(ctx, intensity, size, half_width, theta_guas, jy, static_cast<hipStream_t>(hip_stream));
}

int kt_eht_calibrate(kt_ctx* ctx, const kt_eht_view* view, const kt_eht_flow* flow, double target_jy,
                     double theta_guas, const kt_trace_opts* opts, double* n0_out, void* hip_stream) {
    if (!ctx) return KT_EINVAL;
    if (!view || !n0_out || !(target_jy > 0.0) || !(theta_guas > 0.0))
        return fail(ctx, KT_EINVAL, "calibrate: view, n0, target_jy > 0 and theta_guas > 0 needed");
    try {
        hipStream_t stream = static_cast<hipStream_t>(hip_stream);
        KT_HIP(hipSetDevice(ctx->device));
        if (int st = ensure_eht(ctx)) return st;
        kt_eht_view small = *view;
        small.size = kCalibrateSize;
        kt_eht_flow f;
        kt_eht_flow_default(&f);
        if (flow) f = *flow;
        kt_eht_out out{};
        out.intensity = ctx->d_eht + kEhtImage;
        out.on_device = 1;
        int st = KT_OK;
        auto flux_at = [&](double n0) {
            f.n0 = n0;
            double jy = 0.0;
            if (st == KT_OK) st = eht_run(ctx, &small, &f, nullptr, 0, opts, &out, stream);
            if (st == KT_OK) st = eht_flux(ctx, out.intensity, small.size, small.half_width, theta_guas, &jy, stream);
            return jy;
        };
        // src/eht.cpp:436-456: brighten a dark start tenfold at a time, then a secant in
        // log n0 against log flux, each step held within a factor 30, to 1%
        double n0 = f.n0, f0 = flux_at(n0);
        for (int k = 0; k < 8 && f0 <= 0.0 && st == KT_OK; ++k) f0 = flux_at(n0 *= 10.0);
        if (st == KT_ECANCELLED) *n0_out = n0;  // the estimate so far, as eht::calibrate returns it
        if (st != KT_OK) return st;
        if (f0 <= 0.0) {
            *n0_out = n0;
            return KT_OK;
        }
        double n1 = n0 * target_jy / f0, f1 = flux_at(n1);
        const volatile int* cancel = opts ? opts->cancel : nullptr;
        for (int k = 0; k < 8 && st == KT_OK; ++k) {
            if (cancel && *cancel) {
                st = KT_ECANCELLED;
                break;
            }
            if (f1 <= 0.0 || std::fabs(f1 / target_jy - 1.0) < 0.01) break;
            const double slope = (std::log(f1) - std::log(f0)) / (std::log(n1) - std::log(n0));
            const double next =
                std::exp(std::log(n1) + (std::log(target_jy) - std::log(f1)) / (std::fabs(slope) > 0.05 ? slope : 1.0));
            n0 = n1;
            f0 = f1;
            n1 = std::clamp(next, n0 / 30.0, n0 * 30.0);
            f1 = flux_at(n1);
        }
        if (st == KT_ECANCELLED) *n0_out = n1;
        if (st != KT_OK) return st;
        *n0_out = n1;
        return KT_OK;
    } catch (const std::exception& ex) {
        return fail(ctx, KT_EINVAL, ex.what());
    }
}

double kt_eht_angular_scale(double mass_solar, double distance_pc) {
    return kSolarLengthCm * mass_solar / (distance_pc * kParsecCm) / kMicroarcsecond;
}

}  // extern "C"

extern "C" {

static int ensure_surrogate(kt_ctx* ctx) {
    if (ctx->surrogate) return KT_OK;
    hipError_t e = hipSuccess;
    ctx->surrogate = kts::create(ctx->device, &e);
    if (!ctx->surrogate) return hip_fail(ctx, e, "surrogate create");
    return KT_OK;
}

void kt_surrogate_opts_default(kt_surrogate_opts* o) {
    if (!o) return;
    kts::TrainOpts d;
    o->n_train = d.n_train;
    o->n_holdout = d.n_holdout;
    o->epochs = d.epochs;
    o->batch = d.batch;
    o->lr = d.lr;
    o->w_sky = d.w_sky;
    o->w_disk = d.w_disk;
    o->band = d.band;
    o->band_factor = d.band_factor;
    o->seed = d.seed;
    o->label_precision = d.label_precision;
    o->disk_outer = d.disk_outer;
    o->escape_radius = d.escape_radius;
    o->capture_factor = d.capture_factor;
    o->verbose = 0;
    o->net = KT_NET_WIDE;
    o->edge_band = d.edge_band;
    o->edge_copies = d.edge_copies;
}

long long kt_surrogate_param_count(void) { return ktn::param_count(ktn::net_of(KT_NET_WIDE).hidden); }

long long kt_surrogate_net_param_count(int net) {
    if (net != KT_NET_WIDE && net != KT_NET_FAST) return 0;
    return ktn::param_count(ktn::net_of(net).hidden);
}

int kt_surrogate_set_net(kt_ctx* ctx, int net) {
    if (!ctx) return KT_EINVAL;
    if (net != KT_NET_WIDE && net != KT_NET_FAST) return fail(ctx, KT_EINVAL, "unknown network (KT_NET_*)");
    if (int st = ensure_surrogate(ctx)) return st;
    hipError_t e = kts::set_net(ctx->surrogate, net);
    return e == hipSuccess ? KT_OK : hip_fail(ctx, e, "surrogate set_net");
}

int kt_surrogate_get_net(kt_ctx* ctx) {
    if (!ctx || ensure_surrogate(ctx) != KT_OK) return -1;
    return kts::get_net(ctx->surrogate);
}

int kt_surrogate_init(kt_ctx* ctx, unsigned long long seed) {
    if (!ctx) return KT_EINVAL;
    if (int st = ensure_surrogate(ctx)) return st;
    hipError_t e = kts::init_params(ctx->surrogate, seed);
    return e == hipSuccess ? KT_OK : hip_fail(ctx, e, "surrogate init");
}

int kt_surrogate_load(kt_ctx* ctx, const float* params, long long n) {
    if (!ctx || !params) return KT_EINVAL;
    // the parameter count names the network
    const int net = n == kt_surrogate_net_param_count(KT_NET_WIDE) ? KT_NET_WIDE
                  : n == kt_surrogate_net_param_count(KT_NET_FAST) ? KT_NET_FAST : -1;
    if (net < 0) return fail(ctx, KT_EINVAL, "parameter count must be kt_surrogate_net_param_count() of a network");
    if (int st = ensure_surrogate(ctx)) return st;
    hipError_t e = kts::set_net(ctx->surrogate, net);
    if (e == hipSuccess) e = kts::set_params(ctx->surrogate, params, n);
    return e == hipSuccess ? KT_OK : hip_fail(ctx, e, "surrogate load");
}

int kt_surrogate_save(kt_ctx* ctx, float* params, long long n) {
    if (!ctx || !params) return KT_EINVAL;
    if (int st = ensure_surrogate(ctx)) return st;
    if (n != kt_surrogate_net_param_count(kts::get_net(ctx->surrogate)))
        return fail(ctx, KT_EINVAL, "parameter count must be kt_surrogate_net_param_count() of the loaded network");
    if (!kts::ready(ctx->surrogate)) return fail(ctx, KT_EUNSUPPORTED, "no surrogate parameters");
    hipError_t e = kts::get_params(ctx->surrogate, params, n);
    return e == hipSuccess ? KT_OK : hip_fail(ctx, e, "surrogate save");
}

int kt_surrogate_train(kt_ctx* ctx, const kt_surrogate_opts* opts, kt_surrogate_report* report) {
    if (!ctx) return KT_EINVAL;
    if (int st = ensure_surrogate(ctx)) return st;
    kts::TrainOpts o;
    if (opts) {
        o.n_train = opts->n_train > 0 ? opts->n_train : o.n_train;
        o.n_holdout = opts->n_holdout > 0 ? opts->n_holdout : o.n_holdout;
        o.epochs = opts->epochs > 0 ? opts->epochs : o.epochs;
        o.batch = opts->batch > 0 ? opts->batch : o.batch;
        o.lr = opts->lr > 0 ? opts->lr : o.lr;
        o.w_sky = opts->w_sky > 0 ? opts->w_sky : o.w_sky;
        o.w_disk = opts->w_disk > 0 ? opts->w_disk : o.w_disk;
        o.band = opts->band > 0 ? opts->band : o.band;
        o.band_factor = opts->band_factor > 0 ? opts->band_factor : o.band_factor;
        o.seed = opts->seed ? opts->seed : o.seed;
        o.label_precision = opts->label_precision;
        o.disk_outer = opts->disk_outer > 0 ? opts->disk_outer : o.disk_outer;
        o.escape_radius = opts->escape_radius > 0 ? opts->escape_radius : o.escape_radius;
        o.capture_factor = opts->capture_factor >= 1.0 ? opts->capture_factor : o.capture_factor;
        o.verbose = opts->verbose;
        if (opts->net != KT_NET_WIDE && opts->net != KT_NET_FAST) return fail(ctx, KT_EINVAL, "unknown network");
        o.net = opts->net;
        o.edge_band = opts->edge_band > 0 ? opts->edge_band : o.edge_band;
        o.edge_copies = opts->edge_copies < 0 ? 0 : (opts->edge_copies > 0 ? opts->edge_copies : o.edge_copies);
    }
    if (o.batch % 128 != 0) return fail(ctx, KT_EINVAL, "batch must be a multiple of 128");
    kts::TrainReport r;
    hipError_t e = kts::train(ctx->surrogate, o, r);
    if (e != hipSuccess) return hip_fail(ctx, e, "surrogate train");
    if (report) {
        report->datagen_seconds = r.datagen_seconds;
        report->train_seconds = r.train_seconds;
        report->eval_seconds = r.eval_seconds;
        report->first_loss = r.first_loss;
        report->final_loss = r.final_loss;
        report->train_gemm_tflops = r.train_gemm_tflops;
        report->steps = r.steps;
        report->class_agreement = r.class_agreement;
        report->gated_fraction = r.gated_fraction;
        report->gated_class_agreement = r.gated_class_agreement;
        report->defl_median = r.defl_median;
        report->defl_p99 = r.defl_p99;
        report->defl_p999 = r.defl_p999;
        report->g_rel_median = r.g_rel_median;
        report->g_rel_p99 = r.g_rel_p99;
        report->rhit_rel_median = r.rhit_rel_median;
        report->rhit_rel_p99 = r.rhit_rel_p99;
    }
    return KT_OK;
}

int kt_surrogate_predict(kt_ctx* ctx, const float* inputs, long long n, float* outputs, void* hip_stream) {
    if (!ctx || !inputs || !outputs || n <= 0) return KT_EINVAL;
    if (!kts::ready(ctx->surrogate)) return fail(ctx, KT_EUNSUPPORTED, "no surrogate parameters");
    hipError_t e = kts::predict(ctx->surrogate, inputs, n, outputs, static_cast<hipStream_t>(hip_stream));
    return e == hipSuccess ? KT_OK : hip_fail(ctx, e, "surrogate predict");
}

int kt_surrogate_gradients(kt_ctx* ctx, const float* inputs, const int8_t* cls, const float* targets, int n,
                           float* grads, double* loss) {
    if (!ctx || !inputs || !cls || !targets || n <= 0) return KT_EINVAL;
    if (n % 128 != 0) return fail(ctx, KT_EINVAL, "n must be a multiple of 128");
    if (!kts::ready(ctx->surrogate)) return fail(ctx, KT_EUNSUPPORTED, "no surrogate parameters");
    hipError_t e = kts::gradients(ctx->surrogate, inputs, cls, targets, n, grads, loss, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? KT_OK : hip_fail(ctx, e, "surrogate gradients");
}

int kt_surrogate_adam_step(kt_ctx* ctx, const float* grads, double lr, int t) {
    if (!ctx || t < 1) return KT_EINVAL;
    if (!kts::ready(ctx->surrogate)) return fail(ctx, KT_EUNSUPPORTED, "no surrogate parameters");
    hipError_t e = kts::adam_step(ctx->surrogate, grads, lr, t, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? KT_OK : hip_fail(ctx, e, "surrogate adam");
}

double kt_surrogate_table_radius(kt_ctx* ctx, double a, double inclination, double psi) {
    if (!ctx || ensure_surrogate(ctx) != KT_OK) return 0.0;
    return kts::table_critical_radius(ctx->surrogate, a, inclination, psi);
}

double kt_surrogate_shadow_radius(kt_ctx* ctx, double a, double inclination) {
    if (!ctx || ensure_surrogate(ctx) != KT_OK) return 0.0;
    return kts::table_shadow_radius(ctx->surrogate, a, inclination);
}

void kt_surrogate_domain_default(kt_surrogate_domain* o) {
    if (!o) return;
    const kts::Domain d;
    o->spin_min = d.spin_min;
    o->spin_max = d.spin_max;
    o->inclination_min = d.inc_min;
    o->inclination_max = d.inc_max;
    o->r_cam_min = d.r_cam_min;
    o->r_cam_max = d.r_cam_max;
    o->extent = d.extent;
    o->disk_outer = d.disk_outer;
    o->escape_radius = d.escape_radius;
    o->capture_factor = d.capture_factor;
}

int kt_surrogate_get_domain(kt_ctx* ctx, kt_surrogate_domain* o) {
    if (!ctx || !o) return KT_EINVAL;
    if (int st = ensure_surrogate(ctx)) return st;
    const kts::Domain d = kts::get_domain(ctx->surrogate);
    o->spin_min = d.spin_min;
    o->spin_max = d.spin_max;
    o->inclination_min = d.inc_min;
    o->inclination_max = d.inc_max;
    o->r_cam_min = d.r_cam_min;
    o->r_cam_max = d.r_cam_max;
    o->extent = d.extent;
    o->disk_outer = d.disk_outer;
    o->escape_radius = d.escape_radius;
    o->capture_factor = d.capture_factor;
    return KT_OK;
}

int kt_surrogate_set_domain(kt_ctx* ctx, const kt_surrogate_domain* o) {
    if (!ctx || !o) return KT_EINVAL;
    const bool ok = o->spin_min >= 0.0 && o->spin_min <= o->spin_max && o->spin_max <= 1.0 &&
                    o->inclination_min >= 0.0 && o->inclination_min <= o->inclination_max &&
                    o->inclination_max <= kPi && o->r_cam_min > 0.0 && o->r_cam_min <= o->r_cam_max &&
                    o->extent > 0.0 && o->extent <= 1.5 && o->disk_outer > 0.0 && o->escape_radius > 0.0 &&
                    o->capture_factor >= 1.0;
    if (!ok) return fail(ctx, KT_EINVAL, "invalid surrogate domain (ranges reversed, or extent beyond 1.5)");
    if (int st = ensure_surrogate(ctx)) return st;
    kts::Domain d;
    d.spin_min = o->spin_min;
    d.spin_max = o->spin_max;
    d.inc_min = o->inclination_min;
    d.inc_max = o->inclination_max;
    d.r_cam_min = o->r_cam_min;
    d.r_cam_max = o->r_cam_max;
    d.extent = o->extent;
    d.disk_outer = o->disk_outer;
    d.escape_radius = o->escape_radius;
    d.capture_factor = o->capture_factor;
    kts::set_domain(ctx->surrogate, d);
    return KT_OK;
}

int kt_observer_frame(double a, double r, double theta, const double u[4], double e[16]) {
    if (!u || !e || !(std::fabs(a) <= 1.0) || !(r > 0.0)) return KT_EINVAL;
    double g[4][4], w[4];
    metric_tensor(a, r, theta, g);
    for (int i = 0; i < 4; ++i) w[i] = u[i];
    const double n = dot4(g, w, w);
    if (!(n < 0.0) || !(w[0] > 0.0)) return KT_EINVAL;
    const double s = 1.0 / std::sqrt(-n);
    for (int i = 0; i < 4; ++i) w[i] *= s;
    double f[4][4];
    observer_frame(a, r, theta, w, f);
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) e[4 * j + i] = f[j][i];
    return KT_OK;
}

int kt_geodesic_four_velocity(double a, double E, double L, double Q, double r, double theta, int radial_sign,
                              int polar_sign, double u[4]) {
    if (!u || !(std::fabs(a) <= 1.0) || !(r > 0.0)) return KT_EINVAL;
    // kerr::fourVelocity (src/kerr.cpp:414-418) of a massive body (mu = 1): the Kerr-Schild
    // derivatives of kerr.cpp:331-351 divided by Sigma, with inflowFactor (kerr.cpp:323-329)
    const double sn = hsin(theta), cs = hcos(theta);
    const double sn2 = std::fmax(sn * sn, 1e-18), a2 = a * a;
    const double delta = r * r - 2.0 * r + a2, rra = r * r + a2;
    const double P = E * rra - a * L, x = L - a * E;
    const double K = r * r + x * x + Q;
    const double R = P * P - delta * K;
    const double Th = Q - cs * cs * (a2 * (1.0 - E * E) + L * L / sn2);
    const double tol = 1e-12 * (P * P + std::fabs(delta) * K + std::fabs(Q) + L * L + 1.0);
    if (R < -tol || Th < -tol) return KT_EINVAL;
    const double vr = (radial_sign > 0 ? 1.0 : -1.0) * std::sqrt(std::fmax(R, 0.0));
    const double vth = (polar_sign > 0 ? 1.0 : -1.0) * std::sqrt(std::fmax(Th, 0.0));
    double inflow;
    if (P * vr < 0.0) {
        inflow = K / (P - vr);
    } else {
        const double d = std::fabs(delta) > 1e-12 ? delta : std::copysign(1e-12, delta);
        inflow = (P + vr) / d;
    }
    const double sigma = r * r + a2 * cs * cs;
    u[0] = (rra * inflow - a * (a * E * sn * sn - L) - vr) / sigma;
    u[1] = vr / sigma;
    u[2] = vth / sigma;
    u[3] = (a * inflow - a * E + L / sn2) / sigma;
    return KT_OK;
}

// persistent particle waves: 8 per CU (two blocks of four waves; the kernel holds 2 per SIMD)
static int particle_grid(const kt_ctx* ctx) { return std::max(1, ctx->cu_count) * 2; }

int kt_particle_from_zamo(double a, double r, double theta, const double v[3], double out[5]) {
    if (!v || !out || !(std::fabs(a) <= 1.0) || !(r > horizon(a))) return KT_EINVAL;
    // kerr::launchFromZamo with mu = 1 (src/kerr.cpp:365-406): the ZAMO frame's lapse and
    // circumference radius give E and L, Sigma Delta and Sigma the radial and polar momenta
    const double sn = hsin(theta), cs = hcos(theta), a2 = a * a;
    const double sigma = r * r + a2 * cs * cs, delta = r * r - 2.0 * r + a2;
    const double A = (r * r + a2) * (r * r + a2) - a2 * delta * sn * sn;
    const double omega = 2.0 * a * r / A;
    const double v2 = std::min(v[0] * v[0] + v[1] * v[1] + v[2] * v[2], 1.0 - 1e-12);
    const double gamma = 1.0 / std::sqrt(1.0 - v2);
    const double L = gamma * v[2] * std::sqrt(A / sigma) * sn;
    const double E = gamma * std::sqrt(sigma * delta / A) + omega * L;
    const double vr = gamma * v[0] * std::sqrt(sigma * delta);
    const double vth = gamma * v[1] * std::sqrt(sigma);
    const double Q = vth * vth + cs * cs * (a2 * (1.0 - E * E) + L * L / std::max(sn * sn, 1e-18));
    out[0] = E;
    out[1] = L;
    out[2] = Q;
    out[3] = vr;
    out[4] = -sn * vth;
    return KT_OK;
}

int kt_particles_advance(kt_ctx* ctx, double a, const kt_particles* pt, long long n, double dt, double r_stop,
                         double r_escape, const kt_trace_opts* opts, void* hip_stream) {
    if (!ctx) return KT_EINVAL;
    try {
        if (!pt || n < 0) return fail(ctx, KT_EINVAL, "particles/n missing");
        if (!pt->r || !pt->mu || !pt->phi || !pt->vr || !pt->vmu || !pt->E || !pt->L || !pt->Q || !pt->fate)
            return fail(ctx, KT_EINVAL, "particles need r, mu, phi, vr, vmu, E, L, Q and fate");
        if (!(std::fabs(a) <= 1.0)) return fail(ctx, KT_EINVAL, "spin must lie in [-1, 1]");
        if (!(dt >= 0.0) || !std::isfinite(dt)) return fail(ctx, KT_EINVAL, "dt must be finite and >= 0");
        if (!(r_stop > 0.0)) r_stop = horizon(a) + 0.02;
        if (!(r_escape > r_stop)) return fail(ctx, KT_EINVAL, "r_escape must exceed r_stop");
        if (n == 0) return KT_OK;
        kt_trace_opts o;
        fill_defaults(opts, o);
        hipStream_t stream = static_cast<hipStream_t>(hip_stream);
        KT_HIP(hipSetDevice(ctx->device));
        ktp::ParticleLaunch L{};
        L.n = n;
        L.a = a;
        L.dt = dt;
        L.r_stop = r_stop;
        L.r_escape = r_escape;
        L.rtol = o.rtol;
        L.atol = o.atol;
        L.max_steps = o.max_steps;
        const size_t nn = (size_t)n;
        if (pt->on_device) {
            L.r = pt->r;
            L.mu = pt->mu;
            L.phi = pt->phi;
            L.vr = pt->vr;
            L.vmu = pt->vmu;
            L.E = pt->E;
            L.L = pt->L;
            L.Q = pt->Q;
            L.fate = pt->fate;
            L.steps = pt->steps;
            L.h = pt->h;
            hipError_t e = ktp::advance(L, ctx->d_counter, particle_grid(ctx), stream);
            return e == hipSuccess ? KT_OK : hip_fail(ctx, e, "particle kernel launch");
        }
        // host arrays: nine columns of doubles (the ninth the step sizes), then fate and
        // steps, staged on the device
        if (int st = ensure_stage(ctx, nn * (9 * 8 + 1 + 4) + 64)) return st;
        double* base = static_cast<double*>(ctx->d_stage);
        double* col[9];
        for (int i = 0; i < 9; ++i) col[i] = base + i * nn;
        const double* src[8] = {pt->r, pt->mu, pt->phi, pt->vr, pt->vmu, pt->E, pt->L, pt->Q};
        for (int i = 0; i < 8; ++i) KT_HIP(hipMemcpyAsync(col[i], src[i], 8 * nn, hipMemcpyHostToDevice, stream));
        if (pt->h) KT_HIP(hipMemcpyAsync(col[8], pt->h, 8 * nn, hipMemcpyHostToDevice, stream));
        int8_t* fate = reinterpret_cast<int8_t*>(base + 9 * nn);
        int32_t* steps = pt->steps ? reinterpret_cast<int32_t*>(fate + ((nn + 7) & ~(size_t)7)) : nullptr;
        KT_HIP(hipMemcpyAsync(fate, pt->fate, nn, hipMemcpyHostToDevice, stream));
        L.r = col[0];
        L.mu = col[1];
        L.phi = col[2];
        L.vr = col[3];
        L.vmu = col[4];
        L.E = col[5];
        L.L = col[6];
        L.Q = col[7];
        L.fate = fate;
        L.steps = steps;
        L.h = pt->h ? col[8] : nullptr;
        hipError_t e = ktp::advance(L, ctx->d_counter, particle_grid(ctx), stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "particle kernel launch");
        double* dst[5] = {pt->r, pt->mu, pt->phi, pt->vr, pt->vmu};
        for (int i = 0; i < 5; ++i) KT_HIP(hipMemcpyAsync(dst[i], col[i], 8 * nn, hipMemcpyDeviceToHost, stream));
        KT_HIP(hipMemcpyAsync(pt->fate, fate, nn, hipMemcpyDeviceToHost, stream));
        if (steps) KT_HIP(hipMemcpyAsync(pt->steps, steps, 4 * nn, hipMemcpyDeviceToHost, stream));
        if (pt->h) KT_HIP(hipMemcpyAsync(pt->h, col[8], 8 * nn, hipMemcpyDeviceToHost, stream));
        KT_HIP(hipStreamSynchronize(stream));
        return KT_OK;
    } catch (const std::exception& ex) {
        return fail(ctx, KT_EINVAL, ex.what());
    }
}

double kt_horizon_radius(double a) { return horizon(a); }
double kt_isco_radius(double a, int prograde) { return isco(a, prograde); }
double kt_photon_orbit_radius(double a, int prograde) { return photon_orbit(a, prograde); }
double kt_circular_omega(double a, double r, int prograde) {
    double s = prograde ? 1.0 : -1.0;
    return s / (r * std::sqrt(r) + s * a);
}

int kt_critical_curve(double a, double inclination, int samples, double* alpha, double* beta, int cap) {
    if (!alpha || !beta || cap <= 0 || naked(a)) return 0;  // a naked singularity casts no such shadow edge
    const double sinI = hsin(inclination), cosI = hcos(inclination);
    if (std::fabs(a) < 1e-6) {
        int n = 0;
        for (int k = 0; k < samples && n < cap; ++k, ++n) {
            const double psi = 2.0 * kPi * k / samples;
            alpha[n] = std::sqrt(27.0) * hcos(psi);
            beta[n] = std::sqrt(27.0) * hsin(psi);
        }
        return n;
    }
    const double r1 = photon_orbit(a, 1), r2 = photon_orbit(a, 0);
    const int steps = std::max(samples, 16) * 8;
    std::vector<double> ua, ub;
    for (int k = 0; k <= steps; ++k) {
        const double x = 0.5 - 0.5 * hcos(kPi * k / steps);
        const double r = r1 + (r2 - r1) * x;
        const double xi = (r * r * (3.0 - r) - a * a * (r + 1.0)) / (a * (r - 1.0));
        const double eta = r * r * r * (4.0 * a * a - r * (r - 3.0) * (r - 3.0)) / (a * a * (r - 1.0) * (r - 1.0));
        const double b2 = eta + a * a * cosI * cosI - xi * xi * cosI * cosI / (sinI * sinI);
        if (b2 < 0.0) continue;
        ua.push_back(-xi / sinI);
        ub.push_back(std::sqrt(b2));
    }
    int n = 0;
    for (size_t k = 0; k < ua.size() && n < cap; ++k, ++n) {
        alpha[n] = ua[k];
        beta[n] = ub[k];
    }
    for (size_t k = ua.size(); k-- > 0 && n < cap; ++n) {
        alpha[n] = ua[k];
        beta[n] = -ub[k];
    }
    return n;
}

double kt_critical_radius(const double* alpha, const double* beta, int n, double psi) {
    if (!alpha || !beta || n <= 0) return 0.0;
    const double dx = hcos(psi), dy = hsin(psi);
    double best = 0.0;
    for (int k = 0; k < n; ++k) {
        const int q = (k + 1) % n;
        const double ex = alpha[q] - alpha[k], ey = beta[q] - beta[k];
        const double det = ex * dy - ey * dx;
        if (std::fabs(det) < 1e-15) continue;
        const double t = (beta[k] * dx - alpha[k] * dy) / det;
        const double s = (ex * beta[k] - ey * alpha[k]) / det;
        if (t >= 0.0 && t <= 1.0 && s > 0.0 && s > best) best = s;
    }
    return best;
}

}  // extern "C"
