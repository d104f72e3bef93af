// kerr_trace.hpp -- header-only C++20 face of libkerr_trace.so (over include/kerr_trace.h).
//
// Shaped like the reference's own calls so that a caller swaps one for the other:
//   kt::Tracer::render(view)            per-ray classes, sky angles, disk r and g of a view
//   kt::Tracer::traceRay(view, a, b)    one Bardeen-plane ray
//   kt::Tracer::ehtRender(view, flow, cancel)
//                                       for eht::render(View, Flow, threads, cancel)  (src/eht.h:140)
//   kt::Tracer::ehtTraceRay(view, flow, a, b)
//                                       for eht::traceRay(View, Flow, alpha, beta)   (src/eht.h:130)
//   kt::Tracer::ehtCalibrate(view, flow, jy, theta, cancel)
//                                       for eht::calibrate(...)                      (src/eht.h:152)
//   kt::Tracer::traceLooks(view, looks) for trajectory::traceRay(..., look, scene)   (src/trajectory.h:118)
//   kt::Tracer::inspect(view, ...)      its RayReport: turns, rMin, delay, b, q       (src/trajectory.h:97-113)
//   kt::horizonRadius, iscoRadius, photonOrbitRadius, circularOmega              (src/kerr.h:39-45, 71)
//   kt::criticalCurve, criticalRadius                                             (src/eht.h:66-68)
// Values come back by value in std::vectors (as eht::Image, trajectory::RayReport);
// failures throw kt::Error (the C layer below never throws).
#pragma once

#include <array>
#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "kerr_trace.h"

namespace kt {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

enum class RayEnd : int { Sky = KT_ESCAPED, Horizon = KT_CAPTURED, Disk = KT_DISK, Unfinished = KT_UNFINISHED };

struct View : kt_view {
    View() { kt_view_default(this); }
    long long rays() const { return static_cast<long long>(width) * height; }
};

struct Options : kt_trace_opts {
    Options() { kt_trace_opts_default(this); }
};

// eht::View and eht::Flow (src/eht.h:84-127) with the reference's defaults
struct EhtView : kt_eht_view {
    EhtView() { kt_eht_view_default(this); }
};
struct Flow : kt_eht_flow {
    Flow() { kt_eht_flow_default(this); }
};

// eht::Image (src/eht.h:132-138): rows from the top; cancelled when the flag stopped it
// (rows not traced are zero, as eht::render leaves them)
struct EhtImage {
    int size = 0;
    double halfWidth = 0.0;
    std::vector<double> intensity;
    std::vector<uint8_t> shadow;
    bool cancelled = false;
    double pixel() const { return 2.0 * halfWidth / size; }
};

// eht::Ray (src/eht.h:123-128)
struct EhtRay {
    double intensity = 0.0;
    bool captured = false;
    double minRadius = 0.0;
    int steps = 0;
};

// The cancel flag the C layer reads (const volatile int*): a std::atomic<int> with int's
// representation.
static_assert(sizeof(std::atomic<int>) == sizeof(int) && std::atomic<int>::is_always_lock_free,
              "std::atomic<int> must be a plain int to serve as the cancel flag");

struct Ray {
    RayEnd end = RayEnd::Unfinished;
    bool fp64Restepped = false, surrogate = false;
    double thetaInf = 0.0, phiInf = 0.0;  // asymptotic sky direction (Sky)
    double rHit = 0.0, g = 0.0;           // disk radius and redshift (Disk)
    int steps = 0;
};

// trajectory::RayReport (src/trajectory.h:97-113) of one ray, as App::inspectPixel prints it
// (src/app.cpp:776-777): the ray, where it ended (the disk radius, or r_cap or r_esc, where
// the trace stops), how many times it went around, its closest approach, the Kerr-Schild
// coordinate time it spent, and b = L/E, q = Q/E^2.
struct RayReport {
    Ray ray;
    double rEnd = 0.0, turns = 0.0, rMin = 0.0, delay = 0.0, b = 0.0, q = 0.0;
};

// Per-pixel outputs of one or more views, row-major from the top of each view.
struct Image {
    int width = 0, height = 0;
    std::vector<int8_t> cls;
    std::vector<double> thetaInf, phiInf, rHit, g;
    std::vector<int32_t> steps;
    Ray at(size_t k) const {
        Ray r;
        r.end = static_cast<RayEnd>(cls[k] & 7);
        r.fp64Restepped = (cls[k] & KT_FLAG_FP64_RESTEP) != 0;
        r.surrogate = (cls[k] & KT_FLAG_SURROGATE) != 0;
        r.thetaInf = thetaInf[k];
        r.phiInf = phiInf[k];
        r.rHit = rHit[k];
        r.g = g[k];
        r.steps = steps[k];
        return r;
    }
};

class Tracer {
public:
    explicit Tracer(int device = 0) {
        kt_device_opts o{device};
        int s = kt_create(&o, &ctx_);
        if (s != KT_OK) throw Error(s, "kt_create failed");
    }
    ~Tracer() { kt_destroy(ctx_); }
    Tracer(const Tracer&) = delete;
    Tracer& operator=(const Tracer&) = delete;

    // Every pixel of `views` in one launch, copied back to host memory.
    Image render(const std::vector<View>& views, const Options& opts = Options()) {
        long long n = 0;
        for (const View& v : views) n += v.rays();
        Image img;
        img.width = views.size() == 1 ? views[0].width : 0;
        img.height = views.size() == 1 ? views[0].height : 0;
        resize(img, static_cast<size_t>(n));
        kt_rays_out out = hostOut(img);
        std::vector<kt_view> raw(views.begin(), views.end());
        check(kt_trace_views(ctx_, raw.data(), static_cast<int>(raw.size()), &opts, &out, nullptr));
        return img;
    }
    Image render(const View& view, const Options& opts = Options()) { return render(std::vector<View>{view}, opts); }

    // One Bardeen-plane ray (alpha, beta).
    Ray traceRay(const View& view, double alpha, double beta, const Options& opts = Options()) {
        const double ab[2] = {alpha, beta};
        return traceParams(view, ab, 1, opts).at(0);
    }

    // Pinhole rays arriving from unit directions `looks` of the camera's ZAMO frame.
    Image traceLooks(const View& view, const std::vector<std::array<double, 3>>& looks,
                     const Options& opts = Options()) {
        return traceParams(view, looks.front().data(), static_cast<long long>(looks.size()), opts);
    }

    // The full report of explicit rays (FP64): (alpha, beta) pairs of a Bardeen view, or unit
    // looks of a pinhole or observer view; the counterpart of trajectory::traceRay's return.
    std::vector<RayReport> inspect(const View& view, const double* params, long long n,
                                   Options opts = Options()) {
        opts.precision = KT_PREC_FP64;
        opts.use_surrogate = 0;
        Image img;
        resize(img, static_cast<size_t>(n));
        std::vector<double> turns(n), rmin(n), delay(n), b(n), q(n);
        kt_rays_out out = hostOut(img);
        out.turns = turns.data();
        out.r_min = rmin.data();
        out.delay = delay.data();
        out.b = b.data();
        out.q = q.data();
        check(kt_trace_rays(ctx_, &view, params, n, &opts, &out, nullptr));
        const double r_cap = view.capture_factor * kt_horizon_radius(view.spin);
        const double r_esc = view.escape_radius > 1.5 * view.r_cam ? view.escape_radius : 1.5 * view.r_cam;
        std::vector<RayReport> rep(static_cast<size_t>(n));
        for (size_t k = 0; k < rep.size(); ++k) {
            rep[k].ray = img.at(k);
            const RayEnd e = rep[k].ray.end;
            rep[k].rEnd = e == RayEnd::Disk ? rep[k].ray.rHit : e == RayEnd::Horizon ? r_cap
                        : e == RayEnd::Sky ? r_esc : 0.0;
            rep[k].turns = turns[k];
            rep[k].rMin = rmin[k];
            rep[k].delay = delay[k];
            rep[k].b = b[k];
            rep[k].q = q[k];
        }
        return rep;
    }
    RayReport inspect(const View& view, double alpha, double beta, const Options& opts = Options()) {
        const double ab[2] = {alpha, beta};
        return inspect(view, ab, 1, opts).front();
    }

    // eht::render: the view's size^2 pixels with the light; `cancel` (may be null) is read
    // between slices of the image
    EhtImage ehtRender(const EhtView& view, const Flow& flow, const std::atomic<int>* cancel = nullptr,
                       Options opts = Options()) {
        EhtImage img;
        img.size = view.size;
        img.halfWidth = view.half_width;
        const size_t n = static_cast<size_t>(view.size) * view.size;
        img.intensity.assign(n, 0.0);
        img.shadow.assign(n, 0);
        opts.cancel = reinterpret_cast<const volatile int*>(cancel);
        kt_eht_out out{img.intensity.data(), img.shadow.data(), nullptr, nullptr, 0};
        const int s = kt_eht_render(ctx_, &view, &flow, &opts, &out, nullptr);
        img.cancelled = s == KT_ECANCELLED;
        if (!img.cancelled) check(s);
        return img;
    }

    // eht::traceRay with the light
    EhtRay ehtTraceRay(const EhtView& view, const Flow& flow, double alpha, double beta,
                       const Options& opts = Options()) {
        const double ab[2] = {alpha, beta};
        double intensity = 0.0, rmin = 0.0;
        uint8_t shadow = 0;
        int32_t steps = 0;
        kt_eht_out out{&intensity, &shadow, &rmin, &steps, 0};
        check(kt_eht_trace_rays(ctx_, &view, &flow, ab, 1, &opts, &out, nullptr));
        return EhtRay{intensity, shadow != 0, rmin, steps};
    }

    // eht::calibrate: the density n0 for which a 48 x 48 render has targetJy janskys (the
    // estimate so far when `cancel` turns nonzero)
    double ehtCalibrate(const EhtView& view, const Flow& flow, double targetJy, double thetaGUas,
                        const std::atomic<int>* cancel = nullptr, Options opts = Options()) {
        opts.cancel = reinterpret_cast<const volatile int*>(cancel);
        double n0 = flow.n0;
        const int s = kt_eht_calibrate(ctx_, &view, &flow, targetJy, thetaGUas, &opts, &n0, nullptr);
        if (s != KT_ECANCELLED) check(s);
        return n0;
    }

    kt_stats stats() const {
        kt_stats s{};
        check(kt_last_stats(ctx_, &s));
        return s;
    }

    kt_ctx* handle() { return ctx_; }

private:
    kt_ctx* ctx_ = nullptr;

    void check(int s) const {
        if (s != KT_OK) throw Error(s, kt_last_error_string(ctx_));
    }
    static void resize(Image& img, size_t n) {
        img.cls.resize(n);
        img.thetaInf.resize(n);
        img.phiInf.resize(n);
        img.rHit.resize(n);
        img.g.resize(n);
        img.steps.resize(n);
    }
    static kt_rays_out hostOut(Image& img) {
        return kt_rays_out{img.cls.data(), img.thetaInf.data(), img.phiInf.data(), img.rHit.data(),
                           img.g.data(), img.steps.data(), 0};
    }
    Image traceParams(const View& view, const double* params, long long n, const Options& opts) {
        Image img;
        resize(img, static_cast<size_t>(n));
        kt_rays_out out = hostOut(img);
        check(kt_trace_rays(ctx_, &view, params, n, &opts, &out, nullptr));
        return img;
    }
};

inline double horizonRadius(double a) { return kt_horizon_radius(a); }
inline double iscoRadius(double a, bool prograde) { return kt_isco_radius(a, prograde ? 1 : 0); }
inline double photonOrbitRadius(double a, bool prograde) { return kt_photon_orbit_radius(a, prograde ? 1 : 0); }
inline double circularOmega(double a, double r, bool prograde) { return kt_circular_omega(a, r, prograde ? 1 : 0); }

struct Point {
    double alpha, beta;
};
inline std::vector<Point> criticalCurve(double a, double inclination, int samples) {
    const int cap = 16 * (samples > 16 ? samples : 16) + 8 + samples;
    std::vector<double> al(cap), be(cap);
    const int n = kt_critical_curve(a, inclination, samples, al.data(), be.data(), cap);
    std::vector<Point> c(n);
    for (int k = 0; k < n; ++k) c[k] = {al[k], be[k]};
    return c;
}
inline double criticalRadius(const std::vector<Point>& curve, double psi) {
    std::vector<double> al(curve.size()), be(curve.size());
    for (size_t k = 0; k < curve.size(); ++k) {
        al[k] = curve[k].alpha;
        be[k] = curve[k].beta;
    }
    return kt_critical_radius(al.data(), be.data(), static_cast<int>(curve.size()), psi);
}

}  // namespace kt
