// kt_inspect.cpp -- command-line ray inspector and image driver over include/kerr_trace.hpp.
//
// The C++20 counterpart of the reference's `--inspect X,Y` (App::inspectPixel,
// src/app.cpp:1145-1192, printing the RayReport fields of src/app.cpp:771-784: the end,
// r_end, g, turns, r_min, delay, b and q) and of an eht::render call: one line of JSON per
// request.
//
//   kt_inspect <spin> <r_cam> <incl_deg> <half_width> <width> <height> [col row]
//   kt_inspect eht <spin> <incl_deg> <half_width> <size> [cancel_after_ms]
//       eht::render of M87*'s defaults (230 GHz, 6.5e9 M_sun, the default flow); with
//       cancel_after_ms >= 0 another thread sets the cancel flag after that many ms
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>

#include "../../include/kerr_trace.hpp"

// a JSON number, or null for the NaN of an output that does not apply to the ray's class
static std::string num(double x) {
    if (!std::isfinite(x)) return "null";
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", x);
    return buf;
}

static int eht(int argc, char** argv) {
    kt::EhtView v;
    v.a = std::atof(argv[2]);
    v.inclination = std::atof(argv[3]) * 3.14159265358979323846 / 180.0;
    v.half_width = std::atof(argv[4]);
    v.size = std::atoi(argv[5]);
    const int after = argc > 6 ? std::atoi(argv[6]) : -1;
    kt::Flow flow;
    std::atomic<int> cancel{0};
    std::thread setter;
    if (after >= 0)
        setter = std::thread([&] {
            std::this_thread::sleep_for(std::chrono::milliseconds(after));
            cancel = 1;
        });
    kt::Tracer tracer(0);
    const kt::EhtImage img = tracer.ehtRender(v, flow, after >= 0 ? &cancel : nullptr);
    if (setter.joinable()) setter.join();
    // rows with any light or shadow: a cancelled render leaves the rest zero
    int rows = 0;
    double total = 0.0;
    for (int r = 0; r < img.size; ++r) {
        bool any = false;
        for (int c = 0; c < img.size; ++c) {
            const size_t k = static_cast<size_t>(r) * img.size + c;
            total += img.intensity[k];
            any = any || img.intensity[k] != 0.0 || img.shadow[k] != 0;
        }
        rows += any;
    }
    std::printf("{\"size\": %d, \"cancelled\": %s, \"rows_with_light\": %d, \"intensity_sum\": %s}\n", img.size,
                img.cancelled ? "true" : "false", rows, num(total).c_str());
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 6 && std::strcmp(argv[1], "eht") == 0) {
        try {
            return eht(argc, argv);
        } catch (const kt::Error& e) {
            std::fprintf(stderr, "kt_inspect: %s (status %d)\n", e.what(), e.status);
            return 1;
        }
    }
    if (argc != 7 && argc != 9) {
        std::fprintf(stderr, "usage: %s spin r_cam incl_deg half_width width height [col row]\n", argv[0]);
        return 2;
    }
    kt::View v;
    v.spin = std::atof(argv[1]);
    v.r_cam = std::atof(argv[2]);
    v.theta_cam = std::atof(argv[3]) * 3.14159265358979323846 / 180.0;
    v.width = std::atoi(argv[5]);
    v.height = std::atoi(argv[6]);
    v.half_width = std::atof(argv[4]);
    v.half_height = v.half_width * v.height / v.width;
    try {
        kt::Tracer tracer(0);
        if (argc == 9) {
            const int col = std::atoi(argv[7]), row = std::atoi(argv[8]);
            const double px = 2.0 * v.half_width / v.width, py = 2.0 * v.half_height / v.height;
            const double alpha = -v.half_width + (col + 0.5) * px, beta = v.half_height - (row + 0.5) * py;
            const kt::RayReport rep = tracer.inspect(v, alpha, beta);
            const kt::Ray& r = rep.ray;
            std::printf("{\"alpha\": %s, \"beta\": %s, \"end\": %d, \"theta_inf\": %s, \"phi_inf\": %s, "
                        "\"r_hit\": %s, \"g\": %s, \"steps\": %d, \"r_end\": %s, \"turns\": %s, \"r_min\": %s, "
                        "\"delay\": %s, \"b\": %s, \"q\": %s}\n",
                        num(alpha).c_str(), num(beta).c_str(), static_cast<int>(r.end), num(r.thetaInf).c_str(),
                        num(r.phiInf).c_str(), num(r.rHit).c_str(), num(r.g).c_str(), r.steps, num(rep.rEnd).c_str(),
                        num(rep.turns).c_str(), num(rep.rMin).c_str(), num(rep.delay).c_str(), num(rep.b).c_str(),
                        num(rep.q).c_str());
            return 0;
        }
        const kt::Image img = tracer.render(v);
        long long count[4] = {0, 0, 0, 0}, steps = 0;
        for (size_t k = 0; k < img.cls.size(); ++k) {
            ++count[img.cls[k] & 3];
            steps += img.steps[k];
        }
        const kt_stats s = tracer.stats();
        std::printf("{\"rays\": %zu, \"escaped\": %lld, \"captured\": %lld, \"disk\": %lld, \"unfinished\": %lld, "
                    "\"steps\": %lld, \"kernel_ms\": %.4f}\n",
                    img.cls.size(), count[0], count[1], count[2], count[3], steps, s.kernel_ms);
    } catch (const kt::Error& e) {
        std::fprintf(stderr, "kt_inspect: %s (status %d)\n", e.what(), e.status);
        return 1;
    }
    return 0;
}
