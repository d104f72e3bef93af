// kt_internal.hpp -- what the host side (kt_host.cpp) and the kernels (kt_trace.hip) share.
#pragma once

#include <hip/hip_runtime.h>

#include "kt_device.hpp"

namespace kti {

constexpr int kMaxLdsViews = 64;

// device-side counters written by the trace kernels
enum : int { kStatSteps = 0, kStatRestepRays = 1, kStatFallback = 2, kStatFp64Steps = 3, kStatCount = 4 };

struct TraceLaunch {
    const ktd::ViewDev* views;  // device array
    int n_views;
    long long total;            // rays in the batch
    unsigned long long* counter;  // work-queue head (zeroed by the launch)
    unsigned long long* stats;    // kStatCount counters (zeroed by the launch)
    ktd::Outputs out;
    const long long* ray_list;  // optional: trace only these ray indices (fallback re-queue)
    const long long* total_dev; // optional: the number of rays, read on the device (total = capacity)
    // optional, with ray_list: the list holds kQueueBuckets regions bucket_stride apart, region
    // b holding bucket_count[b] entries; ray q of the launch is taken from them in order
    const unsigned long long* bucket_count;
    long long bucket_stride;
    int n_buckets;              // regions of that queue to read: kQueueBuckets (0 means the same), or 1
    double rtol, atol, rtol32, atol32, restep_drift;
    int max_steps;
    int precision;              // KT_PREC_*
    int mode;                   // KT_MODE_*
    int block, grid, steps_per_dequeue, waves_per_simd;
    int ks;                     // 1: observer cameras (FP64, Kerr-Schild phi)
    int no_reset;               // 1: the caller has zeroed the queue head and counters (reset_counters)
};

constexpr int kQueueBuckets = 3;

// A fallback queue in device memory: kQueueBuckets regions `stride` entries apart, region b
// holding count[b] ray indices, which the integrator reads; count[kQueueBuckets] is a spare
// counter (the surrogate's tile counter); region kQueueBuckets with count[kUnsureCount]
// holds the rays the closed form's FP32 pass leaves to its FP64 one; *total the sum of the
// integrator's regions as it reads it.
constexpr int kUnsureRegion = kQueueBuckets, kUnsureCount = kQueueBuckets + 1, kQueueCounts = kQueueBuckets + 2;
struct RayQueue {
    long long* list;  // (kQueueBuckets + 1) x stride
    unsigned long long* count;  // kQueueCounts
    long long* total;
    long long stride;
};

// The closed-form solution (kt_closed.hpp) after the surrogate's network: its values for the
// rays the network answered as escaped or disk hits (fp32: an FP32 pass, then FP64 for the
// rays it leaves unsure); a ray it declines or classes otherwise joins the region of the
// network's class.
hipError_t launch_closed_form_polish(const TraceLaunch& L, const RayQueue& q, int fp32, hipStream_t stream);
hipError_t trace_closed_form(const TraceLaunch& L, const RayQueue& q, int fp32, hipStream_t stream, hipEvent_t* ev);

// launches the counter reset (unless L.no_reset) and the trace kernel on `stream`; returns a
// hipError_t
hipError_t launch_trace(const TraceLaunch& L, hipStream_t stream);
// the counter reset alone (for a caller that launches other work between it and the trace)
hipError_t reset_counters(const TraceLaunch& L, hipStream_t stream);

// The fixed device scratch of the no-allocation entry point, on one GPU: the counters
// (queue head + kStatCount) and a view table of max_views views.
struct Scratch {
    unsigned long long* counter = nullptr;
    ktd::ViewDev* views = nullptr;
    int max_views = 0;
};
// Resolves the scratch's device address on `device`, which must be the current device
// (once per device; later calls return the same address).  Allocates nothing.
hipError_t scratch_setup(int device, Scratch* out);
size_t scratch_bytes();

}  // namespace kti
