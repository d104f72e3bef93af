// kt_trace.hip -- Kerr photon trace kernels for gfx950.
//
// Two launch forms over the same per-lane integrator (kt_device.hpp):
//   persistent: a grid sized to the chip; each wave pulls ray indices from one
//               device work queue (a 64-bit atomic head).  Every
//               `steps_per_dequeue` DP5 attempts the wave refills the lanes whose
//               rays ended, so that short rays (disk hits, direct captures) do not
//               leave lanes idle behind long ones (photon-ring orbits).
//   one-shot:   one lane per ray, no refill (the simple form, kept for comparison).
// The per-view table is staged in LDS when it fits (<= 64 views).
#include "kt_internal.hpp"

using namespace ktd;

namespace {

__device__ __forceinline__ int find_view(const ViewDev* v, int n, long long idx) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (v[mid].offset <= idx) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ const ViewDev* stage_views(const kti::TraceLaunch& L, ViewDev* sv) {
    if (L.n_views > kti::kMaxLdsViews) return L.views;
    const int words = (int)(sizeof(ViewDev) / sizeof(unsigned long long)) * L.n_views;
    const unsigned long long* src = reinterpret_cast<const unsigned long long*>(L.views);
    unsigned long long* dst = reinterpret_cast<unsigned long long*>(sv);
    for (int i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
    return sv;
}

template <class R>
__device__ __forceinline__ void start_ray(Lane<R>& s, const ViewDev* views, int n_views, long long idx) {
    const int vid = find_view(views, n_views, idx);
    const ViewDev& v = views[vid];
    Photon<double> p;
    double y0[5];
    launch_photon(v, idx - v.offset, p, y0);
    lane_start(s, v, p, y0);
}

template <class R>
__device__ __forceinline__ R tol_r(const kti::TraceLaunch& L);
template <>
__device__ __forceinline__ double tol_r<double>(const kti::TraceLaunch& L) { return L.rtol; }
template <>
__device__ __forceinline__ float tol_r<float>(const kti::TraceLaunch& L) { return (float)L.rtol32; }
template <class R>
__device__ __forceinline__ R tol_a(const kti::TraceLaunch& L);
template <>
__device__ __forceinline__ double tol_a<double>(const kti::TraceLaunch& L) { return L.atol; }
template <>
__device__ __forceinline__ float tol_a<float>(const kti::TraceLaunch& L) { return (float)L.atol32; }

__global__ void k_reset(unsigned long long* counter, unsigned long long* stats) {
    if (threadIdx.x == 0) *counter = 0ull;
    if (threadIdx.x < kti::kStatCount) stats[threadIdx.x] = 0ull;
}

template <class R>
__global__ __launch_bounds__(256) void k_trace_persistent(kti::TraceLaunch L) {
    __shared__ ViewDev sv[kti::kMaxLdsViews];
    const ViewDev* views = stage_views(L, sv);
    const int lane = (int)(threadIdx.x & 63u);
    const R rtol = tol_r<R>(L), atol = tol_a<R>(L);
    const int K = L.steps_per_dequeue;

    Lane<R> s;
    bool active = false;
    bool drained = false;
    long long my = 0;
    unsigned long long my_steps = 0;
    for (;;) {
        if (!drained) {
            const unsigned long long idle = __ballot(!active);
            const int nidle = __popcll(idle);
            if (nidle > 0) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(L.counter, (unsigned long long)nidle);
                base = __shfl(base, 0);
                if (base + (unsigned long long)nidle >= (unsigned long long)L.total) drained = true;
                if (!active) {
                    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
                    const long long q = (long long)(base + (unsigned long long)__popcll(idle & below));
                    if (q < L.total) {
                        my = L.ray_list ? L.ray_list[q] : q;
                        start_ray(s, views, L.n_views, my);
                        active = true;
                    }
                }
            }
        }
        if (__ballot(active) == 0ull) break;
        for (int k = 0; k < K; ++k) {
            if (active) {
                Result r;
                if (lane_step(s, rtol, atol, L.max_steps, r)) {
                    write_result(L.out, my, r, s.steps, 0);
                    my_steps += (unsigned long long)s.steps;
                    active = false;
                }
            }
        }
    }
    if (my_steps) atomicAdd(&L.stats[kti::kStatSteps], my_steps);
}

template <class R>
__global__ __launch_bounds__(256) void k_trace_oneshot(kti::TraceLaunch L) {
    __shared__ ViewDev sv[kti::kMaxLdsViews];
    const ViewDev* views = stage_views(L, sv);
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= L.total) return;
    const R rtol = tol_r<R>(L), atol = tol_a<R>(L);
    const long long my = L.ray_list ? L.ray_list[q] : q;
    Lane<R> s;
    start_ray(s, views, L.n_views, my);
    Result r;
    while (!lane_step(s, rtol, atol, L.max_steps, r)) {
    }
    write_result(L.out, my, r, s.steps, 0);
    atomicAdd(&L.stats[kti::kStatSteps], (unsigned long long)s.steps);
}

}  // namespace

namespace kti {

hipError_t launch_trace(const TraceLaunch& L, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, L.counter, L.stats);
    if (L.total <= 0) return hipGetLastError();
    const bool fp32 = L.precision == 2;
    if (L.mode == 1) {
        const long long blocks = (L.total + L.block - 1) / L.block;
        if (fp32)
            hipLaunchKernelGGL(k_trace_oneshot<float>, dim3((unsigned)blocks), dim3(L.block), 0, stream, L);
        else
            hipLaunchKernelGGL(k_trace_oneshot<double>, dim3((unsigned)blocks), dim3(L.block), 0, stream, L);
    } else {
        if (fp32)
            hipLaunchKernelGGL(k_trace_persistent<float>, dim3(L.grid), dim3(L.block), 0, stream, L);
        else
            hipLaunchKernelGGL(k_trace_persistent<double>, dim3(L.grid), dim3(L.block), 0, stream, L);
    }
    return hipGetLastError();
}

int trace_kernel_vgprs(int) { return 0; }

}  // namespace kti
