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
#include "kt_closed.hpp"

#include <hip/hip_ext.h>

#include <mutex>

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

// The view table in LDS, sized to the launch's views (dynamic shared memory, n_views x
// sizeof(ViewDev) when n_views <= kMaxLdsViews, else none and the table is read from
// global memory).  Sizing it to the views, not to the maximum, leaves a CU's LDS to other
// kernels running beside the integrator: the surrogate's network keeps its weights there.
extern __shared__ ViewDev view_lds[];

__host__ __device__ inline size_t view_lds_bytes(int n_views) {
    return n_views <= kti::kMaxLdsViews ? sizeof(ViewDev) * (size_t)n_views : 0;
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

// sum of x over the 64 lanes of the wave (all of them must execute the call)
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

// the queue slot of a launch's q-th ray: regions of a bucketed queue are read in order
struct QueueMap {
    long long c0, c1, c2, stride;
    bool bucketed;
    __device__ __forceinline__ explicit QueueMap(const kti::TraceLaunch& L)
        : c0(0), c1(0), c2(0), stride(0), bucketed(false) {
        if (L.bucket_count) {
            bucketed = true;
            c0 = (long long)L.bucket_count[0];
            if (L.n_buckets != 1) {
                c1 = (long long)L.bucket_count[1];
                c2 = (long long)L.bucket_count[2];
            }
            stride = L.bucket_stride;
        }
    }
    // the launch's ray count: the queue's length, else the device count, else the host's
    __device__ __forceinline__ long long total(const kti::TraceLaunch& L) const {
        if (bucketed) return c0 + c1 + c2;
        return L.total_dev ? *L.total_dev : L.total;
    }
    __device__ __forceinline__ long long slot(long long q) const {
        if (!bucketed || q < c0) return q;
        if (q < c0 + c1) return stride + (q - c0);
        return 2 * stride + (q - c0 - c1);
    }
};
static_assert(kti::kQueueBuckets == 3, "QueueMap reads three regions");

__device__ __forceinline__ void launch_of(const ViewDev* views, int n_views, long long idx, Launch& la) {
    const ViewDev& v = views[find_view(views, n_views, idx)];
    launch_photon(v, idx - v.offset, la);
}

// How one lane traces one ray, per precision.
template <class R>
struct Plain {
    using LaneT = Lane<R>;
    __device__ static void start(LaneT& s, const Launch& la) { lane_start<R, kPhiOut>(s, la); }
    __device__ static bool step(LaneT& s, const kti::TraceLaunch& L, Result& r);
    __device__ static int steps(const LaneT& s) { return s.steps; }
    __device__ static int fp64_steps(const LaneT&) { return 0; }
};
template <>
__device__ bool Plain<double>::step(Lane<double>& s, const kti::TraceLaunch& L, Result& r) {
    return lane_step<double, kPhiOut>(s, L.rtol, L.atol, L.max_steps, r);
}
template <>
__device__ bool Plain<float>::step(Lane<float>& s, const kti::TraceLaunch& L, Result& r) {
    return lane_step<float, kPhiOut>(s, (float)L.rtol32, (float)L.atol32, L.max_steps, r);
}

// Observer cameras (kind 3): FP64 with the Kerr-Schild phi, inside the horizon allowed.
struct PlainKS {
    using LaneT = Lane<double>;
    __device__ static void start(LaneT& s, const Launch& la) { lane_start<double, kPhiIn>(s, la); }
    __device__ static bool step(LaneT& s, const kti::TraceLaunch& L, Result& r) {
        return lane_step<double, kPhiIn>(s, L.rtol, L.atol, L.max_steps, r);
    }
    __device__ static int steps(const LaneT& s) { return s.steps; }
    __device__ static int fp64_steps(const LaneT&) { return 0; }
};

struct Mixed {
    using LaneT = MixedLane;
    __device__ static void start(LaneT& s, const Launch& la) {
        const Photon<double>& p = la.p;
        lane_start<double, kPhiOut>(s.d, la);
        s.pf.a = (float)p.a;
        s.pf.a2 = (float)p.a2;
        s.pf.E = (float)p.E;
        s.pf.L = (float)p.L;
        s.pf.Q = (float)p.Q;
        s.pf.aE = (float)p.aE;
        s.pf.aL = (float)p.aL;
        s.pf.K = (float)p.K;
        s.pf.C1 = (float)p.C1;
        s.pf.A2E2 = (float)p.A2E2;
        s.pf.twoE = (float)p.twoE;
        s.pf.twoA2E2 = (float)p.twoA2E2;
        s.r_fp64 = (float)la.r_fp64;
        s.fp64_steps = 0;
    }
    __device__ static bool step(LaneT& s, const kti::TraceLaunch& L, Result& r) {
        return lane_step_mixed(s, L.rtol32, L.atol32, (float)L.restep_drift, L.max_steps, r);
    }
    __device__ static int steps(const LaneT& s) { return s.d.steps; }
    __device__ static int fp64_steps(const LaneT& s) { return s.fp64_steps; }
};

__global__ void k_reset(unsigned long long* counter, unsigned long long* stats) {
    if (threadIdx.x == 0) *counter = 0ull;
    if (threadIdx.x < kti::kStatCount) stats[threadIdx.x] = 0ull;
}

// W: the occupancy the register allocator aims for (waves per SIMD); a tuning knob.
template <class Pol, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W))) void k_trace_persistent(kti::TraceLaunch L) {
    const ViewDev* views = stage_views(L, view_lds);
    const int lane = (int)(threadIdx.x & 63u);
    const int K = L.steps_per_dequeue;

    // a device-side count (the surrogate's fallback queue) overrides the launch size
    const QueueMap qm(L);
    const long long total = qm.total(L);
    typename Pol::LaneT s;
    bool active = false;
    bool drained = false;
    long long my = 0;
    unsigned long long my_steps = 0, my_fp64 = 0, my_restep_rays = 0;
    for (;;) {
        if (!drained) {
            const unsigned long long idle = __ballot(!active);
            const int nidle = __popcll(idle);
            if (nidle > 0) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(L.counter, (unsigned long long)nidle);
                base = __shfl(base, 0);
                if (base + (unsigned long long)nidle >= (unsigned long long)total) drained = true;
                if (!active) {
                    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
                    const long long q = (long long)(base + (unsigned long long)__popcll(idle & below));
                    if (q < total) {
                        my = L.ray_list ? L.ray_list[qm.slot(q)] : q;
                        Launch la;
                        launch_of(views, L.n_views, my, la);
                        Pol::start(s, la);
                        active = true;
                    }
                }
            }
        }
        if (__ballot(active) == 0ull) break;
        for (int k = 0; k < K; ++k) {
            if (active) {
                Result r;
                if (Pol::step(s, L, r)) {
                    const int f64 = Pol::fp64_steps(s);
                    write_result(L.out, my, r, Pol::steps(s), f64 > 0 ? kFlagRestep : 0);
                    my_steps += (unsigned long long)Pol::steps(s);
                    my_fp64 += (unsigned long long)f64;
                    my_restep_rays += f64 > 0 ? 1ull : 0ull;
                    active = false;
                }
            }
        }
    }
    if (my_steps) atomicAdd(&L.stats[kti::kStatSteps], my_steps);
    if (my_fp64) atomicAdd(&L.stats[kti::kStatFp64Steps], my_fp64);
    if (my_restep_rays) atomicAdd(&L.stats[kti::kStatRestepRays], my_restep_rays);
}

// One lane per ray.  With a count read on the device (the surrogate's fallback queue) the
// grid is sized for the likely count, not for the batch, and strides over the queue when
// it holds more rays than the grid has lanes.
template <class Pol>
__global__ __launch_bounds__(256) void k_trace_oneshot(kti::TraceLaunch L) {
    // issue priority over other kernels' waves on the same SIMD: a fallback queue's time is
    // the latency of its longest rays, while the surrogate's network beside it (another
    // frame, another stream) is throughput work that fills the issue slots left over
    if (L.total_dev || L.bucket_count) __builtin_amdgcn_s_setprio(3);
    const ViewDev* views = stage_views(L, view_lds);
    const QueueMap qm(L);
    const long long total = qm.total(L);
    const long long stride = (long long)gridDim.x * blockDim.x;
    unsigned long long st = 0, s64 = 0, nre = 0;
    for (long long base = (long long)blockIdx.x * blockDim.x; base < total; base += stride) {  // block-uniform
        const long long q = base + threadIdx.x;
        if (q >= total) break;
        const long long my = L.ray_list ? L.ray_list[qm.slot(q)] : q;
        typename Pol::LaneT s;
        Launch la;
        launch_of(views, L.n_views, my, la);
        Pol::start(s, la);
        Result r;
        while (!Pol::step(s, L, r)) {
        }
        const int f64 = Pol::fp64_steps(s);
        write_result(L.out, my, r, Pol::steps(s), f64 > 0 ? kFlagRestep : 0);
        st += (unsigned long long)Pol::steps(s);
        s64 += (unsigned long long)f64;
        nre += f64 > 0 ? 1ull : 0ull;
    }
    // the counters, summed over the wave first: one atomic per wave each (every lane of
    // the wave reaches this point; lanes past the end add 0)
    st = wave_sum(st);
    s64 = wave_sum(s64);
    nre = wave_sum(nre);
    if ((threadIdx.x & 63u) == 0u) {
        if (st) atomicAdd(&L.stats[kti::kStatSteps], st);
        if (s64) atomicAdd(&L.stats[kti::kStatFp64Steps], s64);
        if (nre) atomicAdd(&L.stats[kti::kStatRestepRays], nre);
    }
}

// The single-ray report: the one-shot FP64 trace (Plain<double>, or PlainKS for observer
// cameras) with report_step, which integrates exactly as lane_step does and accumulates
// turns, r_min and delay; they are written beside the usual outputs.
template <int PHI>
__global__ __launch_bounds__(256) void k_trace_report(kti::TraceLaunch L) {
    const ViewDev* views = stage_views(L, view_lds);
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if ((long long)blockIdx.x * blockDim.x >= L.total) return;  // the whole block is past the end
    unsigned long long st = 0;
    if (q < L.total) {
        Lane<double> s;
        Launch la;
        launch_of(views, L.n_views, q, la);
        lane_start<double, PHI>(s, la);
        ReportAcc acc;
        report_start<PHI>(s, acc);
        Result r;
        while (!report_step<PHI>(s, acc, L.rtol, L.atol, L.max_steps, r)) {
        }
        write_result(L.out, q, r, s.steps, 0);
        write_report(L.out, q, s.p, acc);
        st = (unsigned long long)s.steps;
    }
    st = wave_sum(st);
    if ((threadIdx.x & 63u) == 0u && st) atomicAdd(&L.stats[kti::kStatSteps], st);
}

// ---- the closed-form solution ---------------------------------------------------------------

// the closed form's inputs for launched ray `la` of view v: lambda, eta and the signs of the
// photon's motion at the camera; false for what it does not take: observer cameras inside
// the horizon (their rays start on the far side of r+), spins past 1
__device__ __forceinline__ bool closed_ray(const ViewDev& v, const Launch& la, ktc::Ray& ry) {
    if (la.dark || !(fabs(la.p.a) <= 1.0) || (v.kind == 3 && !(v.r_cam > v.r_plus))) return false;
    const Photon<double>& p = la.p;
    if (!(p.E > 0.0) || la.y0[3] == 0.0 || la.y0[4] == 0.0) return false;
    const double invE = 1.0 / p.E;
    ry.a = p.a;
    ry.E = p.E;
    ry.lam = p.L * invE;
    ry.eta = p.Q * invE * invE;
    ry.r_o = la.y0[0];
    ry.mu_o = la.y0[1];
    // Boyer-Lindquist phi at the camera: an observer view's phi_cam is the ingoing Kerr-Schild
    // one, phi_BL + psi(r)
    ry.phi_o = v.kind == 2 ? 0.0 : v.kind == 3 ? v.phi_cam - ks_shift(p.a, la.y0[0]) : v.phi_cam;
    ry.s_r = la.y0[3] < 0.0 ? -1 : 1;
    ry.s_mu = la.y0[4] < 0.0 ? -1 : 1;
    ry.r_in = la.r_in;
    ry.r_out = la.r_out;
    ry.r_cap = la.r_cap;
    ry.r_esc = la.r_esc;
    ry.disk = la.disk;
    ry.prograde = la.prograde;
    return true;
}

// The same for ray `local` of view v: a pixel of a Bardeen plane straight from its (alpha, beta)
// (src/eht.cpp:262-271: E = 1, L = -alpha sin i, Q = beta^2 + (alpha^2 - a^2) cos^2 i,
// inward, toward smaller theta when beta > 0; declined where R or M vanishes at the camera,
// as a launch with w_r or w_mu zero), other kinds through their launch
__device__ __forceinline__ bool closed_ray_of(const ViewDev& v, long long local, ktc::Ray& ry) {
    if (v.kind == 0 && !v.params && fabs(v.a) <= 1.0) {
        int row, col;
        pixel_rc(local, v.width, row, col);
        const double alpha = -v.half_width + ((double)col + 0.5) * v.px;
        const double beta = v.half_height - ((double)row + 0.5) * v.py;
        const double a = v.a, lam = -alpha * v.sinI;
        const double eta = beta * beta + (alpha * alpha - a * a) * v.cosI * v.cosI;
        const double m2 = v.m0 * v.m0;
        const double M = eta - (eta + lam * lam - a * a) * m2 - a * a * m2 * m2;  // sin^2 theta Theta at the camera
        const double r = v.r_cam, P = r * r + a * a - a * lam, lma = lam - a;
        const double R = P * P - (r * r - 2.0 * r + a * a) * (eta + lma * lma);  // R(r) at the camera
        if (!(M > 0.0) || !(R > 0.0)) return false;
        ry.a = a;
        ry.E = 1.0;
        ry.lam = lam;
        ry.eta = eta;
        ry.r_o = v.r_cam;
        ry.mu_o = v.m0;
        ry.phi_o = v.phi_cam;
        ry.s_r = -1;
        ry.s_mu = beta > 0.0 ? 1 : -1;
        ry.r_in = v.r_in;
        ry.r_out = v.r_out;
        ry.r_cap = v.r_cap;
        ry.r_esc = v.r_esc;
        ry.disk = v.disk;
        ry.prograde = v.prograde;
        return true;
    }
    Launch la;
    launch_photon(v, local, la);
    return closed_ray(v, la, ry);
}

// The closed-form passes, one lane per ray, the grid striding over the rays: over every ray
// of the launch (SRC kEvery), the rays the surrogate answered as escaped or disk hits (SRC
// kAnswered), or the unsure region (SRC kUnsure, its length read on the device), with the
// elliptic integrals in FP32 (FP32) or FP64.  Answered rays get the closed form's outputs
// (class | kFlagClosedForm, 0 steps, and kFlagSurrogate after the network); rays the FP32
// pass is unsure of join the unsure region, rays it declines (or, after the network, classes
// otherwise) the integrator's region 0 (after the network, that of the network's class).
// Joining a region takes one atomic per wave and region.
enum : int { kEvery = 0, kAnswered = 1, kUnsure = 2 };

template <int SRC, bool POLISH, bool FP32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FP32 ? 4 : 2))) void k_closed_form(
    kti::TraceLaunch L, kti::RayQueue rq) {
    const ViewDev* views = stage_views(L, view_lds);
    const int lane = (int)(threadIdx.x & 63u);
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long total = SRC == kUnsure ? (long long)rq.count[kti::kUnsureCount] : L.total;
    for (long long base = (long long)blockIdx.x * blockDim.x; base < total; base += stride) {  // block-uniform
        const long long i = base + threadIdx.x;
        long long q = 0;
        bool queue = false;
        int region = 0;
        if (i < total) {
            q = SRC == kUnsure ? rq.list[kti::kUnsureRegion * rq.stride + i] : i;
            int prior = -1;  // after the network: its class
            bool take = true;
            if (POLISH) {
                const int c = L.out.cls[q];
                prior = c & 7;
                take = (c & kFlagSurrogate) && (prior == kEscaped || prior == kDisk);
            }
            if (take) {
                const ViewDev& v = views[find_view(views, L.n_views, q)];
                ktc::Ray ry;
                ktc::Outcome o{ktc::kDeclined, 0.0, 0.0, 0.0, 0.0};
                if (closed_ray_of(v, q - v.offset, ry)) {
                    if constexpr (FP32) o = ktc::solve_t<float>(ry);
                    else o = ktc::solve_t<double>(ry);
                }
                if (o.cls == ktc::kUncertain) {
                    queue = true;
                    region = kti::kUnsureRegion;
                } else if (o.cls < 0 || (POLISH && o.cls != prior)) {
                    queue = true;
                    region = POLISH ? prior : 0;
                } else {
                    Result r;
                    r.cls = o.cls;
                    r.flags = 0;
                    r.theta = o.theta;
                    r.phi = o.phi;
                    r.r_hit = o.r_hit;
                    r.g = o.g;
                    write_result(L.out, q, r, 0, kFlagClosedForm | (POLISH ? kFlagSurrogate : 0));
                }
            }
        }
#pragma unroll
        for (int b = 0; b <= kti::kQueueBuckets; ++b) {
            const unsigned long long mask = __ballot(queue && region == b);
            if (mask) {
                const int first = __ffsll((long long)mask) - 1;
                unsigned long long at = 0;
                const int ci = b == kti::kUnsureRegion ? kti::kUnsureCount : b;
                if (lane == first) at = atomicAdd(rq.count + ci, (unsigned long long)__popcll(mask));
                at = __shfl(at, first);
                if (queue && region == b)
                    rq.list[b * rq.stride + (long long)(at + __popcll(mask & ((1ull << lane) - 1ull)))] = q;
            }
        }
    }
}

// the queue's region counts, and the launch's queue head and counters
__global__ void k_zero_queue(unsigned long long* count, unsigned long long* head, unsigned long long* stats) {
    if ((int)threadIdx.x < kti::kQueueCounts) count[threadIdx.x] = 0ull;
    if (threadIdx.x == 0) *head = 0ull;
    if (threadIdx.x < kti::kStatCount) stats[threadIdx.x] = 0ull;
}

// the queue's length (the sum of its regions) for the integrator and the stats
__global__ void k_queue_total(const unsigned long long* count, long long* total, unsigned long long* stat) {
    if (threadIdx.x == 0) {
        *total = (long long)count[0];  // the integrator's region
        if (stat) *stat = count[0];
    }
}

}  // namespace

// ---- the scratch of the entry point that allocates nothing (kt_trace_views_noalloc) --------
// The work-queue head and counters, then a view table of kMaxLdsViews views, in a fixed
// device array: each GPU the library runs on holds one copy until the process ends.  Host
// code reaches it only through the address hipGetSymbolAddress gives per device.
namespace {

constexpr size_t kScratchCounterBytes = 256;
constexpr size_t kScratchBytes = kScratchCounterBytes + sizeof(ViewDev) * kti::kMaxLdsViews;
static_assert(sizeof(unsigned long long) * (1 + kti::kStatCount) <= kScratchCounterBytes, "counters fit");

constexpr int kScratchDevices = 64;
std::mutex scratch_mu;
void* scratch_at[kScratchDevices] = {};  // per device ordinal; null until set up

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace kti {

size_t scratch_bytes() { return kScratchBytes; }

hipError_t scratch_setup(int device, Scratch* out) {
    *out = Scratch{};
    if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
    int current = -1;
    hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess) return e;
    if (current != device) return hipErrorInvalidDevice;  // the caller makes `device` current
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_at[device]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        scratch_at[device] = p;
    }
    char* base = static_cast<char*>(scratch_at[device]);
    out->counter = reinterpret_cast<unsigned long long*>(base);
    out->views = reinterpret_cast<ViewDev*>(base + kScratchCounterBytes);
    out->max_views = kMaxLdsViews;
    return hipSuccess;
}

template <class Pol>
void launch_persistent(const TraceLaunch& L, hipStream_t stream) {
    const dim3 grid(L.grid), block(L.block);
    const size_t lds = view_lds_bytes(L.n_views);
    switch (L.waves_per_simd) {
        case 2: k_trace_persistent<Pol, 2><<<grid, block, lds, stream>>>(L); break;
        case 3: k_trace_persistent<Pol, 3><<<grid, block, lds, stream>>>(L); break;
        case 4: k_trace_persistent<Pol, 4><<<grid, block, lds, stream>>>(L); break;
        default: k_trace_persistent<Pol, 1><<<grid, block, lds, stream>>>(L); break;
    }
}

template <class Pol>
void launch_oneshot(const TraceLaunch& L, hipStream_t stream) {
    long long blocks = (L.total + L.block - 1) / L.block;
    // a queue counted on the device: twice the persistent grid's lanes cover the usual
    // share of rejected rays in one pass (the kernel strides over a longer queue)
    if (L.total_dev || L.bucket_count) blocks = std::min<long long>(blocks, 2LL * std::max(1, L.grid));
    const dim3 grid((unsigned)std::max(1LL, blocks)), block(L.block);
    k_trace_oneshot<Pol><<<grid, block, view_lds_bytes(L.n_views), stream>>>(L);
}

hipError_t reset_counters(const TraceLaunch& L, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, L.counter, L.stats);
    return hipGetLastError();
}

// one lane per ray up to `cap` blocks of 256, striding beyond
static dim3 closed_grid(long long n, long long cap) { return dim3((unsigned)std::max(1LL, std::min((n + 255) / 256, cap))); }
// blocks of the FP64 pass over the rays the FP32 pass left unsure: the escapes and the
// crossings near a disk edge, one ray in a thousand on the orbit but one in five where the
// sky fills the frame's edges.  1024 blocks of 256 fill the chip at the kernel's four waves
// per SIMD; blocks past the region's end return at once (on the orbit 1024 blocks cost what
// 32 do; on a frame with 20% escapes they take 0.18 ms where 256 take 0.20 and 32 take 0.64)
constexpr unsigned kRedoBlocks = 1024;

hipError_t launch_closed_form_polish(const TraceLaunch& L, const RayQueue& q, int fp32, hipStream_t stream) {
    if (L.total <= 0) return hipSuccess;
    const size_t lds = view_lds_bytes(L.n_views);
    const dim3 block(256);
    if (fp32) {
        k_closed_form<kEvery, true, true><<<closed_grid(L.total, 4096), block, lds, stream>>>(L, q);
        k_closed_form<kUnsure, true, false><<<dim3(kRedoBlocks), block, lds, stream>>>(L, q);
    } else {
        k_closed_form<kEvery, true, false><<<closed_grid(L.total, 2048), block, lds, stream>>>(L, q);
    }
    return hipGetLastError();
}

hipError_t trace_closed_form(const TraceLaunch& L, const RayQueue& q, int fp32, hipStream_t stream, hipEvent_t* ev) {
    hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(64), 0, stream, q.count, L.counter, L.stats);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (L.total > 0) {
        const size_t lds = view_lds_bytes(L.n_views);
        const dim3 block(256);
        // ev[0] and ev[1] take the first kernel's start and the last one's end from the
        // dispatches themselves (no marker packets between the kernels), so the stats time
        // exactly the closed-form kernels
        hipEvent_t e0 = ev ? ev[0] : nullptr, e1 = ev ? ev[1] : nullptr;
        if (fp32) {
            // the FP32 pass over every ray, then the FP64 pass over the rays it was unsure of
            // (a small grid striding over the unsure region, whose length it reads on the device)
            const dim3 g32 = closed_grid(L.total, 4096), g64(kRedoBlocks);
            if (ev) {
                hipExtLaunchKernelGGL(k_closed_form<kEvery, false, true>, g32, block, lds, stream, e0, nullptr, 0, L, q);
                hipExtLaunchKernelGGL(k_closed_form<kUnsure, false, false>, g64, block, lds, stream, nullptr, e1, 0, L,
                                      q);
            } else {
                k_closed_form<kEvery, false, true><<<g32, block, lds, stream>>>(L, q);
                k_closed_form<kUnsure, false, false><<<g64, block, lds, stream>>>(L, q);
            }
        } else {
            const dim3 g = closed_grid(L.total, 2048);
            if (ev) hipExtLaunchKernelGGL(k_closed_form<kEvery, false, false>, g, block, lds, stream, e0, e1, 0, L, q);
            else k_closed_form<kEvery, false, false><<<g, block, lds, stream>>>(L, q);
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
    } else if (ev) {
        if ((e = hipEventRecord(ev[0], stream)) != hipSuccess) return e;
        if ((e = hipEventRecord(ev[1], stream)) != hipSuccess) return e;
    }
    TraceLaunch F = L;
    F.ray_list = q.list;
    F.total_dev = nullptr;
    F.bucket_count = q.count;
    F.bucket_stride = q.stride;
    F.n_buckets = 1;  // region 0 only
    F.no_reset = 1;
    F.mode = 1;   // the one-shot form over the queue: its grid is capped near the usual queue length,
    F.grid = 128;  // a few hundred rays at most here (an empty launch of a frame-sized grid costs 8 us)
    if ((e = launch_trace(F, stream)) != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev[2], stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_queue_total, dim3(1), dim3(64), 0, stream, q.count, q.total, L.stats + kStatFallback);
    return hipGetLastError();
}

hipError_t launch_trace(const TraceLaunch& L, hipStream_t stream) {
    if (!L.no_reset) hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, L.counter, L.stats);
    if (L.total <= 0) return hipGetLastError();
    if (L.out.turns || L.out.r_min || L.out.delay || L.out.b || L.out.q) {
        const dim3 grid((unsigned)((L.total + L.block - 1) / L.block)), block(L.block);
        if (L.ks) k_trace_report<kPhiIn><<<grid, block, view_lds_bytes(L.n_views), stream>>>(L);
        else k_trace_report<kPhiOut><<<grid, block, view_lds_bytes(L.n_views), stream>>>(L);
    } else if (L.ks) {
        if (L.mode == 1) launch_oneshot<PlainKS>(L, stream);
        else launch_persistent<PlainKS>(L, stream);
    } else if (L.mode == 1) {
        if (L.precision == 2) launch_oneshot<Plain<float>>(L, stream);
        else if (L.precision == 1) launch_oneshot<Mixed>(L, stream);
        else launch_oneshot<Plain<double>>(L, stream);
    } else {
        if (L.precision == 2) launch_persistent<Plain<float>>(L, stream);
        else if (L.precision == 1) launch_persistent<Mixed>(L, stream);
        else launch_persistent<Plain<double>>(L, stream);
    }
    return hipGetLastError();
}

}  // namespace kti
