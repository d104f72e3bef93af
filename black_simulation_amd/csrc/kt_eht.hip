// kt_eht.hip -- what the Event Horizon Telescope sees: every pixel of a distant observer's
// Bardeen plane traced back from r = 1000 M through the FP64 DP5 engine, gathering the
// thermal synchrotron light of a hot accretion flow by radiative transfer (src/eht.cpp:
// 260-319, restated over the DP5 steps).  After every accepted step inside r = 60 M:
//   the gas's n_e, T_e, B (eht.cpp:196-215) and 4-velocity (eht.cpp:217-253) at the step's
//   end; the frequency ratio g = 1 / (-p.u) with p_t = -1, p_phi = L, p_r = -w_r / Delta;
//   j at nu / g (Leung, Gammie and Noble 2011, eht.cpp:163-177), absorption j / B_nu
//   (Kirchhoff); over the step's affine length dl = Sigma_mean dMino (in cm),
//   dtau = alpha dl / g and I += exp(-tau) g^3 S (1 - exp(-dtau)).
// A ray ends captured below r+ (1 + 1e-4), back out beyond 1010 M, or at optical depth 30.
#include "kt_eht.hpp"

#include <algorithm>
#include <cmath>

using namespace ktd;

namespace kte {

namespace {

constexpr double kQe = 4.80320425e-10, kMe = 9.1093837015e-28;
constexpr double kC = 2.99792458e10, kKb = 1.380649e-16, kHp = 6.62607015e-27;
constexpr double kStart = 1000.0, kLight = 60.0, kMaxDepth = 30.0;

constexpr double kSqrt2 = 1.4142135623730951;
constexpr double kTwo11_12 = 1.8877486253633868;  // 2^(11/12)
// the emissivity's prefactor sqrt(2) pi e^2 / (3 c) and nu_c = e B / (2 pi m_e c) per gauss
constexpr double kJ = kSqrt2 * kPi * kQe * kQe / (3.0 * kC);
constexpr double kNuC = kQe / (2.0 * kPi * kMe * kC);
constexpr double kThetaPerK = kKb / (kMe * kC * kC);

// K2(x) e^x for x > 0 from K2 = K0 + (2/x) K1 with Abramowitz and Stegun 9.8.1, 9.8.3,
// 9.8.5-9.8.8 (as eht.cpp:98-152).  Below x = 2 the series of K0 and K1 share one exp(x)
// and one log(x / 2); above, the asymptotic forms share 1 / sqrt(x).
__device__ __forceinline__ double k2_scaled(double x) {
    if (x <= 2.0) {
        const double y = x * x / 4.0;
        const double t = (x / 3.75) * (x / 3.75);
        const double i0 = 1.0 + t * (3.5156229 + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 + t * (0.360768e-1 +
                          t * 0.45813e-2)))));
        const double i1 = x * (0.5 + t * (0.87890594 + t * (0.51498869 + t * (0.15084934 + t * (0.2658733e-1 +
                          t * (0.301532e-2 + t * 0.32411e-3))))));
        const double lx = log(x / 2.0), ex = exp(x);
        const double k0 = ex * ((-lx * i0) + (-0.57721566 + y * (0.42278420 + y * (0.23069756 + y * (0.3488590e-1 +
                          y * (0.262698e-2 + y * (0.10750e-3 + y * 0.74e-5)))))));
        const double k1 = ex * ((lx * i1) + (1.0 / x) * (1.0 + y * (0.15443144 + y * (-0.67278579 + y * (-0.18156897 +
                          y * (-0.1919402e-1 + y * (-0.110404e-2 + y * (-0.4686e-4))))))));
        return k0 + 2.0 / x * k1;
    }
    const double y = 2.0 / x, rs = 1.0 / sqrt(x);
    const double k0 = rs * (1.25331414 + y * (-0.7832358e-1 + y * (0.2189568e-1 + y * (-0.1062446e-1 + y * (0.587872e-2 +
                      y * (-0.251540e-2 + y * 0.53208e-3))))));
    const double k1 = rs * (1.25331414 + y * (0.23498619 + y * (-0.3655620e-1 + y * (0.1504268e-1 + y * (-0.780353e-2 +
                      y * (0.325614e-2 + y * (-0.68245e-3)))))));
    return k0 + 2.0 / x * k1;
}

// |u^r| of the ISCO's plunging geodesic at r (eht.cpp:222-233)
__device__ __forceinline__ double plunge_ur(const EhtLaunch& e, double r) {
    const double a = e.a, delta = r * r - 2.0 * r + a * a, x = e.L - a * e.E;
    const double P = e.E * (r * r + a * a) - a * e.L;
    return sqrt(fmax(P * P - delta * (r * r + x * x), 0.0)) / (r * r);
}

// Radiative transfer over one accepted step ending at y (src/eht.cpp:290-316).  The
// reference's formulas, arranged for the vector units: one log(r) gives r^-1.1 and
// r^-0.84 through exp (and the Gaussian joins the density's exp), X^(1/6) is the square
// root of X^(1/3), K0 and K1 share their exp and log, the divisions by Sigma become one
// reciprocal, and exp(-tau) is carried as the product of the steps' exp(-dtau) (each
// step's expm1 gives it), so a step costs nine transcendental calls instead of eighteen.
// Results differ from the formulas as written by rounding only (1e-13 relative over a
// ray).  Returns true once the optical depth passes kMaxDepth.
__device__ __forceinline__ bool transfer(const EhtLaunch& e, const Photon<double>& p, double r_prev, double mu_prev,
                                         const double y[5], double h, double& intensity, double& depth,
                                         double& seen) {
    const double a = e.a, a2 = a * a, r = y[0], mu = y[1];
    const double rr = r * r, s2 = 1.0 - mu * mu;
    const double sigma = rr + a2 * mu * mu;
    const double dl = 0.5 * (sigma + (r_prev * r_prev + a2 * mu_prev * mu_prev)) * h * e.len_cm;
    // the plasma at (r, mu) (eht.cpp:196-215)
    const double z = r * mu, rho = fmax(r * sqrt(fmax(s2, 0.0)), 1e-6);
    const double hh = e.hfac * rho;
    const double gauss = -z * z / (2.0 * hh * hh);
    const double lr = log(r);
    double ne;
    if (r < e.isco && r < e.start) {  // inside the plunge the gas thins as it speeds up
        double radial = exp(-1.1 * lr);
        const double here = plunge_ur(e, r);
        if (here > 0.0) radial = fmin(radial, e.start_rate / (rr * here));
        ne = e.n0 * radial * exp(gauss);
    } else {
        ne = e.n0 * exp(fma(-1.1, lr, gauss));
    }
    const double Te = e.T0 * exp(-0.84 * lr);
    const double B = sqrt(e.b_coef * ne / r);
    // the gas's 4-velocity (eht.cpp:217-253)
    const double delta = rr - 2.0 * r + a2, rra = rr + a2;
    const double A = rra * rra - a2 * delta * s2;
    const double isig = 1.0 / sigma;
    const double gtt = -(1.0 - 2.0 * r * isig), gtp = -2.0 * a * r * s2 * isig, gpp = A * s2 * isig;
    const double grr = sigma / delta;
    double omega, ur = 0.0;
    if (r >= e.isco) {
        omega = e.keplerian / (r * sqrt(r) + a);
    } else {
        const double x = e.L - a * e.E, P = e.E * rra - a * e.L;
        const double R = fmax(P * P - delta * (rr + x * x), 0.0);
        const double id = 1.0 / delta;
        omega = (a * P * id + x) / (rra * P * id + a * x);  // u^phi / u^t (their 1/r^2 cancels)
        ur = -sqrt(R) / rr;
    }
    double den = gtt + 2.0 * omega * gtp + omega * omega * gpp;
    double num = -1.0 - grr * ur * ur;
    if (!(den < 0.0) || !(num < 0.0)) {  // the local ZAMO
        omega = -gtp / gpp;
        ur = 0.0;
        den = gtt + 2.0 * omega * gtp + omega * omega * gpp;
        num = -1.0;
    }
    const double ut = sqrt(num / den);
    // -p.u with p_t = -1, p_phi = L, p_r = -w_r / Delta (towards the camera)
    const double mpu = ut - p.L * (omega * ut) + y[3] / delta * ur;
    if (!(mpu > 0.0)) return false;
    const double g = 1.0 / mpu, nu_f = e.nu / g;
    // thermal synchrotron emissivity at nu_f (Leung, Gammie and Noble 2011, eq. 72)
    const double thetaE = kThetaPerK * Te;
    if (ne <= 0.0 || thetaE < 0.05 || B <= 0.0 || e.sin_theta_b <= 0.0) return false;
    const double nuS = (2.0 / 9.0) * (kNuC * B) * thetaE * thetaE * e.sin_theta_b;
    const double X = nu_f / nuS;
    const double cx = cbrt(X);
    const double shape = sqrt(X) + kTwo11_12 * sqrt(cx);
    const double inv = 1.0 / thetaE;
    const double expo = inv - cx;
    if (expo < -700.0) return false;
    const double j = ne * kJ * nuS / k2_scaled(inv) * shape * shape * exp(expo);
    if (!(j > 0.0)) return false;
    // Kirchhoff: absorption j / B_nu(T_e)
    const double bnu = 2.0 * kHp * nu_f * nu_f * nu_f / (kC * kC) / expm1(kHp * nu_f / (kKb * Te));
    const double absorb = j / bnu;
    const double dtau = absorb * dl / g;
    double add, fade;
    if (dtau < 1e-6) {
        add = g * g * j * dl;
        fade = 1.0 - dtau * (1.0 - 0.5 * dtau);
    } else {
        const double em = expm1(-dtau);
        add = g * g * g * (j / absorb) * -em;
        fade = 1.0 + em;
    }
    intensity += seen * add;
    seen *= fade;
    depth += dtau;
    return depth > kMaxDepth;
}

// One ray in one lane: started from its pixel (or list entry), stepped one DP5 attempt at
// a time, written when done.
struct ELane {
    Photon<double> p;
    double y[5], k0[5], h, intensity, depth, seen, rmin;
    long long q;
    int captured, steps;
    bool rejected;
};

__device__ __forceinline__ void lane_start(ELane& L, const EhtLaunch& e, long long q) {
    double alpha, beta;
    if (e.rays) {
        alpha = e.rays[2 * q];
        beta = e.rays[2 * q + 1];
    } else {
        // pixel centres, rows from the top (eht.cpp:332-334)
        const int row = (int)(q / e.size), col = (int)(q - (long long)row * e.size);
        alpha = -e.half_width + (col + 0.5) * e.px;
        beta = e.half_width - (row + 0.5) * e.px;
    }
    bardeen_photon(e.a, kStart, 0.0, e.sinI, e.cosI, e.m0, alpha, beta, L.p, L.y);
    L.q = q;
    L.intensity = 0.0;
    L.depth = 0.0;
    L.seen = 1.0;
    L.rmin = L.y[0];
    L.captured = 0;
    L.steps = 0;
    L.rejected = false;
    L.h = 0.01 / kStart;
    rhs_main(L.p, L.y, L.k0);
}

// one attempt; true once the ray is done (captured, back out, opaque or out of steps)
__device__ __forceinline__ bool lane_step(ELane& L, const EhtLaunch& e) {
    if (L.steps >= e.max_steps) return true;
    double k[7][5], yn[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) k[0][i] = L.k0[i];
    const double err = dp5_attempt(L.p, false, 0.0, L.y, L.h, k, yn, e.rtol, e.atol);
    ++L.steps;
    if (!(err <= 1.0)) {
        L.h *= step_factor(err, true);
        L.rejected = true;
        return false;
    }
    const double fac = step_factor(err, L.rejected);
    L.rejected = false;
    if (L.p.L == 0.0 && L.y[1] * L.y[4] > 0.0 && yn[1] * yn[4] < 0.0) yn[2] += kPi;
    project(L.p, yn);
    // commit the step before the event tests (the previous point stays for the transfer's
    // mean Sigma)
    const double r_prev = L.y[0], mu_prev = L.y[1], h_used = L.h;
#pragma unroll
    for (int i = 0; i < 5; ++i) L.y[i] = yn[i];
#pragma unroll
    for (int i = 0; i < 5; ++i) L.k0[i] = k[6][i];
    L.k0[0] = L.y[3];
    L.k0[1] = L.y[4];
    L.h *= fac;
    if (!(fabs(L.y[0]) < 1e300) || !(fabs(L.y[1]) < 1e300) || L.y[0] < e.r_cap) {
        L.captured = 1;
        return true;
    }
    if (L.y[0] < L.rmin) L.rmin = L.y[0];
    if (L.y[0] > kStart * 1.01 && L.y[3] > 0.0) return true;
    return L.y[0] <= kLight && transfer(e, L.p, r_prev, mu_prev, L.y, h_used, L.intensity, L.depth, L.seen);
}

__device__ __forceinline__ void lane_write(const ELane& L, const EhtLaunch& e) {
    if (e.intensity) e.intensity[L.q] = L.intensity;
    if (e.shadow) e.shadow[L.q] = (uint8_t)L.captured;
    if (e.min_radius) e.min_radius[L.q] = L.rmin;
    if (e.steps) e.steps[L.q] = L.steps;
}

__global__ void k_zero_counter(unsigned long long* c) {
    if (threadIdx.x == 0) *c = 0ull;
}

// Persistent waves over the rays [first, n): every kRefill attempts a wave's finished lanes
// take new rays, handed out from 64 indices the wave claims per atomic.  Rays that the hole
// captures end within tens of attempts while rays through the flow take hundreds, so lanes
// of a one-ray-per-lane launch idled a fifth of the time.
constexpr int kRefill = 8;
constexpr int kChunk = 64;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_eht(EhtLaunch e, unsigned long long* counter) {
    const int lane = (int)(threadIdx.x & 63u);
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const long long total = e.n - e.first;
    ELane L;
    bool active = false;
    long long cbase = 0;  // the wave's claimed offsets not yet handed out: [cbase, cbase + cleft)
    int cleft = 0;
    bool exhausted = false;
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (nidle > 0 && (cleft > 0 || !exhausted)) {
            const int rank = __popcll(idle & below);
            long long q = rank < cleft ? cbase + rank : -1;
            if (nidle > cleft && !exhausted) {
                unsigned long long nb = 0;
                if (lane == 0) nb = atomicAdd(counter, (unsigned long long)kChunk);
                nb = __shfl(nb, 0);
                if ((long long)nb + kChunk >= total) exhausted = true;
                const int from_new = nidle - cleft;
                if (rank >= cleft) q = (long long)nb + (rank - cleft);
                cbase = (long long)nb + from_new;
                cleft = kChunk - from_new;
            } else {
                const int used = nidle < cleft ? nidle : cleft;
                cbase += used;
                cleft -= used;
            }
            if (!active && q >= 0 && q < total) {
                lane_start(L, e, e.first + q);
                active = true;
            }
        }
        if (__ballot(active) == 0ull) {
            if (exhausted && cleft == 0) break;
            continue;
        }
        for (int k = 0; k < kRefill; ++k) {
            if (active && lane_step(L, e)) {
                lane_write(L, e);
                active = false;
            }
        }
    }
}

// sum of the intensities (for the flux) in a fixed order, so that it does not change from
// run to run: 256-lane partial sums per block, then one block adds the partials
constexpr int kSumBlock = 256;

__global__ __launch_bounds__(kSumBlock) void k_sum_partial(const double* x, long long n, double* partial) {
    __shared__ double red[kSumBlock];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * kSumBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kSumBlock)
        s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kSumBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kSumBlock) void k_sum_final(const double* partial, int m, double* out) {
    __shared__ double red[kSumBlock];
    double s = 0.0;
    for (int i = threadIdx.x; i < m; i += kSumBlock) s += partial[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kSumBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace

hipError_t render(const EhtLaunch& e, unsigned long long* counter, int grid, hipStream_t st) {
    if (e.n <= e.first) return hipSuccess;
    const long long need = (e.n - e.first + 255) / 256;  // no more workgroups than 256-ray groups
    const int blocks = (int)(need < grid ? need : grid);
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, counter);
    hipLaunchKernelGGL(k_eht, dim3((unsigned)blocks), dim3(256), 0, st, e, counter);
    return hipGetLastError();
}

hipError_t sum(const double* x, long long n, double* partial, double* out, hipStream_t st) {
    const long long blocks = std::min<long long>(std::max<long long>((n + kSumBlock - 1) / kSumBlock, 1), kSumPartials);
    hipLaunchKernelGGL(k_sum_partial, dim3((unsigned)blocks), dim3(kSumBlock), 0, st, x, n, partial);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kSumBlock), 0, st, partial, (int)blocks, out);
    return hipGetLastError();
}

}  // namespace kte
