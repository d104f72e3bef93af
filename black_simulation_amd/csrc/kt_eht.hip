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

constexpr double kQe = 4.80320425e-10, kMe = 9.1093837015e-28, kMp = 1.67262192369e-24;
constexpr double kC = 2.99792458e10, kKb = 1.380649e-16, kHp = 6.62607015e-27;
constexpr double kStart = 1000.0, kLight = 60.0, kMaxDepth = 30.0;

// Abramowitz and Stegun 9.8.1-9.8.8 (as eht.cpp:98-152): I0, I1, and K0, K1 times e^x
__device__ double bessel_i0(double x) {
    const double ax = fabs(x);
    if (ax < 3.75) {
        const double y = (x / 3.75) * (x / 3.75);
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))));
    }
    const double y = 3.75 / ax;
    return (exp(ax) / sqrt(ax)) * (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 +
           y * (-0.2057706e-1 + y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))));
}

__device__ double bessel_i1(double x) {
    const double ax = fabs(x);
    double res;
    if (ax < 3.75) {
        const double y = (x / 3.75) * (x / 3.75);
        res = ax * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 + y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
    } else {
        const double y = 3.75 / ax;
        double t = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
        t = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 + y * (-0.1031555e-1 + y * t))));
        res = t * exp(ax) / sqrt(ax);
    }
    return x < 0.0 ? -res : res;
}

__device__ double k0_scaled(double x) {
    if (x <= 2.0) {
        const double y = x * x / 4.0;
        return exp(x) * ((-log(x / 2.0) * bessel_i0(x)) + (-0.57721566 + y * (0.42278420 + y * (0.23069756 + y * (0.3488590e-1 +
               y * (0.262698e-2 + y * (0.10750e-3 + y * 0.74e-5)))))));
    }
    const double y = 2.0 / x;
    return (1.0 / sqrt(x)) * (1.25331414 + y * (-0.7832358e-1 + y * (0.2189568e-1 + y * (-0.1062446e-1 + y * (0.587872e-2 +
           y * (-0.251540e-2 + y * 0.53208e-3))))));
}

__device__ double k1_scaled(double x) {
    if (x <= 2.0) {
        const double y = x * x / 4.0;
        return exp(x) * ((log(x / 2.0) * bessel_i1(x)) + (1.0 / x) * (1.0 + y * (0.15443144 + y * (-0.67278579 + y * (-0.18156897 +
               y * (-0.1919402e-1 + y * (-0.110404e-2 + y * (-0.4686e-4))))))));
    }
    const double y = 2.0 / x;
    return (1.0 / sqrt(x)) * (1.25331414 + y * (0.23498619 + y * (-0.3655620e-1 + y * (0.1504268e-1 + y * (-0.780353e-2 +
           y * (0.325614e-2 + y * (-0.68245e-3)))))));
}

__device__ double emissivity(double nu, double ne, double thetaE, double B, double sinThetaB) {
    if (ne <= 0.0 || thetaE < 0.05 || B <= 0.0 || sinThetaB <= 0.0) return 0.0;
    const double nuC = kQe * B / (2.0 * kPi * kMe * kC);
    const double nuS = (2.0 / 9.0) * nuC * thetaE * thetaE * sinThetaB;
    const double X = nu / nuS;
    const double shape = sqrt(X) + pow(2.0, 11.0 / 12.0) * pow(X, 1.0 / 6.0);
    const double inv = 1.0 / thetaE;
    const double expo = inv - cbrt(X);
    if (expo < -700.0) return 0.0;
    const double k2 = k0_scaled(inv) + 2.0 / inv * k1_scaled(inv);
    return ne * sqrt(2.0) * kPi * kQe * kQe * nuS / (3.0 * k2 * kC) * shape * shape * exp(expo);
}

__device__ double planck_nu(double nu, double T) {
    const double x = kHp * nu / (kKb * T);
    return 2.0 * kHp * nu * nu * nu / (kC * kC) / expm1(x);
}

__device__ double plunge_ur(const EhtLaunch& e, double r) {
    const double a = e.a, delta = r * r - 2.0 * r + a * a, x = e.L - a * e.E;
    const double P = e.E * (r * r + a * a) - a * e.L;
    return sqrt(fmax(P * P - delta * (r * r + x * x), 0.0)) / (r * r);
}

// Radiative transfer over one accepted step ending at y (src/eht.cpp:290-316): the gas at
// the step's end, the step's affine length from the mean of Sigma at both ends (cm).
// Returns true once the optical depth passes kMaxDepth.
__device__ __forceinline__ bool transfer(const EhtLaunch& e, const Photon<double>& p, double r_prev, double mu_prev,
                                         const double y[5], double h, double& intensity, double& depth) {
    const double a = e.a, r = y[0], mu = y[1];
    const double sig_b = r_prev * r_prev + a * a * mu_prev * mu_prev, sig_a = r * r + a * a * mu * mu;
    const double dl = 0.5 * (sig_a + sig_b) * h * e.len_cm;
    // the plasma at (r, mu)
    const double z = r * mu, rho = fmax(r * sqrt(fmax(1.0 - mu * mu, 0.0)), 1e-6);
    const double hh = fmax(e.height, 0.05) * rho;
    double radial = pow(r, -1.1);
    if (r < e.isco && r < e.start) {
        const double here = plunge_ur(e, r);
        if (here > 0.0)
            radial = fmin(radial, pow(e.start, -1.1) * (e.start * e.start * e.start_speed) / (r * r * here));
    }
    const double ne = e.n0 * radial * exp(-z * z / (2.0 * hh * hh));
    const double Te = e.T0 * pow(r, -0.84);
    const double B = sqrt(8.0 * kPi * ne * kMp * kC * kC / (6.0 * e.beta * r));
    // the gas's 4-velocity
    const double s2 = 1.0 - mu * mu, c2 = mu * mu;
    const double sigma = r * r + a * a * c2, delta = r * r - 2.0 * r + a * a;
    const double A = (r * r + a * a) * (r * r + a * a) - a * a * delta * s2;
    const double gtt = -(1.0 - 2.0 * r / sigma), gtp = -2.0 * a * r * s2 / sigma, gpp = A * s2 / sigma;
    const double grr = sigma / delta;
    double omega, ur = 0.0;
    if (r >= e.isco) {
        omega = e.keplerian / (pow(r, 1.5) + a);
    } else {
        const double x = e.L - a * e.E, P = e.E * (r * r + a * a) - a * e.L;
        const double R = fmax(P * P - delta * (r * r + x * x), 0.0);
        const double ut = ((r * r + a * a) * P / delta + a * x) / (r * r);
        const double up = (a * P / delta + x) / (r * r);
        omega = up / ut;
        ur = -sqrt(R) / (r * r);
    }
    double den = gtt + 2.0 * omega * gtp + omega * omega * gpp;
    double num = -1.0 - grr * ur * ur;
    if (!(den < 0.0) || !(num < 0.0)) {
        omega = -gtp / gpp;
        ur = 0.0;
        den = gtt + 2.0 * omega * gtp + omega * omega * gpp;
        num = -1.0;
    }
    const double ut = sqrt(num / den);
    const double u0 = ut, u1 = ur, u3 = omega * ut;
    const double mpu = u0 - p.L * u3 + y[3] / delta * u1;
    if (mpu > 0.0) {
        const double g = 1.0 / mpu, nu_f = e.nu / g;
        const double thetaE = kKb * Te / (kMe * kC * kC);
        const double j = emissivity(nu_f, ne, thetaE, B, e.sin_theta_b);
        if (j > 0.0) {
            const double absorb = j / planck_nu(nu_f, Te);
            const double dtau = absorb * dl / g;
            const double add = dtau < 1e-6 ? g * g * j * dl : g * g * g * (j / absorb) * -expm1(-dtau);
            intensity += exp(-depth) * add;
            depth += dtau;
            if (depth > kMaxDepth) return true;
        }
    }
    return false;
}

__global__ __launch_bounds__(128) void k_eht(EhtLaunch e) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= e.n) return;
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
    Photon<double> p;
    double y[5], yn[5], k[7][5];
    bardeen_photon(e.a, kStart, 0.0, e.sinI, e.cosI, e.m0, alpha, beta, p, y);
    double intensity = 0.0, depth = 0.0, rmin = y[0];
    int captured = 0, steps = 0;
    bool rejected = false;
    double h = 0.01 / kStart;
    rhs_main(p, y, k[0]);
    while (steps < e.max_steps) {
        const double err = dp5_attempt(p, false, 0.0, y, h, k, yn, e.rtol, e.atol);
        ++steps;
        if (!(err <= 1.0)) {
            h *= step_factor(err, true);
            rejected = true;
            continue;
        }
        const double fac = step_factor(err, rejected);
        rejected = false;
        if (p.L == 0.0 && y[1] * y[4] > 0.0 && yn[1] * yn[4] < 0.0) yn[2] += kPi;
        project(p, yn);
        // commit the step before the event tests (the previous point stays for the
        // transfer's mean Sigma)
        const double r_prev = y[0], mu_prev = y[1], h_used = h;
#pragma unroll
        for (int i = 0; i < 5; ++i) y[i] = yn[i];
#pragma unroll
        for (int i = 0; i < 5; ++i) k[0][i] = k[6][i];
        k[0][0] = y[3];
        k[0][1] = y[4];
        h *= fac;
        if (!(fabs(y[0]) < 1e300) || !(fabs(y[1]) < 1e300) || y[0] < e.r_cap) {
            captured = 1;
            break;
        }
        if (y[0] < rmin) rmin = y[0];
        if (y[0] > kStart * 1.01 && y[3] > 0.0) break;
        if (y[0] <= kLight && transfer(e, p, r_prev, mu_prev, y, h_used, intensity, depth)) break;
    }
    if (e.intensity) e.intensity[q] = intensity;
    if (e.shadow) e.shadow[q] = (uint8_t)captured;
    if (e.min_radius) e.min_radius[q] = rmin;
    if (e.steps) e.steps[q] = steps;
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

hipError_t render(const EhtLaunch& e, hipStream_t st) {
    if (e.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_eht, dim3((unsigned)((e.n + 127) / 128)), dim3(128), 0, st, e);
    return hipGetLastError();
}

hipError_t sum(const double* x, long long n, double* partial, double* out, hipStream_t st) {
    const long long blocks = std::min<long long>(std::max<long long>((n + kSumBlock - 1) / kSumBlock, 1), kSumPartials);
    hipLaunchKernelGGL(k_sum_partial, dim3((unsigned)blocks), dim3(kSumBlock), 0, st, x, n, partial);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kSumBlock), 0, st, partial, (int)blocks, out);
    return hipGetLastError();
}

}  // namespace kte
