// kt_shade.hip -- HDR colour of traced rays: the thin disk's thermal emission seen at the
// traced redshift, the sky at the asymptotic direction, black for the hole.
//
// Disk (shaders/trace.frag:373-403): a thin disk with zero torque at its inner edge has
// T^4 ~ x^-3 (1 - x^-1/2), x = r / r_in, whose maximum is at x = 49/36; T is scaled so the
// maximum equals the disk temperature.  A blackbody seen with frequency ratio g looks like
// a blackbody at g T (trace.frag:330-335, the full-relativity case), and its colour is read
// from a table of linear-sRGB blackbody radiance normalised to luminance 1 at 6500 K
// (src/blackbody.cpp:48-60), divided by the luminance of the disk temperature's colour.
// The edges fade as in trace.frag:396-398; the disk is taken as opaque (the trace stops at
// the first crossing).
// Sky (trace.frag:1419-1458): a 15-degree longitude/latitude grid with quadrant colours,
// or an equirectangular image sampled bilinearly, or black.
#include "kt_shade.hpp"

#include <algorithm>
#include <cmath>

using ktd::ViewDev;

namespace ktsh {

namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ int view_of(const ViewDev* v, int n, long long q) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (v[mid].offset <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ float smoothstep(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
    return t * t * (3.f - 2.f * t);
}

__device__ __forceinline__ float3 table_rgb(const float* tab, int n, float T) {
    float u = (log10f(fmaxf(T, 1.0f)) - (float)kLogTMin) / (float)(kLogTMax - kLogTMin);
    u = u < 0.f ? 0.f : (u > 1.f ? 1.f : u);
    const float x = u * (n - 1);
    int i = (int)x;
    if (i > n - 2) i = n - 2;
    const float f = x - i;
    const float* a = tab + 3 * i;
    return make_float3(a[0] + f * (a[3] - a[0]), a[1] + f * (a[4] - a[1]), a[2] + f * (a[5] - a[2]));
}

__device__ __forceinline__ float3 grid_sky(float dx, float dy, float dz, float w) {
    const float pi = (float)kPi;
    const float lon = atan2f(dy, dx);
    const float lat = asinf(fminf(fmaxf(dz, -1.f), 1.f));
    const float u = lon / pi * 12.f, v = lat / pi * 12.f;  // 15-degree squares
    const float cu = floorf(u), cv = floorf(v);
    const int quadrant = ((int)floorf((lon + pi) / (0.5f * pi))) & 3;
    float3 c = quadrant == 0 ? make_float3(0.95f, 0.3f, 0.25f)
             : quadrant == 1 ? make_float3(0.25f, 0.75f, 0.35f)
             : quadrant == 2 ? make_float3(0.3f, 0.45f, 0.95f)
                             : make_float3(0.95f, 0.8f, 0.25f);
    float k = lat < 0.f ? 0.55f : 1.f;
    const float parity = (cu + cv) - 2.f * floorf((cu + cv) * 0.5f);  // 0 or 1
    k *= 0.45f + 0.55f * parity;
    const float eu = 0.5f - fabsf(u - cu - 0.5f), ev = 0.5f - fabsf(v - cv - 0.5f);
    const float line = fminf(smoothstep(0.f, w, eu), smoothstep(0.f, w, ev));
    k *= (0.15f + 0.85f * line) * 0.6f;
    return make_float3(c.x * k, c.y * k, c.z * k);
}

__device__ __forceinline__ float3 image_sky(const float* img, int W, int H, float dx, float dy, float dz) {
    const float pi = (float)kPi;
    const float lon = atan2f(dy, dx);
    const float lat = asinf(fminf(fmaxf(dz, -1.f), 1.f));
    // texel centres at (i + 0.5) / W; longitude wraps, latitude clamps
    // clamped before the integer conversion (a NaN direction gives texel 0, not an
    // undefined index)
    float x = fminf(fmaxf((lon / (2.f * pi) + 0.5f) * W - 0.5f, -0.5f), (float)W - 0.5f);
    float y = (lat / pi + 0.5f) * H - 0.5f;
    y = fminf(fmaxf(y, 0.f), (float)(H - 1));
    const float fx = floorf(x), fy = floorf(y);
    const float tx = x - fx, ty = y - fy;
    int x0 = (int)fx % W;
    if (x0 < 0) x0 += W;
    const int x1 = (x0 + 1) % W;
    const int y0 = (int)fy, y1 = y0 + 1 < H ? y0 + 1 : H - 1;
    const float* p00 = img + 3 * ((long long)y0 * W + x0);
    const float* p01 = img + 3 * ((long long)y0 * W + x1);
    const float* p10 = img + 3 * ((long long)y1 * W + x0);
    const float* p11 = img + 3 * ((long long)y1 * W + x1);
    float3 r;
    r.x = (1 - ty) * ((1 - tx) * p00[0] + tx * p01[0]) + ty * ((1 - tx) * p10[0] + tx * p11[0]);
    r.y = (1 - ty) * ((1 - tx) * p00[1] + tx * p01[1]) + ty * ((1 - tx) * p10[1] + tx * p11[1]);
    r.z = (1 - ty) * ((1 - tx) * p00[2] + tx * p01[2]) + ty * ((1 - tx) * p10[2] + tx * p11[2]);
    return r;
}

__global__ void k_shade(ShadeLaunch s) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= s.n) return;
    const ViewDev& v = s.views[view_of(s.views, s.n_views, q)];
    const int cls = s.cls[q] & 7;
    float3 c = make_float3(0.f, 0.f, 0.f);
    if (cls == 0) {
        const double th = s.theta ? s.theta[q] : (double)s.theta32[q], ph = s.phi ? s.phi[q] : (double)s.phi32[q];
        const float st = (float)sin(th);
        const float dx = st * (float)cos(ph), dy = st * (float)sin(ph), dz = (float)cos(th);
        if (s.sky == 1) c = grid_sky(dx, dy, dz, s.grid_line);
        else if (s.sky == 2) c = image_sky(s.sky_image, s.sky_w, s.sky_h, dx, dy, dz);
        c.x *= s.sky_brightness;
        c.y *= s.sky_brightness;
        c.z *= s.sky_brightness;
    } else if (cls == 2) {
        const double r = s.r_hit ? s.r_hit[q] : (double)s.r_hit32[q], g = s.g ? s.g[q] : (double)s.g32[q];
        const double x = r / v.r_in;
        const double flux = fmax(1.0 / (x * x * x) * (1.0 - 1.0 / sqrt(x)), 0.0);
        const float T = (float)(s.disk_temperature * sqrt(sqrt(flux / kFluxPeak)));
        const float3 bb = table_rgb(s.table, s.table_n, T * (float)g);
        const float edges = (1.f - smoothstep((float)(0.7 * v.r_out), (float)v.r_out, (float)r)) *
                            smoothstep((float)v.r_in, (float)(1.03 * v.r_in), (float)r);
        const float k = s.disk_scale * edges;
        c = make_float3(bb.x * k, bb.y * k, bb.z * k);
    }
    // as the renderer's output (trace.frag:1905-1906): a non-finite colour is black, the
    // rest is capped at 30000
    if (!(fabsf(c.x) <= 3.4e38f && fabsf(c.y) <= 3.4e38f && fabsf(c.z) <= 3.4e38f)) c = make_float3(0.f, 0.f, 0.f);
    __half* o = s.rgba + 4 * q;
    o[0] = __float2half(fminf(c.x, 3.0e4f));
    o[1] = __float2half(fminf(c.y, 3.0e4f));
    o[2] = __float2half(fminf(c.z, 3.0e4f));
    o[3] = __float2half(1.0f);
}

// Wyman, Sloan and Shirley (2013), "Simple analytic approximations to the CIE XYZ colour
// matching functions", multi-lobe fit to the CIE 1931 2-degree observer (lambda in nm)
double lobe(double x, double mu, double s1, double s2) {
    const double t = (x - mu) / (x < mu ? s1 : s2);
    return std::exp(-0.5 * t * t);
}

void cie_xyz(double nm, double xyz[3]) {
    xyz[0] = 1.056 * lobe(nm, 599.8, 37.9, 31.0) + 0.362 * lobe(nm, 442.0, 16.0, 26.7) -
             0.065 * lobe(nm, 501.1, 20.4, 26.2);
    xyz[1] = 0.821 * lobe(nm, 568.8, 46.9, 40.5) + 0.286 * lobe(nm, 530.9, 16.3, 31.1);
    xyz[2] = 1.217 * lobe(nm, 437.0, 11.8, 36.0) + 0.681 * lobe(nm, 459.0, 26.0, 13.8);
}

}  // namespace

// Planck's law integrated against the matching functions, 360-830 nm in 2 nm steps
void planck_xyz(double T, double xyz[3]) {
    const double h = 6.62607015e-34, c = 2.99792458e8, k = 1.380649e-23;
    xyz[0] = xyz[1] = xyz[2] = 0.0;
    for (int nm = 360; nm <= 830; nm += 2) {
        const double lam = nm * 1e-9;
        const double B = 2.0 * h * c * c / (lam * lam * lam * lam * lam) / std::expm1(h * c / (lam * k * T));
        double m[3];
        cie_xyz((double)nm, m);
        for (int i = 0; i < 3; ++i) xyz[i] += B * m[i];
    }
}

void blackbody_table(int n, float* rgb) {
    double ref[3];
    planck_xyz(6500.0, ref);
    for (int i = 0; i < n; ++i) {
        const double logT = kLogTMin + (kLogTMax - kLogTMin) * i / (n - 1);
        double v[3];
        planck_xyz(std::pow(10.0, logT), v);
        // XYZ -> linear sRGB (IEC 61966-2-1); out-of-gamut negatives clipped to 0
        const double r = 3.2406 * v[0] - 1.5372 * v[1] - 0.4986 * v[2];
        const double g = -0.9689 * v[0] + 1.8758 * v[1] + 0.0415 * v[2];
        const double b = 0.0557 * v[0] - 0.2040 * v[1] + 1.0570 * v[2];
        rgb[3 * i + 0] = (float)(std::max(r, 0.0) / ref[1]);
        rgb[3 * i + 1] = (float)(std::max(g, 0.0) / ref[1]);
        rgb[3 * i + 2] = (float)(std::max(b, 0.0) / ref[1]);
    }
}

float table_luminance(const float* tab, int n, double T) {
    double u = (std::log10(std::max(T, 1.0)) - kLogTMin) / (kLogTMax - kLogTMin);
    u = std::min(std::max(u, 0.0), 1.0);
    const double x = u * (n - 1);
    int i = (int)x;
    if (i > n - 2) i = n - 2;
    const double f = x - i;
    const float* a = tab + 3 * i;
    const double r = a[0] + f * (a[3] - a[0]), g = a[1] + f * (a[4] - a[1]), b = a[2] + f * (a[5] - a[2]);
    return (float)(0.2126 * r + 0.7152 * g + 0.0722 * b);
}

hipError_t shade(const ShadeLaunch& s, hipStream_t st) {
    if (s.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_shade, dim3((unsigned)((s.n + 255) / 256)), dim3(256), 0, st, s);
    return hipGetLastError();
}

}  // namespace ktsh
