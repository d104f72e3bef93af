// kt_surrogate.hip -- the learned surrogate: training data from the integrator, training
// with the BF16 MFMA kernels of kt_nn.hip, evaluation against FP64 labels, and traced
// inference with the gate and a device-side fallback queue for the integrator.
#include "kt_surrogate.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/kerr_trace.h"

using namespace ktm;
using ktd::ViewDev;

namespace kts {

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---- shadow table: critical radius over (a, i, psi) and the mean shadow radius ------------
constexpr int kTabA = 41, kTabI = 45, kTabPsi = 129;
constexpr double kTabA0 = 0.0, kTabA1 = 0.999;
constexpr double kTabI0 = 1.0 * kPi / 180.0, kTabI1 = 89.0 * kPi / 180.0;

struct Table {
    const float* rho;  // [kTabA][kTabI][kTabPsi], psi in [0, pi]
    const float* rsh;  // [kTabA][kTabI]
};

__host__ __device__ inline void cell(double x, double x0, double x1, int n, int& i, float& f) {
    double u = (x - x0) / (x1 - x0) * (n - 1);
    u = fmin(fmax(u, 0.0), (double)(n - 1));  // NaN -> 0: the index stays in the table
    i = (int)u;
    if (i > n - 2) i = n - 2;
    f = (float)(u - i);
}

__host__ __device__ inline float table_rsh(const float* rsh, double a, double inc) {
    int ia, ii;
    float fa, fi;
    cell(a, kTabA0, kTabA1, kTabA, ia, fa);
    cell(inc, kTabI0, kTabI1, kTabI, ii, fi);
    const float v00 = rsh[ia * kTabI + ii], v01 = rsh[ia * kTabI + ii + 1];
    const float v10 = rsh[(ia + 1) * kTabI + ii], v11 = rsh[(ia + 1) * kTabI + ii + 1];
    return (1 - fa) * ((1 - fi) * v00 + fi * v01) + fa * ((1 - fi) * v10 + fi * v11);
}

__host__ __device__ inline float table_rho(const float* rho, double a, double inc, double psi) {
    int ia, ii, ip;
    float fa, fi, fp;
    cell(a, kTabA0, kTabA1, kTabA, ia, fa);
    cell(inc, kTabI0, kTabI1, kTabI, ii, fi);
    cell(psi, 0.0, kPi, kTabPsi, ip, fp);
    auto at = [&](int x, int y, int z) { return rho[((long long)x * kTabI + y) * kTabPsi + z]; };
    float r = 0.f;
    for (int dx = 0; dx < 2; ++dx)
        for (int dy = 0; dy < 2; ++dy) {
            const float w = (dx ? fa : 1 - fa) * (dy ? fi : 1 - fi);
            r += w * ((1 - fp) * at(ia + dx, ii + dy, ip) + fp * at(ia + dx, ii + dy, ip + 1));
        }
    return r;
}

// normalised inputs of one ray and its relative distance from the critical curve
__host__ __device__ inline void features_of(const float* rho, const float* rsh, double a, double r_cam, double inc,
                                            double alpha, double beta, float f[5], float* ring) {
    const float R = table_rsh(rsh, a, inc);
    f[0] = (float)(log10(r_cam) - 2.0);
    f[1] = (float)((cos(inc) - 0.5415) / 0.4545);
    f[2] = (float)((a - 0.499) / 0.499);
    f[3] = (float)(alpha / (1.5 * R));
    f[4] = (float)(beta / (1.5 * R));
    const double psi = atan2(fabs(beta), alpha);
    const double rc = table_rho(rho, a, inc, psi);
    *ring = (float)fabs(sqrt(alpha * alpha + beta * beta) / rc - 1.0);
}

// Frame of a ray's straight-line (M = 0) continuation: e1 the direction the backward-traced
// ray would leave along (launch conventions: inward, -beta/r along e_theta, -alpha/r along
// e_phi at the camera), e2 the unit vector from the line's closest point towards the hole
// (the way gravity bends it), e3 = e1 x e2.  Cartesian, z along the spin axis.
__host__ __device__ inline void ray_frame(double r, double inc, double phc, double al, double be, double e1[3],
                                          double e2[3], double e3[3]) {
    const double si = sin(inc), ci = cos(inc), sp = sin(phc), cp = cos(phc);
    const double rh[3] = {si * cp, si * sp, ci}, th[3] = {ci * cp, ci * sp, -si}, ph[3] = {-sp, cp, 0.0};
    double q = 1.0 - (al * al + be * be) / (r * r);
    q = q > 0.0 ? sqrt(q) : 0.0;
    double n = 0.0;
    for (int k = 0; k < 3; ++k) {
        e1[k] = -q * rh[k] - (be / r) * th[k] - (al / r) * ph[k];
        n += e1[k] * e1[k];
    }
    n = 1.0 / sqrt(n);
    for (int k = 0; k < 3; ++k) e1[k] *= n;
    // -(P - (P.e1) e1) with P = r rh
    const double pe = r * (rh[0] * e1[0] + rh[1] * e1[1] + rh[2] * e1[2]);
    double m = 0.0;
    for (int k = 0; k < 3; ++k) {
        e2[k] = -(r * rh[k] - pe * e1[k]);
        m += e2[k] * e2[k];
    }
    if (!(m > 1e-24)) {  // aimed at the hole: any perpendicular will do
        const double te = th[0] * e1[0] + th[1] * e1[1] + th[2] * e1[2];
        m = 0.0;
        for (int k = 0; k < 3; ++k) {
            e2[k] = th[k] - te * e1[k];
            m += e2[k] * e2[k];
        }
    }
    m = 1.0 / sqrt(m);
    for (int k = 0; k < 3; ++k) e2[k] *= m;
    e3[0] = e1[1] * e2[2] - e1[2] * e2[1];
    e3[1] = e1[2] * e2[0] - e1[0] * e2[2];
    e3[2] = e1[0] * e2[1] - e1[1] * e2[0];
}

// sky angles from the network's components along (e1, e2, e3)
__host__ __device__ inline void sky_from_frame(const double e1[3], const double e2[3], const double e3[3], double c1,
                                               double c2, double c3, double& theta, double& phi) {
    double N[3];
    double n = 0.0;
    for (int k = 0; k < 3; ++k) {
        N[k] = c1 * e1[k] + c2 * e2[k] + c3 * e3[k];
        n += N[k] * N[k];
    }
    n = n > 0.0 ? 1.0 / sqrt(n) : 0.0;
    double z = N[2] * n;
    z = z > 1.0 ? 1.0 : (z < -1.0 ? -1.0 : z);
    theta = acos(z);
    phi = atan2(N[1], N[0]);
    if (phi < 0.0) phi += 2.0 * kPi;
}

// ---- the same in FP32, for the fused inference kernel --------------------------------------
//
// The network reads its inputs through a split-bf16 input layer and carries bf16
// activations, and its sky directions are good to ~1e-2 rad, so its inputs and the angles
// formed from its outputs need float accuracy only (~1e-7; the polar angle comes from
// atan2, well conditioned at the poles).  Training labels keep the FP64 forms above.

__device__ __forceinline__ void cell_f(float x, float x0, float x1, int n, int& i, float& f) {
    float u = (x - x0) * ((float)(n - 1) / (x1 - x0));
    u = fminf(fmaxf(u, 0.f), (float)(n - 1));
    i = (int)u;
    if (i > n - 2) i = n - 2;
    f = u - (float)i;
}

__device__ __forceinline__ float table_rsh_f(const float* rsh, float a, float inc) {
    int ia, ii;
    float fa, fi;
    cell_f(a, (float)kTabA0, (float)kTabA1, kTabA, ia, fa);
    cell_f(inc, (float)kTabI0, (float)kTabI1, kTabI, ii, fi);
    const float v00 = rsh[ia * kTabI + ii], v01 = rsh[ia * kTabI + ii + 1];
    const float v10 = rsh[(ia + 1) * kTabI + ii], v11 = rsh[(ia + 1) * kTabI + ii + 1];
    return (1 - fa) * ((1 - fi) * v00 + fi * v01) + fa * ((1 - fi) * v10 + fi * v11);
}

__device__ __forceinline__ float table_rho_f(const float* rho, float a, float inc, float psi) {
    int ia, ii, ip;
    float fa, fi, fp;
    cell_f(a, (float)kTabA0, (float)kTabA1, kTabA, ia, fa);
    cell_f(inc, (float)kTabI0, (float)kTabI1, kTabI, ii, fi);
    cell_f(psi, 0.f, (float)kPi, kTabPsi, ip, fp);
    auto at = [&](int x, int y, int z) { return rho[(x * kTabI + y) * kTabPsi + z]; };
    float r = 0.f;
    for (int dx = 0; dx < 2; ++dx)
        for (int dy = 0; dy < 2; ++dy) {
            const float w = (dx ? fa : 1 - fa) * (dy ? fi : 1 - fi);
            r += w * ((1 - fp) * at(ia + dx, ii + dy, ip) + fp * at(ia + dx, ii + dy, ip + 1));
        }
    return r;
}

// one ray's camera and image-plane position in float: spin, camera radius, inclination
// (with its sine and cosine), sin/cos of the camera azimuth, Bardeen alpha and beta
struct RayF {
    float a, r_cam, inc, si, ci, sp, cp, alpha, beta;
};

// features_of in float
__device__ __forceinline__ void features_f(const float* rho, const float* rsh, const RayF& p, float f[5],
                                           float* ring) {
    const float R = table_rsh_f(rsh, p.a, p.inc);
    const float inv = __builtin_amdgcn_rcpf(1.5f * R);
    f[0] = __log10f(p.r_cam) - 2.0f;
    f[1] = (p.ci - 0.5415f) * (1.0f / 0.4545f);
    f[2] = (p.a - 0.499f) * (1.0f / 0.499f);
    f[3] = p.alpha * inv;
    f[4] = p.beta * inv;
    const float rc = table_rho_f(rho, p.a, p.inc, atan2f(fabsf(p.beta), p.alpha));
    *ring = fabsf(sqrtf(p.alpha * p.alpha + p.beta * p.beta) * __builtin_amdgcn_rcpf(rc) - 1.0f);
}

// ray_frame in float
__device__ __forceinline__ void ray_frame_f(const RayF& p, float e1[3], float e2[3], float e3[3]) {
    const float r = p.r_cam, al = p.alpha, be = p.beta;
    const float rh[3] = {p.si * p.cp, p.si * p.sp, p.ci}, th[3] = {p.ci * p.cp, p.ci * p.sp, -p.si},
                ph[3] = {-p.sp, p.cp, 0.f};
    float q = 1.0f - (al * al + be * be) / (r * r);
    q = q > 0.f ? sqrtf(q) : 0.f;
    float n = 0.f;
    for (int k = 0; k < 3; ++k) {
        e1[k] = -q * rh[k] - (be / r) * th[k] - (al / r) * ph[k];
        n += e1[k] * e1[k];
    }
    n = __frsqrt_rn(n);
    for (int k = 0; k < 3; ++k) e1[k] *= n;
    const float pe = r * (rh[0] * e1[0] + rh[1] * e1[1] + rh[2] * e1[2]);
    float m = 0.f;
    for (int k = 0; k < 3; ++k) {
        e2[k] = -(r * rh[k] - pe * e1[k]);
        m += e2[k] * e2[k];
    }
    if (!(m > 1e-12f * r * r)) {  // aimed at the hole: any perpendicular will do
        const float te = th[0] * e1[0] + th[1] * e1[1] + th[2] * e1[2];
        m = 0.f;
        for (int k = 0; k < 3; ++k) {
            e2[k] = th[k] - te * e1[k];
            m += e2[k] * e2[k];
        }
    }
    m = __frsqrt_rn(m);
    for (int k = 0; k < 3; ++k) e2[k] *= m;
    e3[0] = e1[1] * e2[2] - e1[2] * e2[1];
    e3[1] = e1[2] * e2[0] - e1[0] * e2[2];
    e3[2] = e1[0] * e2[1] - e1[1] * e2[0];
}

// sky_from_frame in float: theta from atan2 (accurate at the poles, where acos is not)
__device__ __forceinline__ void sky_from_frame_f(const float e1[3], const float e2[3], const float e3[3], float c1,
                                                 float c2, float c3, double& theta, double& phi) {
    float N[3];
    for (int k = 0; k < 3; ++k) N[k] = c1 * e1[k] + c2 * e2[k] + c3 * e3[k];
    theta = (double)atan2f(sqrtf(N[0] * N[0] + N[1] * N[1]), N[2]);
    float f = atan2f(N[1], N[0]);
    if (f < 0.f) f += 2.0f * (float)kPi;
    phi = (double)f;
}

// ---- kernels -------------------------------------------------------------------------------

__device__ __forceinline__ int find_view(const ViewDev* v, int n, long long idx) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (v[mid].offset <= idx) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Bardeen parameters of ray q of the batch, in float (the fused inference kernels): false for rays the surrogate does not
// cover (cameras other than Bardeen planes, views outside the network's domain).  The
// camera azimuth's sine and cosine (sp, cp) only WITH_PHI: only answered escaped rays need
// them, for their sky angles.
template <bool WITH_PHI = true>
__device__ __forceinline__ bool ray_params_f(const ViewDev* views, int n_views, long long q, RayF& p) {
    const ViewDev& v = views[find_view(views, n_views, q)];
    const long long local = q - v.offset;
    if (v.kind == 2) {
        const double* pp = v.params + 5 * local;
        p.a = (float)pp[0];
        p.r_cam = (float)pp[1];
        p.inc = (float)pp[2];
        p.alpha = (float)pp[3];
        p.beta = (float)pp[4];
        sincosf(p.inc, &p.si, &p.ci);
        p.sp = 0.f;
        p.cp = 1.f;
        return true;
    }
    if (v.kind != 0 || !v.sur) return false;
    p.a = (float)v.a;
    p.r_cam = (float)v.r_cam;
    p.si = (float)v.sinI;
    p.ci = (float)v.cosI;
    p.inc = atan2f(p.si, p.ci);
    if (WITH_PHI) sincosf((float)v.phi_cam, &p.sp, &p.cp);
    if (v.params) {
        p.alpha = (float)v.params[2 * local];
        p.beta = (float)v.params[2 * local + 1];
    } else {
        // a view holds at most 2^31 pixels: the row and column in 32 bits
        const unsigned lu = (unsigned)local, w = (unsigned)v.width;
        const int row = (int)(lu / w), col = (int)(lu - (unsigned)row * w);
        p.alpha = (float)(-v.half_width + ((double)col + 0.5) * v.px);
        p.beta = (float)(v.half_height - ((double)row + 0.5) * v.py);
    }
    return true;
}

// features and targets of integrator-labelled rays (kind-2 params)
__global__ void k_targets(const double* params, long long n, Table t, const int8_t* cls, const double* th,
                          const double* ph, const double* rh, const double* gg, float* feat,
                          int8_t* ocls, float* tgt, float* ring) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double* p = params + 5 * q;
    float f[5], rg;
    features_of(t.rho, t.rsh, p[0], p[1], p[2], p[3], p[4], f, &rg);
    for (int k = 0; k < 5; ++k) feat[q * 5 + k] = f[k];
    if (ring) ring[q] = rg;
    int c = cls[q] & 7;
    if (c == 3) c = 1;  // a ray that exhausted its budget orbits on the critical curve: dark
    ocls[q] = (int8_t)c;
    float* o = tgt + q * 5;
    for (int k = 0; k < 5; ++k) o[k] = 0.f;
    if (c == 0) {
        double e1[3], e2[3], e3[3];
        ray_frame(p[1], p[2], 0.0, p[3], p[4], e1, e2, e3);
        const double st = sin(th[q]);
        const double N[3] = {st * cos(ph[q]), st * sin(ph[q]), cos(th[q])};
        o[0] = (float)(N[0] * e1[0] + N[1] * e1[1] + N[2] * e1[2]);
        o[1] = (float)(N[0] * e2[0] + N[1] * e2[1] + N[2] * e2[2]);
        o[2] = (float)(N[0] * e3[0] + N[1] * e3[1] + N[2] * e3[2]);
    } else if (c == 2) {
        o[3] = (float)log(rh[q]);
        o[4] = (float)log(gg[q]);
    }
}

// batch indices: a keyed Feistel bijection on [0, 2^bits), cycle-walked into [0, n)
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// balanced Feistel network on [0, 2^bits), bits even: (L, R) -> (R, L ^ F(R)), four rounds
__device__ __forceinline__ uint64_t feistel(uint64_t x, int bits, uint64_t key) {
    const int half = bits / 2;
    const uint64_t mask = (1ull << half) - 1;
    uint64_t L = x >> half, R = x & mask;
    for (int r = 0; r < 4; ++r) {
        const uint64_t F = mix32((uint32_t)R ^ (uint32_t)(key >> (16 * r)) ^ ((uint32_t)r * 0x9e3779b9u)) & mask;
        const uint64_t nR = L ^ F;
        L = R;
        R = nR;
    }
    return (L << half) | R;
}

__global__ void k_batch_index(long long n, int bits, uint64_t key, long long start, int B, long long* idx) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    uint64_t x = (uint64_t)((start + b) % n);
    do {
        x = feistel(x, bits, key);
    } while (x >= (uint64_t)n);
    idx[b] = (long long)x;
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// the fallback queue's length (the sum over its regions) for the integrator and the stats
__global__ void k_copy_count(const unsigned long long* c, long long* dst, unsigned long long* stat) {
    if (threadIdx.x == 0) {
        unsigned long long n = 0;
        for (int b = 0; b < kti::kQueueBuckets; ++b) n += c[b];
        *dst = (long long)n;
        if (stat) *stat = n;
    }
}

// ---- the fused network: one kernel from features to gated outputs -------------------------
//
// One wave owns 32 rays and keeps their activations in registers from layer to layer; the
// weights stream through LDS in 16 KB slabs shared by the block's 4 waves.  The layers use
// v_mfma_f32_32x32x16_bf16 with the transposed product D[n][m] = sum_k W[k][n] X[m][k]:
// the output feature n is the MFMA row, the ray m the column.  Lane l holds ray m = l % 32
// and, of every 32-feature block fb, the 16 features 32 fb + 8 q + 4 h + r (h = l / 32,
// q, r = 0..3) in accumulator element 4 q + r.  Read as the next layer's B operand, the
// elements 8 j .. 8 j + 7 of block fb form k-block 2 fb + j, so k-slot s of lane half h in
// k-block kb stands for feature
//     F(kb, h, s) = 32 (kb / 2) + 8 (2 (kb % 2) + s / 4) + 4 h + s % 4,
// and the packed weights hold W[F(kb, h, s)][n] in A-operand order (lane = 32 h + n % 32).
// The accumulators therefore never leave the lane: SiLU and the bf16 rounding produce the
// next layer's operand in place.
//
// A hidden layer runs in output-block order: block fb is one chain of sixteen MFMAs over
// the layer's whole input (one 16 KB slab: the weights of its 32 outputs), and the SiLU of
// block fb - 1, which the chain does not wait for, issues between its MFMAs.  The vector
// ALU, the kernel's busiest unit (two transcendentals per activation), so works while the
// matrix cores do.
//
// Every pre-activation is kept scaled by -log2(e): z' = -log2(e) z, so exp(-z) = 2^z' is a
// single v_exp_f32, and the stored activation is z' sigma(z) = -log2(e) silu(z).  A layer's
// weights need no change (its input carries the factor and so must its output), its bias
// carries it, and the output layer multiplies by -ln(2).
//
// The input layer also runs on the matrix cores: one 32x32x16 MFMA per output block with
// split-bf16 operands, f = fh + fl and W = Wh + Wl (each part bf16), summing fh Wh + fl Wh +
// fh Wl over the five inputs in k-slots 0-4, 5-9 and 10-14 (slot 15 zero).  The parts carry
// 16 significant bits and the dropped fl Wl is 2^-18 relative, so its pre-activations are
// good to ~1e-5, far below the bf16 rounding of the activations they feed.
constexpr int kWideH = 256;            // the wide network's hidden width
constexpr int kFusedRays = 128;        // rays per block: 4 waves x 32
constexpr int kSlabElems = 8192;       // bf16 per slab (16 KB): one output block, all 16 k-blocks
constexpr int kHiddenSlabs = 3 * 8;    // layers 1-3, eight slabs each; then one slab for layer 4
constexpr int kPackedElems = (kHiddenSlabs + 1) * kSlabElems;
constexpr int kInputFragElems = 8 * 64 * 8;  // input layer: [fb][lane][8]
// the scaled FP32 copies: input weights [5][256], input bias, biases of layers 1-3
constexpr int kFoldElems = ktn::kInputs * kWideH + 4 * kWideH;
constexpr float kNegLog2e = -1.44269504088896341f, kNegLn2 = -0.693147180559945309f;

typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline int packed_feature(int kb, int h, int s) {
    return 32 * (kb >> 1) + 8 * (2 * (kb & 1) + (s >> 2)) + 4 * h + (s & 3);
}

// packed bf16 weights from the FP32 masters: slab t < 24 is output block t % 8 of layer
// 1 + t / 8 for all sixteen k-blocks, [kb][lane][8] with row n = 32 (t % 8) + lane % 32;
// slab 24 is the output layer, [kb][lane][8] with rows n >= 8 zero
__global__ void k_pack_fused(const float* P, uint16_t* wpk) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kPackedElems) return;
    const int t = e / kSlabElems, w = e % kSlabElems;
    const int kb = w / 512, lane = (w / 8) % 64, s = w % 8;
    const int f = packed_feature(kb, lane >> 5, s);
    float v;
    if (t < kHiddenSlabs) {
        const int l = 1 + t / 8, n = 32 * (t % 8) + (lane & 31);
        v = P[ktn::param_offset(kWideH, l) + (long long)f * kWideH + n];
    } else {
        const int n = lane & 31;
        v = n < ktn::kOutputs ? P[ktn::param_offset(kWideH, 4) + (long long)f * ktn::kOutputs + n] : 0.f;
    }
    wpk[e] = f2bf(v);
}

// the input layer and the biases of layers 0-3 times -log2(e)
__global__ void k_fold(const float* P, float* fold) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kFoldElems) return;
    const int n0 = (ktn::kInputs + 1) * kWideH;  // input layer: weights then bias
    const float v = e < n0 ? P[e]
                           : P[ktn::param_offset(kWideH, 1 + (e - n0) / kWideH) + (long long)kWideH * kWideH +
                               (e - n0) % kWideH];
    fold[e] = v * kNegLog2e;
}

// the input layer's split-bf16 A fragments from the scaled weights: row n = 32 fb +
// lane % 32, k-slot kk = 8 (lane / 32) + s holds Wh[kk] (kk < 5), Wh[kk - 5] (kk < 10),
// Wl[kk - 10] (kk < 15), 0
__global__ void k_pack_input(const float* fold, uint16_t* wpk0) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kInputFragElems) return;
    const int fb = e / 512, lane = (e / 8) % 64, kk = 8 * (lane >> 5) + e % 8;
    const int n = 32 * fb + (lane & 31);
    uint16_t v = 0;
    if (kk < 15) {
        const float w = fold[(kk % 5) * kWideH + n];
        const uint16_t hi = f2bf(w);
        v = kk < 10 ? hi : f2bf(w - bf2f(hi));
    }
    wpk0[e] = v;
}

struct FusedArgs {
    long long n;
    // inputs: either views (features computed per ray) or a feature array
    const ViewDev* views;
    int n_views;
    Table t;
    const float* xf;       // n x 5 when views == nullptr
    const float* P;        // FP32 masters (output bias)
    const float* fold;     // input layer and biases of layers 0-3 times -log2(e)
    const uint16_t* wpk0;  // the input layer's split-bf16 fragments
    const uint16_t* wpk;   // packed bf16 weights of layers 1-4
    // the fast network (k_mlp_fast): its packed weights (hidden layers, output layer, the
    // input layer's split fragments) and biases, all staged into LDS once per block
    const uint4* wfast;
    const float* bfast;
    // outputs: raw network outputs (n x 8), or the gate's decisions
    float* y;
    float gate_prob, gate_ring;
    // 1: the closed form verifies and replaces what the network answers as escaped or disk
    // hits, so those pass the gate whatever their probability and ring distance
    int verify;
    float extent;          // answered rays have |alpha|, |beta| <= extent x 1.5 R_sh
    ktd::Outputs out;
    long long* fallback;             // kti::kQueueBuckets regions fallback_stride apart
    unsigned long long* n_fallback;  // their counts
    long long fallback_stride;
    // k_mlp_fast: tiles past the first gridDim.x are claimed from this counter (zero at
    // launch), so blocks that start late, on CUs still busy with other work, take fewer
    unsigned long long* next_tile;
};

// -log2(e) silu(z) from z' = -log2(e) z
__device__ __forceinline__ float silu_scaled(float zs) {
    return zs * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(zs));
}

// one pair of the next layer's B operand (k-blocks 2 fb and 2 fb + 1) from block fb of a
// layer's scaled pre-activations: SiLU, rounded to bf16
__device__ __forceinline__ void silu_pair(const f32x16& z, bf16x8& p0, bf16x8& p1) {
#pragma unroll
    for (int i = 0; i < 8; ++i) p0[i] = (__bf16)silu_scaled(z[i]);
#pragma unroll
    for (int i = 0; i < 8; ++i) p1[i] = (__bf16)silu_scaled(z[8 + i]);
}

// the 16 values of block fb this lane holds (features 32 fb + 8 q + 4 h + r)
__device__ __forceinline__ f32x16 block_bias(const float* b, int fb, int h) {
    f32x16 z;
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
        const float4 bb = *reinterpret_cast<const float4*>(b + 32 * fb + 8 * qq + 4 * h);
        z[4 * qq + 0] = bb.x;
        z[4 * qq + 1] = bb.y;
        z[4 * qq + 2] = bb.z;
        z[4 * qq + 3] = bb.w;
    }
    return z;
}

__device__ __forceinline__ f32x16 mfma32(const uint4& a, const bf16x8& b, const f32x16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a), b, c, 0, 0, 0);
}

// The gate and the answered rays' outputs, from the network outputs y of ray q (its owner
// lane decides and writes; lanes that hold a ray twice pass owner only on one): a ray is
// answered when its largest class probability exceeds gate_prob and it lies outside
// gate_ring of the critical curve (rg < 0: not coverable), or (verify) when its class is
// escaped or disk; the others join the fallback queue, one atomic per wave.  Every lane of
// the wave calls it.
__device__ __forceinline__ void gate_and_write(const FusedArgs& a, long long q, bool live, bool owner, int lane,
                                               const float y[8], float rg) {
    const float mx = fmaxf(y[0], fmaxf(y[1], y[2]));
    const float e0 = __expf(y[0] - mx), e1 = __expf(y[1] - mx), e2 = __expf(y[2] - mx);
    const float inv = 1.0f / (e0 + e1 + e2);
    const float p0 = e0 * inv, p1 = e1 * inv, p2 = e2 * inv;
    const int c = p0 >= p1 && p0 >= p2 ? 0 : (p1 >= p2 ? 1 : 2);
    const float pm = c == 0 ? p0 : c == 1 ? p1 : p2;
    const bool verified = a.verify && (c == 0 || c == 2) && rg >= 0.0f;
    const bool to_fallback = live && owner && !verified && (!(pm > a.gate_prob) || !(rg > a.gate_ring));
    // the fallback queue in one region per predicted class, so that a wave of the integrator
    // takes rays of like length (escaped rays near the ring run longest, disk rays shortest):
    // one atomic per wave and region
#pragma unroll
    for (int b = 0; b < kti::kQueueBuckets; ++b) {
        const unsigned long long mask = __ballot(to_fallback && c == b);
        if (mask) {
            const int first = __ffsll((long long)mask) - 1;
            unsigned long long base = 0;
            if (lane == first) base = atomicAdd(a.n_fallback + b, (unsigned long long)__popcll(mask));
            base = __shfl(base, first);
            if (to_fallback && c == b)
                a.fallback[b * a.fallback_stride + (long long)(base + __popcll(mask & ((1ull << lane) - 1)))] = q;
        }
    }
    if (!live || !owner || to_fallback) return;
    const double nan = __builtin_nan("");
    double th = nan, ph = nan, rh = nan, g = nan;
    if (c == 0) {
        RayF p;
        ray_params_f(a.views, a.n_views, q, p);
        float u1[3], u2[3], u3[3];
        ray_frame_f(p, u1, u2, u3);
        sky_from_frame_f(u1, u2, u3, y[3], y[4], y[5], th, ph);
    } else if (c == 2) {
        rh = (double)__expf(y[6]);
        g = (double)__expf(y[7]);
    }
    if (a.out.cls) a.out.cls[q] = (int8_t)(c | ktd::kFlagSurrogate);
    if (a.out.theta) a.out.theta[q] = th;
    if (a.out.phi) a.out.phi[q] = ph;
    if (a.out.r_hit) a.out.r_hit[q] = rh;
    if (a.out.g) a.out.g[q] = g;
    if (a.out.steps) a.out.steps[q] = 0;
    if (a.out.theta32) a.out.theta32[q] = (float)th;
    if (a.out.phi32) a.out.phi32[q] = (float)ph;
    if (a.out.r_hit32) a.out.r_hit32[q] = (float)rh;
    if (a.out.g32) a.out.g32[q] = (float)g;
}

template <bool VIEWS>
// two waves per SIMD: 248 VGPRs, accumulators in the VGPR file, no spills
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mlp_fused(FusedArgs a) {
    __shared__ uint4 slab0[kSlabElems / 8];
    __shared__ uint4 slab1[kSlabElems / 8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 31, h = lane >> 5;
    const long long q = (long long)blockIdx.x * kFusedRays + wave * 32 + m;
    const bool live = q < a.n;
    const long long qc = live ? q : a.n - 1;

    // stage slab 0 while the inputs are formed
    const uint4* src = reinterpret_cast<const uint4*>(a.wpk);
    // four named registers, not an array: the compiler would place a private array in LDS
    uint4 w0 = src[tid], w1 = src[tid + 256], w2 = src[tid + 512], w3 = src[tid + 768];

    float f[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    float rg = -1.f;  // relative distance from the critical curve; < 0: not coverable
    if (VIEWS) {
        RayF p;
        if (ray_params_f<false>(a.views, a.n_views, qc, p)) {
            features_f(a.t.rho, a.t.rsh, p, f, &rg);
            // outside the image-plane square the network was trained on: not coverable
            if (!(fabsf(f[3]) <= a.extent) || !(fabsf(f[4]) <= a.extent)) rg = -1.f;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) f[k] = a.xf[qc * 5 + k];
    }

    // input layer: the ray's split features as the B operand (lane half 0: fh0-4, fl0-2;
    // half 1: fl3-4, fh0-4, 0), one MFMA per output block, then layer 1's operand
    bf16x8 B[16];
    {
        float fh[5], fl[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            fh[k] = (float)(__bf16)f[k];
            fl[k] = f[k] - fh[k];  // exact
        }
        const float x0[8] = {fh[0], fh[1], fh[2], fh[3], fh[4], fl[0], fl[1], fl[2]};
        const float x1[8] = {fl[3], fl[4], fh[0], fh[1], fh[2], fh[3], fh[4], 0.f};
        bf16x8 xb;
#pragma unroll
        for (int k = 0; k < 8; ++k) xb[k] = (__bf16)(h ? x1[k] : x0[k]);
        const uint4* a0 = reinterpret_cast<const uint4*>(a.wpk0);
        const float* b0 = a.fold + ktn::kInputs * kWideH;
#pragma unroll
        for (int fb = 0; fb < 8; ++fb) {
            const f32x16 z = mfma32(a0[fb * 64 + lane], xb, block_bias(b0, fb, h));
            silu_pair(z, B[2 * fb], B[2 * fb + 1]);
        }
    }

    slab0[tid] = w0;
    slab0[tid + 256] = w1;
    slab0[tid + 512] = w2;
    slab0[tid + 768] = w3;
    __syncthreads();
    {
        bf16x8 C[16];  // the next layer's operand, a pair per block
        f32x16 acc, prev;
#pragma unroll
        for (int l = 1; l <= 3; ++l) {
            const float* bl = a.fold + (ktn::kInputs + l) * kWideH;
            f32x16 bias = block_bias(bl, 0, h);
#pragma unroll
            for (int fb = 0; fb < 8; ++fb) {
                const int t = 8 * (l - 1) + fb;
                // the next slab (the output layer's after the last hidden one) into registers
                const uint4* nsrc = src + (long long)(t + 1) * (kSlabElems / 8);
                w0 = nsrc[tid];
                w1 = nsrc[tid + 256];
                w2 = nsrc[tid + 512];
                w3 = nsrc[tid + 768];
                acc = bias;
                if (fb < 7) bias = block_bias(bl, fb + 1, h);
                // scheduling barriers: the global loads above stay ahead of the matrix
                // work, all sixteen fragment reads issue before the first MFMA (sunk between
                // the MFMAs each would wait out its LDS latency), and the LDS stores follow
                __builtin_amdgcn_sched_barrier(0);
                // eight slabs per layer: the buffer parity is static (slab0 at even fb)
                const uint4* S = (fb & 1) ? slab1 : slab0;
                uint4* N = (fb & 1) ? slab0 : slab1;
                uint4 af[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) af[i] = S[i * 64 + lane];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < 16; ++i) acc = mfma32(af[i], B[i], acc);
                if (fb > 0) silu_pair(prev, C[2 * fb - 2], C[2 * fb - 1]);
                prev = acc;
                __builtin_amdgcn_sched_barrier(0);
                N[tid] = w0;
                N[tid + 256] = w1;
                N[tid + 512] = w2;
                N[tid + 768] = w3;
                __syncthreads();
            }
            silu_pair(prev, C[14], C[15]);
#pragma unroll
            for (int k = 0; k < 16; ++k) B[k] = C[k];
        }
    }

    // output layer: four partial sums over the sixteen k-blocks
    f32x16 o[4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int i = 0; i < 16; ++i) o[p][i] = 0.f;
    {
        const uint4* S = slab0;  // 24 hidden slabs: the output layer's landed in slab0
#pragma unroll
        for (int kb = 0; kb < 16; ++kb) o[kb & 3] = mfma32(S[kb * 64 + lane], B[kb], o[kb & 3]);
    }
    const float* b4 = a.P + ktn::param_offset(kWideH, 4) + (long long)kWideH * ktn::kOutputs;
    float y4[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) y4[r] = (o[0][r] + o[1][r] + o[2][r] + o[3][r]) * kNegLn2 + b4[4 * h + r];
    // lane m (h = 0) holds outputs 0..3 of ray m, lane m + 32 outputs 4..7
    float y[8];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        y[r] = y4[r];
        y[4 + r] = __shfl_down(y4[r], 32);
    }
    if (!VIEWS) {
        if (live && h == 0)
#pragma unroll
            for (int k = 0; k < 8; ++k) a.y[q * 8 + k] = y[k];
        return;
    }

    gate_and_write(a, q, live, h == 0, lane, y, rg);
}

// ---- the fast network: 5 -> 128 -> 128 -> 128 -> 128 -> 8, hard SiLU --------------------
// The same transposed product and in-lane operand layout as k_mlp_fused, over four 32-feature
// output blocks and eight k-blocks per layer.  Its 16-bit weights, 108 KB, fit in LDS whole:
// a block of eight waves stages them once and then runs over tiles of ray tiles until the
// batch is done (one block per CU), so no weight traffic remains in the loop.  The input
// layer multiplies split-bf16 features by split-bf16 weights; the hidden and output layers
// multiply FP16 activations by FP16 weights (the FP32 masters rounded once; FP16 holds
// three more mantissa bits than bf16, and these values stay far inside its range), which
// lets the activation run on packed halves.
//
// Hard SiLU, a = z clamp(z / 6 + 1/2, 0, 1).  The kernel keeps every activation divided by
// 6: with z' = z / 6 it is z' clamp(z' + 1/2, 0, 1), an add with the clamp modifier and a
// multiply.  The hidden layers' weights are unchanged by this (their input carries the 1/6
// and so does their output), the input layer's weights and the biases of layers 0-3 are
// divided by 6, and the output layer multiplies by 6.
constexpr int kFastH = 128;
constexpr int kFastBlocks = kFastH / 32, kFastK = kFastH / 16;   // output blocks, k-blocks
constexpr int kFastLayerFrags = kFastBlocks * kFastK * 64;       // uint4 per hidden layer
constexpr int kFastOutFrags = kFastK * 64;                       // uint4 of the output layer
constexpr int kFastInFrags = kFastBlocks * 64;                   // uint4 of the input layer
constexpr int kFastFrags = 3 * kFastLayerFrags + kFastOutFrags + kFastInFrags;  // 6912 (108 KB)
constexpr int kFastBias = 4 * kFastH + 8;                        // b0..b3 / 6, then b4
constexpr int kFastThreads = 512, kFastTile = 256;               // 8 waves x 32 rays per network pass

__device__ __forceinline__ uint16_t f2h(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }

// packed fragments from the FP32 masters: hidden layer l (1-3) block fb, k-block kb, lane
// (row n = 32 fb + lane % 32, k-slots F(kb, lane / 32, 0..7)) in FP16; then the output layer
// [kb][lane] (rows n >= 8 zero) in FP16; then the input layer's split-bf16 fragments of
// W0 / 6 [fb][lane] (k-slot kk holds Wh[kk] for kk < 5, Wh[kk - 5] for kk < 10, Wl[kk - 10]
// for kk < 15, 0)

__global__ void k_pack_fast(const float* P, uint16_t* w, float* bias) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = kFastH;
    if (e < kFastBias) {  // biases
        float v;
        if (e < 4 * H) {
            const int l = e / H, n = e % H;
            v = P[ktn::param_offset(H, l) + (long long)ktn::kreal(H, l) * H + n] * (1.0f / 6.0f);
        } else {
            v = P[ktn::param_offset(H, 4) + (long long)H * ktn::kOutputs + (e - 4 * H)];
        }
        bias[e] = v;
    }
    if (e >= kFastFrags * 8) return;
    const int frag = e / 8, s = e % 8;
    uint16_t out = 0;
    if (frag < 3 * kFastLayerFrags) {
        const int l = 1 + frag / kFastLayerFrags, r = frag % kFastLayerFrags;
        const int fb = r / (kFastK * 64), kb = (r / 64) % kFastK, lane = r % 64;
        const int f = packed_feature(kb, lane >> 5, s), n = 32 * fb + (lane & 31);
        out = f2h(P[ktn::param_offset(H, l) + (long long)f * H + n]);
    } else if (frag < 3 * kFastLayerFrags + kFastOutFrags) {
        const int r = frag - 3 * kFastLayerFrags;
        const int kb = r / 64, lane = r % 64, n = lane & 31;
        const int f = packed_feature(kb, lane >> 5, s);
        out = n < ktn::kOutputs ? f2h(P[ktn::param_offset(H, 4) + (long long)f * ktn::kOutputs + n]) : 0;
    } else {
        const int r = frag - 3 * kFastLayerFrags - kFastOutFrags;
        const int fb = r / 64, lane = r % 64, kk = 8 * (lane >> 5) + s, n = 32 * fb + (lane & 31);
        if (kk < 15) {
            const float wv = P[(kk % 5) * H + n] * (1.0f / 6.0f);
            const uint16_t hi = f2bf(wv);
            out = kk < 10 ? hi : f2bf(wv - bf2f(hi));
        }
    }
    w[e] = out;
}

// hard SiLU of a scaled pre-activation z' = z / 6: z' clamp(z' + 1/2, 0, 1) = hsilu(z) / 6.
// The clamp is one v_med3_f32 (a min/max pair compiles to compares and selects).
__device__ __forceinline__ float hsilu_scaled(float zs) {
    return zs * __builtin_amdgcn_fmed3f(zs + 0.5f, 0.0f, 1.0f);
}

// The fast network's activations are FP16 (its hidden and output layers multiply on
// v_mfma_f32_32x32x16_f16): a pair of scaled pre-activations goes to packed halves in one
// convert (round to nearest even), and the hard SiLU x clamp(x + 1/2, 0, 1) is then one
// packed add with the clamp modifier and one packed multiply: three instructions for two
// values, where the FP32 form takes five.  (Written as an instruction because the compiler
// keeps a min/max pair on packed halves instead of the clamp bit.)
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(2 * sizeof(_Float16)))) _Float16 f16x2;

__device__ __forceinline__ unsigned hsilu2_h(float z0, float z1) {
    const f16x2 x = {(_Float16)z0, (_Float16)z1};
    f16x2 t;
    asm("v_pk_add_f16 %0, %1, 0.5 op_sel_hi:[1,0] clamp" : "=v"(t) : "v"(x));
    return __builtin_bit_cast(unsigned, x * t);
}

__device__ __forceinline__ void hsilu_pair(const f32x16& z, f16x8& p0, f16x8& p1) {
    uint4 lo, hi;
    lo.x = hsilu2_h(z[0], z[1]);
    lo.y = hsilu2_h(z[2], z[3]);
    lo.z = hsilu2_h(z[4], z[5]);
    lo.w = hsilu2_h(z[6], z[7]);
    hi.x = hsilu2_h(z[8], z[9]);
    hi.y = hsilu2_h(z[10], z[11]);
    hi.z = hsilu2_h(z[12], z[13]);
    hi.w = hsilu2_h(z[14], z[15]);
    p0 = __builtin_bit_cast(f16x8, lo);
    p1 = __builtin_bit_cast(f16x8, hi);
}

__device__ __forceinline__ f32x16 mfma32h(const uint4& a, const f16x8& b, const f32x16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), b, c, 0, 0, 0);
}

// the input layer's B operand of the 32 rays of a group (lane m and m + 32 hold ray m):
// split-bf16 features, k-slots (hi 0-4, lo 0-2) in the lower half-wave and (lo 3-4, hi 0-4,
// 0) in the upper, as the input layer's fragments expect
__device__ __forceinline__ bf16x8 input_split(const float f[5], int h) {
    float fh[5], fl[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        fh[k] = (float)(__bf16)f[k];
        fl[k] = f[k] - fh[k];
    }
    const float x0[8] = {fh[0], fh[1], fh[2], fh[3], fh[4], fl[0], fl[1], fl[2]};
    const float x1[8] = {fl[3], fl[4], fh[0], fh[1], fh[2], fh[3], fh[4], 0.f};
    bf16x8 xb;
#pragma unroll
    for (int k = 0; k < 8; ++k) xb[k] = (__bf16)(h ? x1[k] : x0[k]);
    return xb;
}

// The network for both 32-ray groups of a wave at once (group g's ray m in lanes m and
// m + 32), layer by layer: each weight fragment read from LDS feeds the two groups' MFMAs,
// which halves the LDS traffic of the single-group form, and the two accumulation chains
// interleave, so an MFMA never waits on the one before it.  Returns the four outputs
// 4 h .. 4 h + 3 of group g's ray m in lane (m, h) of y[g].
__device__ __forceinline__ void fast_network2(const uint4* W, const float* Bs, int lane, int h, const float f0[5],
                                              const float f1[5], float y0[4], float y1[4]) {
    const uint4* Win = W + 3 * kFastLayerFrags + kFastOutFrags;
    const uint4* Wout = W + 3 * kFastLayerFrags;
    f16x8 B0[kFastK], B1[kFastK];
    {
        const bf16x8 xb0 = input_split(f0, h), xb1 = input_split(f1, h);
#pragma unroll
        for (int fb = 0; fb < kFastBlocks; ++fb) {
            const uint4 w = Win[fb * 64 + lane];
            const f32x16 bias = block_bias(Bs, fb, h);
            const f32x16 z0 = mfma32(w, xb0, bias);
            const f32x16 z1 = mfma32(w, xb1, bias);
            hsilu_pair(z0, B0[2 * fb], B0[2 * fb + 1]);
            hsilu_pair(z1, B1[2 * fb], B1[2 * fb + 1]);
        }
    }
#pragma unroll
    for (int l = 1; l <= 3; ++l) {
        const uint4* Wl = W + (l - 1) * kFastLayerFrags;
        const float* bl = Bs + l * kFastH;
        f16x8 C0[kFastK], C1[kFastK];
        f32x16 prev0, prev1;
        // the previous output block's activation runs beside this block's MFMA chains
#pragma unroll
        for (int fb = 0; fb < kFastBlocks; ++fb) {
            f32x16 acc0 = block_bias(bl, fb, h);
            f32x16 acc1 = acc0;
#pragma unroll
            for (int i = 0; i < kFastK; ++i) {
                const uint4 w = Wl[(fb * kFastK + i) * 64 + lane];
                acc0 = mfma32h(w, B0[i], acc0);
                acc1 = mfma32h(w, B1[i], acc1);
            }
            if (fb > 0) {
                hsilu_pair(prev0, C0[2 * fb - 2], C0[2 * fb - 1]);
                hsilu_pair(prev1, C1[2 * fb - 2], C1[2 * fb - 1]);
            }
            prev0 = acc0;
            prev1 = acc1;
        }
        hsilu_pair(prev0, C0[kFastK - 2], C0[kFastK - 1]);
        hsilu_pair(prev1, C1[kFastK - 2], C1[kFastK - 1]);
#pragma unroll
        for (int k = 0; k < kFastK; ++k) {
            B0[k] = C0[k];
            B1[k] = C1[k];
        }
    }
    f32x16 o0[2], o1[2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int i = 0; i < 16; ++i) o0[p][i] = o1[p][i] = 0.f;
#pragma unroll
    for (int kb = 0; kb < kFastK; ++kb) {
        const uint4 w = Wout[kb * 64 + lane];
        o0[kb & 1] = mfma32h(w, B0[kb], o0[kb & 1]);
        o1[kb & 1] = mfma32h(w, B1[kb], o1[kb & 1]);
    }
    const float* b4 = Bs + 4 * kFastH;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        y0[r] = (o0[0][r] + o0[1][r]) * 6.0f + b4[4 * h + r];
        y1[r] = (o1[0][r] + o1[1][r]) * 6.0f + b4[4 * h + r];
    }
}


// The fast kernel's inputs of ray q, false when its view is not covered: the view's
// constants come from the host (kts::view_constants); the critical-curve radius at the ray's
// angle from `slice`, the table's psi row interpolated to the view's spin and inclination
// (staged in LDS for a one-view launch), else from the table itself.
__device__ __forceinline__ bool fast_features(const FusedArgs& a, const float* slice, long long q, float f[5],
                                              float* rg) {
    const ViewDev& v = a.views[find_view(a.views, a.n_views, q)];
    if (v.kind != 0 || !v.sur) return false;
    const long long local = q - v.offset;
    float alpha, beta;
    if (v.params) {
        alpha = (float)v.params[2 * local];
        beta = (float)v.params[2 * local + 1];
    } else {
        const unsigned lu = (unsigned)local, w = (unsigned)v.width;
        const int row = (int)(lu / w), col = (int)(lu - (unsigned)row * w);
        alpha = (float)(-v.half_width + ((double)col + 0.5) * v.px);
        beta = (float)(v.half_height - ((double)row + 0.5) * v.py);
    }
    f[0] = v.sur_c[0];
    f[1] = v.sur_c[1];
    f[2] = v.sur_c[2];
    f[3] = alpha * v.sur_c[3];
    f[4] = beta * v.sur_c[3];
    const float psi = atan2f(fabsf(beta), alpha);
    float rc;
    if (slice) {
        int ip;
        float fp;
        cell_f(psi, 0.f, (float)kPi, kTabPsi, ip, fp);
        rc = (1 - fp) * slice[ip] + fp * slice[ip + 1];
    } else {
        rc = table_rho_f(a.t.rho, (float)v.a, v.sur_c[4], psi);
    }
    *rg = fabsf(sqrtf(alpha * alpha + beta * beta) * __builtin_amdgcn_rcpf(rc) - 1.0f);
    return true;
}

// the psi row of the critical-curve table at spin a and inclination inc (bilinear in them)
__device__ __forceinline__ float rho_row(const float* rho, float a, float inc, int k) {
    int ia, ii;
    float fa, fi;
    cell_f(a, (float)kTabA0, (float)kTabA1, kTabA, ia, fa);
    cell_f(inc, (float)kTabI0, (float)kTabI1, kTabI, ii, fi);
    auto at = [&](int x, int y) { return rho[(x * kTabI + y) * kTabPsi + k]; };
    return (1 - fa) * ((1 - fi) * at(ia, ii) + fi * at(ia, ii + 1)) + fa * ((1 - fi) * at(ia + 1, ii) + fi * at(ia + 1, ii + 1));
}

// A wave takes 64 rays at a time, one per lane for everything but the network: lane l forms
// the inputs of ray l and, after the network has run on rays 0-31 and then on rays 32-63
// (each with the 32x32 layout, in which lanes m and m + 32 share ray m), gates and writes
// ray l.  The inputs, the gate and the stores, which do not scale with the network, so cost
// one pass per 64 rays, not two.
template <bool VIEWS>
// two waves per SIMD leave room on a CU for the integrator's waves beside it (three measured no faster)
__global__ __launch_bounds__(kFastThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mlp_fast(FusedArgs a) {
    __shared__ uint4 W[kFastFrags];
    __shared__ __attribute__((aligned(16))) float Bs[kFastBias];
    __shared__ float slice[kTabPsi];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 31, h = lane >> 5;
    for (int i = tid; i < kFastFrags; i += kFastThreads) W[i] = a.wfast[i];
    for (int i = tid; i < kFastBias; i += kFastThreads) Bs[i] = a.bfast[i];
    const bool one_view = VIEWS && a.n_views == 1 && a.views[0].sur;
    if (one_view)
        for (int k = tid; k < kTabPsi; k += kFastThreads)
            slice[k] = rho_row(a.t.rho, (float)a.views[0].a, a.views[0].sur_c[4], k);
    __syncthreads();
    constexpr int kPair = 2 * kFastTile;  // rays per block and iteration: 8 waves x 64
    const long long tiles = (a.n + kPair - 1) / kPair;
    __shared__ long long claimed;
    for (long long tile = blockIdx.x; tile < tiles;) {
        const long long q = tile * kPair + wave * 64 + lane;  // this lane's ray
        const bool live = q < a.n;
        const long long qc = live ? q : a.n - 1;
        float f[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
        float rg = -1.f;
        if (VIEWS) {
            if (fast_features(a, one_view ? slice : nullptr, qc, f, &rg))
                if (!(fabsf(f[3]) <= a.extent) || !(fabsf(f[4]) <= a.extent)) rg = -1.f;
        } else {
#pragma unroll
            for (int k = 0; k < 5; ++k) f[k] = a.xf[qc * 5 + k];
        }
        // the two groups' inputs in the network's layout: ray m of group g sits in lane
        // 32 g + m
        float g0[5], g1[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            g0[k] = __shfl(f[k], m);
            g1[k] = __shfl(f[k], m + 32);
        }
        float y0[4], y1[4];
        fast_network2(W, Bs, lane, h, g0, g1, y0, y1);
        // outputs 0-3 of group 0's ray m sit in lane m, 4-7 in lane m + 32; group 1's alike:
        // one exchange across the half-waves gives lane l all eight of ray l
        float y[8];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float x = __shfl_xor(h ? y0[r] : y1[r], 32);
            y[r] = h ? x : y0[r];
            y[4 + r] = h ? y1[r] : x;
        }
        if (!VIEWS) {
            if (live)
#pragma unroll
                for (int k = 0; k < 8; ++k) a.y[q * 8 + k] = y[k];
        } else {
            gate_and_write(a, q, live, true, lane, y, rg);
        }
        if (a.next_tile) {
            __syncthreads();  // every wave has read `claimed` for this tile
            if (tid == 0) claimed = (long long)gridDim.x + (long long)atomicAdd(a.next_tile, 1ull);
            __syncthreads();
            tile = claimed;
        } else {
            tile += gridDim.x;
        }
    }
}

template <class T>
hipError_t dalloc(T** p, size_t n) {
    return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1));
}

}  // namespace

// ---- the surrogate object ---------------------------------------------------------------------

struct Surrogate {
    int device = 0;
    int cu_count = 0;
    bool have_params = false;
    int net = ktn::kNetWide;      // which network (ktn::kNet*): its width and activation
    int hid = 256, act = ktn::kActSilu;
    float* P = nullptr;     // FP32 master parameters
    float* G = nullptr;     // gradients
    float* M = nullptr;     // Adam moments
    float* V = nullptr;
    uint16_t* wb[ktn::kLayers] = {};
    uint16_t* wt[ktn::kLayers] = {};
    ktn::LayerTable lt{};
    float* tab_rho = nullptr;
    float* tab_rsh = nullptr;
    std::vector<float> h_rho, h_rsh;
    // work buffers, sized for `cap` rows
    long long cap = 0;
    float* xf = nullptr;        // inputs, B x 5
    float* Z[4] = {};
    uint16_t* H[4] = {};
    float* Y = nullptr;
    float* gy32 = nullptr;
    uint16_t* gyb = nullptr;
    float* g32[2] = {};
    uint16_t* gb[2] = {};
    int8_t* bcls = nullptr;
    float* btgt = nullptr;
    long long* bidx = nullptr;
    float* ws = nullptr;
    float* partial = nullptr;
    float* partial0 = nullptr;  // input-layer gradient partials
    float* colpart = nullptr;   // hidden-layer bias gradient partials, one row per 64 batch rows
    double* loss = nullptr;
    uint16_t* wpk = nullptr;  // layers 1-4 packed for the fused kernel (bf16)
    float* fold = nullptr;    // input layer and biases times -log2(e) for the fused kernel
    uint16_t* wpk0 = nullptr; // the input layer's split-bf16 MFMA fragments
    uint16_t* wfast = nullptr;  // the fast network's packed fragments (k_mlp_fast)
    float* bfast = nullptr;     // and its biases
    Queue queue;              // fallback queue of eager calls
    Domain dom;               // what the parameters cover
    // split-K slices of the weight gradients; row chunks of the bias and input-layer
    // gradient sums (64 rows each at the default batch)
    static constexpr int kSplits = 128, kChunks = 1024;
};

// frees what a surrogate holds; device frees during teardown report nothing actionable
static void free_all(Surrogate* s) {
    auto f = [](void* p) {
        if (p) (void)hipFree(p);
    };
    f(s->P);
    f(s->G);
    f(s->M);
    f(s->V);
    for (int l = 0; l < ktn::kLayers; ++l) {
        f(s->wb[l]);
        f(s->wt[l]);
    }
    f(s->tab_rho);
    f(s->tab_rsh);
    f(s->xf);
    f(s->partial0);
    for (int l = 0; l < 4; ++l) {
        f(s->Z[l]);
        f(s->H[l]);
    }
    f(s->Y);
    f(s->gy32);
    f(s->gyb);
    for (int k = 0; k < 2; ++k) {
        f(s->g32[k]);
        f(s->gb[k]);
    }
    f(s->bcls);
    f(s->btgt);
    f(s->bidx);
    f(s->ws);
    f(s->partial);
    f(s->loss);
    f(s->wpk);
    f(s->fold);
    f(s->wpk0);
    f(s->wfast);
    f(s->bfast);
    queue_release(s->queue);
}

#define KTS_TRY(x)                          \
    do {                                    \
        hipError_t e_ = (x);                \
        if (e_ != hipSuccess) return e_;    \
    } while (0)

static hipError_t build_table(Surrogate* s) {
    s->h_rho.assign((size_t)kTabA * kTabI * kTabPsi, 0.f);
    s->h_rsh.assign((size_t)kTabA * kTabI, 0.f);
    std::vector<double> al(16 * 200 + 8 + 200), be(al.size());
    for (int ia = 0; ia < kTabA; ++ia) {
        const double a = kTabA0 + (kTabA1 - kTabA0) * ia / (kTabA - 1);
        for (int ii = 0; ii < kTabI; ++ii) {
            const double inc = kTabI0 + (kTabI1 - kTabI0) * ii / (kTabI - 1);
            const int n = kt_critical_curve(a < 1e-7 ? 0.0 : a, inc, 200, al.data(), be.data(), (int)al.size());
            double mean = 0.0;
            for (int k = 0; k < 72; ++k) mean += kt_critical_radius(al.data(), be.data(), n, 2.0 * kPi * k / 72);
            s->h_rsh[(size_t)ia * kTabI + ii] = (float)(mean / 72.0);
            for (int ip = 0; ip < kTabPsi; ++ip) {
                const double psi = kPi * ip / (kTabPsi - 1);
                s->h_rho[((size_t)ia * kTabI + ii) * kTabPsi + ip] =
                    (float)kt_critical_radius(al.data(), be.data(), n, psi);
            }
        }
    }
    KTS_TRY(dalloc(&s->tab_rho, s->h_rho.size()));
    KTS_TRY(dalloc(&s->tab_rsh, s->h_rsh.size()));
    KTS_TRY(hipMemcpy(s->tab_rho, s->h_rho.data(), s->h_rho.size() * 4, hipMemcpyHostToDevice));
    KTS_TRY(hipMemcpy(s->tab_rsh, s->h_rsh.data(), s->h_rsh.size() * 4, hipMemcpyHostToDevice));
    return hipSuccess;
}

// the layer table of the current network (the bf16 copies keep their maximal allocation)
static void set_layers(Surrogate* s) {
    const int H = s->hid;
    for (int l = 0; l < ktn::kLayers; ++l) {
        s->lt.off[l] = ktn::param_offset(H, l);
        s->lt.kreal[l] = ktn::kreal(H, l);
        s->lt.kpad[l] = ktn::kpad(H, l);
        s->lt.n[l] = ktn::nout(H, l);
        s->lt.nld[l] = ktn::nld(H, l);
    }
}

hipError_t set_net(Surrogate* s, int net) {
    if (net != ktn::kNetWide && net != ktn::kNetFast) return hipErrorInvalidValue;
    if (net == s->net) return hipSuccess;
    KTS_TRY(hipSetDevice(s->device));
    KTS_TRY(hipDeviceSynchronize());
    s->net = net;
    s->hid = ktn::net_of(net).hidden;
    s->act = ktn::net_of(net).act;
    set_layers(s);
    // the bf16 working copies change their strides: clear them (padding must stay zero)
    for (int l = 0; l < ktn::kLayers; ++l) {
        const int H = ktn::kMaxHidden;
        KTS_TRY(hipMemset(s->wb[l], 0, (size_t)ktn::kpad(H, l) * ktn::nld(H, l) * 2));
        KTS_TRY(hipMemset(s->wt[l], 0, (size_t)ktn::nout(H, l) * ktn::kpad(H, l) * 2));
    }
    s->have_params = false;
    return hipSuccess;
}

int get_net(const Surrogate* s) { return s ? s->net : ktn::kNetWide; }

Surrogate* create(int device, hipError_t* err) {
    Surrogate* s = new Surrogate();
    s->device = device;
    auto fail = [&](hipError_t e) {
        if (err) *err = e;
        free_all(s);
        delete s;
        return (Surrogate*)nullptr;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return fail(e);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) s->cu_count = prop.multiProcessorCount;
    // every buffer sized for the wider network, so that either fits
    const long long n = ktn::kMaxParams;
    if ((e = dalloc(&s->P, n)) != hipSuccess || (e = dalloc(&s->G, n)) != hipSuccess ||
        (e = dalloc(&s->M, n)) != hipSuccess || (e = dalloc(&s->V, n)) != hipSuccess)
        return fail(e);
    for (int l = 0; l < ktn::kLayers; ++l) {
        const int H = ktn::kMaxHidden;
        const size_t nb = (size_t)ktn::kpad(H, l) * ktn::nld(H, l), nt = (size_t)ktn::nout(H, l) * ktn::kpad(H, l);
        if ((e = dalloc(&s->wb[l], nb)) != hipSuccess || (e = dalloc(&s->wt[l], nt)) != hipSuccess) return fail(e);
        if ((e = hipMemset(s->wb[l], 0, nb * 2)) != hipSuccess || (e = hipMemset(s->wt[l], 0, nt * 2)) != hipSuccess)
            return fail(e);
        s->lt.wb[l] = s->wb[l];
        s->lt.wt[l] = s->wt[l];
    }
    set_layers(s);
    if ((e = dalloc(&s->wpk, kPackedElems)) != hipSuccess || (e = dalloc(&s->fold, kFoldElems)) != hipSuccess ||
        (e = dalloc(&s->wpk0, kInputFragElems)) != hipSuccess || (e = dalloc(&s->wfast, (size_t)kFastFrags * 8)) !=
        hipSuccess || (e = dalloc(&s->bfast, kFastBias)) != hipSuccess)
        return fail(e);
    if ((e = build_table(s)) != hipSuccess) return fail(e);
    if ((e = dalloc(&s->loss, 1)) != hipSuccess) return fail(e);
    if (err) *err = hipSuccess;
    return s;
}

void destroy(Surrogate* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    free_all(s);
    delete s;
}

bool ready(const Surrogate* s) { return s && s->have_params; }

void view_constants(const Surrogate* s, double a, double sin_i, double cos_i, double r_cam, float c[6]) {
    const float af = (float)a, ci = (float)cos_i, inc = atan2f((float)sin_i, ci);
    const float R = table_rsh(s->h_rsh.data(), (double)af, (double)inc);
    c[0] = log10f((float)r_cam) - 2.0f;
    c[1] = (ci - 0.5415f) * (1.0f / 0.4545f);
    c[2] = (af - 0.499f) * (1.0f / 0.499f);
    c[3] = 1.0f / (1.5f * R);
    c[4] = inc;
    c[5] = 0.0f;
}

Domain get_domain(const Surrogate* s) { return s ? s->dom : Domain(); }

void set_domain(Surrogate* s, const Domain& d) {
    if (s) s->dom = d;
}

double table_critical_radius(const Surrogate* s, double a, double inc, double psi) {
    return table_rho(s->h_rho.data(), a, inc, psi);
}
double table_shadow_radius(const Surrogate* s, double a, double inc) { return table_rsh(s->h_rsh.data(), a, inc); }

static hipError_t pack(Surrogate* s, hipStream_t st) {
    if (s->net == ktn::kNetFast) {
        hipLaunchKernelGGL(k_pack_fast, dim3((kFastFrags * 8 + 255) / 256), dim3(256), 0, st, s->P, s->wfast,
                           s->bfast);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_fold, dim3((kFoldElems + 255) / 256), dim3(256), 0, st, s->P, s->fold);
    hipLaunchKernelGGL(k_pack_input, dim3(kInputFragElems / 256), dim3(256), 0, st, s->fold, s->wpk0);
    hipLaunchKernelGGL(k_pack_fused, dim3((kPackedElems + 255) / 256), dim3(256), 0, st, s->P, s->wpk);
    return hipGetLastError();
}

hipError_t set_params(Surrogate* s, const float* host, long long n) {
    if (n != ktn::param_count(s->hid)) return hipErrorInvalidValue;
    KTS_TRY(hipSetDevice(s->device));
    KTS_TRY(hipMemcpy(s->P, host, sizeof(float) * n, hipMemcpyHostToDevice));
    KTS_TRY(hipMemset(s->M, 0, sizeof(float) * n));
    KTS_TRY(hipMemset(s->V, 0, sizeof(float) * n));
    KTS_TRY(ktn::refresh_bf16(s->P, n, s->lt, nullptr));
    KTS_TRY(pack(s, nullptr));
    KTS_TRY(hipDeviceSynchronize());
    s->have_params = true;
    return hipSuccess;
}

hipError_t get_params(Surrogate* s, float* host, long long n) {
    if (n != ktn::param_count(s->hid)) return hipErrorInvalidValue;
    KTS_TRY(hipSetDevice(s->device));
    return hipMemcpy(host, s->P, sizeof(float) * n, hipMemcpyDeviceToHost);
}

hipError_t init_params(Surrogate* s, uint64_t seed) {
    const int H = s->hid;
    std::vector<float> p(ktn::param_count(H), 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < ktn::kLayers; ++l) {
        const double lim = std::sqrt(6.0 / ktn::kreal(H, l));
        std::uniform_real_distribution<double> u(-lim, lim);
        const long long off = ktn::param_offset(H, l);
        const long long nw = (long long)ktn::kreal(H, l) * ktn::nout(H, l);
        for (long long i = 0; i < nw; ++i) p[off + i] = (float)u(rng);
        // biases start at 0; the output layer starts small
        if (l == ktn::kLayers - 1)
            for (long long i = 0; i < nw; ++i) p[off + i] *= 0.1f;
    }
    return set_params(s, p.data(), (long long)p.size());
}

static hipError_t ensure_train_buffers(Surrogate* s, long long B) {
    if (B <= s->cap) return hipSuccess;
    auto f = [](void* p) {
        if (p) (void)hipFree(p);
    };
    f(s->xf);
    f(s->partial0);
    for (int l = 0; l < 4; ++l) {
        f(s->Z[l]);
        f(s->H[l]);
    }
    f(s->Y);
    f(s->gy32);
    f(s->gyb);
    for (int k = 0; k < 2; ++k) {
        f(s->g32[k]);
        f(s->gb[k]);
    }
    f(s->bcls);
    f(s->btgt);
    f(s->bidx);
    f(s->ws);
    f(s->partial);
    f(s->colpart);
    s->cap = 0;
    KTS_TRY(dalloc(&s->xf, (size_t)B * 5));
    KTS_TRY(dalloc(&s->partial0, (size_t)Surrogate::kChunks * 6 * 256));
    for (int l = 0; l < 4; ++l) {
        KTS_TRY(dalloc(&s->Z[l], (size_t)B * 256));
        KTS_TRY(dalloc(&s->H[l], (size_t)B * 256));
    }
    KTS_TRY(dalloc(&s->Y, (size_t)B * 8));
    KTS_TRY(dalloc(&s->gy32, (size_t)B * 8));
    KTS_TRY(dalloc(&s->gyb, (size_t)B * 32));
    for (int k = 0; k < 2; ++k) {
        KTS_TRY(dalloc(&s->g32[k], (size_t)B * 256));
        KTS_TRY(dalloc(&s->gb[k], (size_t)B * 256));
    }
    KTS_TRY(dalloc(&s->bcls, (size_t)B));
    KTS_TRY(dalloc(&s->btgt, (size_t)B * 5));
    KTS_TRY(dalloc(&s->bidx, (size_t)B));
    KTS_TRY(dalloc(&s->ws, (size_t)Surrogate::kSplits * 256 * 256));
    KTS_TRY(dalloc(&s->partial, (size_t)Surrogate::kChunks * 256));
    KTS_TRY(dalloc(&s->colpart, (size_t)((B + 127) / 128) * 2 * 256));
    s->cap = B;
    return hipSuccess;
}

// forward over B rows of inputs xf (B x 5) into Z / H / Y (z stored when keep_z)
static hipError_t forward(Surrogate* s, const float* xf, uint16_t** H, float** Z, float* Y, int B, bool keep_z,
                          hipStream_t st) {
    const int W = s->hid;
    KTS_TRY(ktn::layer0(xf, B, s->P, W, s->act, keep_z ? Z[0] : nullptr, H[0], st));
    for (int l = 1; l < 4; ++l) {
        ktn::GemmArgs g{};
        g.M = B;
        g.N = W;
        g.K = ktn::kpad(W, l);
        g.A = H[l - 1];
        g.lda = ktn::kpad(W, l);
        g.B = s->wt[l];
        g.ldb = ktn::kpad(W, l);
        g.ldc = W;
        g.bias = s->P + ktn::param_offset(W, l) + (long long)ktn::kreal(W, l) * ktn::nout(W, l);
        g.out32 = keep_z ? Z[l] : nullptr;
        g.outb = H[l];
        g.k_split = g.K;
        g.act = s->act;
        KTS_TRY(ktn::gemm(g, false, true, ktn::kEpiFwd, 1, st));
    }
    ktn::GemmArgs g{};
    g.M = B;
    g.N = 8;
    g.K = W;
    g.A = H[3];
    g.lda = W;
    g.B = s->wt[4];
    g.ldb = W;
    g.ldc = 8;
    g.bias = s->P + ktn::param_offset(W, 4) + (long long)W * 8;
    g.out32 = Y;
    g.k_split = W;
    return ktn::gemm(g, false, true, ktn::kEpiLin, 1, st);
}

// backward of the mean loss; gradients into s->G
static hipError_t backward(Surrogate* s, int B, hipStream_t st) {
    const int W = s->hid;
    const int splits = std::max(1, std::min(Surrogate::kSplits, B / 512));
    const int ksplit = ((B / splits + 31) / 32) * 32;
    auto dW = [&](const uint16_t* A, int lda, int M, const uint16_t* Bm, int ldb, int N, int l) -> hipError_t {
        ktn::GemmArgs g{};
        g.M = M;
        g.N = N;
        g.K = B;
        g.A = A;
        g.lda = lda;
        g.B = Bm;
        g.ldb = ldb;
        g.ldc = N;
        g.out32 = s->ws;
        g.k_split = ksplit;
        g.ws_stride = (long long)M * N;
        const int nsplit = (B + ksplit - 1) / ksplit;
        KTS_TRY(ktn::gemm(g, true, false, ktn::kEpiWs, nsplit, st));
        return ktn::reduce_ws(s->ws, nsplit, (long long)M * N, ktn::kreal(W, l), N, N,
                              s->G + ktn::param_offset(W, l), st);
    };
    auto db = [&](const float* Gm, int N, int l) -> hipError_t {
        return ktn::colsum(Gm, B, N, s->partial, Surrogate::kChunks,
                           s->G + ktn::param_offset(W, l) + (long long)ktn::kreal(W, l) * ktn::nout(W, l), st);
    };
    // the bias gradients of hidden layers 3..1 come from the column sums the Bwd GEMM that
    // forms g leaves in colpart; layer 0 takes g in FP32 instead
    const int parts = 2 * ((B + 127) / 128);
    auto bwd = [&](const uint16_t* A, int lda, int K, const uint16_t* Wt, int ldw, const float* Z, int dst_l,
                   int out) -> hipError_t {
        ktn::GemmArgs g{};
        g.M = B;
        g.N = W;
        g.K = K;
        g.A = A;
        g.lda = lda;
        g.B = Wt;
        g.ldb = ldw;
        g.ldc = W;
        g.zin = Z;
        g.ldz = W;
        g.out32 = dst_l == 0 ? s->g32[out] : nullptr;
        g.outb = s->gb[out];
        g.colpart = dst_l == 0 ? nullptr : s->colpart;
        g.k_split = K;
        g.act = s->act;
        KTS_TRY(ktn::gemm(g, false, true, ktn::kEpiBwd, 1, st));
        if (dst_l == 0) return hipSuccess;
        return ktn::sum_chunks(s->colpart, parts, W,
                               s->G + ktn::param_offset(W, dst_l) + (long long)ktn::kreal(W, dst_l) * W, st);
    };
    // output layer
    KTS_TRY(dW(s->H[3], W, W, s->gyb, 32, 8, 4));
    KTS_TRY(db(s->gy32, 8, 4));
    KTS_TRY(bwd(s->gyb, 32, 32, s->wb[4], 32, s->Z[3], 3, 1));
    int cur = 1;
    for (int l = 3; l >= 1; --l) {
        KTS_TRY(dW(s->H[l - 1], W, W, s->gb[cur], W, W, l));
        KTS_TRY(bwd(s->gb[cur], W, W, s->wb[l], W, s->Z[l - 1], l - 1, cur ^ 1));
        cur ^= 1;
    }
    // input layer in FP32 (its weights and bias are the first 6 x W parameters)
    return ktn::layer0_grad(s->xf, s->g32[cur], B, W, s->partial0, Surrogate::kChunks, s->G, st);
}

static FusedArgs fused_args(Surrogate* s, long long n) {
    FusedArgs a{};
    a.n = n;
    a.t = Table{s->tab_rho, s->tab_rsh};
    a.P = s->P;
    a.fold = s->fold;
    a.wpk0 = s->wpk0;
    a.wpk = s->wpk;
    a.wfast = reinterpret_cast<const uint4*>(s->wfast);
    a.bfast = s->bfast;
    return a;
}

// the network over n rays: k_mlp_fast (one block per CU, over tiles) or k_mlp_fused
static void launch_network(Surrogate* s, const FusedArgs& a, bool views, hipStream_t st) {
    if (s->net == ktn::kNetFast) {
        const long long tiles = (a.n + 2 * kFastTile - 1) / (2 * kFastTile);
        const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(tiles, std::max(1, s->cu_count)));
        if (views) hipLaunchKernelGGL(k_mlp_fast<true>, dim3(grid), dim3(kFastThreads), 0, st, a);
        else hipLaunchKernelGGL(k_mlp_fast<false>, dim3(grid), dim3(kFastThreads), 0, st, a);
        return;
    }
    const dim3 grid((unsigned)((a.n + kFusedRays - 1) / kFusedRays));
    if (views) hipLaunchKernelGGL(k_mlp_fused<true>, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_mlp_fused<false>, grid, dim3(256), 0, st, a);
}

hipError_t predict(Surrogate* s, const float* feat, long long n, float* out, hipStream_t st) {
    KTS_TRY(hipSetDevice(s->device));
    if (n <= 0) return hipSuccess;
    FusedArgs a = fused_args(s, n);
    a.xf = feat;
    a.y = out;
    launch_network(s, a, false, st);
    return hipGetLastError();
}

hipError_t gradients(Surrogate* s, const float* feat, const int8_t* cls, const float* tgt, int n, float* grads,
                     double* loss, hipStream_t st) {
    KTS_TRY(hipSetDevice(s->device));
    KTS_TRY(ensure_train_buffers(s, n));
    KTS_TRY(ktn::gather(feat, cls, tgt, nullptr, n, s->xf, s->bcls, s->btgt, st));
    KTS_TRY(forward(s, s->xf, s->H, s->Z, s->Y, n, true, st));
    KTS_TRY(hipMemsetAsync(s->loss, 0, sizeof(double), st));
    KTS_TRY(ktn::loss_head(s->Y, s->bcls, s->btgt, n, 1.0f, 1.0f, s->gy32, s->gyb, s->loss, st));
    KTS_TRY(backward(s, n, st));
    if (grads)
        KTS_TRY(hipMemcpyAsync(grads, s->G, sizeof(float) * ktn::param_count(s->hid), hipMemcpyDeviceToDevice, st));
    if (loss) {
        KTS_TRY(hipMemcpyAsync(loss, s->loss, sizeof(double), hipMemcpyDeviceToHost, st));
        KTS_TRY(hipStreamSynchronize(st));
        *loss /= n;
    }
    return hipSuccess;
}

static hipError_t adam_update(Surrogate* s, const float* grads, double lr, int t, hipStream_t st) {
    const double b1 = 0.9, b2 = 0.999;
    const double c1 = 1.0 / (1.0 - std::pow(b1, t)), c2 = 1.0 / (1.0 - std::pow(b2, t));
    return ktn::adam(s->P, grads ? grads : s->G, s->M, s->V, ktn::param_count(s->hid), (float)lr, (float)b1,
                     (float)b2, 1e-8f, (float)c1, (float)c2, s->lt, st);
}

hipError_t adam_step(Surrogate* s, const float* grads, double lr, int t, hipStream_t st) {
    KTS_TRY(adam_update(s, grads, lr, t, st));
    return pack(s, st);
}

// ---- data generation ----------------------------------------------------------------------

struct Dataset {
    long long n = 0;
    double* params = nullptr;  // 5 per ray
    float* feat = nullptr;
    int8_t* cls = nullptr;
    float* tgt = nullptr;
    float* ring = nullptr;
    // integrator outputs (kept for the held-out judgement)
    int8_t* ocls = nullptr;
    double *th = nullptr, *ph = nullptr, *rh = nullptr, *gg = nullptr;
    unsigned long long* counter = nullptr;
    ViewDev* view = nullptr;
    void release() {
        for (void* p : {(void*)params, (void*)feat, (void*)cls, (void*)tgt, (void*)ring, (void*)ocls, (void*)th,
                        (void*)ph, (void*)rh, (void*)gg, (void*)counter, (void*)view})
            if (p) (void)hipFree(p);
        *this = Dataset();
    }
};

// seeded uniform draws over the batch-view ranges, with the band around the critical
// curve `band_factor` times denser (by thinning the rest)
static void draw_rays(const Surrogate* s, long long n, uint64_t seed, double band, double factor,
                      std::vector<double>& p) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    p.resize((size_t)n * 5);
    long long k = 0;
    const double keep_out = factor > 1.0 ? 1.0 / factor : 1.0;
    while (k < n) {
        const double a = 0.998 * U(rng);
        const double inc = (5.0 + 80.0 * U(rng)) * kPi / 180.0;
        const double rc = std::pow(10.0, 1.0 + 2.0 * U(rng));
        const double R = table_rsh(s->h_rsh.data(), a, inc);
        const double al = (3.0 * U(rng) - 1.5) * R, be = (3.0 * U(rng) - 1.5) * R;
        const double rho_c = table_rho(s->h_rho.data(), a, inc, std::atan2(std::fabs(be), al));
        const bool in_band = std::fabs(std::sqrt(al * al + be * be) / rho_c - 1.0) < band;
        if (!in_band && U(rng) >= keep_out) continue;
        double* q = &p[(size_t)k * 5];
        q[0] = a;
        q[1] = rc;
        q[2] = inc;
        q[3] = al;
        q[4] = be;
        ++k;
    }
}

static hipError_t label(Surrogate* s, Dataset& d, const std::vector<double>& hp, int precision, const TrainOpts& o) {
    const long long n = (long long)hp.size() / 5;
    d.n = n;
    KTS_TRY(dalloc(&d.params, (size_t)n * 5));
    KTS_TRY(dalloc(&d.feat, (size_t)n * 5));
    KTS_TRY(dalloc(&d.cls, (size_t)n));
    KTS_TRY(dalloc(&d.tgt, (size_t)n * 5));
    KTS_TRY(dalloc(&d.ring, (size_t)n));
    KTS_TRY(dalloc(&d.ocls, (size_t)n));
    KTS_TRY(dalloc(&d.th, (size_t)n));
    KTS_TRY(dalloc(&d.ph, (size_t)n));
    KTS_TRY(dalloc(&d.rh, (size_t)n));
    KTS_TRY(dalloc(&d.gg, (size_t)n));
    KTS_TRY(dalloc(&d.counter, 1 + kti::kStatCount));
    KTS_TRY(dalloc(&d.view, 1));
    KTS_TRY(hipMemcpy(d.params, hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice));
    ViewDev v;
    std::memset(&v, 0, sizeof v);
    v.kind = 2;
    v.width = (int)std::min<long long>(n, 0x7fffffff);
    v.height = 1;
    v.disk = 1;
    v.prograde = 1;
    v.r_cap = o.capture_factor;  // kind 2: the capture factor; r+ is per ray
    v.r_out = o.disk_outer;
    v.escape_radius_ = o.escape_radius;
    v.params = d.params;
    KTS_TRY(hipMemcpy(d.view, &v, sizeof v, hipMemcpyHostToDevice));
    kti::TraceLaunch L{};
    L.views = d.view;
    L.n_views = 1;
    L.total = n;
    L.counter = d.counter;
    L.stats = d.counter + 1;
    L.out = ktd::Outputs{d.ocls, d.th, d.ph, d.rh, d.gg, nullptr};
    L.rtol = o.rtol;
    L.atol = o.atol;
    L.rtol32 = o.rtol32;
    L.atol32 = o.atol32;
    L.restep_drift = o.restep_drift;
    L.max_steps = 200000;
    L.precision = precision;
    L.mode = KT_MODE_ONESHOT;
    L.block = 256;
    KTS_TRY(kti::launch_trace(L, nullptr));
    Table t{s->tab_rho, s->tab_rsh};
    hipLaunchKernelGGL(k_targets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, d.params, n, t, d.ocls,
                       d.th, d.ph, d.rh, d.gg, d.feat, d.cls, d.tgt, d.ring);
    KTS_TRY(hipGetLastError());
    return hipDeviceSynchronize();
}

// Rays that hit the disk just outside its inner edge, where the disk's image meets the
// hole's shadow away from the critical curve: the class changes there over a thin band
// the uniform draw samples sparsely, and a network trained on it is confidently wrong
// along that edge.  Every training ray that hits within `band` of the inner edge gets
// `copies` neighbours, moved by up to +-0.005 mean shadow radii in alpha and beta, to be
// labelled by the integrator like the rest.
static hipError_t edge_rays(const Surrogate* s, const Dataset& d, const std::vector<double>& hp, double band,
                            int copies, uint64_t seed, std::vector<double>& out) {
    out.clear();
    if (copies <= 0 || !(band > 0.0)) return hipSuccess;
    const long long n = d.n;
    std::vector<int8_t> cls((size_t)n);
    std::vector<double> rh((size_t)n);
    KTS_TRY(hipMemcpy(cls.data(), d.ocls, (size_t)n, hipMemcpyDeviceToHost));
    KTS_TRY(hipMemcpy(rh.data(), d.rh, (size_t)n * 8, hipMemcpyDeviceToHost));
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (long long q = 0; q < n; ++q) {
        if ((cls[q] & 7) != 2) continue;
        const double* p = &hp[(size_t)q * 5];
        if (!(rh[q] < (1.0 + band) * kt_isco_radius(p[0], 1))) continue;
        const double R = table_rsh(s->h_rsh.data(), p[0], p[2]);
        for (int c = 0; c < copies; ++c) {
            out.insert(out.end(), p, p + 5);
            out[out.size() - 2] += 0.005 * R * U(rng);
            out[out.size() - 1] += 0.005 * R * U(rng);
        }
    }
    return hipSuccess;
}

// appends b's training inputs, classes and targets to a's (b is released)
static hipError_t append(Dataset& a, Dataset& b) {
    const long long n = a.n + b.n;
    float *feat = nullptr, *tgt = nullptr;
    int8_t* cls = nullptr;
    hipError_t e = dalloc(&feat, (size_t)n * 5);
    if (e == hipSuccess) e = dalloc(&tgt, (size_t)n * 5);
    if (e == hipSuccess) e = dalloc(&cls, (size_t)n);
    if (e == hipSuccess) e = hipMemcpy(feat, a.feat, (size_t)a.n * 5 * 4, hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(feat + a.n * 5, b.feat, (size_t)b.n * 5 * 4, hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(tgt, a.tgt, (size_t)a.n * 5 * 4, hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(tgt + a.n * 5, b.tgt, (size_t)b.n * 5 * 4, hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(cls, a.cls, (size_t)a.n, hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(cls + a.n, b.cls, (size_t)b.n, hipMemcpyDeviceToDevice);
    if (e != hipSuccess) {
        for (void* p : {(void*)feat, (void*)tgt, (void*)cls})
            if (p) (void)hipFree(p);
        return e;
    }
    (void)hipFree(a.feat);
    (void)hipFree(a.tgt);
    (void)hipFree(a.cls);
    a.feat = feat;
    a.tgt = tgt;
    a.cls = cls;
    a.n = n;
    b.release();
    return hipSuccess;
}

static double quantile(std::vector<double> v, double q) {
    if (v.empty()) return std::nan("");
    const size_t k = std::min(v.size() - 1, (size_t)(q * (v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + k, v.end());
    return v[k];
}

// held-out judgement: predictions against the FP64 integrator's outputs
static hipError_t evaluate(Surrogate* s, Dataset& d, const TrainOpts& o, TrainReport& rep) {
    const long long n = d.n;
    float* y = nullptr;
    KTS_TRY(dalloc(&y, (size_t)n * 8));
    hipError_t e = predict(s, d.feat, n, y, nullptr);
    std::vector<float> hy((size_t)n * 8), ring((size_t)n), tgt((size_t)n * 5);
    std::vector<int8_t> cls((size_t)n);
    std::vector<double> hp((size_t)n * 5), th((size_t)n), ph((size_t)n), rh((size_t)n), gg((size_t)n);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(hy.data(), y, hy.size() * 4, hipMemcpyDeviceToHost);
    (void)hipFree(y);
    KTS_TRY(e);
    KTS_TRY(hipMemcpy(ring.data(), d.ring, ring.size() * 4, hipMemcpyDeviceToHost));
    KTS_TRY(hipMemcpy(cls.data(), d.cls, cls.size(), hipMemcpyDeviceToHost));
    KTS_TRY(hipMemcpy(th.data(), d.th, th.size() * 8, hipMemcpyDeviceToHost));
    KTS_TRY(hipMemcpy(ph.data(), d.ph, ph.size() * 8, hipMemcpyDeviceToHost));
    KTS_TRY(hipMemcpy(rh.data(), d.rh, rh.size() * 8, hipMemcpyDeviceToHost));
    KTS_TRY(hipMemcpy(gg.data(), d.gg, gg.size() * 8, hipMemcpyDeviceToHost));
    KTS_TRY(hipMemcpy(hp.data(), d.params, hp.size() * 8, hipMemcpyDeviceToHost));
    long long agree = 0, gated = 0, gated_agree = 0;
    std::vector<double> defl, gre, rre;
    for (long long q = 0; q < n; ++q) {
        const float* yy = &hy[(size_t)q * 8];
        const float m = std::max(yy[0], std::max(yy[1], yy[2]));
        const double e0 = std::exp(yy[0] - m), e1 = std::exp(yy[1] - m), e2 = std::exp(yy[2] - m);
        const double sum = e0 + e1 + e2;
        const int c = e0 >= e1 && e0 >= e2 ? 0 : (e1 >= e2 ? 1 : 2);
        const double pm = std::max(e0, std::max(e1, e2)) / sum;
        agree += c == cls[q];
        if (!(pm > 0.999) || !(ring[q] > 0.01)) continue;
        ++gated;
        gated_agree += c == cls[q];
        if (c != cls[q]) continue;
        if (c == 0) {
            const double* pq = &hp[(size_t)q * 5];
            double e1[3], e2[3], e3[3], tt, pt;
            ray_frame(pq[1], pq[2], 0.0, pq[3], pq[4], e1, e2, e3);
            sky_from_frame(e1, e2, e3, yy[3], yy[4], yy[5], tt, pt);
            const double v1[3] = {std::sin(tt) * std::cos(pt), std::sin(tt) * std::sin(pt), std::cos(tt)};
            const double v2[3] = {std::sin(th[q]) * std::cos(ph[q]), std::sin(th[q]) * std::sin(ph[q]), std::cos(th[q])};
            const double dx = v1[0] - v2[0], dy = v1[1] - v2[1], dz = v1[2] - v2[2];
            defl.push_back(2.0 * std::asin(std::min(1.0, std::sqrt(dx * dx + dy * dy + dz * dz) / 2.0)));
        } else if (c == 2) {
            rre.push_back(std::fabs(std::exp((double)yy[6]) / rh[q] - 1.0));
            gre.push_back(std::fabs(std::exp((double)yy[7]) / gg[q] - 1.0));
        }
    }
    rep.class_agreement = (double)agree / n;
    rep.gated_fraction = (double)gated / n;
    rep.gated_class_agreement = gated ? (double)gated_agree / gated : 0.0;
    rep.defl_median = quantile(defl, 0.5);
    rep.defl_p99 = quantile(defl, 0.99);
    rep.defl_p999 = quantile(defl, 0.999);
    rep.g_rel_median = quantile(gre, 0.5);
    rep.g_rel_p99 = quantile(gre, 0.99);
    rep.rhit_rel_median = quantile(rre, 0.5);
    rep.rhit_rel_p99 = quantile(rre, 0.99);
    (void)o;
    return hipSuccess;
}

hipError_t train(Surrogate* s, const TrainOpts& o, TrainReport& rep) {
    using clk = std::chrono::steady_clock;
    KTS_TRY(hipSetDevice(s->device));
    if (o.net != s->net) KTS_TRY(set_net(s, o.net));
    if (!s->have_params) KTS_TRY(init_params(s, o.seed ^ 0x1234567ull));
    // the ranges draw_rays samples and the scene the labels come from
    s->dom = Domain();
    s->dom.disk_outer = o.disk_outer;
    s->dom.escape_radius = o.escape_radius;
    s->dom.capture_factor = o.capture_factor;
    // data: the training set labelled by the mixed integrator, the held-out set by FP64
    auto t0 = clk::now();
    Dataset tr, ho;
    std::vector<double> hp;
    draw_rays(s, o.n_train, o.seed, o.band, o.band_factor, hp);
    hipError_t e = label(s, tr, hp, o.label_precision, o);
    if (e == hipSuccess && o.edge_copies > 0) {  // more rays along the disk's inner edge
        std::vector<double> ep;
        e = edge_rays(s, tr, hp, o.edge_band, o.edge_copies, o.seed ^ 0xED6Eull, ep);
        if (e == hipSuccess && !ep.empty()) {
            Dataset ed;
            e = label(s, ed, ep, o.label_precision, o);
            if (e == hipSuccess) e = append(tr, ed);
            ed.release();
        }
    }
    if (e == hipSuccess) {
        draw_rays(s, o.n_holdout, o.seed ^ 0xABCDEFull, o.band, 1.0, hp);
        e = label(s, ho, hp, KT_PREC_FP64, o);
    }
    if (e != hipSuccess) {
        tr.release();
        ho.release();
        return e;
    }
    rep.datagen_seconds = std::chrono::duration<double>(clk::now() - t0).count();
    if (o.verbose) {
        std::fprintf(stderr, "surrogate: labelled %lld training + %lld held-out rays in %.1f s\n", tr.n, ho.n,
                     rep.datagen_seconds);
        std::fflush(stderr);
    }

    // training
    const int B = (int)std::min<long long>(o.batch, tr.n);
    e = ensure_train_buffers(s, B);
    const long long steps_per_epoch = std::max<long long>(1, tr.n / B);
    const long long total_steps = steps_per_epoch * o.epochs;
    int bits = 2;
    while ((1ll << bits) < tr.n) bits += 2;
    std::mt19937_64 rng(o.seed ^ 0x77ull);
    auto t1 = clk::now();
    long long t = 0;
    double loss_sum = 0.0;
    for (int ep = 0; ep < o.epochs && e == hipSuccess; ++ep) {
        const uint64_t key = rng();
        double ep_loss = 0.0;
        KTS_TRY(hipMemsetAsync(s->loss, 0, sizeof(double), nullptr));
        for (long long k = 0; k < steps_per_epoch && e == hipSuccess; ++k) {
            hipLaunchKernelGGL(k_batch_index, dim3((B + 255) / 256), dim3(256), 0, nullptr, tr.n, bits, key,
                               k * (long long)B, B, s->bidx);
            e = ktn::gather(tr.feat, tr.cls, tr.tgt, s->bidx, B, s->xf, s->bcls, s->btgt, nullptr);
            if (e == hipSuccess) e = forward(s, s->xf, s->H, s->Z, s->Y, B, true, nullptr);
            if (e == hipSuccess)
                e = ktn::loss_head(s->Y, s->bcls, s->btgt, B, (float)o.w_sky, (float)o.w_disk, s->gy32, s->gyb,
                                   s->loss, nullptr);
            if (e == hipSuccess) e = backward(s, B, nullptr);
            ++t;
            const double lr = 0.5 * o.lr * (1.0 + std::cos(kPi * (double)(t - 1) / (double)total_steps));
            if (e == hipSuccess) e = adam_update(s, nullptr, lr, (int)t, nullptr);
        }
        if (e == hipSuccess) e = hipMemcpy(&ep_loss, s->loss, sizeof(double), hipMemcpyDeviceToHost);
        ep_loss /= (double)steps_per_epoch * B;
        if (ep == 0) rep.first_loss = ep_loss;
        loss_sum = ep_loss;
        if (o.verbose && e == hipSuccess) {
            std::fprintf(stderr, "surrogate: epoch %d/%d loss %.6f (%.1f s)\n", ep + 1, o.epochs, ep_loss,
                         std::chrono::duration<double>(clk::now() - t1).count());
            std::fflush(stderr);
        }
    }
    if (e == hipSuccess) e = pack(s, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    rep.train_seconds = std::chrono::duration<double>(clk::now() - t1).count();
    rep.final_loss = loss_sum;
    rep.steps = t;
    // BF16 MFMA work per step: forward (2 B sum K N) + backward (2x that)
    double macs = 0;
    for (int l = 0; l < ktn::kLayers; ++l) macs += (double)ktn::kpad(s->hid, l) * ktn::nout(s->hid, l);
    rep.train_gemm_tflops = rep.train_seconds > 0 ? 6.0 * macs * B * (double)t / rep.train_seconds / 1e12 : 0.0;
    auto t2 = clk::now();
    if (e == hipSuccess) e = evaluate(s, ho, o, rep);
    rep.eval_seconds = std::chrono::duration<double>(clk::now() - t2).count();
    tr.release();
    ho.release();
    return e;
}

// ---- traced inference ------------------------------------------------------------------------

hipError_t queue_reserve(Queue& q, long long n) {
    if (!q.count) {
        KTS_TRY(dalloc(&q.count, kti::kQueueCounts));
        KTS_TRY(dalloc(&q.total, 1));
    }
    if (n <= q.cap) return hipSuccess;
    if (q.list) (void)hipFree(q.list);
    q.list = nullptr;
    q.cap = 0;
    KTS_TRY(dalloc(&q.list, (size_t)n * (kti::kQueueBuckets + 1)));
    q.cap = n;
    return hipSuccess;
}

void queue_release(Queue& q) {
    for (void* p : {(void*)q.list, (void*)q.count, (void*)q.total})
        if (p) (void)hipFree(p);
    q = Queue();
}

hipError_t trace_with_surrogate(Surrogate* s, const kti::TraceLaunch& L, double gate_prob, double gate_ring,
                                int closed_form, Queue* qu, hipStream_t st, hipEvent_t* ev) {
    if (!s->have_params) return hipErrorNotReady;
    const long long n = L.total;
    Queue& Q = qu ? *qu : s->queue;
    if (n > Q.cap || !Q.count) {
        if (!ev) return hipErrorInvalidValue;  // under capture: the queue must exist already
        KTS_TRY(queue_reserve(Q, n));
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, Q.count, kti::kQueueCounts);
    KTS_TRY(kti::reset_counters(L, st));
    // the network and the gate in one kernel: answered rays get their outputs, the
    // others join the fallback queue, whose length the integrator reads on the device:
    // the two kernels run back to back, bracketed by the events the stats read
    if (ev) KTS_TRY(hipEventRecord(ev[0], st));
    FusedArgs a = fused_args(s, n);
    a.views = L.views;
    a.n_views = L.n_views;
    a.gate_prob = (float)gate_prob;
    a.gate_ring = (float)gate_ring;
    a.verify = closed_form;
    a.extent = (float)(std::min(s->dom.extent, 1.5) / 1.5);
    a.out = L.out;
    a.fallback = Q.list;
    a.n_fallback = Q.count;
    a.fallback_stride = Q.cap;
    a.next_tile = Q.count + kti::kQueueBuckets;  // zeroed above with the fallback counts
    launch_network(s, a, true, st);
    KTS_TRY(hipGetLastError());
    // the closed form's values for the network's escaped and disk answers: the network
    // picks the class and the rays the gate trusts, the elliptic integrals give the numbers
    if (closed_form)
        KTS_TRY(kti::launch_closed_form_polish(L, kti::RayQueue{Q.list, Q.count, Q.total, Q.cap}, closed_form == 2,
                                               st));
    // the integrator on the queue the gate left, its regions read in order and its length
    // read on the device (the counters were zeroed before the network).  The rays queued
    // are the slowest of the frame (photon ring, class boundaries), so the launch is bound
    // by their latency: the one-shot form's grid is capped near the usual queue length and
    // strides over a longer one; the persistent form measured slower on the 1080p orbit
    // (its periodic refill costs, and the few rays leave little to refill)
    kti::TraceLaunch F = L;
    F.ray_list = Q.list;
    F.total_dev = nullptr;
    F.bucket_count = Q.count;
    F.bucket_stride = Q.cap;
    F.no_reset = 1;
    F.mode = 1;
    if (ev) KTS_TRY(hipEventRecord(ev[1], st));
    KTS_TRY(kti::launch_trace(F, st));
    if (ev) KTS_TRY(hipEventRecord(ev[2], st));
    // the fallback count into the context's counters (after launch_trace zeroed them)
    hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(64), 0, st, Q.count, Q.total, L.stats + kti::kStatFallback);
    return hipGetLastError();
}

}  // namespace kts
