// kt_nn.hpp -- interface of the surrogate network's kernels (kt_nn.hip).
//
// Two networks share the kernels: the 5 -> 256 -> 256 -> 256 -> 256 -> 8 network with
// SiLU (the default), and a 5 -> 128 -> 128 -> 128 -> 128 -> 8 network with hard SiLU,
// x clamp(x / 6 + 1/2, 0, 1), two full-rate vector instructions per activation instead of
// SiLU's two transcendentals.  Layer l maps x (rows) to x W_l + b_l, W_l stored K_l x N_l
// row-major in FP32 master copies.  The input layer runs in FP32 on the vector units in
// training (its K = 5 is too narrow for the matrix cores and its inputs need more than
// bf16's 8 bits); layers 1-4 run on the matrix cores from bf16 working copies W_l
// (K_pad x N_ld) and W_l^T (N x K_pad), the output layer's N padded 8 -> 32 with zeros.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_mfma.hpp"

namespace ktn {

constexpr int kLayers = 5;
constexpr int kInputs = 5, kOutputs = 8;
constexpr int kMaxHidden = 256;

enum : int { kActSilu = 0, kActHard = 1 };
enum : int { kNetWide = 0, kNetFast = 1 };  // KT_NET_* of include/kerr_trace.h

struct Net {
    int hidden, act;
};
constexpr Net net_of(int id) { return id == kNetFast ? Net{128, kActHard} : Net{256, kActSilu}; }

constexpr int kreal(int H, int l) { return l == 0 ? kInputs : H; }
constexpr int kpad(int H, int l) { return l == 0 ? 32 : H; }
constexpr int nout(int H, int l) { return l == kLayers - 1 ? kOutputs : H; }
constexpr int nld(int H, int l) { return l == kLayers - 1 ? 32 : H; }  // row stride of the bf16 W copy

// parameters in FP32, layer by layer: W_l (kreal x nout row-major) then b_l
constexpr long long param_offset(int H, int l) {
    long long o = 0;
    for (int i = 0; i < l; ++i) o += (long long)kreal(H, i) * nout(H, i) + nout(H, i);
    return o;
}
constexpr long long param_count(int H) { return param_offset(H, kLayers); }
constexpr long long kMaxParams = param_count(kMaxHidden);
static_assert(param_count(256) == 200968, "5-256-256-256-256-8 has 200968 parameters");
static_assert(param_count(128) == 51336, "5-128-128-128-128-8 has 51336 parameters");

enum : int { kEpiFwd = 0, kEpiLin = 1, kEpiBwd = 2, kEpiWs = 3 };

struct GemmArgs {
    int M, N, K, lda, ldb, ldc;
    int act;             // Fwd, Bwd: kAct*
    const uint16_t* A;
    const uint16_t* B;
    const float* bias;   // Fwd, Lin
    float* out32;        // Fwd: z (optional); Lin: y; Bwd: g; Ws: split-K partials
    uint16_t* outb;      // Fwd: silu(z) in bf16; Bwd: g in bf16
    const float* zin;    // Bwd: the pre-activation whose SiLU derivative multiplies
    int ldz;
    int k_split;         // reduction length per blockIdx.z (a multiple of 32)
    long long ws_stride; // Ws: elements per split
    float* colpart;      // Bwd (optional): column sums of g, one row of N per 64 rows of C
};

struct LayerTable {
    long long off[kLayers];
    int kreal[kLayers], kpad[kLayers], n[kLayers], nld[kLayers];
    uint16_t* wb[kLayers];
    uint16_t* wt[kLayers];
};

// C = A B with A reduction-contiguous unless `at`, B given as B^T ([n][k]) when `bt`.  N and
// ldc are multiples of 4 (four consecutive columns leave one lane together).
hipError_t gemm(const GemmArgs& g, bool at, bool bt, int epi, int splits, hipStream_t s);
hipError_t reduce_ws(const float* ws, int splits, long long stride, int rows, int N, int ld, float* dst,
                     hipStream_t s);
hipError_t colsum(const float* G, int B, int N, float* partial, int chunks, float* dst, hipStream_t s);
// dst[j] = sum over c < chunks of partial[c * ncols + j], in a fixed order
hipError_t sum_chunks(const float* partial, int chunks, int ncols, float* dst, hipStream_t s);
hipError_t loss_head(const float* Y, const int8_t* cls, const float* tgt, int B, float w_sky, float w_disk,
                     float* gy32, uint16_t* gyb, double* loss_sum, hipStream_t s);
hipError_t adam(float* P, const float* G, float* M, float* V, long long n, float lr, float b1, float b2, float eps,
                float c1, float c2, const LayerTable& lt, hipStream_t s);
hipError_t refresh_bf16(const float* P, long long n, const LayerTable& lt, hipStream_t s);
hipError_t gather(const float* feat, const int8_t* cls, const float* tgt, const long long* idx, int B, float* xf,
                  int8_t* bcls, float* btgt, hipStream_t s);
// input layer (width H) in FP32 from the master weights (P at layer 0): z0 (optional) and
// the bf16 activation
hipError_t layer0(const float* xf, int B, const float* P, int H, int act, float* z0, uint16_t* h0, hipStream_t s);
// its gradients (W0, b0) in FP32 from the inputs and dL/dz0 (fp32 B x H); partial holds
// chunks x 6 x H floats
hipError_t layer0_grad(const float* xf, const float* g, int B, int H, float* partial, int chunks, float* G,
                       hipStream_t s);

}  // namespace ktn
