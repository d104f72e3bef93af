// kt_nn.hip -- the surrogate network's GPU kernels: BF16 MFMA GEMMs with fused
// epilogues, the loss head, reductions and Adam.  Host orchestration lives in
// kt_surrogate.hip.
//
// GEMM: C[M x N] = A[M x K] * B[K x N], bf16 in, fp32 accumulate.  A block of 256 threads
// (4 waves in 2 x 2) owns a 128 x 128 tile of C; each wave a 64 x 64 sub-tile = 4 x 4
// MFMA 16x16x32 tiles.  K advances 32 at a time through LDS; tiles whose source layout
// is not reduction-contiguous are transposed on the way into LDS.
#include "kt_nn.hpp"

namespace ktn {

using namespace ktm;

namespace {

constexpr int kLdsStride = 40;  // bf16 per LDS row: 32 + 8 pad (80 B, keeps 16-byte alignment)

// 128 rows x 32 k of a tile, from global memory into registers (fetch) and from there into
// LDS [row][k] (put).  RC = false: element (row, k) at P[row * ld + k] (reduction-
// contiguous), two 16-byte chunks per thread.  RC = true: at P[k * ld + row]; an 8 (k) x 8
// (row) block per thread of the first wave, eight 16-byte loads along the rows, transposed
// in registers (v_perm_b32 pairs the halves of two words) into eight 16-byte stores along
// k.  Rows >= R read as zero.
template <bool RC>
__device__ __forceinline__ void fetch(uint4 (&v)[8], const uint16_t* P, int ld, int r0, int k0, int R, int tid) {
    if (!RC) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int c = tid + 256 * i, row = c >> 2, kc = (c & 3) * 8;
            v[i] = uint4{0u, 0u, 0u, 0u};
            if (r0 + row < R) v[i] = *reinterpret_cast<const uint4*>(P + (long long)(r0 + row) * ld + k0 + kc);
        }
    } else if (tid < 64) {
        const int kb = (tid >> 4) * 8, rc = (tid & 15) * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            v[i] = uint4{0u, 0u, 0u, 0u};
            if (r0 + rc < R) v[i] = *reinterpret_cast<const uint4*>(P + (long long)(k0 + kb + i) * ld + r0 + rc);
        }
    }
}

template <bool RC>
__device__ __forceinline__ void put(const uint4 (&v)[8], uint16_t* T, int tid) {
    if (!RC) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int c = tid + 256 * i, row = c >> 2, kc = (c & 3) * 8;
            *reinterpret_cast<uint4*>(T + row * kLdsStride + kc) = v[i];
        }
    } else if (tid < 64) {
        const int kb = (tid >> 4) * 8, rc = (tid & 15) * 8;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            // row rc + j: element j of each of the eight k-rows
            const unsigned sel = (j & 1) ? 0x07060302u : 0x05040100u;  // high halves : low halves
            const int w = j >> 1;
            auto word = [w](const uint4& x) -> unsigned { return w == 0 ? x.x : w == 1 ? x.y : w == 2 ? x.z : x.w; };
            uint4 o;
            o.x = __builtin_amdgcn_perm(word(v[1]), word(v[0]), sel);
            o.y = __builtin_amdgcn_perm(word(v[3]), word(v[2]), sel);
            o.z = __builtin_amdgcn_perm(word(v[5]), word(v[4]), sel);
            o.w = __builtin_amdgcn_perm(word(v[7]), word(v[6]), sel);
            *reinterpret_cast<uint4*>(T + (rc + j) * kLdsStride + kb) = o;
        }
    }
}

// Double-buffered: the next k-slice's global loads are in flight while the matrix cores work
// on the current one from LDS; one barrier per slice.
template <bool AT, bool BT, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[2][128 * kLdsStride];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][128 * kLdsStride];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.x * 128, n0 = blockIdx.y * 128;
    const int kb = blockIdx.z * g.k_split;
    const int ke = g.K < kb + g.k_split ? g.K : kb + g.k_split;
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    uint4 va[8], vb[8];
    if (kb < ke) {
        fetch<AT>(va, g.A, g.lda, m0, kb, g.M, tid);
        fetch<!BT>(vb, g.B, g.ldb, n0, kb, g.N, tid);
        put<AT>(va, As[0], tid);
        put<!BT>(vb, Bs[0], tid);
    }
    __syncthreads();
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += 32, buf ^= 1) {
        const bool more = k0 + 32 < ke;
        if (more) {
            fetch<AT>(va, g.A, g.lda, m0, k0 + 32, g.M, tid);
            fetch<!BT>(vb, g.B, g.ldb, n0, k0 + 32, g.N, tid);
        }
        bf16x8 af[4], bfr[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            af[i] = as_bf16x8(*reinterpret_cast<const uint4*>(As[buf] + (64 * wm + 16 * i + fr) * kLdsStride + fk));
#pragma unroll
        for (int j = 0; j < 4; ++j)
            bfr[j] = as_bf16x8(*reinterpret_cast<const uint4*>(Bs[buf] + (64 * wn + 16 * j + fr) * kLdsStride + fk));
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = mfma16(bfr[j], af[i], acc[i][j]);  // C^T: columns in-lane
        if (more) {
            put<AT>(va, As[buf ^ 1], tid);
            put<!BT>(vb, Bs[buf ^ 1], tid);
        }
        __syncthreads();
    }
    // the tile was formed transposed, so lane l holds C[row0 + 16 i][col0 + 16 j + q], q < 4:
    // four consecutive columns, stored as one 16-byte (FP32) or 8-byte (BF16) access
    const int row0 = m0 + 64 * wm + fr, col0 = n0 + 64 * wn + (lane >> 4) * 4;
    f32x4 cs[4];  // Bwd: this lane's column sums of g over its 4 rows
#pragma unroll
    for (int j = 0; j < 4; ++j) cs[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = row0 + 16 * i;
        if (r >= g.M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = col0 + 16 * j;  // N is a multiple of 4: the four columns are in or out together
            if (c >= g.N) continue;
            const f32x4 v = acc[i][j];
            const long long o = (long long)r * g.ldc + c;
            if (EPI == kEpiFwd || EPI == kEpiLin) {
                f32x4 z;
#pragma unroll
                for (int q = 0; q < 4; ++q) z[q] = v[q] + g.bias[c + q];
                if (EPI == kEpiLin || g.out32) *reinterpret_cast<f32x4*>(g.out32 + o) = z;
                if (EPI == kEpiFwd) {
                    uint2 h;
                    h.x = (unsigned)f2bf(activate(z[0], g.act)) | ((unsigned)f2bf(activate(z[1], g.act)) << 16);
                    h.y = (unsigned)f2bf(activate(z[2], g.act)) | ((unsigned)f2bf(activate(z[3], g.act)) << 16);
                    *reinterpret_cast<uint2*>(g.outb + o) = h;
                }
            } else if (EPI == kEpiBwd) {
                const f32x4 z = *reinterpret_cast<const f32x4*>(g.zin + (long long)r * g.ldz + c);
                f32x4 gz;
#pragma unroll
                for (int q = 0; q < 4; ++q) gz[q] = v[q] * dactivate(z[q], g.act);
                if (g.out32) *reinterpret_cast<f32x4*>(g.out32 + o) = gz;
                uint2 h;
                h.x = (unsigned)f2bf(gz[0]) | ((unsigned)f2bf(gz[1]) << 16);
                h.y = (unsigned)f2bf(gz[2]) | ((unsigned)f2bf(gz[3]) << 16);
                *reinterpret_cast<uint2*>(g.outb + o) = h;
                cs[j] += gz;
            } else {  // kEpiWs: split-K partial
                *reinterpret_cast<f32x4*>(g.out32 + (long long)blockIdx.z * g.ws_stride + o) = v;
            }
        }
    }
    if (EPI == kEpiBwd && g.colpart) {
        // sum over the 16 rows the lanes of one column group hold (lanes differing in bits
        // 0-3), in a fixed order; one partial row per (row tile, wave row): 2 per 128 rows
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float x = cs[j][q];
                x += __shfl_xor(x, 1);
                x += __shfl_xor(x, 2);
                x += __shfl_xor(x, 4);
                x += __shfl_xor(x, 8);
                cs[j][q] = x;
            }
        if (fr == 0) {
            float* dst = g.colpart + (long long)(2 * blockIdx.x + wm) * g.N;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = col0 + 16 * j;
                if (c < g.N) *reinterpret_cast<f32x4*>(dst + c) = cs[j];
            }
        }
    }
}

// sum over splits: dst[i * N + j] = sum_s ws[s][i][j] for i < rows
// dst[i * N + j] = sum over k < slices of src[k * stride + i * ld + j], for i < rows, j < N
// (N, ld and stride multiples of 4, src 16-byte aligned; dst dense).  A block of 256 threads
// takes 16 consecutive outputs as four 16-byte groups, 64 threads per group each summing
// every 64th slice; then one thread per output adds the 64 sums in a fixed order, so the
// result does not depend on scheduling.
__global__ __launch_bounds__(256) void k_sum_slices(const float* src, int slices, long long stride, int rows, int N,
                                                    int ld, float* dst) {
    __shared__ f32x4 red[64][4];
    const int cg = threadIdx.x & 3, sg = threadIdx.x >> 2;
    const long long total = (long long)rows * N;
    const long long o = (long long)blockIdx.x * 16 + cg * 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    if (o < total) {
        const int i = (int)(o / N), j = (int)(o - (long long)i * N);
        const float* p = src + (long long)i * ld + j;
#pragma unroll 4
        for (int k = sg; k < slices; k += 64) acc += *reinterpret_cast<const f32x4*>(p + (long long)k * stride);
    }
    red[sg][cg] = acc;
    __syncthreads();
    const long long oo = (long long)blockIdx.x * 16 + threadIdx.x;
    if (threadIdx.x < 16 && oo < total) {
        float t = 0.f;
        for (int k = 0; k < 64; ++k) t += red[k][threadIdx.x >> 2][threadIdx.x & 3];
        dst[oo] = t;
    }
}

// column sums of G (B x N fp32, row-major) in two passes: partial[chunk][j], then dst[j]
__global__ void k_colsum_partial(const float* G, int B, int N, int rows_per_chunk, float* partial) {
    const int j = threadIdx.x;
    if (j >= N) return;
    const int b0 = blockIdx.x * rows_per_chunk;
    const int b1 = b0 + rows_per_chunk < B ? b0 + rows_per_chunk : B;
    float s = 0.f;
    for (int b = b0; b < b1; ++b) s += G[(long long)b * N + j];
    partial[(long long)blockIdx.x * N + j] = s;
}

__global__ void k_loss_head(const float* Y, const int8_t* cls, const float* tgt, int B, float w_sky, float w_disk,
                            float* gy32, uint16_t* gyb, double* loss_sum) {
    __shared__ float red[256];
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float loss = 0.f;
    if (b < B) {
        const float* y = Y + (long long)b * 8;
        const int c = cls[b];
        const float m = fmaxf(y[0], fmaxf(y[1], y[2]));
        const float e0 = __expf(y[0] - m), e1 = __expf(y[1] - m), e2 = __expf(y[2] - m);
        const float inv = 1.0f / (e0 + e1 + e2);
        float g[8] = {e0 * inv, e1 * inv, e2 * inv, 0.f, 0.f, 0.f, 0.f, 0.f};
        const float pc = c == 0 ? g[0] : c == 1 ? g[1] : g[2];
        loss = -__logf(fmaxf(pc, 1e-30f));
        g[c] -= 1.0f;
        const float* t = tgt + (long long)b * 5;
        if (c == 0) {
            for (int k = 0; k < 3; ++k) {
                const float d = y[3 + k] - t[k];
                loss += w_sky * d * d;
                g[3 + k] = 2.f * w_sky * d;
            }
        } else if (c == 2) {
            for (int k = 0; k < 2; ++k) {
                const float d = y[6 + k] - t[3 + k];
                loss += w_disk * d * d;
                g[6 + k] = 2.f * w_disk * d;
            }
        }
        const float invB = 1.0f / (float)B;
        for (int k = 0; k < 8; ++k) {
            gy32[(long long)b * 8 + k] = g[k] * invB;
            gyb[(long long)b * 32 + k] = f2bf(g[k] * invB);
        }
        for (int k = 8; k < 32; ++k) gyb[(long long)b * 32 + k] = 0;
    }
    red[threadIdx.x] = loss;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss_sum, (double)red[0]);
}

// Adam on the FP32 master parameters, then the bf16 working copies of each weight:
// Wb (K_pad x N_ld, row-major) and Wt (N x K_pad).
__global__ void k_adam(float* P, const float* G, float* Mm, float* Vv, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2, LayerTable lt) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float g = G[p];
    const float m = b1 * Mm[p] + (1.f - b1) * g;
    const float v = b2 * Vv[p] + (1.f - b2) * g * g;
    Mm[p] = m;
    Vv[p] = v;
    // bias-corrected moments: c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
    const float w = P[p] - lr * (m * c1) / (sqrtf(v * c2) + eps);
    P[p] = w;
    // which layer / weight element
    int l = 0;
    while (l + 1 < kLayers && p >= lt.off[l + 1]) ++l;
    const long long q = p - lt.off[l];
    const long long nw = (long long)lt.kreal[l] * lt.n[l];
    if (q < nw) {
        const int i = (int)(q / lt.n[l]), j = (int)(q - (long long)i * lt.n[l]);
        const uint16_t h = f2bf(w);
        lt.wb[l][(long long)i * lt.nld[l] + j] = h;
        lt.wt[l][(long long)j * lt.kpad[l] + i] = h;
    }
}

// bf16 copies of the master parameters (after a load), same mapping as k_adam
__global__ void k_refresh_bf16(const float* P, long long n, LayerTable lt) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int l = 0;
    while (l + 1 < kLayers && p >= lt.off[l + 1]) ++l;
    const long long q = p - lt.off[l];
    const long long nw = (long long)lt.kreal[l] * lt.n[l];
    if (q < nw) {
        const int i = (int)(q / lt.n[l]), j = (int)(q - (long long)i * lt.n[l]);
        const uint16_t h = f2bf(P[p]);
        lt.wb[l][(long long)i * lt.nld[l] + j] = h;
        lt.wt[l][(long long)j * lt.kpad[l] + i] = h;
    }
}

// gather a batch: inputs (B x 5 fp32), class and targets
__global__ void k_gather(const float* feat, const int8_t* cls, const float* tgt, const long long* idx, int B,
                         float* xf, int8_t* bcls, float* btgt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const long long s = idx ? idx[b] : b;
    for (int k = 0; k < 5; ++k) xf[(long long)b * 5 + k] = feat[s * 5 + k];
    if (bcls) bcls[b] = cls[s];
    if (btgt)
        for (int k = 0; k < 5; ++k) btgt[(long long)b * 5 + k] = tgt[s * 5 + k];
}

// Input layer in FP32 on the vector units: z = x W0 + b0 with the master weights, then
// the activation rounded to bf16 for the matrix-core layers.  Thread j owns hidden unit j
// (blockDim.x = H).
constexpr int kL0Rows = 16;
__global__ __launch_bounds__(256) void k_layer0(const float* xf, int B, const float* P, int act, float* z0,
                                                uint16_t* h0) {
    const int j = threadIdx.x, H = blockDim.x;
    float w[kInputs];
#pragma unroll
    for (int k = 0; k < kInputs; ++k) w[k] = P[k * H + j];
    const float bias = P[kInputs * H + j];
    const int r0 = blockIdx.x * kL0Rows;
    for (int r = r0; r < r0 + kL0Rows && r < B; ++r) {
        const float* x = xf + (long long)r * kInputs;
        float z = bias;
#pragma unroll
        for (int k = 0; k < kInputs; ++k) z = fmaf(x[k], w[k], z);
        if (z0) z0[(long long)r * H + j] = z;
        h0[(long long)r * H + j] = f2bf(activate(z, act));
    }
}

// Input-layer gradients in FP32: partial[c][k][j] = sum over chunk c of x[b][k] g[b][j]
// (k < 5) and of g[b][j] (k = 5), then the sum over chunks into G (W0 then b0).  Thread j
// owns hidden unit j (blockDim.x = H).
__global__ __launch_bounds__(256) void k_layer0_grad_partial(const float* xf, const float* g, int B, int rows,
                                                             float* partial) {
    const int j = threadIdx.x, H = blockDim.x;
    const int b0 = blockIdx.x * rows;
    const int b1 = b0 + rows < B ? b0 + rows : B;
    float s[kInputs + 1] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int b = b0; b < b1; ++b) {
        const float gj = g[(long long)b * H + j];
        const float* x = xf + (long long)b * kInputs;
#pragma unroll
        for (int k = 0; k < kInputs; ++k) s[k] = fmaf(x[k], gj, s[k]);
        s[kInputs] += gj;
    }
    float* o = partial + (long long)blockIdx.x * (kInputs + 1) * H;
#pragma unroll
    for (int k = 0; k <= kInputs; ++k) o[k * H + j] = s[k];
}


}  // namespace

hipError_t layer0(const float* xf, int B, const float* P, int H, int act, float* z0, uint16_t* h0, hipStream_t s) {
    if (H < 64 || H > kMaxHidden || H % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_layer0, dim3((B + kL0Rows - 1) / kL0Rows), dim3(H), 0, s, xf, B, P, act, z0, h0);
    return hipGetLastError();
}

hipError_t layer0_grad(const float* xf, const float* g, int B, int H, float* partial, int chunks, float* G,
                       hipStream_t s) {
    if (H < 64 || H > kMaxHidden || H % 64) return hipErrorInvalidValue;
    const int rows = (B + chunks - 1) / chunks;
    hipLaunchKernelGGL(k_layer0_grad_partial, dim3(chunks), dim3(H), 0, s, xf, g, B, rows, partial);
    // W0 (5 x H row-major) is followed directly by b0: one (k, j) column per parameter
    const int ncols = (kInputs + 1) * H;
    const hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : sum_chunks(partial, chunks, ncols, G, s);
}

hipError_t gemm(const GemmArgs& g, bool at, bool bt, int epi, int splits, hipStream_t stream) {
    if (g.N % 4 || g.ldc % 4 || (epi == kEpiBwd && g.ldz % 4)) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((g.M + 127) / 128), (unsigned)((g.N + 127) / 128), (unsigned)splits);
    const dim3 block(256);
#define KTN_CASE(A_, B_, E_)                                                          \
    if (at == A_ && bt == B_ && epi == E_) {                                          \
        hipLaunchKernelGGL((k_gemm<A_, B_, E_>), grid, block, 0, stream, g);          \
        return hipGetLastError();                                                     \
    }
    KTN_CASE(false, true, kEpiFwd)
    KTN_CASE(false, true, kEpiLin)
    KTN_CASE(false, true, kEpiBwd)
    KTN_CASE(true, false, kEpiWs)
    KTN_CASE(false, false, kEpiFwd)
    KTN_CASE(false, false, kEpiLin)
#undef KTN_CASE
    return hipErrorInvalidValue;
}

hipError_t reduce_ws(const float* ws, int splits, long long stride, int rows, int N, int ld, float* dst,
                     hipStream_t s) {
    if (N % 4 || ld % 4 || stride % 4) return hipErrorInvalidValue;
    const long long n = (long long)rows * N;
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, s, ws, splits, stride, rows, N, ld,
                       dst);
    return hipGetLastError();
}

hipError_t sum_chunks(const float* partial, int chunks, int ncols, float* dst, hipStream_t s) {
    return reduce_ws(partial, chunks, ncols, 1, ncols, 0, dst, s);
}

hipError_t colsum(const float* G, int B, int N, float* partial, int chunks, float* dst, hipStream_t s) {
    const int rpc = (B + chunks - 1) / chunks;
    hipLaunchKernelGGL(k_colsum_partial, dim3(chunks), dim3(256), 0, s, G, B, N, rpc, partial);
    const hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : sum_chunks(partial, chunks, N, dst, s);
}

hipError_t loss_head(const float* Y, const int8_t* cls, const float* tgt, int B, float w_sky, float w_disk,
                     float* gy32, uint16_t* gyb, double* loss_sum, hipStream_t s) {
    hipLaunchKernelGGL(k_loss_head, dim3((B + 255) / 256), dim3(256), 0, s, Y, cls, tgt, B, w_sky, w_disk, gy32, gyb,
                       loss_sum);
    return hipGetLastError();
}

hipError_t adam(float* P, const float* G, float* M, float* V, long long n, float lr, float b1, float b2, float eps,
                float c1, float c2, const LayerTable& lt, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, G, M, V, n, lr, b1, b2, eps,
                       c1, c2, lt);
    return hipGetLastError();
}

hipError_t refresh_bf16(const float* P, long long n, const LayerTable& lt, hipStream_t s) {
    hipLaunchKernelGGL(k_refresh_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, n, lt);
    return hipGetLastError();
}

hipError_t gather(const float* feat, const int8_t* cls, const float* tgt, const long long* idx, int B, float* xf,
                  int8_t* bcls, float* btgt, hipStream_t s) {
    hipLaunchKernelGGL(k_gather, dim3((B + 255) / 256), dim3(256), 0, s, feat, cls, tgt, idx, B, xf, bcls, btgt);
    return hipGetLastError();
}

}  // namespace ktn
