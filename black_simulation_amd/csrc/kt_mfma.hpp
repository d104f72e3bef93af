// kt_mfma.hpp -- BF16 matrix-core building blocks for gfx950.
//
// v_mfma_f32_16x16x32_bf16 (one wave): D[16x16] += A[16x32] B[32x16], FP32 accumulate.
// Operand layout (lane l of 64):
//   A: row l % 16, k = 8 (l / 16) .. 8 (l / 16) + 7   (8 bf16 in one 16-byte register quad)
//   B: col l % 16, k = 8 (l / 16) .. 8 (l / 16) + 7
//   D: col l % 16, rows 4 (l / 16) .. 4 (l / 16) + 3
// So both operands are read with the reduction index contiguous: A row-major [m][k],
// B column-major [n][k] ("B transposed").
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ktm {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 as_bf16x8(const uint4& v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// hard SiLU, z clamp(z / 6 + 1/2, 0, 1), and its derivative (one-sided at z = +-3)
__device__ __forceinline__ float hsilu(float z) { return z * fminf(fmaxf(z * (1.0f / 6.0f) + 0.5f, 0.f), 1.f); }

__device__ __forceinline__ float dhsilu(float z) {
    return z <= -3.0f ? 0.f : (z >= 3.0f ? 1.f : z * (1.0f / 3.0f) + 0.5f);
}

// the activation of kind `act` (0: SiLU, 1: hard SiLU) and its derivative
__device__ __forceinline__ float activate(float z, int act) { return act ? hsilu(z) : silu(z); }
__device__ __forceinline__ float dactivate(float z, int act) { return act ? dhsilu(z) : dsilu(z); }

}  // namespace ktm
