// bf16 MFMA helpers for gfx950.  Lane layouts of the 16x16 output tile
// (64 lanes, lane l, element i):
//   v_mfma_f32_16x16x16_bf16  A[m = l%16][k = 4(l/16)+i], B[k = 4(l/16)+i][n = l%16], i < 4
//   v_mfma_f32_16x16x32_bf16  A[m = l%16][k = 8(l/16)+i], B[k = 8(l/16)+i][n = l%16], i < 8
//   accumulator (both)        D[m = 4(l/16)+i][n = l%16], i < 4
// (checked on the device by kym_mfma_layout_check).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef short kym_bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 kym_bf16x8 __attribute__((ext_vector_type(8)));
typedef float kym_f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even float -> bf16 bits
__host__ __device__ __forceinline__ uint16_t kym_f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ __forceinline__ float kym_bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ kym_f32x4 kym_mfma16(kym_bf16x4 a, kym_bf16x4 b, kym_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ kym_f32x4 kym_mfma32(kym_bf16x8 a, kym_bf16x8 b, kym_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
