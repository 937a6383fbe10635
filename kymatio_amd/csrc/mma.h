// One wave's FP32 matrix tile on the matrix cores (v_mfma_f32_16x16x4f32):
//   D (16 x 16) = sum over k-steps s of A_s (16 x 4) B_s (4 x 16)
// Lane l of the wave supplies A_s[l % 16][l / 16] = fa(l, s) and
// B_s[l / 16][l % 16] = fb(l, s); D[4 (l / 16) + v][l % 16] is delivered to
// fd(l, v, value).  Every lane of the wave must call it (operands that fall
// outside a problem are passed as zeros, never skipped).
//
// The CPU test harness runs a workgroup as one thread; there the function
// plays all 64 lanes itself with the same operand functors, so the index
// arithmetic that builds the operands is exercised by the CPU tests.
#pragma once
#include "common.h"

typedef float kym_f4 __attribute__((ext_vector_type(4)));

template <class FA, class FB, class FD>
KYM_HD KYM_INL void wave_mma16(int lane, int ksteps, FA fa, FB fb, FD fd) {
#if defined(__HIP_DEVICE_COMPILE__)
    kym_f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < ksteps; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(fa(lane, s), fb(lane, s), acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) fd(lane, v, acc[v]);
#else
    (void)lane;
    std::vector<float> acc(64 * 4, 0.f);
    float a[64], b[64];
    for (int s = 0; s < ksteps; ++s) {
        for (int l = 0; l < 64; ++l) {
            a[l] = fa(l, s);
            b[l] = fb(l, s);
        }
        for (int m = 0; m < 16; ++m)
            for (int n = 0; n < 16; ++n) {
                float t = 0.f;
                for (int k = 0; k < 4; ++k) t += a[m + 16 * k] * b[n + 16 * k];
                acc[(n + 16 * (m / 4)) * 4 + (m % 4)] += t;
            }
    }
    for (int l = 0; l < 64; ++l)
        for (int v = 0; v < 4; ++v) fd(l, v, acc[l * 4 + v]);
#endif
}

// Wave index / wave count of a workgroup body (the CPU harness is one wave).
// (readfirstlane: the compiler then knows the value, and everything derived
// from it, is uniform across the wave, so tile loops run on scalar registers.)
KYM_HD KYM_INL int wave_id(int tid) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(tid >> 6);
#else
    (void)tid;
    return 0;
#endif
}
KYM_HD KYM_INL int wave_count(int nth) {
#if defined(__HIP_DEVICE_COMPILE__)
    return nth >> 6;
#else
    (void)nth;
    return 1;
#endif
}
KYM_HD KYM_INL int lane_id(int tid) { return tid & 63; }
