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

// The same tile accumulated in pieces: mma16_strided runs k-steps whose
// operands sit at a fixed stride from a per-lane base (plain loads, no
// conditions, unrolled by four so the loads of four steps are in flight
// together), mma16_step runs one step from operand functors (boundaries:
// wrap-around and tails), mma16_store hands D to fd(l, v, value).
struct Mma16Acc {
#if defined(__HIP_DEVICE_COMPILE__)
    kym_f4 v;
#else
    float v[256];  // [lane * 4 + v], the device layout
#endif
};

KYM_HD KYM_INL void mma16_zero(Mma16Acc& c) {
#if defined(__HIP_DEVICE_COMPILE__)
    c.v = kym_f4{0.f, 0.f, 0.f, 0.f};
#else
    for (int i = 0; i < 256; ++i) c.v[i] = 0.f;
#endif
}

#if !defined(__HIP_DEVICE_COMPILE__)
// host: one k-step from the 64 lanes' A and B operands
inline void mma16_host_step(Mma16Acc& c, const float* a, const float* b) {
    for (int m = 0; m < 16; ++m)
        for (int n = 0; n < 16; ++n) {
            float t = 0.f;
            for (int k = 0; k < 4; ++k) t += a[m + 16 * k] * b[n + 16 * k];
            c.v[(n + 16 * (m / 4)) * 4 + (m % 4)] += t;
        }
}
#endif

// n k-steps; at step s lane l supplies A = pa(l)[s * sa] and B = pb(l)[s * sb].
// SA / SB > 0 fix the strides at compile time (immediate load offsets),
// otherwise sa / sb are used.
template <int SA, int SB, class PA, class PB>
KYM_HD KYM_INL void mma16_strided(Mma16Acc& c, int lane, PA pa, PB pb, int n, int sa = SA, int sb = SB) {
    if (SA > 0) sa = SA;
    if (SB > 0) sb = SB;
#if defined(__HIP_DEVICE_COMPILE__)
    const float* a = pa(lane);
    const float* b = pb(lane);
    int s = 0;
    for (; s + 4 <= n; s += 4) {
        const float a0 = a[0], a1 = a[sa], a2 = a[2 * sa], a3 = a[3 * sa];
        const float b0 = b[0], b1 = b[sb], b2 = b[2 * sb], b3 = b[3 * sb];
        c.v = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, c.v, 0, 0, 0);
        c.v = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, c.v, 0, 0, 0);
        c.v = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b2, c.v, 0, 0, 0);
        c.v = __builtin_amdgcn_mfma_f32_16x16x4f32(a3, b3, c.v, 0, 0, 0);
        a += 4 * sa;
        b += 4 * sb;
    }
    for (; s < n; ++s) {
        c.v = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], c.v, 0, 0, 0);
        a += sa;
        b += sb;
    }
#else
    (void)lane;
    const float* ap[64];
    const float* bp[64];
    for (int l = 0; l < 64; ++l) {
        ap[l] = pa(l);
        bp[l] = pb(l);
    }
    float a[64], b[64];
    for (int s = 0; s < n; ++s) {
        for (int l = 0; l < 64; ++l) {
            a[l] = ap[l][(long long)s * sa];
            b[l] = bp[l][(long long)s * sb];
        }
        mma16_host_step(c, a, b);
    }
#endif
}

template <class FA, class FB>
KYM_HD KYM_INL void mma16_step(Mma16Acc& c, int lane, FA fa, FB fb) {
#if defined(__HIP_DEVICE_COMPILE__)
    c.v = __builtin_amdgcn_mfma_f32_16x16x4f32(fa(lane), fb(lane), c.v, 0, 0, 0);
#else
    (void)lane;
    float a[64], b[64];
    for (int l = 0; l < 64; ++l) {
        a[l] = fa(l);
        b[l] = fb(l);
    }
    mma16_host_step(c, a, b);
#endif
}

template <class FD>
KYM_HD KYM_INL void mma16_store(const Mma16Acc& c, int lane, FD fd) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
    for (int v = 0; v < 4; ++v) fd(lane, v, c.v[v]);
#else
    (void)lane;
    for (int l = 0; l < 64; ++l)
        for (int v = 0; v < 4; ++v) fd(l, v, c.v[l * 4 + v]);
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
