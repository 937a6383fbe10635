// Kernel lookup across the separately compiled kernel families (kinst2d.hip,
// kspec2d.hip, kinst1d.hip).  Host-side only.
#pragma once
#include "scat1d_body.h"
#include "scat2d_bwd.h"

template <class T> using Scat2DKernel = void (*)(Scat2DArgs<T>);
template <class T> using Scat1DKernel = void (*)(Scat1DArgs<T>);

Scat2DKernel<float> kym_pick2d_f32_p7(int kind, int threads);
Scat2DKernel<float> kym_pick2d_f32_p17(int kind, int threads);
Scat2DKernel<float> kym_pick2d_f32_p31(int kind, int threads);
Scat2DKernel<double> kym_pick2d_f64_p7(int kind, int threads);
Scat2DKernel<double> kym_pick2d_f64_p17(int kind, int threads);
Scat2DKernel<double> kym_pick2d_f64_p31(int kind, int threads);
Scat2DKernel<float> kym_pick_node_spec_f32(int id);
Scat2DKernel<float> kym_pick_bwd_node_spec_f32(int id);
Scat2DKernel<float> kym_pick_pass_spec_f32(int kind, int id);
Scat1DKernel<float> kym_pick1d_f32(int threads, int kind);
Scat1DKernel<float> kym_pick_node1d_spec_f32(int id);
Scat1DKernel<double> kym_pick1d_f64(int threads, int kind);

// pclass: 7, 17 or 31 (kym_prime_class of the plan's largest odd prime);
// spec >= 0 selects a compile-time specialisation of the node or pass kernel.
template <class T>
inline Scat2DKernel<T> pick_scat2d(int kind, int threads, int pclass, int spec = -1) {
    if constexpr (sizeof(T) == 4) {
        if (spec >= 0 && kind == K2_LDS_NODE) return kym_pick_node_spec_f32(spec);
        if (spec >= 0 && kind == K2B_NODE) return kym_pick_bwd_node_spec_f32(spec);
        if (spec >= 0) return kym_pick_pass_spec_f32(kind, spec);
        if (pclass <= 7) return kym_pick2d_f32_p7(kind, threads);
        if (pclass <= 17) return kym_pick2d_f32_p17(kind, threads);
        return kym_pick2d_f32_p31(kind, threads);
    } else {
        if (pclass <= 7) return kym_pick2d_f64_p7(kind, threads);
        if (pclass <= 17) return kym_pick2d_f64_p17(kind, threads);
        return kym_pick2d_f64_p31(kind, threads);
    }
}

// spec >= 0 (FP32, forward nodes) selects a compile-time specialisation.
template <class T>
inline Scat1DKernel<T> pick_scat1d(int threads, int kind = 0, int spec = -1) {
    if constexpr (sizeof(T) == 4) {
        if (spec >= 0 && kind == 0) return kym_pick_node1d_spec_f32(spec);
        return kym_pick1d_f32(threads, kind);
    } else {
        return kym_pick1d_f64(threads, kind);
    }
}
