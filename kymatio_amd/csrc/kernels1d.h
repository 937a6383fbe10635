// __global__ wrapper around the 1D node body.
#pragma once
#include "kernels2d.h"
#include "scat1d_body.h"

template <class T, int NT, int KIND>
__global__ void __launch_bounds__(NT) k_scat1d(Scat1DArgs<T> a) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem1d[];
    const int task = xcd_task(blockIdx.x, gridDim.x);
    if (task >= a.ntask) return;
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(kym_smem1d);
    if constexpr (KIND == 0) body_node1d<T, 7>(a, task, lds, threadIdx.x, NT);
    else if constexpr (KIND == 1) body_bwd_node1d<T, 7>(a, task, lds, threadIdx.x, NT);
    else body_bwd_fold1d<T>(a, task, threadIdx.x, NT);
}

template <class T, int SPEC, int NT>
__global__ void __launch_bounds__(NT) k_node1d_spec(Scat1DArgs<T> a) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem1d_ns[];
    const int task = xcd_task(blockIdx.x, gridDim.x);
    if (task >= a.ntask) return;
    node1d_spec_body<T, SPEC, false>(a, task, reinterpret_cast<cpx<T>*>(kym_smem1d_ns), threadIdx.x);
}

template <class T>
using Scat1DKernel = void (*)(Scat1DArgs<T>);

template <class T>
Scat1DKernel<T> pick_node1d_spec(int id) {
    if constexpr (sizeof(T) != 4) return nullptr;  // specialisations are FP32 only
    else switch (id) {
#define KYM_N1_K(sid, nn, nt) \
    case sid: return k_node1d_spec<T, sid, nt>;
        KYM_NODE1D_SPECS(KYM_N1_K)
#undef KYM_N1_K
        default: return nullptr;
    }
}

template <class T, int KIND>
Scat1DKernel<T> pick_scat1d_k(int threads) {
    if (threads >= 1024) return k_scat1d<T, 1024, KIND>;
    if (threads >= 512) return k_scat1d<T, 512, KIND>;
    if (threads >= 256) return k_scat1d<T, 256, KIND>;
    if (threads >= 128) return k_scat1d<T, 128, KIND>;
    return k_scat1d<T, 64, KIND>;
}

template <class T>
Scat1DKernel<T> pick_scat1d_nt(int threads, int kind) {
    if (kind == 1) return pick_scat1d_k<T, 1>(threads);
    if (kind == 2) return k_scat1d<T, 256, 2>;
    return pick_scat1d_k<T, 0>(threads);
}
