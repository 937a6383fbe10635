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

template <class T>
using Scat1DKernel = void (*)(Scat1DArgs<T>);

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
