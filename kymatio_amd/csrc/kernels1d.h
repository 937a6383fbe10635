// __global__ wrapper around the 1D node body.
#pragma once
#include "kernels2d.h"
#include "scat1d_body.h"

template <class T, int NT>
__global__ void __launch_bounds__(NT) k_scat1d(Scat1DArgs<T> a) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem1d[];
    const int task = xcd_task(blockIdx.x, gridDim.x);
    if (task >= a.ntask) return;
    body_node1d<T, 7>(a, task, reinterpret_cast<cpx<T>*>(kym_smem1d), threadIdx.x, NT);
}

template <class T>
using Scat1DKernel = void (*)(Scat1DArgs<T>);

template <class T>
Scat1DKernel<T> pick_scat1d_nt(int threads) {
    if (threads >= 1024) return k_scat1d<T, 1024>;
    if (threads >= 512) return k_scat1d<T, 512>;
    if (threads >= 256) return k_scat1d<T, 256>;
    if (threads >= 128) return k_scat1d<T, 128>;
    return k_scat1d<T, 64>;
}
