// __global__ wrappers around the 2D workgroup bodies.
#pragma once
#include "scat2d_bwd.h"

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Give each XCD a contiguous range of tasks so that all consumers of one
// image's parent spectrum run on one XCD and share its L2.  The grid is a
// multiple of 8; tasks >= ntask exit before any barrier.
__device__ KYM_INL int xcd_task(int bid, int nblocks) {
    const int per = nblocks >> 3;
    return (bid & 7) * per + (bid >> 3);
}

template <class T, int KIND, int NT, int MAXP>
__global__ void __launch_bounds__(NT, 2) k_scat2d(Scat2DArgs<T> a) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem2d[];
    const int task = xcd_task(blockIdx.x, gridDim.x);
    if (task >= live_tasks(a)) return;
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(kym_smem2d);
    if constexpr (KIND >= K2B_NODE) dispatch_bwd2d<T, MAXP>(KIND, a, task, lds, threadIdx.x, NT);
    else dispatch_body2d<T, MAXP>(KIND, a, task, lds, threadIdx.x, NT);
}

template <class T, int SPEC, int NT>
__global__ void __launch_bounds__(NT) k_node2d_spec(Scat2DArgs<T> a) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem2d_ns[];
    const int task = xcd_task(blockIdx.x, gridDim.x);
    if (task >= live_tasks(a)) return;
    node_spec_body<T, SPEC, false>(a, task, reinterpret_cast<cpx<T>*>(kym_smem2d_ns), threadIdx.x);
}

template <class T, int KIND, int SPEC, int NT>
__global__ void __launch_bounds__(NT, 2) k_pass2d_spec(Scat2DArgs<T> a) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem2d_ps[];
    const int task = xcd_task(blockIdx.x, gridDim.x);
    if (task >= live_tasks(a)) return;
    cpx<T>* lds = reinterpret_cast<cpx<T>*>(kym_smem2d_ps);
    if constexpr (KIND >= K2B_NODE) bwd_pass_spec_body<T, KIND, SPEC, false>(a, task, lds, threadIdx.x);
    else pass_spec_body<T, KIND, SPEC, false>(a, task, lds, threadIdx.x);
}

template <class T, int SPEC, int NT>
__global__ void __launch_bounds__(NT) k_bwd_node2d_spec(Scat2DArgs<T> a) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem2d_bns[];
    const int task = xcd_task(blockIdx.x, gridDim.x);
    if (task >= a.ntask) return;
    bwd_node_spec_body<T, SPEC, false>(a, task, reinterpret_cast<cpx<T>*>(kym_smem2d_bns), threadIdx.x);
}

template <class T>
using Scat2DKernel = void (*)(Scat2DArgs<T>);

template <class T>
Scat2DKernel<T> pick_bwd_node_spec(int id) {
    if constexpr (sizeof(T) != 4) return nullptr;  // specialisations are FP32 only
    else switch (id) {
#define KYM_BNS_K(sid, h, w, hj, wj, nt, pr) \
    case sid: return k_bwd_node2d_spec<T, sid, nt>;
        KYM_NODE_SPECS(KYM_BNS_K)
#undef KYM_BNS_K
        default: return nullptr;
    }
}

template <class T>
Scat2DKernel<T> pick_node_spec(int id) {
    if constexpr (sizeof(T) != 4) return nullptr;  // specialisations are FP32 only
    else switch (id) {
#define KYM_NS_K(sid, h, w, hj, wj, nt, pr) \
    case sid: return k_node2d_spec<T, sid, nt>;
        KYM_NODE_SPECS(KYM_NS_K)
#undef KYM_NS_K
        default: return nullptr;
    }
}

template <class T, int KIND>
Scat2DKernel<T> pick_pass_spec_k(int id) {
    if constexpr (sizeof(T) != 4) return nullptr;
    else switch (id) {
#define KYM_PS_K(sid, h, w, c, r, nt) \
    case sid: return k_pass2d_spec<T, KIND, sid, nt>;
        KYM_PASS_SPECS(KYM_PS_K)
#undef KYM_PS_K
        default: return nullptr;
    }
}

template <class T>
Scat2DKernel<T> pick_pass_spec(int kind, int id) {
    switch (kind) {
        case K2_ROOT_ROWS: return pick_pass_spec_k<T, K2_ROOT_ROWS>(id);
        case K2_ICOL: return pick_pass_spec_k<T, K2_ICOL>(id);
        case K2_ICOL_PAIR: return pick_pass_spec_k<T, K2_ICOL_PAIR>(id);
        case K2_ROWS_MOD: return pick_pass_spec_k<T, K2_ROWS_MOD>(id);
        case K2_FWD_COLS: return pick_pass_spec_k<T, K2_FWD_COLS>(id);
        case K2B_GCOLS: return pick_pass_spec_k<T, K2B_GCOLS>(id);
        case K2B_GROWS: return pick_pass_spec_k<T, K2B_GROWS>(id);
        case K2B_GFCOLS: return pick_pass_spec_k<T, K2B_GFCOLS>(id);
        default: return nullptr;
    }
}

template <class T, int KIND, int MAXP>
Scat2DKernel<T> pick_scat2d_nt(int threads) {
    if (threads >= 1024) return k_scat2d<T, KIND, 1024, MAXP>;
    if (threads >= 512) return k_scat2d<T, KIND, 512, MAXP>;
    return k_scat2d<T, KIND, 256, MAXP>;
}

template <class T, int MAXP>
Scat2DKernel<T> pick_scat2d_p(int kind, int threads) {
    switch (kind) {
        case K2_LDS_NODE: return pick_scat2d_nt<T, K2_LDS_NODE, MAXP>(threads);
        case K2_LEAF: return pick_scat2d_nt<T, K2_LEAF, MAXP>(threads);
        case K2_ROOT_ROWS: return k_scat2d<T, K2_ROOT_ROWS, 256, MAXP>;
        case K2_ICOL: return k_scat2d<T, K2_ICOL, 256, MAXP>;
        case K2_ROWS_MOD: return k_scat2d<T, K2_ROWS_MOD, 256, MAXP>;
        case K2_FWD_COLS: return k_scat2d<T, K2_FWD_COLS, 256, MAXP>;
        case K2_LEAF_ICOL: return k_scat2d<T, K2_LEAF_ICOL, 256, MAXP>;
        case K2_LEAF_ROWS: return k_scat2d<T, K2_LEAF_ROWS, 256, MAXP>;
        case K2_FIR_COLS: return k_scat2d<T, K2_FIR_COLS, 256, MAXP>;
        case K2B_NODE: return pick_scat2d_nt<T, K2B_NODE, MAXP>(threads);
        case K2B_LEAFG: return pick_scat2d_nt<T, K2B_LEAFG, MAXP>(threads);
        case K2B_GCOLS: return k_scat2d<T, K2B_GCOLS, 256, MAXP>;
        case K2B_GROWS: return k_scat2d<T, K2B_GROWS, 256, MAXP>;
        case K2B_GFCOLS: return k_scat2d<T, K2B_GFCOLS, 256, MAXP>;
        case K2B_FOLD: return k_scat2d<T, K2B_FOLD, 256, MAXP>;
        default: return nullptr;
    }
}
