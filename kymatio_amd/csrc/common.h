// Shared types for the scattering HIP library (gfx950).
//
// Kernel bodies are written as workgroup functions of (tid, nthreads): every
// parallel loop is `for (i = tid; i < n; i += nthreads)` followed by a
// workgroup barrier, and every butterfly reads and writes only its own
// elements.  The same body therefore gives the same result when a CPU test
// harness runs it with one thread (tid = 0, nthreads = 1), which is how the
// index arithmetic is checked without a GPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#define KYM_HD __host__ __device__
#define KYM_INL __forceinline__

template <class T>
struct alignas(2 * sizeof(T)) cpx {
    T re, im;
};

template <class T> KYM_HD KYM_INL cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> KYM_HD KYM_INL cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> KYM_HD KYM_INL cpx<T> operator*(cpx<T> a, T s) { return {a.re * s, a.im * s}; }
// a * b
template <class T> KYM_HD KYM_INL cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <class T> KYM_HD KYM_INL cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
// multiply by -i (forward transform sign) or +i (inverse)
template <bool INV, class T> KYM_HD KYM_INL cpx<T> mul_mi(cpx<T> a) {
    if (INV) return {-a.im, a.re};
    return {a.im, -a.re};
}

// Load through a pointer that is known to address global memory.  Table
// pointers read from the plan's blob are generic to the compiler, which then
// emits flat loads (tied to the LDS counter); this keeps them global.
// (Scalars and vector types only: a cast of a struct pointer to the global
// address space is dropped by the compiler, so complex values load as a
// two-element vector.)
template <class U>
KYM_HD KYM_INL U kym_ldg(const U* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(1))) U*)p;
#else
    return *p;
#endif
}
template <class T>
KYM_HD KYM_INL cpx<T> kym_ldg(const cpx<T>* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = *(const __attribute__((address_space(1))) v2*)(const void*)p;
    return {v.x, v.y};
#else
    return *p;
#endif
}

// for i < n: put(i, get(i)), with E values per thread loaded before any of
// them is stored, so E memory requests are in flight at once (a plain copy
// loop into LDS waits for each load before its store).
template <int E, class GET, class PUT>
KYM_HD KYM_INL void copy_batched(int n, GET get, PUT put, int tid, int nth) {
    int i = tid;
    for (; i + (E - 1) * nth < n; i += E * nth) {
        decltype(get(0)) v[E];
#pragma unroll
        for (int u = 0; u < E; ++u) v[u] = get(i + u * nth);
#pragma unroll
        for (int u = 0; u < E; ++u) put(i + u * nth, v[u]);
    }
    for (; i < n; i += nth) put(i, get(i));
}

// The same with the last batch partial (predicated) instead of a serial
// tail: every element's load is issued with its batch.  Faster in the
// backward passes, slower in the forward ones (measured), so both exist.
template <int E, class GET, class PUT>
KYM_HD KYM_INL void copy_batched_pred(int n, GET get, PUT put, int tid, int nth) {
    for (int i0 = tid; i0 < n; i0 += E * nth) {
        decltype(get(0)) v[E];
#pragma unroll
        for (int u = 0; u < E; ++u)
            if (i0 + u * nth < n) v[u] = get(i0 + u * nth);
#pragma unroll
        for (int u = 0; u < E; ++u)
            if (i0 + u * nth < n) put(i0 + u * nth, v[u]);
    }
}

// Workgroup barrier; a no-op when a body runs on the host with one thread.
KYM_HD KYM_INL void wg_barrier() {
#if defined(__HIP_DEVICE_COMPILE__)
    __syncthreads();
#endif
}

template <class T> KYM_HD KYM_INL T kym_sqrt(T v);
// FP32: the hardware square root (v_sqrt_f32, within 1 ulp) rather than the
// correctly rounded expansion, which costs a dozen instructions per call.
template <> KYM_HD KYM_INL float kym_sqrt<float>(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(v);
#else
    return sqrtf(v);
#endif
}
template <> KYM_HD KYM_INL double kym_sqrt<double>(double v) { return sqrt(v); }

// numpy-style 'reflect' index (edge sample not repeated), any pad width.
KYM_HD KYM_INL int reflect_index(int i, int n) {
    if (n == 1) return 0;
    const int period = 2 * (n - 1);
    i %= period;
    if (i < 0) i += period;
    return i < n ? i : period - i;
}

enum KymDtype { KYM_F32 = 0, KYM_F64 = 1 };
