// Per-op entry points: the backend namespace of the reference
// (kymatio/scattering2d/backend/torch_backend.py, kymatio/scattering1d/
// backend/torch_backend.py, kymatio/backend/torch_backend.py) as single HIP
// kernels, for the plugin route and for op-level tests.  The fused pipelines
// in capi.hip do not use these.
#include <algorithm>
#include <string>

#include "../../include/kymatio_hip.h"
#include "fft_core.h"
#include "plan_common.h"

int kym_ops_fail(int code, const std::string& msg);  // capi.hip

#define OPS_HIP_TRY(expr)                                                                          \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) return kym_ops_fail(KYM_ERR_HIP, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
    } while (0)

static inline unsigned grid_for(long long n, int block) {
    long long g = (n + block - 1) / block;
    if (g > 65535ll * 64) g = 65535ll * 64;
    return (unsigned)(g < 1 ? 1 : g);
}

// ------------------------------------------------------------ elementwise
template <class T>
__global__ void k_pad2d(const T* x, T* y, long long batch, int M, int N, int top, int left, int P, int Q) {
    const long long total = batch * P * Q;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const long long b = i / ((long long)P * Q);
        const int r = (int)(i - b * P * Q);
        const int yy = r / Q, xx = r - (r / Q) * Q;
        y[i] = x[b * M * N + (long long)reflect_index(yy - top, M) * N + reflect_index(xx - left, N)];
    }
}

template <class T>
__global__ void k_pad1d(const T* x, T* y, long long batch, int N, int left, int P) {
    const long long total = batch * P;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const long long b = i / P;
        const int t = (int)(i - b * P);
        y[i] = x[b * N + reflect_index(t - left, N)];
    }
}

template <class T>
__global__ void k_cdgmm(const cpx<T>* A, const T* B, cpx<T>* C, long long total, long long n, int bcomplex) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const long long j = i % n;
        const cpx<T> a = A[i];
        if (bcomplex) {
            const cpx<T> b = reinterpret_cast<const cpx<T>*>(B)[j];
            C[i] = cmul(a, b);
        } else {
            const T b = B[j];
            C[i] = {a.re * b, a.im * b};
        }
    }
}

template <class T>
__global__ void k_modulus(const cpx<T>* x, T* y, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const cpx<T> v = x[i];
        y[i] = kym_sqrt(v.re * v.re + v.im * v.im);
    }
}

// Rotation-covariant modulus accumulated over m (3D solid harmonics):
// y[i] = sqrt(prev[i]^2 + |x[i]|^2), prev absent for the first m.
template <class T>
__global__ void k_modulus_rot(const cpx<T>* x, const T* prev, T* y, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const cpx<T> v = x[i];
        const T t = v.re * v.re + v.im * v.im;
        const T p = prev ? prev[i] : T(0);
        y[i] = kym_sqrt(p * p + t);
    }
}

// Solid-harmonic products for every m of one (l, j) at once:
// out[b, m, i] = U[b, i] * F[m, i], the reference's cdgmm3d(U, F[m]) for all
// m (same complex product as k_cdgmm, U the left operand).
template <class T>
__global__ void k_harmonic_products(const cpx<T>* U, const cpx<T>* F, cpx<T>* out, long long batch, int nm,
                                    long long n) {
    const long long total = batch * nm * n;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const long long bm = i / n, e = i - bm * n;
        const long long b = bm / nm, m = bm - b * nm;
        out[i] = cmul(U[b * n + e], F[m * n + e]);
    }
}

// Rotation-covariant modulus of all m at once: for each (b, i) the running
// value p = sqrt(p^2 + |x[b, m, i]|^2) over m = 0..nm-1 (p = 0 before m = 0),
// the order of the reference's modulus_rotation calls; y[b * ystride + i].
template <class T>
__global__ void k_modulus_rot_sum(const cpx<T>* x, T* y, long long batch, int nm, long long n, long long ystride) {
    const long long total = batch * n;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const long long b = i / n, e = i - b * n;
        const cpx<T>* xb = x + b * nm * n + e;
        T p = T(0);
        for (int m = 0; m < nm; ++m) {
            const cpx<T> v = xb[m * n];
            const T t = v.re * v.re + v.im * v.im;
            p = kym_sqrt(p * p + t);
        }
        y[b * ystride + e] = p;
    }
}

// out[b * npow + q] = sum_i x[b, i]^pw[q]: one workgroup per b, FP64
// partial sums reduced through LDS.
#define KYM_MAX_POWERS 8
struct PowerList {
    double p[KYM_MAX_POWERS];
    int n;
};

template <class T>
__global__ void __launch_bounds__(256) k_integrals(const T* x, T* out, long long n, PowerList pw) {
    __shared__ double part[KYM_MAX_POWERS][256];
    const long long b = blockIdx.x;
    const T* xb = x + b * n;
    double acc[KYM_MAX_POWERS];
    for (int q = 0; q < KYM_MAX_POWERS; ++q) acc[q] = 0.0;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = (double)xb[i];
        for (int q = 0; q < pw.n; ++q) acc[q] += pow(v, pw.p[q]);
    }
    for (int q = 0; q < pw.n; ++q) part[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < pw.n; ++q) part[q][threadIdx.x] += part[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < pw.n; ++q) out[b * pw.n + q] = (T)part[q][0];
}

// out[b, i, j] = mean over a, c < k of x[b, i + a H/k, j + c W/k]   (W = 1 for 1D)
template <class T>
__global__ void k_subsample(const cpx<T>* x, cpx<T>* y, long long batch, int H, int W, int kh, int kw) {
    const int h = H / kh, w = W / kw;
    const long long total = batch * h * w;
    const T inv = T(1) / T(kh * kw);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const long long b = i / ((long long)h * w);
        const int r = (int)(i - b * h * w);
        const int yy = r / w, xx = r - (r / w) * w;
        cpx<T> acc = {T(0), T(0)};
        for (int a = 0; a < kh; ++a)
            for (int c = 0; c < kw; ++c) acc = acc + x[b * H * W + (long long)(yy + a * h) * W + xx + c * w];
        y[i] = acc * inv;
    }
}

template <class T>
__global__ void k_unpad2d(const T* x, T* y, long long batch, int H, int W) {
    const int h = H - 2, w = W - 2;
    const long long total = batch * h * w;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const long long b = i / ((long long)h * w);
        const int r = (int)(i - b * h * w);
        const int yy = r / w, xx = r - (r / w) * w;
        y[i] = x[b * H * W + (long long)(yy + 1) * W + xx + 1];
    }
}

#include "ops_body.h"

template <class T, int MAXP>
__global__ void __launch_bounds__(256) k_fft2_lds(FftOpArgs o, const FftLen<T>* lens, int lH, int lW) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem_op[];
    if ((long long)blockIdx.x >= o.batch) return;
    fftop_lds_body<T, MAXP>(o, lens, lH, lW, blockIdx.x, reinterpret_cast<cpx<T>*>(kym_smem_op), threadIdx.x,
                            blockDim.x);
}

template <class T, int MAXP>
__global__ void __launch_bounds__(256) k_fft1_multi(FftOpArgs o, const FftLen<T>* lens, int lH) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem_op1[];
    fftop_multi_body<T, MAXP>(o, lens, lH, blockIdx.x, reinterpret_cast<cpx<T>*>(kym_smem_op1), threadIdx.x,
                              blockDim.x);
}

template <class T, int NT, int MAXP>
__global__ void __launch_bounds__(NT) k_filter_ifft(FilterIfftArgs o, const FftLen<T>* lens, int l) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem_fi[];
    filter_ifft_body<T, MAXP>(o, lens[l], blockIdx.x, reinterpret_cast<cpx<T>*>(kym_smem_fi), threadIdx.x, NT);
}

template <class T, int MAXP>
__global__ void __launch_bounds__(256) k_lowpass(LowpassArgs o, const FftLen<T>* lens, int lN, int ln) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem_lp[];
    lowpass_body<T, MAXP>(o, lens[lN], lens[ln], blockIdx.x, reinterpret_cast<cpx<T>*>(kym_smem_lp), threadIdx.x, 256);
}

template <class T, int MAXP>
__global__ void __launch_bounds__(256) k_freq_scatter(FreqScatterArgs o) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem_fs[];
    freq_scatter_body<T, MAXP>(o, blockIdx.x, reinterpret_cast<cpx<T>*>(kym_smem_fs), threadIdx.x, 256);
}

template <class T>
__global__ void __launch_bounds__(256) k_freq_scatter_s88(FreqScatterArgs o) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem_fss[];
    freq_scatter_body_s<T, 88, 32, 256>(o, blockIdx.x, reinterpret_cast<cpx<T>*>(kym_smem_fss), threadIdx.x);
}

template <class T, int NN, int M>
__global__ void __launch_bounds__(256) k_filter_ifft_s(FilterIfftArgs o, const FftLen<T>* lens, int l) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem_fis[];
    filter_ifft_body_s<T, NN, M, 256>(o, lens[l], blockIdx.x, reinterpret_cast<cpx<T>*>(kym_smem_fis), threadIdx.x);
}

template <class T, int SPEC>
__global__ void __launch_bounds__(256) k_lowpass_s(LowpassArgs o, const FftLen<T>* lens, int lN, int ln) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem_lps[];
    lowpass_spec_body<T, SPEC, false>(o, lens[lN], lens[ln], blockIdx.x, reinterpret_cast<cpx<T>*>(kym_smem_lps),
                                      threadIdx.x);
}

template <class T, int MAXP>
__global__ void __launch_bounds__(256) k_fft2_pass(FftOpArgs o, FftPassIO io, const FftLen<T>* lens, int lH, int lW,
                                                   int axis) {
    extern __shared__ __attribute__((aligned(16))) char kym_smem_op2[];
    fftop_pass_body<T, MAXP>(o, io, lens, lH, lW, axis, blockIdx.x, reinterpret_cast<cpx<T>*>(kym_smem_op2),
                             threadIdx.x, blockDim.x);
}

// Prime class (7, 17 or 31) of a set of FFT lengths.
static int op_pclass(int a, int b = 1) {
    const int mp = kym_max_odd_prime(a) > kym_max_odd_prime(b) ? kym_max_odd_prime(a) : kym_max_odd_prime(b);
    return kym_prime_class(mp);
}

// FFT length tables, built once per (length set, dtype) and kept for the
// process: a small cache keyed by (H, W, dtype, device).
#include <map>
#include <mutex>
#include <tuple>

struct OpTables {
    void* dev = nullptr;
    size_t lens_off = 0;
    int lH = 0, lW = 0;
};

static std::mutex g_op_mu;
static std::map<std::tuple<int, int, int, int>, OpTables> g_op_tables;

template <class T>
static int op_tables(int H, int W, OpTables* out) {
    int dev = 0;
    OPS_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(g_op_mu);
    auto key = std::make_tuple(H, W, (int)sizeof(T), dev);
    auto it = g_op_tables.find(key);
    if (it != g_op_tables.end()) {
        *out = it->second;
        return KYM_OK;
    }
    LenTable<T> lt;
    OpTables t;
    t.lH = lt.add(H);
    t.lW = lt.add(W);
    Blob b;
    std::string err;
    if (!lt.emit(b, &t.lens_off, &err)) return kym_ops_fail(KYM_ERR_UNSUPPORTED, err);
    OPS_HIP_TRY(hipMalloc(&t.dev, b.bytes.size()));
    b.relocate((uintptr_t)t.dev);
    OPS_HIP_TRY(hipMemcpy(t.dev, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice));
    g_op_tables[key] = t;
    *out = t;
    return KYM_OK;
}

static size_t op_lds_limit() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && v > 0) return v;
    return 65536;
}

template <class T, int MAXP>
static int fft_op_p(const void* x, void* y, void* work, long long batch, int H, int W, int kind, hipStream_t s,
                    long long lds_cap) {
    OpTables t;
    int rc = op_tables<T>(H, W, &t);
    if (rc != KYM_OK) return rc;
    const FftLen<T>* lens = reinterpret_cast<const FftLen<T>*>((char*)t.dev + t.lens_off);
    size_t lim = op_lds_limit();
    if (lds_cap > 0 && (size_t)lds_cap < lim) lim = (size_t)lds_cap;
    // short 1D transforms: several per workgroup, about 4096 points each
    if (const long long m = fftop_multi_count(H, W, batch)) {
        const long long nblk = (batch + m - 1) / m;
        if (nblk > 0x7fffffffll) return kym_ops_fail(KYM_ERR_VALUE, "batch too large");
        FftOpArgs o = {x, y, batch, H, W, kind, 0, (int)m, (int)nblk};
        hipLaunchKernelGGL((k_fft1_multi<T, MAXP>), dim3((unsigned)nblk), dim3(256), (size_t)(m * H) * sizeof(cpx<T>), s, o,
                           lens, t.lH);
        OPS_HIP_TRY(hipGetLastError());
        return KYM_OK;
    }
    FftOpStep steps[2];
    int nsteps = 0;
    const char* err = nullptr;
    const int code = fftop_steps(x, y, work, H, W, kind, lim, sizeof(cpx<T>), steps, &nsteps, &err);
    if (code) return kym_ops_fail(code, err);
    for (int q = 0; q < nsteps; ++q) {
        const FftOpStep& st = steps[q];
        FftOpArgs o = {x, y, batch, H, W, kind, 0, st.nrow, st.nblk};
        if (st.route == 0) {
            if (st.lds > 65536)
                OPS_HIP_TRY(hipFuncSetAttribute((const void*)k_fft2_lds<T, MAXP>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                (int)st.lds));
            hipLaunchKernelGGL((k_fft2_lds<T, MAXP>), dim3((unsigned)batch), dim3(256), st.lds, s, o, lens, t.lH, t.lW);
        } else {
            if (st.lds > 65536)
                OPS_HIP_TRY(hipFuncSetAttribute((const void*)k_fft2_pass<T, MAXP>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                (int)st.lds));
            hipLaunchKernelGGL((k_fft2_pass<T, MAXP>), dim3((unsigned)(batch * st.nblk)), dim3(256), st.lds, s, o, st.io, lens,
                               t.lH, t.lW, st.axis);
        }
        OPS_HIP_TRY(hipGetLastError());
    }
    return KYM_OK;
}

template <class T>
static int fft_op(const void* x, void* y, void* work, long long batch, int H, int W, int kind, hipStream_t s,
                  long long lds_cap) {
    switch (op_pclass(H, W)) {
        case 7: return fft_op_p<T, 7>(x, y, work, batch, H, W, kind, s, lds_cap);
        case 17: return fft_op_p<T, 17>(x, y, work, batch, H, W, kind, s, lds_cap);
        default: return fft_op_p<T, 31>(x, y, work, batch, H, W, kind, s, lds_cap);
    }
}

// Row layout of a filter_ifft launch (FilterIfftArgs): transposed rows,
// filter bank, output row map.
struct FilterIfftRows {
    int tlen, nf, grp, gtot, goff;
};

template <class T, int MAXP>
static int filter_ifft_op_p(const void* x, const void* f, void* y, long long batch, long long N, int k, int mode,
                            FilterIfftRows rl, hipStream_t s) {
    const int tlen = rl.tlen;
    const int n = (int)(N / k);
    OpTables t;
    int rc = op_tables<T>(n, 1, &t);
    if (rc != KYM_OK) return rc;
    const FftLen<T>* lens = reinterpret_cast<const FftLen<T>*>((char*)t.dev + t.lens_off);
    int nt = 256, m = 1, split = 1;
    const size_t lds = (size_t)filter_ifft_shape(n, k, batch, &nt, &m, &split, tlen) * sizeof(cpx<T>);
    if (lds > op_lds_limit()) return kym_ops_fail(KYM_ERR_UNSUPPORTED, "transform length too large for LDS");
    const void* kern =
        nt == 1024 ? (const void*)k_filter_ifft<T, 1024, MAXP> : (const void*)k_filter_ifft<T, 256, MAXP>;
    if (lds > 65536) OPS_HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long nblk = (batch + m - 1) / m;
    if (nblk > 0x7fffffffll) return kym_ops_fail(KYM_ERR_VALUE, "batch too large");
    FilterIfftArgs o = {x, f, y, batch, (int)N, k, n, mode, m, split, tlen, rl.nf, rl.grp, rl.gtot, rl.goff};
    if (nt == 1024)
        hipLaunchKernelGGL((k_filter_ifft<T, 1024, MAXP>), dim3((unsigned)nblk), dim3(1024), lds, s, o, lens, t.lH);
    else
        hipLaunchKernelGGL((k_filter_ifft<T, 256, MAXP>), dim3((unsigned)nblk), dim3(256), lds, s, o, lens, t.lH);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

template <class T>
static const void* filter_ifft_spec_kernel(int id) {
    if constexpr (sizeof(T) != 4) return nullptr;  // FP32 only
    else switch (id) {
#define KYM_FI_K(id, nn, mm) \
    case id: return (const void*)k_filter_ifft_s<T, nn, mm>;
        KYM_FILTER_IFFT_SPECS(KYM_FI_K)
#undef KYM_FI_K
        default: return nullptr;
    }
}

template <class T>
static int filter_ifft_op(const void* x, const void* f, void* y, long long batch, long long N, int k, int mode,
                          FilterIfftRows rl, hipStream_t s) {
    {
        const int n = (int)(N / k);
        FilterIfftArgs o = {x, f, y, batch, (int)N, k, n, mode, 0, 1, rl.tlen, rl.nf, rl.grp, rl.gtot, rl.goff};
        int M = 0;
        if (const void* kern = filter_ifft_spec_kernel<T>(filter_ifft_spec_id(n, batch, o, &M))) {
            OpTables t;
            int rc = op_tables<T>(n, 1, &t);
            if (rc != KYM_OK) return rc;
            const FftLen<T>* lens = reinterpret_cast<const FftLen<T>*>((char*)t.dev + t.lens_off);
            const size_t lds = (size_t)M * n * sizeof(cpx<T>);
            if (lds > 65536) OPS_HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            const long long nblk = (batch + M - 1) / M;
            if (nblk > 0x7fffffffll) return kym_ops_fail(KYM_ERR_VALUE, "batch too large");
            o.nrow = M;
            void* args[] = {&o, (void*)&lens, &t.lH};
            OPS_HIP_TRY(hipLaunchKernel(kern, dim3((unsigned)nblk), dim3(256), args, lds, s));
            return KYM_OK;
        }
    }
    switch (op_pclass((int)(N / k))) {
        case 7: return filter_ifft_op_p<T, 7>(x, f, y, batch, N, k, mode, rl, s);
        case 17: return filter_ifft_op_p<T, 17>(x, f, y, batch, N, k, mode, rl, s);
        default: return filter_ifft_op_p<T, 31>(x, f, y, batch, N, k, mode, rl, s);
    }
}

template <class T, int MAXP>
static int lowpass_op_p(const void* x, const void* f, void* y, long long outer, int N, int S, int k, hipStream_t s) {
    const int n = N / k;
    OpTables t;
    int rc = op_tables<T>(N, n, &t);
    if (rc != KYM_OK) return rc;
    const FftLen<T>* lens = reinterpret_cast<const FftLen<T>*>((char*)t.dev + t.lens_off);
    int m = 1, nsb = 1;
    long long nblk = 0;
    const size_t lds = (size_t)lowpass_shape(N, n, S, outer, &m, &nsb, &nblk) * sizeof(cpx<T>);
    if (lds > op_lds_limit()) return kym_ops_fail(KYM_ERR_UNSUPPORTED, "transform length too large for LDS");
    if (lds > 65536)
        OPS_HIP_TRY(hipFuncSetAttribute((const void*)k_lowpass<T, MAXP>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (nblk > 0x7fffffffll) return kym_ops_fail(KYM_ERR_VALUE, "batch too large");
    LowpassArgs o = {x, f, y, outer, N, S, k, n, m, nsb};
    hipLaunchKernelGGL((k_lowpass<T, MAXP>), dim3((unsigned)nblk), dim3(256), lds, s, o, lens, t.lH, t.lW);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

// The compile-time low-pass for (N, n), or nullptr.
template <class T>
static const void* lowpass_spec_kernel(int id) {
    if constexpr (sizeof(T) != 4) return nullptr;  // FP32 only
    else switch (id) {
#define KYM_LP_K(id, nn, n2, mm, sv) \
    case id: return (const void*)k_lowpass_s<T, id>;
        KYM_LOWPASS_SPECS(KYM_LP_K)
#undef KYM_LP_K
        default: return nullptr;
    }
}

template <class T>
static int lowpass_op(const void* x, const void* f, void* y, long long outer, int N, int S, int k, hipStream_t s) {
    int M = 0;
    if (const void* kern = lowpass_spec_kernel<T>(lowpass_spec_id(N, N / k, S, &M))) {
        const int n = N / k;
        OpTables t;
        int rc = op_tables<T>(N, n, &t);
        if (rc != KYM_OK) return rc;
        const FftLen<T>* lens = reinterpret_cast<const FftLen<T>*>((char*)t.dev + t.lens_off);
        const size_t lds = (size_t)lowpass_spec_lds(N, n, S, M) * sizeof(cpx<T>);
        if (lds > 65536) OPS_HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const long long nblk = S == 1 ? (outer + M - 1) / M : outer * (S / M);
        if (nblk > 0x7fffffffll) return kym_ops_fail(KYM_ERR_VALUE, "batch too large");
        LowpassArgs o = {x, f, y, outer, N, S, k, n, M, S == 1 ? 1 : S / M};
        void* args[] = {&o, (void*)&lens, &t.lH, &t.lW};
        OPS_HIP_TRY(hipLaunchKernel(kern, dim3((unsigned)nblk), dim3(256), args, lds, s));
        return KYM_OK;
    }
    switch (op_pclass(N, N / k)) {
        case 7: return lowpass_op_p<T, 7>(x, f, y, outer, N, S, k, s);
        case 17: return lowpass_op_p<T, 17>(x, f, y, outer, N, S, k, s);
        default: return lowpass_op_p<T, 31>(x, f, y, outer, N, S, k, s);
    }
}

template <class T>
static int freq_scatter_op(const void* x, void* const* ys, const void* const* fs, const int* ks, int nf, long long B,
                           int A, int T_, int c, int N, hipStream_t s) {
    FreqScatterArgs o = {};
    o.x = x;
    o.B = B;
    o.A = A;
    o.T = T_;
    o.c = c;
    o.N = N;
    o.nf = nf;
    o.m = freq_scatter_rows(N, T_);
    o.nt_blk = (T_ + o.m - 1) / o.m;
    int mp = kym_max_odd_prime(N);
    for (int i = 0; i < nf; ++i) {
        OpTables t;
        int rc = op_tables<T>(N, N / ks[i], &t);
        if (rc != KYM_OK) return rc;
        const FftLen<T>* lens = reinterpret_cast<const FftLen<T>*>((char*)t.dev + t.lens_off);
        o.lenN = lens + t.lH;
        o.lenn[i] = lens + t.lW;
        o.f[i] = fs[i];
        o.y[i] = ys[i];
        o.k[i] = ks[i];
        mp = std::max(mp, kym_max_odd_prime(N / ks[i]));
    }
    const size_t lds = (size_t)2 * o.m * N * sizeof(cpx<T>);
    if (lds > op_lds_limit()) return kym_ops_fail(KYM_ERR_UNSUPPORTED, "transform length too large for LDS");
    const long long nblk = B * o.nt_blk;
    if (nblk > 0x7fffffffll) return kym_ops_fail(KYM_ERR_VALUE, "batch too large");
    const void* kern = kym_prime_class(mp) == 7    ? (const void*)k_freq_scatter<T, 7>
                       : kym_prime_class(mp) == 17 ? (const void*)k_freq_scatter<T, 17>
                                                   : (const void*)k_freq_scatter<T, 31>;
    if (sizeof(T) == 4 && freq_scatter_spec_ok(o)) kern = (const void*)k_freq_scatter_s88<T>;
    if (lds > 65536) OPS_HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    void* args[] = {&o};
    OPS_HIP_TRY(hipLaunchKernel(kern, dim3((unsigned)nblk), dim3(256), args, lds, s));
    return KYM_OK;
}

// ------------------------------------------------------------ C ABI
template <class T>
static int launch_pad2d(const void* x, void* y, long long batch, int M, int N, int top, int left, int P, int Q,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_pad2d<T>, dim3(grid_for(batch * P * Q, 256)), dim3(256), 0, s, (const T*)x, (T*)y, batch, M,
                       N, top, left, P, Q);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

template <class T>
static int launch_pad1d(const void* x, void* y, long long batch, int N, int left, int P, hipStream_t s) {
    hipLaunchKernelGGL(k_pad1d<T>, dim3(grid_for(batch * P, 256)), dim3(256), 0, s, (const T*)x, (T*)y, batch, N, left,
                       P);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

template <class T>
static int launch_cdgmm(const void* A, const void* B, void* C, long long total, long long n, int bc, hipStream_t s) {
    hipLaunchKernelGGL(k_cdgmm<T>, dim3(grid_for(total, 256)), dim3(256), 0, s, (const cpx<T>*)A, (const T*)B,
                       (cpx<T>*)C, total, n, bc);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

template <class T>
static int launch_modulus(const void* x, void* y, long long n, hipStream_t s) {
    hipLaunchKernelGGL(k_modulus<T>, dim3(grid_for(n, 256)), dim3(256), 0, s, (const cpx<T>*)x, (T*)y, n);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

template <class T>
static int launch_subsample(const void* x, void* y, long long batch, int H, int W, int kh, int kw, hipStream_t s) {
    const long long total = batch * (H / kh) * (W / kw);
    hipLaunchKernelGGL(k_subsample<T>, dim3(grid_for(total, 256)), dim3(256), 0, s, (const cpx<T>*)x, (cpx<T>*)y, batch,
                       H, W, kh, kw);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

template <class T>
static int launch_unpad2d(const void* x, void* y, long long batch, int H, int W, hipStream_t s) {
    hipLaunchKernelGGL(k_unpad2d<T>, dim3(grid_for(batch * (H - 2) * (W - 2), 256)), dim3(256), 0, s, (const T*)x,
                       (T*)y, batch, H, W);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

template <class T>
static int launch_modulus_rot(const void* x, const void* prev, void* y, long long n, hipStream_t s) {
    hipLaunchKernelGGL(k_modulus_rot<T>, dim3(grid_for(n, 256)), dim3(256), 0, s, (const cpx<T>*)x, (const T*)prev,
                       (T*)y, n);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

template <class T>
static int launch_integrals(const void* x, void* out, long long batch, long long n, const PowerList& pw,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_integrals<T>, dim3((unsigned)batch), dim3(256), 0, s, (const T*)x, (T*)out, n, pw);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

template <class T>
static int launch_harmonic_products(const void* U, const void* F, void* out, long long batch, int nm, long long n,
                                    hipStream_t s) {
    hipLaunchKernelGGL(k_harmonic_products<T>, dim3(grid_for(batch * nm * n, 256)), dim3(256), 0, s,
                       (const cpx<T>*)U, (const cpx<T>*)F, (cpx<T>*)out, batch, nm, n);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

template <class T>
static int launch_modulus_rot_sum(const void* x, void* y, long long batch, int nm, long long n, long long ystride,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_modulus_rot_sum<T>, dim3(grid_for(batch * n, 256)), dim3(256), 0, s, (const cpx<T>*)x,
                       (T*)y, batch, nm, n, ystride);
    OPS_HIP_TRY(hipGetLastError());
    return KYM_OK;
}

static int bad_dtype() { return kym_ops_fail(KYM_ERR_TYPE, "dtype must be float32 or float64"); }

extern "C" int kym_op_modulus_rotation(const void* x, const void* prev, void* y, long long n, int dtype,
                                       void* stream) {
    if (n <= 0) return KYM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return launch_modulus_rot<float>(x, prev, y, n, s);
    if (dtype == KYM_F64) return launch_modulus_rot<double>(x, prev, y, n, s);
    return bad_dtype();
}

extern "C" int kym_op_harmonic_products(const void* U, const void* F, void* out, long long batch, int nm,
                                        long long n, int dtype, void* stream) {
    if (nm < 1) return kym_ops_fail(KYM_ERR_VALUE, "at least one filter per (l, j)");
    if (batch <= 0 || n <= 0) return KYM_OK;
    if (!U || !F || !out) return kym_ops_fail(KYM_ERR_VALUE, "null device pointer");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return launch_harmonic_products<float>(U, F, out, batch, nm, n, s);
    if (dtype == KYM_F64) return launch_harmonic_products<double>(U, F, out, batch, nm, n, s);
    return bad_dtype();
}

extern "C" int kym_op_modulus_rotation_sum(const void* x, void* y, long long batch, int nm, long long n,
                                           long long ystride, int dtype, void* stream) {
    if (nm < 1) return kym_ops_fail(KYM_ERR_VALUE, "at least one filter per (l, j)");
    if (ystride < n) return kym_ops_fail(KYM_ERR_VALUE, "output stride smaller than the volume");
    if (batch <= 0 || n <= 0) return KYM_OK;
    if (!x || !y) return kym_ops_fail(KYM_ERR_VALUE, "null device pointer");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return launch_modulus_rot_sum<float>(x, y, batch, nm, n, ystride, s);
    if (dtype == KYM_F64) return launch_modulus_rot_sum<double>(x, y, batch, nm, n, ystride, s);
    return bad_dtype();
}

extern "C" int kym_op_integrals(const void* x, void* out, long long batch, long long n, const double* powers,
                                int npow, int dtype, void* stream) {
    if (npow < 1 || npow > KYM_MAX_POWERS || !powers)
        return kym_ops_fail(KYM_ERR_VALUE, "between 1 and 8 integral powers are supported");
    if (batch <= 0) return KYM_OK;
    if (batch > 0x7fffffffll) return kym_ops_fail(KYM_ERR_VALUE, "batch too large");
    PowerList pw{};
    pw.n = npow;
    for (int q = 0; q < npow; ++q) pw.p[q] = powers[q];
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return launch_integrals<float>(x, out, batch, n, pw, s);
    if (dtype == KYM_F64) return launch_integrals<double>(x, out, batch, n, pw, s);
    return bad_dtype();
}

extern "C" int kym_op_pad2d(const void* x, void* y, long long batch, int M, int N, int top, int bottom, int left,
                            int right, int dtype, void* stream) {
    if (batch <= 0) return KYM_OK;
    const int P = M + top + bottom, Q = N + left + right;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return launch_pad2d<float>(x, y, batch, M, N, top, left, P, Q, s);
    if (dtype == KYM_F64) return launch_pad2d<double>(x, y, batch, M, N, top, left, P, Q, s);
    return bad_dtype();
}

extern "C" int kym_op_pad1d(const void* x, void* y, long long batch, int N, int left, int right, int dtype,
                            void* stream) {
    if (left >= N || right >= N) return kym_ops_fail(KYM_ERR_VALUE, "Indefinite padding size (larger than tensor).");
    if (batch <= 0) return KYM_OK;
    const int P = N + left + right;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return launch_pad1d<float>(x, y, batch, N, left, P, s);
    if (dtype == KYM_F64) return launch_pad1d<double>(x, y, batch, N, left, P, s);
    return bad_dtype();
}

extern "C" int kym_op_cdgmm(const void* A, const void* B, void* C, long long batch, long long n, int b_complex,
                            int dtype, void* stream) {
    const long long total = batch * n;
    if (total <= 0) return KYM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return launch_cdgmm<float>(A, B, C, total, n, b_complex, s);
    if (dtype == KYM_F64) return launch_cdgmm<double>(A, B, C, total, n, b_complex, s);
    return bad_dtype();
}

extern "C" int kym_op_modulus(const void* x, void* y, long long n, int dtype, void* stream) {
    if (n <= 0) return KYM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return launch_modulus<float>(x, y, n, s);
    if (dtype == KYM_F64) return launch_modulus<double>(x, y, n, s);
    return bad_dtype();
}

extern "C" int kym_op_subsample(const void* x, void* y, long long batch, int H, int W, int kh, int kw, int dtype,
                                void* stream) {
    if (kh < 1 || kw < 1 || H % kh || W % kw)
        return kym_ops_fail(KYM_ERR_RUNTIME, "subsampling factor must divide the size");
    if (batch <= 0) return KYM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return launch_subsample<float>(x, y, batch, H, W, kh, kw, s);
    if (dtype == KYM_F64) return launch_subsample<double>(x, y, batch, H, W, kh, kw, s);
    return bad_dtype();
}

extern "C" int kym_op_unpad2d(const void* x, void* y, long long batch, int H, int W, int dtype, void* stream) {
    if (batch <= 0 || H < 3 || W < 3) return KYM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return launch_unpad2d<float>(x, y, batch, H, W, s);
    if (dtype == KYM_F64) return launch_unpad2d<double>(x, y, batch, H, W, s);
    return bad_dtype();
}

extern "C" int kym_op_filter_ifft_t(const void* x, const void* f, void* y, long long batch, long long N, int k,
                                    int mode, long long tlen, int dtype, void* stream) {
    if (k < 1 || N < 1 || N % k) return kym_ops_fail(KYM_ERR_RUNTIME, "subsampling factor must divide the size");
    if (N / k > (1 << 20)) return kym_ops_fail(KYM_ERR_UNSUPPORTED, "transform length too large for LDS");
    if (mode < 0 || mode > 2) return kym_ops_fail(KYM_ERR_VALUE, "mode must be 0 (complex), 1 (real) or 2 (modulus)");
    if (tlen < 0 || tlen > 0x7fffffffll || (tlen > 0 && batch % tlen))
        return kym_ops_fail(KYM_ERR_VALUE, "the rows must be a whole number of lines of tlen");
    if (batch <= 0) return KYM_OK;
    if (!x || !f || !y) return kym_ops_fail(KYM_ERR_VALUE, "null device pointer");
    hipStream_t s = (hipStream_t)stream;
    const FilterIfftRows rl = {(int)tlen, 1, 0, 0, 0};
    if (dtype == KYM_F32) return filter_ifft_op<float>(x, f, y, batch, N, k, mode, rl, s);
    if (dtype == KYM_F64) return filter_ifft_op<double>(x, f, y, batch, N, k, mode, rl, s);
    return bad_dtype();
}

extern "C" int kym_op_filter_ifft_g(const void* x, const void* f, void* y, long long rows, long long N, int k,
                                    int mode, int nf, int grp, int gtot, int goff, int dtype, void* stream) {
    if (k < 1 || N < 1 || N % k) return kym_ops_fail(KYM_ERR_RUNTIME, "subsampling factor must divide the size");
    if (N / k > (1 << 20)) return kym_ops_fail(KYM_ERR_UNSUPPORTED, "transform length too large for LDS");
    if (mode < 0 || mode > 2) return kym_ops_fail(KYM_ERR_VALUE, "mode must be 0 (complex), 1 (real) or 2 (modulus)");
    if (nf < 1 || rows % nf) return kym_ops_fail(KYM_ERR_VALUE, "rows must be a whole number of filter banks");
    if (grp < 0 || (grp > 0 && (rows % grp || goff < 0 || gtot < goff + grp)))
        return kym_ops_fail(KYM_ERR_VALUE, "invalid output row map");
    if (rows <= 0) return KYM_OK;
    if (!x || !f || !y) return kym_ops_fail(KYM_ERR_VALUE, "null device pointer");
    hipStream_t s = (hipStream_t)stream;
    const FilterIfftRows rl = {0, nf, grp, gtot, goff};
    if (dtype == KYM_F32) return filter_ifft_op<float>(x, f, y, rows, N, k, mode, rl, s);
    if (dtype == KYM_F64) return filter_ifft_op<double>(x, f, y, rows, N, k, mode, rl, s);
    return bad_dtype();
}

extern "C" int kym_op_filter_ifft(const void* x, const void* f, void* y, long long batch, long long N, int k,
                                  int mode, int dtype, void* stream) {
    return kym_op_filter_ifft_t(x, f, y, batch, N, k, mode, 0, dtype, stream);
}

extern "C" int kym_op_freq_scatter(const void* x, void* const* ys, const void* const* fs, const int* ks, int nf,
                                   long long B, int A, int T, int c, int N, int dtype, void* stream) {
    if (nf < 1 || nf > KYM_FS_MAX) return kym_ops_fail(KYM_ERR_VALUE, "between 1 and 16 wavelets per call");
    if (c != 1 && c != 2) return kym_ops_fail(KYM_ERR_VALUE, "input must be real (c = 1) or complex (c = 2)");
    if (A < 1 || A > N || T < 1 || N < 1) return kym_ops_fail(KYM_ERR_VALUE, "invalid sizes");
    if (!ys || !fs || !ks) return kym_ops_fail(KYM_ERR_VALUE, "null pointer array");
    for (int i = 0; i < nf; ++i) {
        if (ks[i] < 1 || N % ks[i]) return kym_ops_fail(KYM_ERR_RUNTIME, "subsampling factor must divide the size");
        if (!ys[i] || !fs[i]) return kym_ops_fail(KYM_ERR_VALUE, "null device pointer");
    }
    if (B <= 0) return KYM_OK;
    if (!x) return kym_ops_fail(KYM_ERR_VALUE, "null device pointer");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return freq_scatter_op<float>(x, ys, fs, ks, nf, B, A, T, c, N, s);
    if (dtype == KYM_F64) return freq_scatter_op<double>(x, ys, fs, ks, nf, B, A, T, c, N, s);
    return bad_dtype();
}

extern "C" int kym_op_lowpass(const void* x, const void* f, void* y, long long outer, long long N, long long S,
                              int k, int dtype, void* stream) {
    if (k < 1 || N < 1 || N % k) return kym_ops_fail(KYM_ERR_RUNTIME, "subsampling factor must divide the size");
    if (N > (1 << 20) || S < 1 || S > 0x7fffffffll) return kym_ops_fail(KYM_ERR_VALUE, "invalid line length or stride");
    if (outer <= 0) return KYM_OK;
    if (!x || !f || !y) return kym_ops_fail(KYM_ERR_VALUE, "null device pointer");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return lowpass_op<float>(x, f, y, outer, (int)N, (int)S, k, s);
    if (dtype == KYM_F64) return lowpass_op<double>(x, f, y, outer, (int)N, (int)S, k, s);
    return bad_dtype();
}

extern "C" int kym_op_fft(const void* x, void* y, void* work, long long batch, int H, int W, int kind, int dtype,
                          long long lds_limit, void* stream) {
    if (kind < 0 || kind > 3) return kym_ops_fail(KYM_ERR_VALUE, "kind must be 0..3");
    if (batch <= 0) return KYM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == KYM_F32) return fft_op<float>(x, y, work, batch, H, W, kind, s, lds_limit);
    if (dtype == KYM_F64) return fft_op<double>(x, y, work, batch, H, W, kind, s, lds_limit);
    return bad_dtype();
}
