// Learned warm start for 1D reconstruction: a 1D conv decoder from
// scattering coefficients to signals, trained and run with hand-written
// kernels.  Convolutions are GEMMs with implicit im2col on bf16 MFMA
// (v_mfma_f32_16x16x16_bf16, FP32 accumulate); weights are kept in FP32 for
// Adam and packed to bf16 for each step.
//
// Layer: out[b, co, t] = bias[co] + sum_{ci, j<9} W[co, ci, j] x(b, ci, t + j - 4),
//   x(b, ci, s) = scale[ci] * in[b, ci, s / up] for 0 <= s < T (nearest
//   upsampling by `up`), 0 outside ('same' zero padding); optional ReLU.
// GEMM view: M = Cout, N = (b, t), K = (j, ci) with ci padded to Cin_pad.
#include <stdint.h>

#include <string>

#include "../../include/kymatio_hip.h"
#include "common.h"
#include "mfma.h"

int kym_ops_fail(int code, const std::string& msg);  // capi.hip

#define DEC_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t _e = (expr);                                                                           \
        if (_e != hipSuccess) return kym_ops_fail(4, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
    } while (0)

#define DEC_K 9
#define DEC_HALO 4
#define DEC_NT 64  // time positions per workgroup tile

// ---------------------------------------------------------------- forward / backward-data
// One workgroup: sample b, time tile [t0, t0 + 64), all M rows (Cout_pad,
// multiple of 16, <= 128).  Waves take 16x16 tiles round-robin.
struct ConvFwdArgs {
    const float* in;
    const uint16_t* w;  // (Cout_pad, 9, Cin_pad) bf16
    const float* bias;  // (Cout) or null
    const float* scale; // (Cin) or null
    float* out;         // (B, Cout, T)
    int B, Cin, Cin_pad, Cout, Cout_pad, T, up, relu, Tin;
};

__global__ void __launch_bounds__(256) k_conv_fwd(ConvFwdArgs a) {
    extern __shared__ __attribute__((aligned(16))) char dec_smem[];
    uint16_t* slab = reinterpret_cast<uint16_t*>(dec_smem);  // [72][Cin_pad]
    const int ntile = a.T / DEC_NT;
    const int b = blockIdx.x / ntile, t0 = (blockIdx.x - b * ntile) * DEC_NT;
    if (b >= a.B) return;
    const int tid = threadIdx.x;
    const int SW = DEC_NT + 2 * DEC_HALO;
    for (int i = tid; i < SW * a.Cin_pad; i += 256) {
        const int tp = i / a.Cin_pad, ci = i - tp * a.Cin_pad;
        const int s = t0 - DEC_HALO + tp;
        float v = 0.f;
        if (ci < a.Cin && s >= 0 && s < a.T) {
            v = a.in[((long long)b * a.Cin + ci) * a.Tin + s / a.up];
            if (a.scale) v *= a.scale[ci];
        }
        slab[i] = kym_f2bf(v);
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    const int mb_count = a.Cout_pad / 16;
    const int ntiles = mb_count * 4;
    const int lr = lane & 15, lq = lane >> 4;
    for (int tile = wave; tile < ntiles; tile += 4) {
        const int mb = tile >> 2, nb = tile & 3;
        kym_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const uint16_t* wrow = a.w + (long long)(mb * 16 + lr) * DEC_K * a.Cin_pad;
        const int n = nb * 16 + lr;
        for (int j = 0; j < DEC_K; ++j) {
            const uint16_t* wj = wrow + j * a.Cin_pad;
            const uint16_t* sj = slab + (n + j) * a.Cin_pad;
            for (int c0 = 0; c0 < a.Cin_pad; c0 += 16) {
                const int c = c0 + 4 * lq;
                kym_bf16x4 av = *reinterpret_cast<const kym_bf16x4*>(wj + c);
                kym_bf16x4 bv = *reinterpret_cast<const kym_bf16x4*>(sj + c);
                acc = kym_mfma16(av, bv, acc);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = mb * 16 + 4 * lq + i;
            if (m < a.Cout) {
                float v = acc[i] + (a.bias ? a.bias[m] : 0.f);
                if (a.relu && v < 0.f) v = 0.f;
                a.out[((long long)b * a.Cout + m) * a.T + t0 + nb * 16 + lr] = v;
            }
        }
    }
}

// ---------------------------------------------------------------- backward-weight
// dW[co][j][ci] += sum_t dOut[b, co, t] x(b, ci, t + j - 4) over a chunk of
// (b, t); one workgroup: one (j, 64-wide ci block) column tile, all Cout_pad
// rows, time tiles [kt0, kt0 + nkt) of 64.  Partial sums are added to dw
// with atomics.
struct ConvBwdWArgs {
    const float* dout;   // (B, Cout, T)
    const float* in;     // (B, Cin, Tin)
    const float* scale;  // (Cin) or null
    float* dw;           // (Cout, 9, Cin) fp32, accumulated
    int B, Cin, Cin_pad, Cout, Cout_pad, T, up, Tin;
    int ktiles;          // time tiles per workgroup
};

__global__ void __launch_bounds__(256) k_conv_bwd_w(ConvBwdWArgs a) {
    extern __shared__ __attribute__((aligned(16))) char dec_smem_w[];
    // xs[ci_local (64)][t (64)] bf16 and ds[co (Cout_pad)][t (64)] bf16
    uint16_t* xs = reinterpret_cast<uint16_t*>(dec_smem_w);
    uint16_t* ds = xs + 64 * DEC_NT;
    const int ncb = a.Cin_pad / 64;
    const int col_tiles = DEC_K * ncb;
    const int ct = blockIdx.x % col_tiles, chunk = blockIdx.x / col_tiles;
    const int j = ct / ncb, cb = ct - j * ncb;
    const int per_b = a.T / DEC_NT;
    const int total_tiles = a.B * per_b;
    const int tile0 = chunk * a.ktiles;
    if (tile0 >= total_tiles) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
    const int mb_count = a.Cout_pad / 16;
    const int ntiles = mb_count * 4;  // 16x16 output tiles (4 ci blocks of 16)
    kym_f32x4 acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = {0.f, 0.f, 0.f, 0.f};
    const int tile1 = (tile0 + a.ktiles < total_tiles) ? tile0 + a.ktiles : total_tiles;
    for (int tt = tile0; tt < tile1; ++tt) {
        const int b = tt / per_b, t0 = (tt - b * per_b) * DEC_NT;
        __syncthreads();
        for (int i = tid; i < 64 * DEC_NT; i += 256) {
            const int cl = i / DEC_NT, t = i - cl * DEC_NT;
            const int ci = cb * 64 + cl, s = t0 + t + j - DEC_HALO;
            float v = 0.f;
            if (ci < a.Cin && s >= 0 && s < a.T) {
                v = a.in[((long long)b * a.Cin + ci) * a.Tin + s / a.up];
                if (a.scale) v *= a.scale[ci];
            }
            xs[i] = kym_f2bf(v);
        }
        for (int i = tid; i < a.Cout_pad * DEC_NT; i += 256) {
            const int co = i / DEC_NT, t = i - co * DEC_NT;
            ds[i] = kym_f2bf(co < a.Cout ? a.dout[((long long)b * a.Cout + co) * a.T + t0 + t] : 0.f);
        }
        __syncthreads();
        int q = 0;
        for (int tile = wave; tile < ntiles; tile += 4, ++q) {
            const int mb = tile >> 2, nb = tile & 3;
            kym_f32x4 c = acc[q];
            for (int k0 = 0; k0 < DEC_NT; k0 += 16) {
                const int k = k0 + 4 * lq;
                kym_bf16x4 av = *reinterpret_cast<const kym_bf16x4*>(ds + (mb * 16 + lr) * DEC_NT + k);
                kym_bf16x4 bv = *reinterpret_cast<const kym_bf16x4*>(xs + (nb * 16 + lr) * DEC_NT + k);
                c = kym_mfma16(av, bv, c);
            }
            acc[q] = c;
        }
    }
    int q = 0;
    for (int tile = wave; tile < ntiles; tile += 4, ++q) {
        const int mb = tile >> 2, nb = tile & 3;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = mb * 16 + 4 * lq + i, ci = cb * 64 + nb * 16 + lr;
            if (co < a.Cout && ci < a.Cin) atomicAdd(a.dw + ((long long)co * DEC_K + j) * a.Cin + ci, acc[q][i]);
        }
    }
}

// ---------------------------------------------------------------- element-wise
// dIn[b, c, s] = (sum_{u < up} dX[b, c, s up + u]) * (a[b, c, s] > 0 if mask)
__global__ void k_up_relu_adjoint(const float* dx, const float* act, float* din, long long total, int Tin, int up,
                                  int mask) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const long long row = i / Tin;
        const int s = (int)(i - row * Tin);
        float acc = 0.f;
        for (int u = 0; u < up; ++u) acc += dx[row * Tin * up + (long long)s * up + u];
        if (mask && !(act[i] > 0.f)) acc = 0.f;
        din[i] = acc;
    }
}

// db[c] += sum over (b, t) of dOut[b, c, t]; one workgroup per (c, b-block)
__global__ void __launch_bounds__(256) k_bias_grad(const float* dout, float* db, int B, int C, int T) {
    __shared__ float red[256];
    const int c = blockIdx.x % C, bb = blockIdx.x / C;
    float acc = 0.f;
    for (int b = bb; b < B; b += gridDim.x / C)
        for (int t = threadIdx.x; t < T; t += 256) acc += dout[((long long)b * C + c) * T + t];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(db + c, red[0]);
}

// mean squared error over n elements: loss += sum (y-x)^2 / n ; dy = 2 (y-x) / n
__global__ void __launch_bounds__(256) k_mse(const float* y, const float* x, float* dy, float* loss, long long n) {
    __shared__ float red[256];
    float acc = 0.f;
    const float inv = 1.f / (float)n;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const float d = y[i] - x[i];
        if (dy) dy[i] = 2.f * d * inv;
        acc += d * d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] * inv);
}

// Adam (Kingma & Ba 2015) with bias correction, FP32 state
__global__ void k_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// FP32 master weights (Cout, 9, Cin) -> bf16 forward layout (Cout_pad, 9,
// Cin_pad) and the backward-data layout (Cin_pad, 9, Cout_pad) with j reversed.
__global__ void k_pack_w(const float* w, uint16_t* wf, uint16_t* wt, int Cout, int Cout_pad, int Cin, int Cin_pad) {
    const int total = Cout_pad * DEC_K * Cin_pad;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int co = i / (DEC_K * Cin_pad), r = i - co * DEC_K * Cin_pad, j = r / Cin_pad, ci = r - j * Cin_pad;
        const float v = (co < Cout && ci < Cin) ? w[((long long)co * DEC_K + j) * Cin + ci] : 0.f;
        wf[i] = kym_f2bf(v);
        if (wt) wt[((long long)ci * DEC_K + (DEC_K - 1 - j)) * Cout_pad + co] = kym_f2bf(v);
    }
}

// Synthetic chirp signals, one workgroup per signal: 1-8 chirps
// A cos(2 pi (f0 t + s t^2 / 2) + phase), f0 ~ U[0.01, 0.4], s ~ U[-1e-4, 1e-4],
// A log-uniform in [0.1, 1], phase ~ U[0, 2 pi), plus noise * N(0, 1).
// Randomness: splitmix64 of (seed, signal index, draw index); normals by Box-Muller.
__device__ __forceinline__ uint64_t dec_mix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double dec_unif(uint64_t seed, uint64_t sig, uint64_t k) {
    const uint64_t h = dec_mix(seed ^ dec_mix(sig * 0x100000001b3ull + k));
    return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__global__ void __launch_bounds__(256) k_chirps(float* x, int N, unsigned long long seed, long long index0,
                                                float noise) {
    const long long sig = index0 + blockIdx.x;
    __shared__ double par[8][4];
    __shared__ int nc;
    if (threadIdx.x == 0) {
        nc = 1 + (int)(dec_unif(seed, sig, 0) * 8.0);
        if (nc > 8) nc = 8;
        for (int c = 0; c < 8; ++c) {
            par[c][0] = 0.01 + 0.39 * dec_unif(seed, sig, 1 + 4 * c);
            par[c][1] = -1e-4 + 2e-4 * dec_unif(seed, sig, 2 + 4 * c);
            par[c][2] = exp(log(0.1) * (1.0 - dec_unif(seed, sig, 3 + 4 * c)));
            par[c][3] = 6.283185307179586 * dec_unif(seed, sig, 4 + 4 * c);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < N; t += 256) {
        double acc = 0.0;
        for (int c = 0; c < nc; ++c) {
            const double ph = par[c][0] * t + 0.5 * par[c][1] * (double)t * t;
            const double fr = ph - floor(ph);
            acc += par[c][2] * cos(6.283185307179586 * fr + par[c][3]);
        }
        const double u1 = dec_unif(seed, sig, 1000003ull + 2ull * t), u2 = dec_unif(seed, sig, 1000004ull + 2ull * t);
        const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        x[(long long)blockIdx.x * N + t] = (float)(acc + noise * g);
    }
}

// Lane-layout check of the bf16 MFMA builtins.  With the layouts of mfma.h,
// A[m][k] = (7m + 3k) % 17 and B[k][n] = (5k + 11n) % 13 (small integers,
// exact in bf16 and in the FP32 sums) give D = A B; the host compares each
// 16x16 D (row-major in out16 / out32) with its own product.  One wave.
__global__ void k_mfma_layout(float* out16, float* out32) {
    const int l = threadIdx.x, lr = l & 15, lq = l >> 4;
    kym_bf16x4 a4, b4;
    kym_bf16x8 a8, b8;
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * lq + i;
        a4[i] = (short)kym_f2bf((float)((7 * lr + 3 * k) % 17));
        b4[i] = (short)kym_f2bf((float)((5 * k + 11 * lr) % 13));
    }
    for (int i = 0; i < 8; ++i) {
        const int k = 8 * lq + i;
        a8[i] = __builtin_bit_cast(__bf16, kym_f2bf((float)((7 * lr + 3 * k) % 17)));
        b8[i] = __builtin_bit_cast(__bf16, kym_f2bf((float)((5 * k + 11 * lr) % 13)));
    }
    const kym_f32x4 z = {0.f, 0.f, 0.f, 0.f};
    const kym_f32x4 d4 = kym_mfma16(a4, b4, z);
    const kym_f32x4 d8 = kym_mfma32(a8, b8, z);
    for (int i = 0; i < 4; ++i) {
        out16[(4 * lq + i) * 16 + lr] = d4[i];
        out32[(4 * lq + i) * 16 + lr] = d8[i];
    }
}

// ---------------------------------------------------------------- C ABI
static unsigned dec_grid(long long n) {
    long long g = (n + 255) / 256;
    if (g > 65535ll * 16) g = 65535ll * 16;
    return (unsigned)(g < 1 ? 1 : g);
}

extern "C" int kym_dec_conv_fwd(const void* in_, const void* w_, const void* bias_, const void* scale_, void* out_,
                                int B, int Cin, int Cin_pad, int Cout, int Cout_pad, int T, int up, int relu,
                                void* stream) {
    const float* in = (const float*)in_;
    const uint16_t* w = (const uint16_t*)w_;
    const float* bias = (const float*)bias_;
    const float* scale = (const float*)scale_;
    float* out = (float*)out_;
    if (B <= 0) return 0;
    if (T % DEC_NT || Cin_pad % 16 || Cin_pad < Cin || Cout_pad % 16 || Cout_pad < Cout || Cout_pad > 128 || up < 1 ||
        T % up)
        return kym_ops_fail(3, "kym_dec_conv_fwd: unsupported shape");
    ConvFwdArgs a = {in, w, bias, scale, out, B, Cin, Cin_pad, Cout, Cout_pad, T, up, relu, T / up};
    const size_t lds = (size_t)(DEC_NT + 2 * DEC_HALO) * Cin_pad * 2;
    hipLaunchKernelGGL(k_conv_fwd, dim3((unsigned)(B * (T / DEC_NT))), dim3(256), lds, (hipStream_t)stream, a);
    DEC_TRY(hipGetLastError());
    return 0;
}

extern "C" int kym_dec_conv_bwd_weight(const void* dout_, const void* in_, const void* scale_, void* dw_, int B,
                                       int Cin, int Cin_pad, int Cout, int Cout_pad, int T, int up, int ktiles,
                                       void* stream) {
    const float* dout = (const float*)dout_;
    const float* in = (const float*)in_;
    const float* scale = (const float*)scale_;
    float* dw = (float*)dw_;
    if (B <= 0) return 0;
    if (T % DEC_NT || Cin_pad % 64 || Cin_pad < Cin || Cout_pad % 16 || Cout_pad < Cout || Cout_pad > 128 ||
        ktiles < 1 || up < 1 || T % up)
        return kym_ops_fail(3, "kym_dec_conv_bwd_weight: unsupported shape");
    ConvBwdWArgs a = {dout, in, scale, dw, B, Cin, Cin_pad, Cout, Cout_pad, T, up, T / up, ktiles};
    const long long total_tiles = (long long)B * (T / DEC_NT);
    const long long chunks = (total_tiles + ktiles - 1) / ktiles;
    const long long grid = chunks * DEC_K * (Cin_pad / 64);
    const size_t lds = (size_t)(64 + Cout_pad) * DEC_NT * 2;
    hipLaunchKernelGGL(k_conv_bwd_w, dim3((unsigned)grid), dim3(256), lds, (hipStream_t)stream, a);
    DEC_TRY(hipGetLastError());
    return 0;
}

extern "C" int kym_dec_up_relu_adjoint(const void* dx_, const void* act_, void* din_, int B, int C, int Tin, int up,
                                       int mask, void* stream) {
    const float* dx = (const float*)dx_;
    const float* act = (const float*)act_;
    float* din = (float*)din_;
    const long long total = (long long)B * C * Tin;
    if (total <= 0) return 0;
    hipLaunchKernelGGL(k_up_relu_adjoint, dim3(dec_grid(total)), dim3(256), 0, (hipStream_t)stream, dx, act, din,
                       total, Tin, up, mask);
    DEC_TRY(hipGetLastError());
    return 0;
}

extern "C" int kym_dec_bias_grad(const void* dout_, void* db_, int B, int C, int T, void* stream) {
    const float* dout = (const float*)dout_;
    float* db = (float*)db_;
    if (B <= 0) return 0;
    const int bblocks = B < 32 ? B : 32;
    hipLaunchKernelGGL(k_bias_grad, dim3(C * bblocks), dim3(256), 0, (hipStream_t)stream, dout, db, B, C, T);
    DEC_TRY(hipGetLastError());
    return 0;
}

extern "C" int kym_dec_mse(const void* y_, const void* x_, void* dy_, void* loss_, long long n, void* stream) {
    const float* y = (const float*)y_;
    const float* x = (const float*)x_;
    float* dy = (float*)dy_;
    float* loss = (float*)loss_;
    if (n <= 0) return 0;
    long long g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_mse, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, y, x, dy, loss, n);
    DEC_TRY(hipGetLastError());
    return 0;
}

extern "C" int kym_dec_adam(void* p_, const void* g_, void* m_, void* v_, long long n, float lr, float b1, float b2,
                            float eps, int step, void* stream) {
    float* p = (float*)p_;
    const float* g = (const float*)g_;
    float* m = (float*)m_;
    float* v = (float*)v_;
    if (n <= 0) return 0;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(dec_grid(n)), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, b1, b2, eps,
                       c1, c2);
    DEC_TRY(hipGetLastError());
    return 0;
}

extern "C" int kym_dec_pack_weights(const void* w_, void* wf_, void* wt_, int Cout, int Cout_pad, int Cin,
                                    int Cin_pad, void* stream) {
    const float* w = (const float*)w_;
    uint16_t* wf = (uint16_t*)wf_;
    uint16_t* wt = (uint16_t*)wt_;
    const long long total = (long long)Cout_pad * DEC_K * Cin_pad;
    hipLaunchKernelGGL(k_pack_w, dim3(dec_grid(total)), dim3(256), 0, (hipStream_t)stream, w, wf, wt, Cout, Cout_pad,
                       Cin, Cin_pad);
    DEC_TRY(hipGetLastError());
    return 0;
}

extern "C" int kym_chirps(void* x_, int B, int N, unsigned long long seed, long long index0, float noise,
                          void* stream) {
    float* x = (float*)x_;
    if (B <= 0) return 0;
    hipLaunchKernelGGL(k_chirps, dim3(B), dim3(256), 0, (hipStream_t)stream, x, N, seed, index0, noise);
    DEC_TRY(hipGetLastError());
    return 0;
}

extern "C" int kym_mfma_layout_check(void* out16_, void* out32_, void* stream) {
    float* out16 = (float*)out16_;
    float* out32 = (float*)out32_;
    hipLaunchKernelGGL(k_mfma_layout, dim3(1), dim3(64), 0, (hipStream_t)stream, out16, out32);
    DEC_TRY(hipGetLastError());
    return 0;
}
