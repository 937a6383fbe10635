// Compile-time specialised variants of the in-place LDS FFT passes: the
// length, the radix sequence, the batch layout and the thread count are
// template parameters, so every index division is by a constant and the
// pass loops have known trip counts.  Same arithmetic, same radix order
// (kym_factor_list) and same digit-reversal permutation as fft_core.h.
#pragma once
#include "fft_core.h"

// Full unrolling of per-thread iteration loops on the device only (the CPU
// harness runs a workgroup as one thread, where the trip count is the whole
// batch).
#if defined(__HIP_DEVICE_COMPILE__)
#define KYM_UNROLL_DEV _Pragma("unroll")
#else
#define KYM_UNROLL_DEV
#endif

// Digit reversal of a compile-time length (the map of kym_rev_map, i.e. the
// rev table): rev(p) = d_{K-1} + R_{K-1} (d_{K-2} + R_{K-2} (... + R_1 d_0)),
// with p's mixed-radix digits d_q = (p / m_q) mod R_q, m_q = R_0 ... R_{q-1}.
// Constant divisors, no table load.
template <int N>
KYM_HD KYM_INL int kym_rev_s(int p) {
    constexpr RadList rl = kym_factor_list(N);
    static_assert(rl.ok, "unsupported static FFT length");
    int value = 0, scale = 1, m = N;
#pragma unroll
    for (int q = rl.nrad - 1; q >= 0; --q) {
        m /= rl.rad[q] > 1 ? rl.rad[q] : 1;
        const int d = p / m;
        p -= d * m;
        value += d * scale;
        scale *= rl.rad[q] > 1 ? rl.rad[q] : 1;
    }
    return value;
}

// TWL: the twiddle table was copied to LDS (plain loads), else it is read
// from global memory.
template <bool TWL, class T>
KYM_HD KYM_INL cpx<T> tw_at(const cpx<T>* tw, int i) {
    if constexpr (TWL) return tw[i];
    else return kym_ldg(tw + i);
}

template <int R, bool INV, int N, int LP, int NTR, int ES, int TS, bool TFAST, int NT, class T, bool TWL = false>
KYM_HD KYM_INL void dit_pass_s(cpx<T>* buf, const cpx<T>* tw, int tid) {
    constexpr int NB = N / R, L = LP * R, TSTEP = N / L, TOTAL = NB * NTR, ST = LP * ES;
    for (int f = tid; f < TOTAL; f += NT) {
        int u, t;
        if constexpr (TFAST) { t = f % NTR; u = f / NTR; }
        else { u = f % NB; t = f / NB; }
        const int b = u / LP, j = u - b * LP;
        cpx<T>* p = buf + t * TS + (b * L + j) * ES;
        cpx<T> v[R];
#pragma unroll
        for (int m = 0; m < R; ++m) v[m] = p[m * ST];
        if constexpr (LP > 1) {
            if (j != 0) {
#pragma unroll
                for (int m = 1; m < R; ++m) {
                    const cpx<T> w = tw_at<TWL>(tw, j * m * TSTEP);
                    v[m] = INV ? cmulc(v[m], w) : cmul(v[m], w);
                }
            }
        }
        dft_r<R, INV>(v);
#pragma unroll
        for (int m = 0; m < R; ++m) p[m * ST] = v[m];
    }
}

template <int R, bool INV, int N, int L, int NTR, int ES, int TS, bool TFAST, int NT, class T, bool TWL = false>
KYM_HD KYM_INL void dif_pass_s(cpx<T>* buf, const cpx<T>* tw, int tid) {
    constexpr int NB = N / R, LS = L / R, TSTEP = N / L, TOTAL = NB * NTR, ST = LS * ES;
    for (int f = tid; f < TOTAL; f += NT) {
        int u, t;
        if constexpr (TFAST) { t = f % NTR; u = f / NTR; }
        else { u = f % NB; t = f / NB; }
        const int b = u / LS, j = u - b * LS;
        cpx<T>* p = buf + t * TS + (b * L + j) * ES;
        cpx<T> v[R];
#pragma unroll
        for (int m = 0; m < R; ++m) v[m] = p[m * ST];
        dft_r<R, INV>(v);
        if constexpr (LS > 1) {
            if (j != 0) {
#pragma unroll
                for (int m = 1; m < R; ++m) {
                    const cpx<T> w = tw_at<TWL>(tw, j * m * TSTEP);
                    v[m] = INV ? cmulc(v[m], w) : cmul(v[m], w);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < R; ++m) p[m * ST] = v[m];
    }
}

// Full DIT (digit-reversed in, natural out) over a batch, barrier after each pass.
template <bool INV, int N, int NTR, int ES, int TS, bool TFAST, int NT, bool TWL = false, class T, int Q = 0,
          int LP = 1>
KYM_HD KYM_INL void fft_dit_s(cpx<T>* buf, const cpx<T>* tw, int tid) {
    constexpr RadList rl = kym_factor_list(N);
    static_assert(rl.ok, "unsupported static FFT length");
    if constexpr (Q < rl.nrad) {
        constexpr int R = rl.rad[Q];
        dit_pass_s<R, INV, N, LP, NTR, ES, TS, TFAST, NT, T, TWL>(buf, tw, tid);
        wg_barrier();
        fft_dit_s<INV, N, NTR, ES, TS, TFAST, NT, TWL, T, Q + 1, LP * R>(buf, tw, tid);
    }
}

// Full DIF (natural in, digit-reversed out); passes in reverse radix order.
// QMIN = 1 stops before the last pass (see dif_last_pass_s).
template <bool INV, int N, int NTR, int ES, int TS, bool TFAST, int NT, bool TWL = false, class T, int Q = -1,
          int L = N, int QMIN = 0>
KYM_HD KYM_INL void fft_dif_s(cpx<T>* buf, const cpx<T>* tw, int tid) {
    constexpr RadList rl = kym_factor_list(N);
    static_assert(rl.ok, "unsupported static FFT length");
    constexpr int q = (Q < 0) ? rl.nrad - 1 : Q;
    if constexpr (q >= QMIN && L > 1) {
        constexpr int R = rl.rad[q];
        dif_pass_s<R, INV, N, L, NTR, ES, TS, TFAST, NT, T, TWL>(buf, tw, tid);
        wg_barrier();
        if constexpr (q > QMIN) fft_dif_s<INV, N, NTR, ES, TS, TFAST, NT, TWL, T, q - 1, L / R, QMIN>(buf, tw, tid);
    }
}

// The last DIF pass (radix r_1, no twiddles) of a batch whose results leave
// the buffer: each thread keeps pre(t, b, m, X) for its butterflies in
// registers, and after a barrier (every input of the pass has been read)
// hands them to post(t, b, m, value), which may overwrite the buffer.
// t is the transform, b the butterfly, m its element: storage index
// b r_1 + m along the transform, whose natural index is
// kym_rev_s<N>(b r_1) + (N / r_1) m.  Ends with a barrier.
// STAGE = false: post may not touch the buffer; each value goes to post
// right away (no registers held across a barrier).
template <bool INV, int N, int NTR, int ES, int TS, bool TFAST, int NT, class T, bool STAGE = true, class PRE,
          class POST>
KYM_HD KYM_INL void dif_last_pass_s(cpx<T>* buf, int tid, PRE pre, POST post) {
    constexpr RadList rl = kym_factor_list(N);
    constexpr int R = rl.rad[0], NB = N / R, TOTAL = NB * NTR, IT = (TOTAL + NT - 1) / NT;
    if constexpr (!STAGE) {
        for (int f = tid; f < TOTAL; f += NT) {
            int u, t;
            if constexpr (TFAST) { t = f % NTR; u = f / NTR; }
            else { u = f % NB; t = f / NB; }
            const cpx<T>* p = buf + t * TS + u * R * ES;
            cpx<T> v[R];
#pragma unroll
            for (int m = 0; m < R; ++m) v[m] = p[m * ES];
            dft_r<R, INV>(v);
#pragma unroll
            for (int m = 0; m < R; ++m) post(t, u, m, pre(t, u, m, v[m]));
        }
        wg_barrier();
        return;
    }
    T keep[IT][R];
    KYM_UNROLL_DEV
    for (int i = 0; i < IT; ++i) {
        const int f = tid + i * NT;
        if (f < TOTAL) {
            int u, t;
            if constexpr (TFAST) { t = f % NTR; u = f / NTR; }
            else { u = f % NB; t = f / NB; }
            const cpx<T>* p = buf + t * TS + u * R * ES;
            cpx<T> v[R];
#pragma unroll
            for (int m = 0; m < R; ++m) v[m] = p[m * ES];
            dft_r<R, INV>(v);
#pragma unroll
            for (int m = 0; m < R; ++m) keep[i][m] = pre(t, u, m, v[m]);
        }
    }
    wg_barrier();
    KYM_UNROLL_DEV
    for (int i = 0; i < IT; ++i) {
        const int f = tid + i * NT;
        if (f < TOTAL) {
            int u, t;
            if constexpr (TFAST) { t = f % NTR; u = f / NTR; }
            else { u = f % NB; t = f / NB; }
#pragma unroll
            for (int m = 0; m < R; ++m) post(t, u, m, keep[i][m]);
        }
    }
    wg_barrier();
}

// The first DIT pass (radix r_1, no twiddles) of a batch of real inputs
// that do not live in the buffer: element m of butterfly b of transform t
// (storage index b r_1 + m, natural index kym_rev_s<N>(b r_1) + (N / r_1) m)
// is load(t, b, m).  Every load happens before a barrier, then the
// butterflies are written to the buffer (which may overlap the source).
// Ends with a barrier.
// STAGE = false: the source does not overlap the buffer (no barrier
// between the loads and the writes).
template <bool INV, int N, int NTR, int ES, int TS, bool TFAST, int NT, class T, bool STAGE = true, class LOAD>
KYM_HD KYM_INL void dit_first_pass_real_s(cpx<T>* buf, int tid, LOAD load) {
    constexpr RadList rl = kym_factor_list(N);
    constexpr int R = rl.rad[0], NB = N / R, TOTAL = NB * NTR, IT = (TOTAL + NT - 1) / NT;
    if constexpr (!STAGE) {
        for (int f = tid; f < TOTAL; f += NT) {
            int u, t;
            if constexpr (TFAST) { t = f % NTR; u = f / NTR; }
            else { u = f % NB; t = f / NB; }
            cpx<T> v[R];
#pragma unroll
            for (int m = 0; m < R; ++m) v[m] = {load(t, u, m), T(0)};
            dft_r<R, INV>(v);
            cpx<T>* p = buf + t * TS + u * R * ES;
#pragma unroll
            for (int m = 0; m < R; ++m) p[m * ES] = v[m];
        }
        wg_barrier();
        return;
    }
    T keep[IT][R];
    KYM_UNROLL_DEV
    for (int i = 0; i < IT; ++i) {
        const int f = tid + i * NT;
        if (f < TOTAL) {
            int u, t;
            if constexpr (TFAST) { t = f % NTR; u = f / NTR; }
            else { u = f % NB; t = f / NB; }
#pragma unroll
            for (int m = 0; m < R; ++m) keep[i][m] = load(t, u, m);
        }
    }
    wg_barrier();
    KYM_UNROLL_DEV
    for (int i = 0; i < IT; ++i) {
        const int f = tid + i * NT;
        if (f < TOTAL) {
            int u, t;
            if constexpr (TFAST) { t = f % NTR; u = f / NTR; }
            else { u = f % NB; t = f / NB; }
            cpx<T> v[R];
#pragma unroll
            for (int m = 0; m < R; ++m) v[m] = {keep[i][m], T(0)};
            dft_r<R, INV>(v);
            cpx<T>* p = buf + t * TS + u * R * ES;
#pragma unroll
            for (int m = 0; m < R; ++m) p[m * ES] = v[m];
        }
    }
    wg_barrier();
}
