// Bodies of the per-op FFT kernels (ops.hip), written as workgroup
// functions of (tid, nthreads) like the scattering bodies so the CPU test
// harness can run them.  MAXP: the largest odd prime radix their run-time
// FFTs handle (7, 17 or 31, kym_prime_class of the lengths): the butterfly
// registers, and with them occupancy and spills, follow it.
#pragma once
#include "fft_core.h"
#include "fft_static.h"

// ------------------------------------------------------------ FFT op
// kind: 0 c2c forward, 1 c2c inverse (1/n), 2 r2c forward (real input),
// 3 inverse, real part out (1/n).  A batch of 2D transforms of H x W (W = 1
// gives 1D along H).  Small sizes: one workgroup per transform in LDS.
// Large sizes: row pass then column pass through `y` (in place).
struct FftOpArgs {
    const void* x;
    void* y;
    long long batch;
    int H, W, kind;
    int r0;     // row/column block start (pass kernels)
    int nrow;   // rows or columns per task
    int nblk;   // tasks per transform (pass kernels)
};

template <class T, bool INV>
KYM_HD void load_2d(const FftOpArgs& o, const FftLen<T>& fH, const FftLen<T>& fW, long long b, cpx<T>* lds, int tid,
                        int nth) {
    const int H = o.H, W = o.W, n = H * W;
    for (int p = tid; p < n; p += nth) {
        const int py = p / W, px = p - (p / W) * W;
        // forward (DIT) wants digit-reversed storage; inverse (DIF) natural
        const int y = INV ? py : fH.rev[py];
        const int x = INV ? px : fW.rev[px];
        const long long src = b * n + (long long)y * W + x;
        if (o.kind == 2) lds[p] = {static_cast<const T*>(o.x)[src], T(0)};
        else lds[p] = static_cast<const cpx<T>*>(o.x)[src];
    }
}

// One transform per workgroup, whole in LDS.
template <class T, int MAXP = 31>
KYM_HD void fftop_lds_body(const FftOpArgs& o, const FftLen<T>* lens, int lH, int lW, long long b, cpx<T>* lds,
                           int tid, int nth) {
    const FftLen<T>& fH = lens[lH];
    const FftLen<T>& fW = lens[lW];
    const int H = o.H, W = o.W, n = H * W;
    const bool inv = o.kind == 1 || o.kind == 3;
    if (W == 1 && !inv) {
        // 1D forward: coalesced reads, each value placed at its digit-reversed slot
        for (int p = tid; p < n; p += nth) {
            const long long src = b * n + p;
            lds[fH.irev[p]] = o.kind == 2 ? cpx<T>{static_cast<const T*>(o.x)[src], T(0)}
                                          : static_cast<const cpx<T>*>(o.x)[src];
        }
    } else if (inv) {
        load_2d<T, true>(o, fH, fW, b, lds, tid, nth);
    } else {
        load_2d<T, false>(o, fH, fW, b, lds, tid, nth);
    }
    wg_barrier();
    if (inv) {
        if (W > 1) fft_dif<true, MAXP>(lds, fW, Batch{H, 1, W, false}, tid, nth);
        fft_dif<true, MAXP>(lds, fH, Batch{W, W, 1, true}, tid, nth);
        const T inv_n = T(1) / T(n);
        if (W == 1) {
            // 1D inverse: coalesced writes, each read from its digit-reversed slot
            for (int p = tid; p < n; p += nth) {
                const cpx<T> v = lds[fH.irev[p]];
                if (o.kind == 3) static_cast<T*>(o.y)[b * n + p] = v.re * inv_n;
                else static_cast<cpx<T>*>(o.y)[b * n + p] = v * inv_n;
            }
            return;
        }
        for (int p = tid; p < n; p += nth) {
            const int py = p / W, px = p - (p / W) * W;
            const long long dst = b * n + (long long)fH.rev[py] * W + fW.rev[px];
            if (o.kind == 3) static_cast<T*>(o.y)[dst] = lds[p].re * inv_n;
            else static_cast<cpx<T>*>(o.y)[dst] = lds[p] * inv_n;
        }
    } else {
        if (W > 1) fft_dit<false, MAXP>(lds, fW, Batch{H, 1, W, false}, tid, nth);
        fft_dit<false, MAXP>(lds, fH, Batch{W, W, 1, true}, tid, nth);
        for (int p = tid; p < n; p += nth) static_cast<cpx<T>*>(o.y)[b * n + p] = lds[p];
    }
}

// Transforms per workgroup of the multi-transform route for (H, W) over a
// batch: about 4096 points per workgroup for 1D lengths up to 1024; 0 where
// the one-transform-per-workgroup route applies.
inline long long fftop_multi_count(int H, int W, long long batch) {
    if (W != 1 || H > 1024 || H < 1 || batch < 1) return 0;
    const long long per = 4096 / H;
    return per < batch ? per : batch;
}

// Short 1D transforms (W = 1), several per workgroup: transforms
// b0 .. b0 + m - 1 (m = o.nrow, fewer at the end of the batch) sit one after
// the other in LDS and run as one batch of m rows, so a workgroup keeps its
// threads busy on lengths far below its thread count (the frequency axis of
// time-frequency scattering: tens of points, batches of 10^5).
template <class T, int MAXP = 31>
KYM_HD void fftop_multi_body(const FftOpArgs& o, const FftLen<T>* lens, int lH, long long blk, cpx<T>* lds, int tid,
                             int nth) {
    const FftLen<T>& f = lens[lH];
    const int n = o.H;
    const long long b0 = blk * o.nrow;
    const int m = (int)((o.batch - b0 < o.nrow) ? o.batch - b0 : o.nrow);
    const bool inv = o.kind == 1 || o.kind == 3;
    const int total = m * n;
    // reads and writes walk the rows in natural order (coalesced); the
    // forward (DIT) stores each value at its digit-reversed slot, the inverse
    // (DIF) reads its results from there
    for (int q = tid; q < total; q += nth) {
        const int t = q / n, p = q - t * n;
        const long long src = (b0 + t) * n + p;
        const int slot = t * n + (inv ? p : f.irev[p]);
        if (o.kind == 2) lds[slot] = {static_cast<const T*>(o.x)[src], T(0)};
        else lds[slot] = static_cast<const cpx<T>*>(o.x)[src];
    }
    wg_barrier();
    if (inv) {
        fft_dif<true, MAXP>(lds, f, Batch{m, 1, n, false}, tid, nth);
        const T inv_n = T(1) / T(n);
        for (int q = tid; q < total; q += nth) {
            const int t = q / n, p = q - t * n;
            const cpx<T> v = lds[t * n + f.irev[p]];
            if (o.kind == 3) static_cast<T*>(o.y)[b0 * n + q] = v.re * inv_n;
            else static_cast<cpx<T>*>(o.y)[b0 * n + q] = v * inv_n;
        }
    } else {
        fft_dit<false, MAXP>(lds, f, Batch{m, 1, n, false}, tid, nth);
        for (int q = tid; q < total; q += nth) static_cast<cpx<T>*>(o.y)[b0 * n + q] = lds[q];
    }
}

// Filter, periodise and inverse-transform in one pass (1D):
//   y[b] = ifft_n(P_k(x[b] * f)),  n = N / k,  P_k(v)[i] = (1/k) sum_a v[i + a n]
// for rows b of x (batch, N) complex and a real filter f (N); the result is
// kept complex (mode 0), its real part (mode 1) or its modulus (mode 2).  The
// reference's ifft(subsample_fourier(cdgmm(x, f), k)) (and irfft / modulus of
// it) as three ops and two intermediate tensors; here the periodised product
// is formed while loading (natural order, as the inverse DIF wants it) and
// o.nrow rows share a workgroup (the last one may hold fewer).
struct FilterIfftArgs {
    const void* x;
    const void* f;
    void* y;
    long long batch;
    int N, k, n, mode;
    int nrow;    // rows per workgroup
    int split;   // alias subsets summed by separate threads (a power of two dividing k)
    // 0: y is (batch, n); tlen > 0: the rows are (b, t) with t < tlen and y is
    // laid out (batch / tlen, n, tlen), the transform axis before t (the
    // workgroup's rows then lie within one b: nrow divides tlen)
    int tlen;
    // nf > 1: a bank of nf filters (nf, N); row r filters x row r / nf with
    // filter r mod nf.  grp > 0 (tlen = 0): row r is written to output row
    // (r / grp) gtot + goff + r mod grp (rows of a (B, grp) block placed in a
    // (B, gtot) output at goff).
    int nf, grp, gtot, goff;
};

KYM_HD KYM_INL long long filter_ifft_out_row(const FilterIfftArgs& o, long long r) {
    if (o.grp <= 0) return r;
    const long long b = r / o.grp;
    return b * o.gtot + o.goff + (r - b * o.grp);
}

// LDS: the m x n transform buffer, then (split > 1) split x m x n partial sums.
// Thread q < split m n sums the aliases a = s, s + split, ... (s = q / (m n))
// of point q mod (m n); the partials are added in the order s = 0, 1, ...
template <class T, int MAXP = 31>
KYM_HD void filter_ifft_body(const FilterIfftArgs& o, const FftLen<T>& fl, long long blk, cpx<T>* lds, int tid,
                             int nth) {
    const int n = o.n, N = o.N, k = o.k, S = o.split;
    const long long b0 = blk * o.nrow;
    const int m = (int)((o.batch - b0 < o.nrow) ? o.batch - b0 : o.nrow);
    const cpx<T>* x = static_cast<const cpx<T>*>(o.x);
    const T* f = static_cast<const T*>(o.f);
    const T inv_k = T(1) / T(k);
    const int total = m * n;
    cpx<T>* part = lds + (long long)o.nrow * n;
    for (int q = tid; q < S * total; q += nth) {
        const int s = q / total, e = q - s * total;
        const int t = e / n, p = e - t * n;
        const long long row = b0 + t;
        const cpx<T>* xr = x + (o.nf > 1 ? row / o.nf : row) * N + p;
        const T* fr = o.nf > 1 ? f + (row % o.nf) * N : f;
        cpx<T> acc = {T(0), T(0)};
        int a = s;
        // eight aliases per round: their loads are issued before the first
        // use (the sum keeps the order a = s, s + S, ...)
        for (; a + 7 * S < k; a += 8 * S) {
            T w[8];
            cpx<T> v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                w[u] = kym_ldg(fr + p + (a + u * S) * n);
                v[u] = kym_ldg(xr + (long long)(a + u * S) * n);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                acc.re += v[u].re * w[u];
                acc.im += v[u].im * w[u];
            }
        }
        for (; a < k; a += S) {
            const T w = fr[p + a * n];
            const cpx<T> v = xr[(long long)a * n];
            acc.re += v.re * w;
            acc.im += v.im * w;
        }
        if (S == 1) lds[e] = acc * inv_k;
        else part[q] = acc;
    }
    if (S > 1) {
        wg_barrier();
        for (int e = tid; e < total; e += nth) {
            cpx<T> acc = part[e];
            for (int s = 1; s < S; ++s) acc = acc + part[s * total + e];
            lds[e] = acc * inv_k;
        }
    }
    wg_barrier();
    fft_dif<true, MAXP>(lds, fl, Batch{m, 1, n, false}, tid, nth);
    const T inv_n = T(1) / T(n);
    for (int q = tid; q < total; q += nth) {  // coalesced writes from the digit-reversed slots
        int t, p;
        long long dst;
        if (o.tlen > 0) {  // row t of this block fastest: (b, p, t0 + t)
            p = q / m;
            t = q - p * m;
            const long long r0 = b0 / o.tlen, t0 = b0 - r0 * o.tlen;
            dst = (r0 * n + p) * o.tlen + t0 + t;
        } else {
            t = q / n;
            p = q - t * n;
            dst = filter_ifft_out_row(o, b0 + t) * n + p;
        }
        const cpx<T> v = lds[t * n + fl.irev[p]];
        if (o.mode == 0) static_cast<cpx<T>*>(o.y)[dst] = v * inv_n;
        else if (o.mode == 1) static_cast<T*>(o.y)[dst] = v.re * inv_n;
        else static_cast<T*>(o.y)[dst] = kym_sqrt(v.re * v.re + v.im * v.im) * inv_n;
    }
}

// filter_ifft_body with the transform length n = NN and the rows per
// workgroup M fixed at compile time (no alias split, one filter, natural or
// transposed rows as FilterIfftArgs says; with tlen > 0, M divides tlen).
// Rows beyond the batch gather zeros and are not stored.  Same arithmetic
// as filter_ifft_body without a split.
template <class T, int NN, int M, int NT>
KYM_HD void filter_ifft_body_s(const FilterIfftArgs& o, const FftLen<T>& fl, long long blk, cpx<T>* lds, int tid) {
    const int N = o.N, k = o.k;
    const long long b0 = blk * M;
    const int m = (int)((o.batch - b0 < M) ? o.batch - b0 : M);
    const cpx<T>* x = static_cast<const cpx<T>*>(o.x);
    const T* f = static_cast<const T*>(o.f);
    const T inv_k = T(1) / T(k);
    for (int q = tid; q < M * NN; q += NT) {
        const int t = q / NN, p = q - t * NN;
        cpx<T> acc = {T(0), T(0)};
        if (t < m) {
            const cpx<T>* xr = x + (b0 + t) * N + p;
            int a = 0;
            for (; a + 7 < k; a += 8) {
                T w[8];
                cpx<T> v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    w[u] = kym_ldg(f + p + (a + u) * NN);
                    v[u] = kym_ldg(xr + (long long)(a + u) * NN);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    acc.re += v[u].re * w[u];
                    acc.im += v[u].im * w[u];
                }
            }
            for (; a < k; ++a) {
                const T w = kym_ldg(f + p + a * NN);
                const cpx<T> v = kym_ldg(xr + (long long)a * NN);
                acc.re += v.re * w;
                acc.im += v.im * w;
            }
        }
        lds[q] = acc * inv_k;
    }
    wg_barrier();
    fft_dif_s<true, NN, M, 1, NN, false, NT>(lds, fl.tw, tid);
    const T inv_n = T(1) / T(NN);
    for (int q = tid; q < m * NN; q += NT) {
        int t, p;
        long long dst;
        if (o.tlen > 0) {
            p = q / m;
            t = q - p * m;
            const long long r0 = b0 / o.tlen, t0 = b0 - r0 * o.tlen;
            dst = (r0 * NN + p) * o.tlen + t0 + t;
        } else {
            t = q / NN;
            p = q - t * NN;
            dst = b0 * NN + q;
        }
        const cpx<T> v = lds[t * NN + kym_ldg(fl.irev + p)];
        if (o.mode == 0) static_cast<cpx<T>*>(o.y)[dst] = v * inv_n;
        else if (o.mode == 1) static_cast<T*>(o.y)[dst] = v.re * inv_n;
        else static_cast<T*>(o.y)[dst] = kym_sqrt(v.re * v.re + v.im * v.im) * inv_n;
    }
}

// Compile-time filter_ifft shapes X(id, n, rows per workgroup): the
// frequency axis of the time-frequency transform (padded to 88, and 22
// after a subsampling by 4).
#define KYM_FILTER_IFFT_SPECS(X) \
    X(0, 88, 32)                 \
    X(1, 22, 128)

// The compile-time shape for output length n over `batch` rows (one filter,
// no row map; tlen a multiple of its rows), or -1.
inline int filter_ifft_spec_id(int n, long long batch, const FilterIfftArgs& o, int* rows) {
    if (o.nf > 1 || o.grp > 0) return -1;
#define KYM_FI_MATCH(id, nn, mm)                                                   \
    if (n == nn && batch >= mm && (o.tlen == 0 || o.tlen % mm == 0)) {             \
        *rows = mm;                                                                \
        return id;                                                                 \
    }
    KYM_FILTER_IFFT_SPECS(KYM_FI_MATCH)
#undef KYM_FI_MATCH
    return -1;
}

// Frequency scattering of one (b, t-block) in one workgroup: the rows x[b][a][t]
// (a < A along the log-frequency axis, t < T; real (c = 1) or complex
// (c = 2)), zero-padded to N along a, are transformed along a; then for every
// wavelet i: Y = |ifft_n(P_k(X f_i))|, n = N / k_i, stored as y_i[b][p][t].
// That is the swap, pad_frequency, cfft / rfft, and per wavelet filter_ifft,
// swap and modulus of kymatio/scattering1d/core/timefrequency_scattering.py:
// 244-267 with the modulus of :409, in one pass over x.
#define KYM_FS_MAX 16
struct FreqScatterArgs {
    const void* x;
    long long B;
    int A, T, c, N, nf, m, nt_blk;  // m t-values per workgroup, nt_blk workgroups per b
    const void* f[KYM_FS_MAX];
    void* y[KYM_FS_MAX];
    int k[KYM_FS_MAX];
    const void* lenN;               // FftLen<T> of N
    const void* lenn[KYM_FS_MAX];   // FftLen<T> of N / k[i]
};

// LDS: m x N spectrum, then m x N work rows.
template <class T, int MAXP>
KYM_HD void freq_scatter_body(const FreqScatterArgs& o, long long blk, cpx<T>* lds, int tid, int nth) {
    const FftLen<T>& fN = *static_cast<const FftLen<T>*>(o.lenN);
    const int N = o.N, A = o.A, TT = o.T, m = o.m;
    const long long b = blk / o.nt_blk;
    const int t0 = (int)(blk - b * o.nt_blk) * m;
    const int mm = (TT - t0 < m) ? TT - t0 : m;
    cpx<T>* X = lds;
    cpx<T>* P = lds + (long long)m * N;
    // rows t of the block along a, each value at its digit-reversed slot
    for (int i = tid; i < m * N; i += nth) {
        const int a = i / m, t = i - a * m;
        cpx<T> v = {T(0), T(0)};
        if (a < A && t < mm) {
            const long long src = (b * A + a) * (long long)TT + t0 + t;
            if (o.c == 2) v = kym_ldg(static_cast<const cpx<T>*>(o.x) + src);
            else v = {kym_ldg(static_cast<const T*>(o.x) + src), T(0)};
        }
        X[t * N + kym_ldg(fN.irev + a)] = v;
    }
    wg_barrier();
    fft_dit<false, MAXP>(X, fN, Batch{m, 1, N, false}, tid, nth);
    wg_barrier();
    for (int w = 0; w < o.nf; ++w) {
        const int k = o.k[w], n = N / k;
        const T* f = static_cast<const T*>(o.f[w]);
        const T inv_k = T(1) / T(k);
        for (int q = tid; q < m * n; q += nth) {
            const int t = q / n, p = q - t * n;
            const cpx<T>* xr = X + t * N + p;
            cpx<T> acc = {T(0), T(0)};
            for (int a = 0; a < k; ++a) {
                const T fv = kym_ldg(f + p + a * n);
                const cpx<T> v = xr[a * n];
                acc.re += v.re * fv;
                acc.im += v.im * fv;
            }
            P[q] = acc * inv_k;
        }
        wg_barrier();
        const FftLen<T>& fl = *static_cast<const FftLen<T>*>(o.lenn[w]);
        fft_dif<true, MAXP>(P, fl, Batch{m, 1, n, false}, tid, nth);
        wg_barrier();
        const T inv_n = T(1) / T(n);
        T* y = static_cast<T*>(o.y[w]);
        for (int q = tid; q < mm * n; q += nth) {
            const int p = q / mm, t = q - p * mm;
            const cpx<T> v = P[t * n + kym_ldg(fl.irev + p)];
            y[(b * n + p) * (long long)TT + t0 + t] = kym_sqrt(v.re * v.re + v.im * v.im) * inv_n;
        }
        wg_barrier();
    }
}

// freq_scatter_body with N and the t-values per workgroup M fixed at compile
// time (static passes) and the output lengths N / k among 88, 44 and 22 as
// compiled (freq_scatter_spec_ok).  Same arithmetic.
template <class T, int N, int M, int NT>
KYM_HD void freq_scatter_body_s(const FreqScatterArgs& o, long long blk, cpx<T>* lds, int tid) {
    const FftLen<T>& fN = *static_cast<const FftLen<T>*>(o.lenN);
    const int A = o.A, TT = o.T;
    const long long b = blk / o.nt_blk;
    const int t0 = (int)(blk - b * o.nt_blk) * M;
    const int mm = (TT - t0 < M) ? TT - t0 : M;
    cpx<T>* X = lds;
    cpx<T>* P = lds + M * N;
    for (int i = tid; i < M * N; i += NT) {
        const int a = i / M, t = i - a * M;
        cpx<T> v = {T(0), T(0)};
        if (a < A && t < mm) {
            const long long src = (b * A + a) * (long long)TT + t0 + t;
            if (o.c == 2) v = kym_ldg(static_cast<const cpx<T>*>(o.x) + src);
            else v = {kym_ldg(static_cast<const T*>(o.x) + src), T(0)};
        }
        X[t * N + kym_ldg(fN.irev + a)] = v;
    }
    wg_barrier();
    fft_dit_s<false, N, M, 1, N, false, NT>(X, fN.tw, tid);
    for (int w = 0; w < o.nf; ++w) {
        const int k = o.k[w], n = N / k;
        const T* f = static_cast<const T*>(o.f[w]);
        const T inv_k = T(1) / T(k);
        for (int q = tid; q < M * n; q += NT) {
            const int t = q / n, p = q - t * n;
            const cpx<T>* xr = X + t * N + p;
            cpx<T> acc = {T(0), T(0)};
            for (int a = 0; a < k; ++a) {
                const T fv = kym_ldg(f + p + a * n);
                const cpx<T> v = xr[a * n];
                acc.re += v.re * fv;
                acc.im += v.im * fv;
            }
            P[q] = acc * inv_k;
        }
        wg_barrier();
        const FftLen<T>& fl = *static_cast<const FftLen<T>*>(o.lenn[w]);
        if constexpr (N == 88) {
            if (n == 88) fft_dif_s<true, 88, M, 1, 88, false, NT>(P, fl.tw, tid);
            else if (n == 44) fft_dif_s<true, 44, M, 1, 44, false, NT>(P, fl.tw, tid);
            else fft_dif_s<true, 22, M, 1, 22, false, NT>(P, fl.tw, tid);
        }
        const T inv_n = T(1) / T(n);
        T* y = static_cast<T*>(o.y[w]);
        for (int q = tid; q < mm * n; q += NT) {
            const int p = q / mm, t = q - p * mm;
            const cpx<T> v = P[t * n + kym_ldg(fl.irev + p)];
            y[(b * n + p) * (long long)TT + t0 + t] = kym_sqrt(v.re * v.re + v.im * v.im) * inv_n;
        }
        wg_barrier();
    }
}

// The compile-time freq_scatter applies: N = 88, 32 t-values per workgroup,
// every output length 88, 44 or 22.
inline bool freq_scatter_spec_ok(const FreqScatterArgs& o) {
    if (o.N != 88 || o.m != 32) return false;
    for (int i = 0; i < o.nf; ++i)
        if (o.k[i] != 1 && o.k[i] != 2 && o.k[i] != 4) return false;
    return true;
}

// t-values per workgroup of freq_scatter: about 4096 points, a power of two,
// at most T.
inline int freq_scatter_rows(int N, int T) {
    int m = 1;
    while (m * 2 * N <= 4096 && m * 2 <= T) m *= 2;
    return m;
}

// Launch shape of filter_ifft for output length n over `batch` rows: threads
// (1024 for n >= 2048, else 256), rows per workgroup (about 4096 points, but
// fewer when that leaves under 512 workgroups) and the alias split (threads
// beyond the points sum alias subsets).  Returns the complex elements of LDS.
// With tlen > 0 the rows per workgroup divide tlen (FilterIfftArgs).
inline long long filter_ifft_shape(int n, int k, long long batch, int* threads, int* rows, int* split, int tlen = 0) {
    const int nt = n >= 2048 ? 1024 : 256;
    long long m = n >= 4096 ? 1 : 4096 / n;
    const long long want = (batch + 511) / 512;
    if (m > want) m = want;
    if (m > batch) m = batch;
    if (m < 1) m = 1;
    if (tlen > 0)
        while (tlen % m) --m;
    int S = 1;
    while (S * 2 <= k && (long long)S * 2 * m * n <= nt) S *= 2;
    *threads = nt;
    *rows = (int)m;
    *split = S;
    return m * n * (S > 1 ? 1 + S : 1);
}

// Low-pass of real lines: y = Re ifft_n(P_k(fft_N(x) f)), the irfft of the
// periodised product of the line's spectrum with the real filter f (what
// rfft, cdgmm, subsample_fourier and irfft compute one after the other,
// kymatio/scattering1d/backend/torch_backend.py:118-136), n = N / k.  x is
// (outer, N, S) and y (outer, n, S): the lines run along the middle axis
// (S = 1: contiguous lines).  A workgroup takes nrow lines: consecutive
// outer indices (S = 1) or consecutive s of one outer index (S > 1).
struct LowpassArgs {
    const void* x;
    const void* f;
    void* y;
    long long outer;
    int N, S, k, n;
    int nrow;  // lines per workgroup
    int nsb;   // workgroups per outer index (S > 1), else 1
};

// LDS: nrow x N spectrum, then nrow x n periodised spectrum.
template <class T, int MAXP = 31>
KYM_HD void lowpass_body(const LowpassArgs& o, const FftLen<T>& fN, const FftLen<T>& fn, long long blk, cpx<T>* lds,
                         int tid, int nth) {
    const int N = o.N, n = o.n, k = o.k, S = o.S;
    long long ob;
    int s0, m;
    if (S == 1) {
        ob = blk * o.nrow;
        s0 = 0;
        m = (int)((o.outer - ob < o.nrow) ? o.outer - ob : o.nrow);
    } else {
        ob = blk / o.nsb;
        s0 = (int)(blk - ob * o.nsb) * o.nrow;
        m = (S - s0 < o.nrow) ? S - s0 : o.nrow;
    }
    const T* x = static_cast<const T*>(o.x);
    const T* f = static_cast<const T*>(o.f);
    T* y = static_cast<T*>(o.y);
    // line t, sample e: x[(ob + t) N + e] (S = 1) or x[(ob N + e) S + s0 + t];
    // stored at its digit-reversed slot for the forward DIT
    for (int i = tid; i < m * N; i += nth) {
        int t, e;
        long long src;
        if (S == 1) {
            t = i / N;
            e = i - t * N;
            src = (ob + t) * N + e;
        } else {
            e = i / m;
            t = i - e * m;
            src = (ob * N + e) * S + s0 + t;
        }
        lds[t * N + fN.irev[e]] = {kym_ldg(x + src), T(0)};
    }
    wg_barrier();
    fft_dit<false, MAXP>(lds, fN, Batch{m, 1, N, false}, tid, nth);
    wg_barrier();
    cpx<T>* P = lds + (long long)o.nrow * N;
    const T inv_k = T(1) / T(k);
    for (int q = tid; q < m * n; q += nth) {
        const int t = q / n, p = q - t * n;
        const cpx<T>* X = lds + t * N + p;
        cpx<T> acc = {T(0), T(0)};
        int a = 0;
        for (; a + 7 < k; a += 8) {  // the filter loads of a round issued together
            T w[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) w[u] = kym_ldg(f + p + (a + u) * n);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const cpx<T> v = X[(a + u) * n];
                acc.re += v.re * w[u];
                acc.im += v.im * w[u];
            }
        }
        for (; a < k; ++a) {
            const T w = kym_ldg(f + p + a * n);
            const cpx<T> v = X[a * n];
            acc.re += v.re * w;
            acc.im += v.im * w;
        }
        P[q] = acc * inv_k;
    }
    wg_barrier();
    fft_dif<true, MAXP>(P, fn, Batch{m, 1, n, false}, tid, nth);
    wg_barrier();
    const T inv_n = T(1) / T(n);
    for (int q = tid; q < m * n; q += nth) {
        int t, p;
        long long dst;
        if (S == 1) {
            t = q / n;
            p = q - t * n;
            dst = (ob + t) * n + p;
        } else {
            p = q / m;
            t = q - p * m;
            dst = (ob * n + p) * S + s0 + t;
        }
        y[dst] = P[t * n + fn.irev[p]].re * inv_n;
    }
}

// lowpass_body for contiguous lines (S = 1) with the lengths fixed at
// compile time: M = 8192 / N real lines of N per workgroup, two per complex
// transform (line u and line u + M / 2 as z = a + i b; their spectra are
// A[q] = (Z[q] + conj(Z[-q])) / 2 and B[q] = (Z[q] - conj(Z[-q])) / (2 i),
// formed inside the alias sums), by the static passes.  Lines beyond
// `outer` load zeros and are not stored.  The same sums as lowpass_body, the
// line spectra to rounding.
template <class T, int N, int NN, int NT>
KYM_HD void lowpass_body_s(const LowpassArgs& o, const FftLen<T>& fN, const FftLen<T>& fn, long long blk,
                           cpx<T>* lds, int tid) {
    constexpr int M = 8192 / N, M2 = M / 2, K = N / NN;
    const long long ob = blk * M;
    const int m = (int)((o.outer - ob < M) ? o.outer - ob : M);
    const T* x = static_cast<const T*>(o.x);
    const T* f = static_cast<const T*>(o.f);
    T* y = static_cast<T*>(o.y);
    for (int i = tid; i < M2 * N; i += NT) {
        const int u = i / N, e = i - u * N;
        const T va = u < m ? kym_ldg(x + (ob + u) * N + e) : T(0);
        const T vb = u + M2 < m ? kym_ldg(x + (ob + u + M2) * N + e) : T(0);
        lds[u * N + kym_ldg(fN.irev + e)] = {va, vb};
    }
    wg_barrier();
    fft_dit_s<false, N, M2, 1, N, false, NT>(lds, fN.tw, tid);
    cpx<T>* P = lds + M2 * N;
    const T inv_2k = T(1) / T(2 * K);
    for (int q = tid; q < M * NN; q += NT) {
        const int t = q / NN, p = q - t * NN;
        const bool second = t >= M2;
        const cpx<T>* Z = lds + (second ? t - M2 : t) * N;
        cpx<T> acc = {T(0), T(0)};
#pragma unroll 8
        for (int a = 0; a < K; ++a) {
            const int i = p + a * NN;
            const T w = kym_ldg(f + i);
            const cpx<T> z = Z[i], zm = Z[i == 0 ? 0 : N - i];
            // 2 A = (z.re + zm.re, z.im - zm.im); 2 B = (z.im + zm.im, zm.re - z.re)
            const T re = second ? z.im + zm.im : z.re + zm.re;
            const T im = second ? zm.re - z.re : z.im - zm.im;
            acc.re += re * w;
            acc.im += im * w;
        }
        P[q] = acc * inv_2k;
    }
    wg_barrier();
    fft_dif_s<true, NN, M, 1, NN, false, NT>(P, fn.tw, tid);
    const T inv_n = T(1) / T(NN);
    for (int q = tid; q < m * NN; q += NT) {
        const int t = q / NN, p = q - t * NN;
        y[(ob + t) * NN + p] = P[t * NN + kym_ldg(fn.irev + p)].re * inv_n;
    }
}

// lowpass_body for strided lines (S > 1) with the lengths and the lines per
// workgroup M fixed at compile time: M consecutive s of one outer index (S a
// multiple of M).  Same arithmetic as lowpass_body.
template <class T, int N, int NN, int M, int NT>
KYM_HD void lowpass_body_sv(const LowpassArgs& o, const FftLen<T>& fN, const FftLen<T>& fn, long long blk,
                            cpx<T>* lds, int tid) {
    constexpr int K = N / NN;
    const int S = o.S, nsb = S / M;
    const long long ob = blk / nsb;
    const int s0 = (int)(blk - ob * nsb) * M;
    const T* x = static_cast<const T*>(o.x);
    const T* f = static_cast<const T*>(o.f);
    T* y = static_cast<T*>(o.y);
    for (int i = tid; i < M * N; i += NT) {
        const int e = i / M, t = i - e * M;
        lds[t * N + kym_ldg(fN.irev + e)] = {kym_ldg(x + (ob * N + e) * S + s0 + t), T(0)};
    }
    wg_barrier();
    fft_dit_s<false, N, M, 1, N, false, NT>(lds, fN.tw, tid);
    cpx<T>* P = lds + M * N;
    const T inv_k = T(1) / T(K);
    for (int q = tid; q < M * NN; q += NT) {
        const int t = q / NN, p = q - t * NN;
        const cpx<T>* X = lds + t * N + p;
        cpx<T> acc = {T(0), T(0)};
#pragma unroll
        for (int a = 0; a < K; ++a) {
            const T w = kym_ldg(f + p + a * NN);
            const cpx<T> v = X[a * NN];
            acc.re += v.re * w;
            acc.im += v.im * w;
        }
        P[q] = acc * inv_k;
    }
    wg_barrier();
    fft_dif_s<true, NN, M, 1, NN, false, NT>(P, fn.tw, tid);
    const T inv_n = T(1) / T(NN);
    for (int q = tid; q < M * NN; q += NT) {
        const int p = q / M, t = q - p * M;
        y[(ob * NN + p) * S + s0 + t] = P[t * NN + kym_ldg(fn.irev + p)].re * inv_n;
    }
}

// Compile-time low-pass shapes X(id, N, n, lines per workgroup M, strided):
// time averaging of the time-frequency transform at the default averaging
// (outputs of 128; M = 8192 / N contiguous lines, two per transform) and
// frequency averaging along the padded n1 axis (88 or 22 points, M lines of
// stride S, S a multiple of M).
#define KYM_LOWPASS_SPECS(X)  \
    X(0, 4096, 128, 2, 0)     \
    X(1, 2048, 128, 4, 0)     \
    X(2, 1024, 128, 8, 0)     \
    X(3, 512, 128, 16, 0)     \
    X(4, 256, 128, 32, 0)     \
    X(5, 88, 22, 32, 1)       \
    X(6, 22, 22, 64, 1)

inline int lowpass_spec_id(int N, int n, int S, int* rows) {
#define KYM_LP_MATCH(id, nn, n2, mm, sv)                                       \
    if (N == nn && n == n2 && (sv ? (S > 1 && S % mm == 0) : S == 1)) {        \
        *rows = mm;                                                            \
        return id;                                                             \
    }
    KYM_LOWPASS_SPECS(KYM_LP_MATCH)
#undef KYM_LP_MATCH
    return -1;
}

// Complex elements of LDS of a compile-time low-pass: two lines per
// transform for contiguous lines (S = 1), one per line along a stride.
inline long long lowpass_spec_lds(int N, int n, int S, int M) {
    return S == 1 ? (long long)(M / 2) * N + (long long)M * n : (long long)M * (N + n);
}

template <class T, int SPEC, bool EMU>
KYM_HD KYM_INL void lowpass_spec_body(const LowpassArgs& o, const FftLen<T>& fN, const FftLen<T>& fn, long long blk,
                                      cpx<T>* lds, int tid) {
#define KYM_LP_BODY(id, nn, n2, mm, sv)                                                          \
    if constexpr (SPEC == id) {                                                                  \
        if constexpr (sv) lowpass_body_sv<T, nn, n2, mm, (EMU ? 1 : 256)>(o, fN, fn, blk, lds, tid); \
        else lowpass_body_s<T, nn, n2, (EMU ? 1 : 256)>(o, fN, fn, blk, lds, tid);               \
    }
    KYM_LOWPASS_SPECS(KYM_LP_BODY)
#undef KYM_LP_BODY
}

// Launch shape of lowpass: lines per workgroup (about 4096 points, at most
// S, a power of two for S > 1, fewer when that leaves under 512 workgroups)
// and workgroups per outer index.  Returns the complex elements of LDS.
inline long long lowpass_shape(int N, int n, int S, long long outer, int* rows, int* nsb, long long* nblk) {
    long long m = N >= 4096 ? 1 : 4096 / N;
    const long long lines = outer * S;
    const long long want = (lines + 511) / 512;
    if (m > want) m = want;
    if (m < 1) m = 1;
    if (S > 1) {
        if (m > S) m = S;
        long long p2 = 1;
        while (p2 * 2 <= m) p2 *= 2;
        m = p2;
        *nsb = (int)((S + m - 1) / m);
        *nblk = outer * *nsb;
    } else {
        if (m > outer) m = outer;
        *nsb = 1;
        *nblk = (outer + m - 1) / m;
    }
    *rows = (int)m;
    return m * (N + n);
}

// Pass kernels for transforms that do not fit: axis 1 = rows (along W),
// axis 0 = columns (along H).  Each task reads whole lines of `src` into LDS,
// transforms them and writes them to `dst` (complex unless dst_real; src
// real only for the first pass of kind 2).  A real output (kind 3) keeps its
// complex intermediate in a separate work buffer, never in the output.
struct FftPassIO {
    const void* src;
    void* dst;
    int src_real, dst_real;
    int scale_last;  // multiply by 1/(H W) on this pass
};

template <class T, int MAXP = 31>
KYM_HD void fftop_pass_body(const FftOpArgs& o, const FftPassIO& io, const FftLen<T>* lens, int lH, int lW, int axis,
                            long long task, cpx<T>* lds, int tid, int nth) {
    const long long b = task / o.nblk;
    const int blk = (int)(task - b * o.nblk);
    if (b >= o.batch) return;
    const int H = o.H, W = o.W;
    const long long n = (long long)H * W;
    const bool inv = o.kind == 1 || o.kind == 3;
    const FftLen<T>& f = lens[axis == 1 ? lW : lH];
    const int len = f.n, c0 = blk * o.nrow;
    const int lines = axis == 1 ? H : W;
    const int nt = lines - c0 < o.nrow ? lines - c0 : o.nrow;
    // lds[t * len + e]: line t of this task, element e
    for (int i = tid; i < nt * len; i += nth) {
        const int t = i / len, e = i - (i / len) * len;
        const int src_e = inv ? e : f.rev[e];
        const long long idx = axis == 1 ? (long long)(c0 + t) * W + src_e : (long long)src_e * W + c0 + t;
        if (io.src_real) lds[i] = {static_cast<const T*>(io.src)[b * n + idx], T(0)};
        else lds[i] = static_cast<const cpx<T>*>(io.src)[b * n + idx];
    }
    wg_barrier();
    if (inv) fft_dif<true, MAXP>(lds, f, Batch{nt, 1, len, false}, tid, nth);
    else fft_dit<false, MAXP>(lds, f, Batch{nt, 1, len, false}, tid, nth);
    const T sc = io.scale_last ? T(1) / T(n) : T(1);
    for (int i = tid; i < nt * len; i += nth) {
        const int t = i / len, e = i - (i / len) * len;
        const int dst_e = inv ? f.rev[e] : e;
        const long long idx = axis == 1 ? (long long)(c0 + t) * W + dst_e : (long long)dst_e * W + c0 + t;
        const cpx<T> v = lds[i] * sc;
        if (io.dst_real) static_cast<T*>(io.dst)[b * n + idx] = v.re;
        else static_cast<cpx<T>*>(io.dst)[b * n + idx] = v;
    }
}


// Launch sequence of one per-op FFT (shared by the GPU entry point and the
// CPU harness).  route 0: whole transform in LDS (grid = batch); route 1: a
// line pass along `axis` (grid = batch * nblk).
struct FftOpStep {
    int route, axis;
    FftPassIO io;
    int nrow, nblk;
    size_t lds;
};

// Returns 0, or 3 (bad argument) / 5 (unsupported) with *err set.
inline int fftop_steps(const void* x, void* y, void* work, int H, int W, int kind, size_t lim, size_t cs,
                       FftOpStep* steps, int* nsteps, const char** err) {
    *nsteps = 0;
    const size_t need = (size_t)H * W * cs;
    if (need <= lim) {
        FftOpStep st = {};
        st.route = 0;
        st.lds = need;
        steps[(*nsteps)++] = st;
        return 0;
    }
    int axes[2], npass = 0;
    if (W > 1) axes[npass++] = 1;
    if (H > 1) axes[npass++] = 0;
    void* mid = y;
    if (kind == 3 && npass > 1) {
        if (!work) { *err = "kind 3 at this size needs a complex work buffer"; return 3; }
        mid = work;
    }
    for (int q = 0; q < npass; ++q) {
        const int axis = axes[q];
        const int len = axis == 1 ? W : H, lines = axis == 1 ? H : W;
        if ((size_t)len * cs > lim) { *err = "FFT length too large for LDS"; return 5; }
        FftOpStep st = {};
        st.route = 1;
        st.axis = axis;
        size_t per = lim / ((size_t)len * cs);
        if (per > 16) per = 16;
        if (per < 1) per = 1;
        st.nrow = (int)per;
        st.nblk = (lines + st.nrow - 1) / st.nrow;
        st.lds = (size_t)st.nrow * len * cs;
        const bool first = q == 0, last = q == npass - 1;
        st.io.src = first ? x : mid;
        st.io.dst = last ? y : mid;
        st.io.src_real = (first && kind == 2) ? 1 : 0;
        st.io.dst_real = (last && kind == 3) ? 1 : 0;
        st.io.scale_last = (last && (kind == 1 || kind == 3)) ? 1 : 0;
        steps[(*nsteps)++] = st;
    }
    return 0;
}
