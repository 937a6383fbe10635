// Mixed-radix in-place FFT passes over data held in LDS.
//
// Length n = r_1 r_2 ... r_K (radices 2, 4, 8, 16 and odd primes 3..31).
// Two in-place Cooley-Tukey variants share one digit-reversal permutation
// `rev` (rev[p] = natural index of the element at storage position p):
//   * DIT, passes r_1 .. r_K: input stored digit-reversed, output natural.
//   * DIF, passes r_K .. r_1: input natural, output stored digit-reversed.
// So an inverse DIF, a point-wise op and a forward DIT chain with no
// permutation at all, and loads/stores from HBM fold the permutation into
// their index math.  Each butterfly reads and writes the same R slots, so a
// pass needs one barrier and no second buffer.
//
// Twiddles come from a per-length table tw[t] = exp(-2 pi i t / n) computed
// in double on the host; the inverse uses the conjugate.  Transforms are
// unnormalised in both directions.
#pragma once
#include "common.h"
#include "radix_consts.h"

#define KYM_MAX_RADICES 12

template <class T>
struct FftLen {
    int n;
    int nrad;
    int rad[KYM_MAX_RADICES];  // DIT order (r_1 first)
    const cpx<T>* tw;          // n entries
    const int* rev;            // n entries: storage position -> natural index
    const int* irev;           // n entries: natural index -> storage position (inverse of rev)
};

// ---------------------------------------------------------------- butterflies
template <bool INV, class T> KYM_HD KYM_INL void dft2(cpx<T>* v) {
    cpx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <bool INV, class T> KYM_HD KYM_INL void dft4(cpx<T>& x0, cpx<T>& x1, cpx<T>& x2, cpx<T>& x3) {
    cpx<T> a0 = x0 + x2, a1 = x0 - x2, a2 = x1 + x3, a3 = mul_mi<INV>(x1 - x3);
    x0 = a0 + a2;
    x2 = a0 - a2;
    x1 = a1 + a3;
    x3 = a1 - a3;
}

template <bool INV, class T> KYM_HD KYM_INL void dft4(cpx<T>* v) { dft4<INV>(v[0], v[1], v[2], v[3]); }

// multiply by exp(-/+ 2 pi i m / 16) with the trivial angles special-cased
template <int M, bool INV, class T> KYM_HD KYM_INL cpx<T> tw16(cpx<T> a) {
    constexpr int m = M % 16;
    if constexpr (m == 0) return a;
    else if constexpr (m == 4) return mul_mi<INV>(a);
    else if constexpr (m == 8) return {-a.re, -a.im};
    else if constexpr (m == 12) return mul_mi<!INV>(a);
    else {
        // cos/sin(2 pi m / 16) for m in 0..15
        constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
        constexpr double cs[16] = {1, c1, h, s1, 0, -s1, -h, -c1, -1, -c1, -h, -s1, 0, s1, h, c1};
        constexpr double sn[16] = {0, s1, h, c1, 1, c1, h, s1, 0, -s1, -h, -c1, -1, -c1, -h, -s1};
        const cpx<T> w = {T(cs[m]), T(INV ? sn[m] : -sn[m])};
        return cmul(a, w);
    }
}

template <bool INV, class T> KYM_HD KYM_INL void dft8(cpx<T>* v) {
    cpx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cpx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    o1 = tw16<2, INV>(o1);
    o2 = tw16<4, INV>(o2);
    o3 = tw16<6, INV>(o3);
    v[0] = e0 + o0; v[4] = e0 - o0;
    v[1] = e1 + o1; v[5] = e1 - o1;
    v[2] = e2 + o2; v[6] = e2 - o2;
    v[3] = e3 + o3; v[7] = e3 - o3;
}

template <bool INV, class T> KYM_HD KYM_INL void dft16(cpx<T>* v) {
    // 4 x 4: inner DFTs over n1 for each n2, twiddle W16^(n2 k1), outer over n2
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) dft4<INV>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
    // after this, v[4*k1 + n2] holds y[n2][k1]
    v[5] = tw16<1, INV>(v[5]);
    v[6] = tw16<2, INV>(v[6]);
    v[7] = tw16<3, INV>(v[7]);
    v[9] = tw16<2, INV>(v[9]);
    v[10] = tw16<4, INV>(v[10]);
    v[11] = tw16<6, INV>(v[11]);
    v[13] = tw16<3, INV>(v[13]);
    v[14] = tw16<6, INV>(v[14]);
    v[15] = tw16<9, INV>(v[15]);
    cpx<T> t[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cpx<T> a = v[4 * k1], b = v[4 * k1 + 1], c = v[4 * k1 + 2], d = v[4 * k1 + 3];
        dft4<INV>(a, b, c, d);
        t[k1] = a;
        t[k1 + 4] = b;
        t[k1 + 8] = c;
        t[k1 + 12] = d;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = t[i];
}

// Odd prime P: pair inputs n and P-n, X_k and X_{P-k} share the cos/sin sums.
template <int P, bool INV, class T> KYM_HD KYM_INL void dft_odd(cpx<T>* v) {
    constexpr int H = (P - 1) / 2;
    cpx<T> a[H], b[H];
    cpx<T> x0 = v[0];
    cpx<T> s = x0;
#pragma unroll
    for (int n = 1; n <= H; ++n) {
        a[n - 1] = v[n] + v[P - n];
        b[n - 1] = v[n] - v[P - n];
        s = s + a[n - 1];
    }
    v[0] = s;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        T rc = x0.re, ic = x0.im, rs = T(0), is = T(0);
#pragma unroll
        for (int n = 1; n <= H; ++n) {
            const int m = (n * k) % P;
            const T c = T(KymCos<P>::v[m]);
            const T sn = T(KymSin<P>::v[m]);
            rc += a[n - 1].re * c;
            ic += a[n - 1].im * c;
            rs += b[n - 1].re * sn;
            is += b[n - 1].im * sn;
        }
        // forward: X_k = sum a cos - i b sin ; inverse flips the sin sign
        if (INV) {
            v[k] = {rc - is, ic + rs};
            v[P - k] = {rc + is, ic - rs};
        } else {
            v[k] = {rc + is, ic - rs};
            v[P - k] = {rc - is, ic + rs};
        }
    }
}

template <int R, bool INV, class T> KYM_HD KYM_INL void dft_r(cpx<T>* v) {
    if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v);
    else if constexpr (R == 8) dft8<INV>(v);
    else if constexpr (R == 16) dft16<INV>(v);
    else dft_odd<R, INV>(v);
}

// ---------------------------------------------------------------- passes
// A batch of `ntr` transforms of length n; element e of transform t lives at
// buf[t * ts + e * es].  `tfast` selects which index varies fastest across
// threads: the transform (column batches, ts == 1) or the butterfly (rows).
struct Batch {
    int ntr, es, ts;
    bool tfast;
};

template <int R, bool INV, class T>
KYM_HD void dit_pass(cpx<T>* buf, const FftLen<T>& fl, int Lp, const Batch& bt, int tid, int nth) {
    const int n = fl.n, nb = n / R, L = Lp * R, tstep = n / L;
    const int total = nb * bt.ntr;
    const cpx<T>* tw = fl.tw;
    for (int f = tid; f < total; f += nth) {
        int u, t;
        if (bt.tfast) { t = f % bt.ntr; u = f / bt.ntr; }
        else { u = f % nb; t = f / nb; }
        const int b = u / Lp, j = u - b * Lp;
        cpx<T>* p = buf + (size_t)t * bt.ts + (size_t)(b * L + j) * bt.es;
        const int st = Lp * bt.es;
        cpx<T> v[R];
#pragma unroll
        for (int m = 0; m < R; ++m) v[m] = p[m * st];
        if (j != 0) {
#pragma unroll
            for (int m = 1; m < R; ++m) {
                const cpx<T> w = kym_ldg(tw + j * m * tstep);
                v[m] = INV ? cmulc(v[m], w) : cmul(v[m], w);
            }
        }
        dft_r<R, INV>(v);
#pragma unroll
        for (int m = 0; m < R; ++m) p[m * st] = v[m];
    }
}

template <int R, bool INV, class T>
KYM_HD void dif_pass(cpx<T>* buf, const FftLen<T>& fl, int L, const Batch& bt, int tid, int nth) {
    const int n = fl.n, nb = n / R, Ls = L / R, tstep = n / L;
    const int total = nb * bt.ntr;
    const cpx<T>* tw = fl.tw;
    for (int f = tid; f < total; f += nth) {
        int u, t;
        if (bt.tfast) { t = f % bt.ntr; u = f / bt.ntr; }
        else { u = f % nb; t = f / nb; }
        const int b = u / Ls, j = u - b * Ls;
        cpx<T>* p = buf + (size_t)t * bt.ts + (size_t)(b * L + j) * bt.es;
        const int st = Ls * bt.es;
        cpx<T> v[R];
#pragma unroll
        for (int m = 0; m < R; ++m) v[m] = p[m * st];
        dft_r<R, INV>(v);
        if (j != 0) {
#pragma unroll
            for (int m = 1; m < R; ++m) {
                const cpx<T> w = kym_ldg(tw + j * m * tstep);
                v[m] = INV ? cmulc(v[m], w) : cmul(v[m], w);
            }
        }
#pragma unroll
        for (int m = 0; m < R; ++m) p[m * st] = v[m];
    }
}

// Radix dispatch.  MAXP is the largest odd prime the kernel is built for, so a
// kernel only carries the butterflies its plan can use (large odd radices
// cost registers for every path through the kernel).
#define KYM_RADIX_SWITCH(R_, CALL)                                   \
    switch (R_) {                                                    \
        case 2: CALL(2); break;                                      \
        case 4: CALL(4); break;                                      \
        case 8: CALL(8); break;                                      \
        case 16: CALL(16); break;                                    \
        case 3: if constexpr (MAXP >= 3) { CALL(3); } break;         \
        case 5: if constexpr (MAXP >= 5) { CALL(5); } break;         \
        case 7: if constexpr (MAXP >= 7) { CALL(7); } break;         \
        case 11: if constexpr (MAXP >= 11) { CALL(11); } break;      \
        case 13: if constexpr (MAXP >= 13) { CALL(13); } break;      \
        case 17: if constexpr (MAXP >= 17) { CALL(17); } break;      \
        case 19: if constexpr (MAXP >= 19) { CALL(19); } break;      \
        case 23: if constexpr (MAXP >= 23) { CALL(23); } break;      \
        case 29: if constexpr (MAXP >= 29) { CALL(29); } break;      \
        case 31: if constexpr (MAXP >= 31) { CALL(31); } break;      \
        default: break;                                              \
    }

// Largest odd prime factor of n (1 if n is a power of two).
KYM_HD KYM_INL int kym_max_odd_prime(int n) {
    while (n % 2 == 0 && n > 1) n /= 2;
    int best = 1;
    for (int p = 3; p * p <= n; p += 2)
        while (n % p == 0) { best = p; n /= p; }
    if (n > 1) best = n > best ? n : best;
    return best;
}

// Kernel families by prime set: 2-only/3/5/7 -> 7, up to 17 -> 17, up to 31 -> 31.
KYM_HD KYM_INL int kym_prime_class(int maxp) { return maxp <= 7 ? 7 : maxp <= 17 ? 17 : 31; }

// Full DIT transform (digit-reversed in, natural out); barrier after each pass.
template <bool INV, int MAXP, class T>
KYM_HD void fft_dit(cpx<T>* buf, const FftLen<T>& fl, const Batch& bt, int tid, int nth) {
    int Lp = 1;
    for (int q = 0; q < fl.nrad; ++q) {
        const int R = fl.rad[q];
#define KYM_DIT_CALL(RR) dit_pass<RR, INV, T>(buf, fl, Lp, bt, tid, nth)
        KYM_RADIX_SWITCH(R, KYM_DIT_CALL)
#undef KYM_DIT_CALL
        Lp *= R;
        wg_barrier();
    }
}

// Full DIF transform (natural in, digit-reversed out); barrier after each pass.
template <bool INV, int MAXP, class T>
KYM_HD void fft_dif(cpx<T>* buf, const FftLen<T>& fl, const Batch& bt, int tid, int nth) {
    int L = fl.n;
    for (int q = fl.nrad - 1; q >= 0; --q) {
        const int R = fl.rad[q];
#define KYM_DIF_CALL(RR) dif_pass<RR, INV, T>(buf, fl, L, bt, tid, nth)
        KYM_RADIX_SWITCH(R, KYM_DIF_CALL)
#undef KYM_DIF_CALL
        L /= R;
        wg_barrier();
    }
}

// Factor n into supported radices (DIT order): odd primes first (their
// butterflies then need no twiddles), then powers of two as 16/8/4/2.
// max_pow2 (16, 8, 4 or 2) caps the power-of-two radices (a tuning choice;
// 16 is the default every compile-time specialisation assumes).
// constexpr so compile-time specialisations use exactly the runtime order.
struct RadList {
    int n = 0;
    int nrad = 0;
    int rad[KYM_MAX_RADICES] = {0};
    bool ok = false;
};

KYM_HD constexpr RadList kym_factor_list(int n, int max_pow2 = 16) {
    RadList rl;
    rl.n = n;
    int k = 0;
    int m = n;
    const int odd[10] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31};
    for (int i = 0; i < 10; ++i) {
        const int p = odd[i];
        while (m % p == 0) {
            if (k >= KYM_MAX_RADICES) return rl;
            rl.rad[k++] = p;
            m /= p;
        }
    }
    int a = 0;
    while (m % 2 == 0) { m /= 2; ++a; }
    if (m != 1) return rl;
    const int dmax = max_pow2 >= 16 ? 4 : max_pow2 >= 8 ? 3 : max_pow2 >= 4 ? 2 : 1;
    while (a > 0) {
        int d = 1;
        if (dmax == 4) {
            if (a >= 8 || a == 4 || a == 7) d = 4;
            else if (a == 6 || a == 3 || a == 5) d = 3;
            else if (a == 2) d = 2;
        } else {
            d = a < dmax ? a : dmax;
            if (a - d == 1 && d > 2) d -= 1;  // 4 x 4 rather than 8 x 2
        }
        if (k >= KYM_MAX_RADICES) return rl;
        rl.rad[k++] = 1 << d;
        a -= d;
    }
    if (n == 1) rl.rad[k++] = 1;
    rl.nrad = k;
    rl.ok = true;
    return rl;
}

inline bool kym_factor(int n, int* rad, int* nrad, int max_pow2 = 16) {
    const RadList rl = kym_factor_list(n, max_pow2);
    if (!rl.ok) return false;
    for (int i = 0; i < rl.nrad; ++i) rad[i] = rl.rad[i];
    *nrad = rl.nrad;
    return true;
}
