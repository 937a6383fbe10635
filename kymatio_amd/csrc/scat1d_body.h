// Scattering1D: node tree, kernel arguments and workgroup bodies.
//
// Same tree as the 2D path, on the generator of
// kymatio/scattering1d/core/scattering1d.py:2-107:
//   root     U_0 = fft(pad(x))                                  length N_pad
//   order 1  U_1 = |ifft(P_2^k1(U_0 * psi1[n1].levels[0]))|        res k1
//   order 2  U_2 = |ifft(P_2^k2(fft(U_1) * psi2[n2].levels[k1]))|  res k1 + k2
// with k1 = min(j1, log2_stride) (j1 without local averaging) and
// k2 = max(min(j2, log2_stride) - k1, 0).  Each node emits one output path:
//   local average  S = Re(ifft(P_kl(fft(U) * phi.levels[res]))), kl = 2^max(log2_stride-res, 0),
//                  unpadded to [ind_start[log2_stride], ind_end[log2_stride])
//   no average     U itself, unpadded at [ind_start[r], ind_end[r]), r = max(j - oversampling, 0)
//   global         the sum of U over time (the DC bin of fft(U))
// written at out[signal * out_row + out_off + t].
#pragma once
#include "fft_static.h"

#define KYM_MAX_RES1 24

enum Avg1D { AVG1_LOCAL = 0, AVG1_NONE = 1, AVG1_GLOBAL = 2 };

struct Node1D {
    int order;       // 0 root, 1, 2
    int res;         // resolution index (length N_pad >> res)
    int parent;      // node index, -1 for the root
    int filt;        // packed-filter index of the wavelet at the parent's resolution
    int k;           // periodisation factor from the parent
    int leaf_filt;   // phi.levels[res] (local average)
    int leaf_k;      // low-pass periodisation factor (local average)
    int out_off;     // offset of this path in a signal's output row
    int unpad_res;   // border-index resolution of the unpad
    int need_fwd;    // forward FFT of |U| needed (children or local leaf or global sum)
    long long slot;  // spectrum slot in the per-signal workspace, -1 none
    // training
    int child_begin, child_end;  // node-index range of the children
    long long save;              // saved u (digit-reversed, unnormalised), read-only in the backward
    long long gsave;             // dL/dX (natural) written by the backward
    long long gpad;              // root: dL/dpad(x), real
};

template <class T>
struct Scat1DArgs {
    const Node1D* nodes;
    const FftLen<T>* lens;
    const long long* filt_off;
    const int* group;
    int ngroup;
    int len[KYM_MAX_RES1], n[KYM_MAX_RES1];
    int ind_start[KYM_MAX_RES1], ind_end[KYM_MAX_RES1];
    int mode, log2_stride;
    const T* x;
    int N, pad_left;
    const T* filt;
    cpx<T>* ws;
    long long ws_sig;
    T* out;
    long long out_row;
    int img0, ntask;
    // training
    int save;
    const T* dS;
    T* gx;
};

// Output of one path from the spectrum / modulus in LDS.
template <class T, int MAXP>
KYM_HD void leaf1d(const Scat1DArgs<T>& a, const Node1D& nd, const cpx<T>* U, int n, int img, cpx<T>* B, int tid,
                   int nth) {
    T* o = a.out + (long long)img * a.out_row + nd.out_off;
    if (a.mode == AVG1_GLOBAL) {
        if (tid == 0) o[0] = U[0].re;
        return;
    }
    const int kk = nd.leaf_k, nl = n / kk;
    const T* phi = a.filt + a.filt_off[nd.leaf_filt];
    const T inv_k = T(1) / T(kk);
    for (int i = tid; i < nl; i += nth) {
        cpx<T> acc = {T(0), T(0)};
#pragma unroll 8
        for (int q = 0; q < kk; ++q) {  // unrolled: several filter loads in flight
            const cpx<T> u = U[i + q * nl];
            const T f = kym_ldg(phi + i + q * nl);
            acc.re += u.re * f;
            acc.im += u.im * f;
        }
        B[i] = acc * inv_k;
    }
    wg_barrier();
    // the low-passed path has length N_pad >> max(res, log2_stride) = nl
    const int lr = nd.res > a.log2_stride ? nd.res : a.log2_stride;
    const FftLen<T>& fB = a.lens[a.len[lr]];
    fft_dif<true, MAXP>(B, fB, Batch{1, 1, nl, false}, tid, nth);
    const int s = a.ind_start[a.log2_stride], e = a.ind_end[a.log2_stride];
    const T inv_n = T(1) / T(nl);
    for (int p = tid; p < nl; p += nth) {
        const int t = fB.rev[p];
        if (t >= s && t < e) o[t - s] = B[p].re * inv_n;
    }
    wg_barrier();
}

// FFT of a node's length: run-time (Dyn1D) or compile-time (Stat1D, length
// N and NT threads; the same factorisation and digit-reversed order).
template <class T, int MAXP>
struct Dyn1D {
    KYM_HD KYM_INL void inv(cpx<T>* b, const FftLen<T>& fl, int tid, int nth) const {
        fft_dif<true, MAXP>(b, fl, Batch{1, 1, fl.n, false}, tid, nth);
    }
    KYM_HD KYM_INL void fwd(cpx<T>* b, const FftLen<T>& fl, int tid, int nth) const {
        fft_dit<false, MAXP>(b, fl, Batch{1, 1, fl.n, false}, tid, nth);
    }
};

template <class T, int N, int NT>
struct Stat1D {
    KYM_HD KYM_INL void inv(cpx<T>* b, const FftLen<T>& fl, int tid, int) const {
        fft_dif_s<true, N, 1, 1, N, false, NT>(b, fl.tw, tid);
    }
    KYM_HD KYM_INL void fwd(cpx<T>* b, const FftLen<T>& fl, int tid, int) const {
        fft_dit_s<false, N, 1, 1, N, false, NT>(b, fl.tw, tid);
    }
};

template <class T, int MAXP, class F>
KYM_HD KYM_INL void node1d_body(const F& ff, const Scat1DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int il = task / a.ngroup, gi = task - il * a.ngroup;
    const int img = a.img0 + il;
    const Node1D& nd = a.nodes[a.group[gi]];
    const int r = nd.res, n = a.n[r];
    const FftLen<T>& fl = a.lens[a.len[r]];
    cpx<T>* wsi = a.ws + il * a.ws_sig;
    T* o = a.out + (long long)img * a.out_row + nd.out_off;
    if (nd.order == 0) {
        const T* xin = a.x + (long long)img * a.N;
        copy_batched<8>(
            n, [&](int p) { return kym_ldg(xin + reflect_index(fl.rev[p] - a.pad_left, a.N)); },
            [&](int p, T v) { lds[p] = {v, T(0)}; }, tid, nth);
        if (a.mode == AVG1_NONE) {
            // S_0 = U_0 unpadded = the input itself
            const int s = a.ind_start[0], e = a.ind_end[0];
            for (int t = s + tid; t < e; t += nth) o[t - s] = xin[reflect_index(t - a.pad_left, a.N)];
        }
        wg_barrier();
        ff.fwd(lds, fl, tid, nth);
    } else {
        const Node1D& pn = a.nodes[nd.parent];
        const cpx<T>* U = wsi + pn.slot;
        const T* f = a.filt + a.filt_off[nd.filt];
        const int k = nd.k;
        const T inv_k = T(1) / T(k);
        // mirror pairs: the parent is the spectrum of a real signal, so point
        // n - i reads the conjugates of point i's aliases in reverse order
        // (alias q of n - i is conj(U[i + (k-1-q) n])); one load serves both.
        // i = 0 and i = n/2 are their own mirrors.
        const int half = n / 2;
        for (int i = tid; i <= half; i += nth) {
            const int im = i == 0 ? 0 : n - i;
            cpx<T> d = {T(0), T(0)}, m = {T(0), T(0)};
            if (k == 1) {
                const cpx<T> u = kym_ldg(U + i);
                const T fd = kym_ldg(f + i), fm = kym_ldg(f + im);
                d = {u.re * fd, u.im * fd};
                m = {u.re * fm, -u.im * fm};
            } else {  // k is a power of two: alias pairs (q, k-1-q), several in flight
#pragma unroll 4
                for (int q = 0; q < k / 2; ++q) {
                    const int q2 = k - 1 - q;
                    const cpx<T> u = kym_ldg(U + i + q * n), u2 = kym_ldg(U + i + q2 * n);
                    const T fd = kym_ldg(f + i + q * n), fd2 = kym_ldg(f + i + q2 * n);
                    const T fm = kym_ldg(f + im + q * n), fm2 = kym_ldg(f + im + q2 * n);
                    d.re += u.re * fd;
                    d.im += u.im * fd;
                    d.re += u2.re * fd2;
                    d.im += u2.im * fd2;
                    m.re += u2.re * fm;  // alias q of the mirror is conj(u2)
                    m.im -= u2.im * fm;
                    m.re += u.re * fm2;
                    m.im -= u.im * fm2;
                }
            }
            lds[i] = d * inv_k;
            if (im != i) lds[im] = m * inv_k;
        }
        wg_barrier();
        ff.inv(lds, fl, tid, nth);
        const T inv_n = T(1) / T(n);
        const int s = a.ind_start[nd.unpad_res], e = a.ind_end[nd.unpad_res];
        cpx<T>* sv = (a.save && nd.save >= 0) ? wsi + nd.save : nullptr;
        for (int p = tid; p < n; p += nth) {
            const cpx<T> v = lds[p];
            if (sv) sv[p] = v;
            const T m = kym_sqrt(v.re * v.re + v.im * v.im) * inv_n;
            lds[p] = {m, T(0)};
            if (a.mode == AVG1_NONE) {
                const int t = fl.rev[p];
                if (t >= s && t < e) o[t - s] = m;
            }
        }
        wg_barrier();
        if (nd.need_fwd) ff.fwd(lds, fl, tid, nth);
    }
    if (nd.slot >= 0) {
        cpx<T>* dst = wsi + nd.slot;
        for (int i = tid; i < n; i += nth) dst[i] = lds[i];
    }
    if (a.mode != AVG1_NONE) leaf1d<T, MAXP>(a, nd, lds, n, img, lds + n, tid, nth);
}

template <class T, int MAXP>
KYM_HD void body_node1d(const Scat1DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    node1d_body<T, MAXP>(Dyn1D<T, MAXP>(), a, task, lds, tid, nth);
}

// Compile-time specialisations X(id, N, threads) of the forward node (FP32):
// a launch uses one when its nodes' length matches (threads follow the
// plan's rule max(64, min(1024, N / 16))).
#define KYM_NODE1D_SPECS(X) \
    X(0, 16384, 1024)       \
    X(1, 8192, 512)         \
    X(2, 4096, 256)         \
    X(3, 2048, 128)         \
    X(4, 1024, 64)          \
    X(5, 512, 64)           \
    X(6, 256, 64)           \
    X(7, 128, 64)

inline int node1d_spec_id(int n, int threads) {
#define KYM_N1_MATCH(id, nn, nt) \
    if (n == nn && threads == nt) return id;
    KYM_NODE1D_SPECS(KYM_N1_MATCH)
#undef KYM_N1_MATCH
    return -1;
}

// Body of specialisation SPEC; EMU = true: one thread (CPU harness).
template <class T, int SPEC, bool EMU>
KYM_HD KYM_INL void node1d_spec_body(const Scat1DArgs<T>& a, int task, cpx<T>* lds, int tid) {
#define KYM_N1_BODY(sid, nn, nt) \
    if constexpr (SPEC == sid) node1d_body<T, 7>(Stat1D<T, nn, (EMU ? 1 : nt)>(), a, task, lds, tid, EMU ? 1 : nt);
    KYM_NODE1D_SPECS(KYM_N1_BODY)
#undef KYM_N1_BODY
}

template <class T, bool EMU>
KYM_HD void dispatch_node1d_spec(int id, const Scat1DArgs<T>& a, int task, cpx<T>* lds, int tid) {
    switch (id) {
#define KYM_N1_CASE(sid, nn, nt) \
    case sid: node1d_spec_body<T, sid, EMU>(a, task, lds, tid); break;
        KYM_NODE1D_SPECS(KYM_N1_CASE)
#undef KYM_N1_CASE
        default: break;
    }
}

// ------------------------------------------------------------------ backward
// Adjoint of body_node1d (same conventions as the 2D backward):
//   local leaf  dL/dY = fft(unpad^T dS) / nl, dL/dU += phi * tile(dL/dY) / kl
//   children    dL/dU += f_c * tile(dL/dX_c) / kc
//   dL/dm = Re(n ifft(dL/dU)) (+ dS on the unpadded window without averaging,
//           + dS everywhere with global averaging)
//   dL/du = dL/dm u / |u|,  dL/dX = fft(dL/du) / n;  root: dL/dpad(x) = dL/dm.
template <class T, int MAXP>
KYM_HD void body_bwd_node1d(const Scat1DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int il = task / a.ngroup, gi = task - il * a.ngroup;
    const int img = a.img0 + il;
    const Node1D& nd = a.nodes[a.group[gi]];
    const int r = nd.res, n = a.n[r];
    const FftLen<T>& fl = a.lens[a.len[r]];
    cpx<T>* wsi = a.ws + il * a.ws_sig;
    const T* d = a.dS + (long long)img * a.out_row + nd.out_off;
    // leaf gradient (local average) into B
    cpx<T>* B = lds + n;
    int nl = 1, kk = 1;
    if (a.mode == AVG1_LOCAL) {
        kk = nd.leaf_k;
        nl = n / kk;
        const int lr = nd.res > a.log2_stride ? nd.res : a.log2_stride;
        const FftLen<T>& fB = a.lens[a.len[lr]];
        const int s = a.ind_start[a.log2_stride], e = a.ind_end[a.log2_stride];
        for (int p = tid; p < nl; p += nth) {
            const int t = fB.rev[p];
            B[p] = {(t >= s && t < e) ? d[t - s] : T(0), T(0)};
        }
        wg_barrier();
        fft_dit<false, MAXP>(B, fB, Batch{1, 1, nl, false}, tid, nth);
    }
    const bool has_u_grad = (a.mode == AVG1_LOCAL) || (nd.child_end > nd.child_begin);
    const T* phi = a.mode == AVG1_LOCAL ? a.filt + a.filt_off[nd.leaf_filt] : nullptr;
    const T sc_leaf = T(1) / (T(nl) * T(kk));
    for (int i = tid; i < n; i += nth) {
        cpx<T> acc = {T(0), T(0)};
        if (phi) {
            const cpx<T> gy = B[i % nl];
            acc = {gy.re * phi[i] * sc_leaf, gy.im * phi[i] * sc_leaf};
        }
#pragma unroll 4
        for (int c = nd.child_begin; c < nd.child_end; ++c) {  // unrolled: several loads in flight
            const Node1D& cn = a.nodes[c];
            const int nc = a.n[cn.res];
            const T fv = kym_ldg(a.filt + a.filt_off[cn.filt] + i) / T(cn.k);
            const cpx<T> gx = kym_ldg(wsi + cn.gsave + (i % nc));
            acc.re += gx.re * fv;
            acc.im += gx.im * fv;
        }
        lds[i] = acc;
    }
    wg_barrier();
    if (has_u_grad) fft_dif<true, MAXP>(lds, fl, Batch{1, 1, n, false}, tid, nth);
    // direct terms on dL/dm (positions p hold natural index rev[p])
    const int us = a.ind_start[nd.unpad_res], ue = a.ind_end[nd.unpad_res];
    for (int p = tid; p < n; p += nth) {
        T gm = has_u_grad ? lds[p].re : T(0);
        const int t = fl.rev[p];
        if (a.mode == AVG1_NONE && t >= us && t < ue) gm += d[t - us];
        if (a.mode == AVG1_GLOBAL) gm += d[0];
        if (nd.order == 0) {
            reinterpret_cast<T*>(wsi + nd.gpad)[t] = gm;
        } else {
            const cpx<T> u = wsi[nd.save + p];
            const T mag = kym_sqrt(u.re * u.re + u.im * u.im);
            lds[p] = mag > T(0) ? cpx<T>{u.re * (gm / mag), u.im * (gm / mag)} : cpx<T>{T(0), T(0)};
        }
    }
    if (nd.order == 0) return;
    wg_barrier();
    fft_dit<false, MAXP>(lds, fl, Batch{1, 1, n, false}, tid, nth);
    const T inv_n = T(1) / T(n);
    for (int i = tid; i < n; i += nth) wsi[nd.gsave + i] = lds[i] * inv_n;
}

// Reflect-pad adjoint (1D).  task -> signal.
template <class T>
KYM_HD void body_bwd_fold1d(const Scat1DArgs<T>& a, int task, int tid, int nth) {
    const int img = a.img0 + task;
    const T* gp = reinterpret_cast<const T*>(a.ws + task * a.ws_sig + a.nodes[0].gpad);
    T* gx = a.gx + (long long)img * a.N;
    const int P = a.n[0];
    for (int i = tid; i < a.N; i += nth) {
        T acc = T(0);
        if (a.N == 1) {
            for (int t = 0; t < P; ++t) acc += gp[t];
        } else {
            const int T_ = 2 * (a.N - 1);
            for (int t = (a.pad_left + i) % T_; t < P; t += T_) acc += gp[t];
            if (i != 0 && i != a.N - 1)
                for (int t = ((a.pad_left - i) % T_ + T_) % T_; t < P; t += T_) acc += gp[t];
        }
        gx[i] = acc;
    }
}
