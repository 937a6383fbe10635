// Scattering2D: node tree, kernel arguments and workgroup bodies.
//
// The reference's path (kymatio/scattering2d/core/scattering2d.py:1-88) is a
// tree of "nodes".  A node is a complex spectrum at resolution r
// (H_r x W_r = P_M/2^r x P_N/2^r):
//   root      U_0 = fft2(pad(x))                               r = 0
//   order 1   U_1 = fft2(|ifft2(P_k(U_0 * psi[n1].levels[0]))|)  r = j1, k = 2^j1
//   order 2   U_2 = fft2(|ifft2(P_k(U_1 * psi[n2].levels[j1]))|) r = j2, k = 2^(j2-j1)
// where P_k is the Fourier periodisation (mean of the k x k aliases,
// kymatio/scattering2d/backend/torch_backend.py:92-129).  Every node has one
// low-pass leaf S = unpad(Re(ifft2(P_kl(U * phi.levels[r])))), kl = 2^(J-r),
// written straight into output channel `leaf_ch` of (batch, C, H_J-2, W_J-2).
//
// Consumers pull: a child gathers its periodised input from the parent's
// spectrum (kept in a workspace slot, usually resident in L2/MALL) while
// loading, so the filter multiply and the periodisation are fused into the
// load of the next transform.  Nodes that fit in LDS run as one workgroup
// (gather, inverse FFT, modulus, leaf, forward FFT when the node has
// children); bigger ones run as column / row / column passes through their
// workspace slot.
//
// Leaves in space (FP32 plans).  With m = |u| the real field whose spectrum
// is U (m = pad(x) at the root), the periodisation identity
//   ifft2_{HJ,WJ}(P_kl(fft2(m) * phi_hat))[y, x] = (m (*) phi)[kl y, kl x],
//   phi = Re(ifft2(phi_hat)) (circular convolution on H_r x W_r),
// turns the leaf into a filter-and-subsample of m.  phi is a periodised
// Gaussian (kymatio/scattering2d/filter_bank.py: gabor_2d with xi = 0 and an
// isotropic envelope, then a separable crop), so phi = g (x) h and
//   S[y][x] = sum_a g[a] sum_b h[b] m[kl (y+1) + a][kl (x+1) + b]   (indices mod H_r, W_r)
// a row pass and a column pass of short real FIRs ("taps", filled by the
// frontend from the same filter bank).  The node with no children then needs
// no forward FFT at all.  FP64 plans keep the spectral leaf, so they follow
// the reference's arithmetic to FP64 rounding.
#pragma once
#include "fft_static.h"
#include "mma.h"

#define KYM_MAX_RES 16

struct Node2D {
    int order;      // 0 root, 1, 2
    int res;        // resolution index r
    int parent;     // node index, -1 for the root
    int filt;       // packed-filter index of the wavelet at the parent's resolution
    int k;          // periodisation factor from the parent (1 = none)
    int leaf_ch;    // output channel of this node's low-pass leaf
    int leaf_filt;  // packed-filter index of phi.levels[res]
    int leaf_k;     // 2^(J - res)
    int leaf_mode;  // 0 in the node kernel, 1 separate small-leaf kernel, 2 big-leaf passes
    int big;        // 1 = does not fit in LDS (column/row passes)
    long long slot;       // offset (complex elements) of U in the per-image workspace, -1 none
    long long leaf_slot;  // scratch for big leaves, -1 none
    // training (forward with saving + backward)
    int child_begin, child_end;  // node-index range of this node's children
    long long save;              // slot of the saved pre-modulus field u, -1 none
    long long gsave;             // slot of dL/dX written by the backward, -1 none
    long long gslot;             // leaf dL/dY (natural, scaled) when not formed in the node kernel
    long long gpad;              // root: dL/d(padded input), real
    // leaf in space (FP32 plans)
    int fir;         // 0 spectral leaf; 1 FIR leaf in the LDS node kernel; 2 FIR row pass + K2_FIR_COLS (big nodes)
    int fwd;         // 1 = forward FFT of the modulus is formed (children or spectral leaf need it)
    long long tslot; // fir == 2: row-filtered field (H_r x ow reals), -1 none
};

// FP32 plans form every leaf in space, FP64 plans spectrally (compile-time,
// so each kernel carries one of the two code paths).
template <class T>
inline constexpr bool kym_fir = sizeof(T) == 4;

// Taps of one axis of a leaf filter: coefficient c[i] (at filt + off) for the
// offset lo + i from the output's centre sample, and the band table of that
// axis (at filt + band_off, kmax x 16 reals, see fir_band) that the matrix-core
// passes read.
struct FirAxis {
    long long off;
    int lo, cnt;
    long long band_off;
    int kmax, pad_;
};

template <class T>
struct Scat2DArgs {
    const Node2D* nodes;
    const FftLen<T>* lens;
    const long long* filt_off;  // element offsets into `filt`
    const int* group;           // node ids of this launch
    const FirAxis* fir;         // [2 r]: along y (H_r), [2 r + 1]: along x (W_r), r < J
    int ngroup, nblk, blk;      // nodes per image, tasks per node, rows/cols per task
    int lenH[KYM_MAX_RES], lenW[KYM_MAX_RES], H[KYM_MAX_RES], W[KYM_MAX_RES];
    int J;
    // input image(s): (batch, M, N); with pre_pad, M x N is the padded size
    const T* x;
    int M, N, top, left, pre_pad;
    const T* filt;
    // workspace: image i of the chunk starts at ws + i * ws_img
    cpx<T>* ws;
    long long ws_img;
    // output (batch, C, oh, ow)
    T* out;
    int C, oh, ow;
    int img0;   // first image of this chunk
    int ntask;  // tasks in this launch
    // training
    int save;        // forward: store u into each node's save slot
    const T* dS;     // backward: dL/dS, same layout as out
    T* gx;           // backward: dL/dx, same layout as x
    // counted graphs: when dev_count is set, the image count n = *dev_count
    // is read at run time (clamped to max_count) and this chunk, which starts
    // at image img0 and was launched for up to ntask / (ngroup nblk) images,
    // runs only images img0 .. n - 1 of it (live_tasks)
    const int* dev_count;
    int max_count;
};

// Tasks of this launch that have work: ntask, or with a run-time image count
// the tasks of the chunk's images below that count (the rest exit at once).
// `a` is not modified: kernels pass their argument on by reference.
template <class T>
KYM_HD KYM_INL int live_tasks(const Scat2DArgs<T>& a) {
    if (!a.dev_count) return a.ntask;
    int n = *a.dev_count;
    n = n < 0 ? 0 : (n > a.max_count ? a.max_count : n);
    const int per = a.ngroup * a.nblk;
    const int cap = a.ntask / per;
    int live = n - a.img0;
    live = live < 0 ? 0 : (live > cap ? cap : live);
    return live * per;
}

// ------------------------------------------------------------------ helpers
template <class T>
KYM_HD KYM_INL cpx<T> gather_periodized(const cpx<T>* U, const T* f, int Wp, int H, int W, int k, int y,
                                        int x) {
    cpx<T> acc = {T(0), T(0)};
    for (int a = 0; a < k; ++a) {
        const long long row = (long long)(y + a * H) * Wp + x;
        for (int b = 0; b < k; ++b) {
            const long long idx = row + b * W;
            const T fv = f[idx];
            const cpx<T> u = U[idx];
            acc.re += u.re * fv;
            acc.im += u.im * fv;
        }
    }
    return acc;
}

// The same with the alias count K fixed at compile time (unrolled loads).
template <int K, class T>
KYM_HD KYM_INL cpx<T> gather_periodized_k(const cpx<T>* U, const T* f, int Wp, int H, int W, int y, int x) {
    cpx<T> acc = {T(0), T(0)};
#pragma unroll
    for (int a = 0; a < K; ++a) {
        const int row = (y + a * H) * Wp + x;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            const T fv = kym_ldg(f + row + b * W);
            const cpx<T> u = kym_ldg(U + row + b * W);
            acc.re += u.re * fv;
            acc.im += u.im * fv;
        }
    }
    return acc;
}

// dst[y * ncols + c] = scale * P_k(U * f)[y][c0 + c] for y < H, c < ncols.
// E = 8 / K elements per thread and round: their 2 K^2 loads each are all
// issued before the first use, so the memory latency is paid once per round.
template <int K, class T>
KYM_HD KYM_INL void gather_block_k(cpx<T>* dst, int ncols, int c0, const cpx<T>* U, const T* f, int H, int W,
                                   T scale, int tid, int nth) {
    constexpr int E = K >= 8 ? 1 : 8 / K;
    const int Wp = W * K;  // the parent's row length
    const int n = H * ncols;
    int i = tid;
    for (; i + (E - 1) * nth < n; i += E * nth) {
        cpx<T> v[E];
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const int e = i + u * nth, y = e / ncols, c = e - y * ncols;
            v[u] = gather_periodized_k<K>(U, f, Wp, H, W, y, c0 + c);
        }
#pragma unroll
        for (int u = 0; u < E; ++u) dst[i + u * nth] = v[u] * scale;
    }
    for (; i < n; i += nth) {
        const int y = i / ncols, c = i - y * ncols;
        dst[i] = gather_periodized_k<K>(U, f, Wp, H, W, y, c0 + c) * scale;
    }
}

// Gathers of a whole spectrum (ncols == W, c0 == 0) by mirror pairs.  Every
// parent is the spectrum of a real field (the padded input, or a modulus), so
// U[-u] = conj(U[u]) on the parent's K H x K W grid, and for 0 < y, 0 < x the
// child's points (y, x) and (H - y, W - x) sum the same K^2 parent values:
// the mirror's alias (a, b) is conj(U[y + (K-1-a) H][x + (K-1-b) W]).  One
// load of each parent value then serves both points; the filter is read at
// both.  Pair index p (H and W even): rows 0 < y < H/2 with 0 < x < W, then
// row H/2 with 0 < x <= W/2 (x = W/2 pairs with itself); the remaining
// pairs (row 0 with 0 <= x <= W/2, column 0 with 0 < y <= H/2) gather both
// points directly.  Results equal the plain gather's up to rounding.
KYM_HD KYM_INL constexpr int herm_pairs(int H, int W) { return (H / 2 - 1) * (W - 1) + W + 1 + H / 2; }

// Point of pair p; returns false for the pairs gathered directly.
KYM_HD KYM_INL bool herm_point(int p, int H, int W, int* y, int* x) {
    const int na = (H / 2 - 1) * (W - 1);
    if (p < na) {
        *y = 1 + p / (W - 1);
        *x = 1 + p - (*y - 1) * (W - 1);
        return true;
    }
    p -= na;
    if (p < W / 2) {
        *y = H / 2;
        *x = 1 + p;
        return true;
    }
    p -= W / 2;
    if (p <= W / 2) {
        *y = 0;
        *x = p;
    } else {
        *y = p - W / 2;
        *x = 0;
    }
    return false;
}

// For NF filters fs[i] and 0 < y, 0 < x: d[i] += sum u fs[i] at (y, x) and
// m[i] += sum conj(u') fs[i] at the mirror (H - y, W - x).  The mirror of
// alias (a, b) reads the point's alias (K-1-a, K-1-b), so alias rows
// {a, K-1-a} and column chunks {c, NC-1-c} (CH = min(K, 4) columns each) are
// taken together: at most 16 parent values are live at a time.
template <int K, int NF, class T>
KYM_HD KYM_INL void herm_block(cpx<T>* d, cpx<T>* m, const cpx<T>* U, const T* const* fs, int H, int W, int y, int x,
                               int a0, int c0) {
    constexpr int CH = K < 4 ? K : 4, NC = K / CH;
    const int Wp = W * K, ym = H - y, xm = W - x, a1 = K - 1 - a0, c1 = NC - 1 - c0;
    const bool two_a = a1 != a0, two_c = c1 != c0;
    cpx<T> u[2][2][CH];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int e = 0; e < CH; ++e) {
                if ((r && !two_a) || (q && !two_c)) {
                    u[r][q][e] = u[r && two_a][q && two_c][e];
                } else {
                    const int al = r ? a1 : a0, b = (q ? c1 : c0) * CH + e;
                    u[r][q][e] = kym_ldg(U + (y + al * H) * Wp + x + b * W);
                }
            }
#pragma unroll
    for (int i = 0; i < NF; ++i) {
        const T* f = fs[i];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int e = 0; e < CH; ++e) {
                    if ((r && !two_a) || (q && !two_c)) continue;
                    const int al = r ? a1 : a0, b = (q ? c1 : c0) * CH + e;
                    const cpx<T> v = u[r][q][e];                   // alias (al, b) of (y, x)
                    const cpx<T> w = u[1 - r][1 - q][CH - 1 - e];  // alias (al, b) of the mirror, conjugated
                    const T fd = kym_ldg(f + (y + al * H) * Wp + x + b * W);
                    const T fm = kym_ldg(f + (ym + al * H) * Wp + xm + b * W);
                    d[i].re += v.re * fd;
                    d[i].im += v.im * fd;
                    m[i].re += w.re * fm;
                    m[i].im -= w.im * fm;
                }
    }
}

template <int K, int NF, class T>
KYM_HD KYM_INL void herm_gather(cpx<T>* d, cpx<T>* m, const cpx<T>* U, const T* const* fs, int H, int W, int y,
                                int x) {
    constexpr int NC = K < 4 ? 1 : K / 4;
    if constexpr (K <= 4) {
#pragma unroll
        for (int a0 = 0; a0 < (K + 1) / 2; ++a0) herm_block<K, NF>(d, m, U, fs, H, W, y, x, a0, 0);
    } else {  // rolled loops over the row and chunk pairs: the body is already 32 alias terms
        for (int a0 = 0; a0 < K / 2; ++a0)
            for (int c0 = 0; c0 < NC / 2; ++c0) herm_block<K, NF>(d, m, U, fs, H, W, y, x, a0, c0);
    }
}

// dst[y W + x] = scale P_K(U f)[y][x] for the whole H x W child.
template <int K, class T>
KYM_HD KYM_INL void gather_herm_k(cpx<T>* dst, const cpx<T>* U, const T* f, int H, int W, T scale, int tid,
                                  int nth) {
    // pairs per thread and round: one from K = 2 on (A/B on MI355X at 256x256
    // J=3: two pairs per round 266 us per 136x136 launch, one 257 us, four 287 us)
    constexpr int E = K >= 2 ? 1 : 2;
    const int np = herm_pairs(H, W), Wp = W * K;
    const T* fs[1] = {f};
    for (int p0 = tid; p0 < np; p0 += E * nth) {
        cpx<T> d[E], m[E];
        int y[E], x[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            d[e] = m[e] = {T(0), T(0)};
            const bool h = herm_point(p0 + e * nth < np ? p0 + e * nth : p0, H, W, &y[e], &x[e]);
            if (h) {
                herm_gather<K, 1>(&d[e], &m[e], U, fs, H, W, y[e], x[e]);
            } else {
                const int ym = y[e] == 0 ? 0 : H - y[e], xm = x[e] == 0 ? 0 : W - x[e];
                if constexpr (K <= 4) {
                    d[e] = gather_periodized_k<K>(U, f, Wp, H, W, y[e], x[e]);
                    m[e] = gather_periodized_k<K>(U, f, Wp, H, W, ym, xm);
                } else {  // rolled: K^2 terms
                    d[e] = gather_periodized(U, f, Wp, H, W, K, y[e], x[e]);
                    m[e] = gather_periodized(U, f, Wp, H, W, K, ym, xm);
                }
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (p0 + e * nth >= np) continue;
            const int ym = y[e] == 0 ? 0 : H - y[e], xm = x[e] == 0 ? 0 : W - x[e];
            dst[y[e] * W + x[e]] = d[e] * scale;
            if (ym != y[e] || xm != x[e]) dst[ym * W + xm] = m[e] * scale;
        }
    }
}

// Wp = W k is the parent's row length.
template <class T>
KYM_HD KYM_INL void gather_block(cpx<T>* dst, int ncols, int c0, const cpx<T>* U, const T* f, int Wp, int H, int W, int k,
                         int tid, int nth) {
    const T scale = T(1) / T(k * k);
    switch (k) {
        case 1: gather_block_k<1>(dst, ncols, c0, U, f, H, W, scale, tid, nth); break;
        case 2: gather_block_k<2>(dst, ncols, c0, U, f, H, W, scale, tid, nth); break;
        case 4: gather_block_k<4>(dst, ncols, c0, U, f, H, W, scale, tid, nth); break;
        default:
            for (int i = tid; i < H * ncols; i += nth) {
                const int y = i / ncols, c = i - y * ncols;
                dst[i] = gather_periodized(U, f, Wp, H, W, k, y, c0 + c) * scale;
            }
    }
}

// Two siblings' periodised products at (y, x): the same parent values
// with the filters fa and fb (one load of each parent value for both).
template <int K, class T>
KYM_HD KYM_INL void gather_periodized_pair_k(const cpx<T>* U, const T* fa, const T* fb, int Wp, int H, int W, int y,
                                             int x, cpx<T>* va, cpx<T>* vb) {
    cpx<T> sa = {T(0), T(0)}, sb = {T(0), T(0)};
#pragma unroll
    for (int a = 0; a < K; ++a) {
        const int row = (y + a * H) * Wp + x;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            const cpx<T> u = kym_ldg(U + row + b * W);
            const T ga = kym_ldg(fa + row + b * W), gb = kym_ldg(fb + row + b * W);
            sa.re += u.re * ga;
            sa.im += u.im * ga;
            sb.re += u.re * gb;
            sb.im += u.im * gb;
        }
    }
    *va = sa;
    *vb = sb;
}

// gather_block for two siblings (same parent, same k) into dsta and dstb.
template <int K, class T>
KYM_HD KYM_INL void gather_block_pair_k(cpx<T>* dsta, cpx<T>* dstb, int ncols, int c0, const cpx<T>* U, const T* fa,
                                        const T* fb, int H, int W, T scale, int tid, int nth) {
    constexpr int E = K >= 4 ? 1 : 4 / K;
    const int Wp = W * K;  // the parent's row length
    const int n = H * ncols;
    int i = tid;
    for (; i + (E - 1) * nth < n; i += E * nth) {
        cpx<T> va[E], vb[E];
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const int e = i + u * nth, y = e / ncols, c = e - y * ncols;
            gather_periodized_pair_k<K>(U, fa, fb, Wp, H, W, y, c0 + c, &va[u], &vb[u]);
        }
#pragma unroll
        for (int u = 0; u < E; ++u) {
            dsta[i + u * nth] = va[u] * scale;
            dstb[i + u * nth] = vb[u] * scale;
        }
    }
    for (; i < n; i += nth) {
        const int y = i / ncols, c = i - y * ncols;
        cpx<T> va, vb;
        gather_periodized_pair_k<K>(U, fa, fb, Wp, H, W, y, c0 + c, &va, &vb);
        dsta[i] = va * scale;
        dstb[i] = vb * scale;
    }
}

// Returns false for a k without a pair gather (the plan pairs k = 2 and 4).
template <class T>
KYM_HD KYM_INL bool gather_block_pair(cpx<T>* dsta, cpx<T>* dstb, int ncols, int c0, const cpx<T>* U, const T* fa,
                                      const T* fb, int H, int W, int k, int tid, int nth) {
    const T scale = T(1) / T(k * k);
    switch (k) {
        case 1: gather_block_pair_k<1>(dsta, dstb, ncols, c0, U, fa, fb, H, W, scale, tid, nth); return true;
        case 2: gather_block_pair_k<2>(dsta, dstb, ncols, c0, U, fa, fb, H, W, scale, tid, nth); return true;
        case 4: gather_block_pair_k<4>(dsta, dstb, ncols, c0, U, fa, fb, H, W, scale, tid, nth); return true;
        default: return false;
    }
}

// The whole H x W child spectrum (a node in LDS): by mirror pairs where the
// sizes are even and K is a power of two up to 32, else gather_block.
template <class T>
KYM_HD KYM_INL void gather_whole(cpx<T>* dst, const cpx<T>* U, const T* f, int H, int W, int k, int tid, int nth) {
    const T scale = T(1) / T(k * k);
    if (H % 2 == 0 && W % 2 == 0 && H >= 2 && W >= 2) {
        switch (k) {
            case 1: gather_herm_k<1>(dst, U, f, H, W, scale, tid, nth); return;
            case 2: gather_herm_k<2>(dst, U, f, H, W, scale, tid, nth); return;
            case 4: gather_herm_k<4>(dst, U, f, H, W, scale, tid, nth); return;
            case 8: gather_herm_k<8>(dst, U, f, H, W, scale, tid, nth); return;
            case 16: gather_herm_k<16>(dst, U, f, H, W, scale, tid, nth); return;
            case 32: gather_herm_k<32>(dst, U, f, H, W, scale, tid, nth); return;
            default: break;
        }
    }
    gather_block(dst, W, 0, U, f, W * k, H, W, k, tid, nth);
}

template <class T>
KYM_HD KYM_INL T input_pixel(const Scat2DArgs<T>& a, const T* xin, int y, int x) {
    if (a.pre_pad) return xin[(long long)y * a.N + x];
    return xin[(long long)reflect_index(y - a.top, a.M) * a.N + reflect_index(x - a.left, a.N)];
}

// ------------------------------------------------------------------ leaf FIR
// sum_i c[i] src[((base + lo + i) mod n) * stride], with base + lo in (-n, n).
// c is uniform across the workgroup (scalar loads).
template <class T>
KYM_HD KYM_INL T fir_sum(const T* src, int stride, int n, int base, const T* c, int lo, int cnt) {
    int q = base + lo;
    if (q < 0) q += n;
    T acc = T(0);
    for (int i = 0; i < cnt; ++i) {
        acc += c[i] * src[q * stride];
        if (++q == n) q = 0;
    }
    return acc;
}

// ------------------------------------------------------------------ geometry
// A node of H x W with its leaf of HJ x WJ.  DynGeom reads the sizes and
// FFT plans at run time; StatGeom fixes them at compile time (constant
// divisors, unrolled passes).  Both use the same twiddle/rev tables.
// Registers per thread when a body stages the whole H x W field through
// registers (see stage_modulus): the plan keeps H W <= KYM_STAGE_MAX * threads.
#define KYM_STAGE_MAX 20

template <class T, int MAXP>
struct DynGeom {
    static constexpr int STAGE = KYM_STAGE_MAX;
    // modulus / restaging as separate sweeps (stage_modulus, restage_complex)
    static constexpr bool FUSED_INV = false, FUSED_FWD = false;
    int H, W, HJ, WJ;
    const FftLen<T>* fH;
    const FftLen<T>* fW;
    const FftLen<T>* fHJ;
    const FftLen<T>* fWJ;
    KYM_HD const int* revH() const { return fH->rev; }
    KYM_HD const int* revW() const { return fW->rev; }
    KYM_HD const int* revHJ() const { return fHJ->rev; }
    KYM_HD const int* revWJ() const { return fWJ->rev; }
    KYM_HD int rh(int p) const { return kym_ldg(fH->rev + p); }
    KYM_HD int rw(int p) const { return kym_ldg(fW->rev + p); }
    KYM_HD KYM_INL void inv2d(cpx<T>* b, int tid, int nth) const {
        fft_dif<true, MAXP>(b, *fW, Batch{H, 1, W, false}, tid, nth);
        fft_dif<true, MAXP>(b, *fH, Batch{W, W, 1, true}, tid, nth);
    }
    KYM_HD KYM_INL void fwd2d(cpx<T>* b, int tid, int nth) const {
        fft_dit<false, MAXP>(b, *fW, Batch{H, 1, W, false}, tid, nth);
        fft_dit<false, MAXP>(b, *fH, Batch{W, W, 1, true}, tid, nth);
    }
    KYM_HD KYM_INL void leaf_inv2d(cpx<T>* b, int tid, int nth) const {
        fft_dif<true, MAXP>(b, *fWJ, Batch{HJ, 1, WJ, false}, tid, nth);
        fft_dif<true, MAXP>(b, *fHJ, Batch{WJ, WJ, 1, true}, tid, nth);
    }
    KYM_HD KYM_INL void leaf_fwd2d(cpx<T>* b, int tid, int nth) const {
        fft_dit<false, MAXP>(b, *fWJ, Batch{HJ, 1, WJ, false}, tid, nth);
        fft_dit<false, MAXP>(b, *fHJ, Batch{WJ, WJ, 1, true}, tid, nth);
    }
};

template <class T, int H_, int W_, int HJ_, int WJ_, int NT>
struct StatGeom {
    static constexpr int H = H_, W = W_, HJ = HJ_, WJ = WJ_;
    static constexpr int STAGE = (H_ * W_ + NT - 1) / NT;
    // The modulus rides on the inverse's last pass (inv2d_modulus) and the
    // forward's first pass reads the real field (fwd2d_real) when those
    // passes have a butterfly for every thread; otherwise the separate
    // sweeps spread the work over the whole workgroup and are faster.
    static constexpr bool FUSED_INV = (H_ / kym_factor_list(H_).rad[0]) * W_ >= NT;
    static constexpr bool FUSED_FWD = (W_ / kym_factor_list(W_).rad[0]) * H_ >= NT;
    const FftLen<T>* fH;
    const FftLen<T>* fW;
    const FftLen<T>* fHJ;
    const FftLen<T>* fWJ;
    KYM_HD const int* revH() const { return fH->rev; }
    KYM_HD const int* revW() const { return fW->rev; }
    KYM_HD const int* revHJ() const { return fHJ->rev; }
    KYM_HD const int* revWJ() const { return fWJ->rev; }
    KYM_HD int rh(int p) const { return kym_rev_s<H>(p); }
    KYM_HD int rw(int p) const { return kym_rev_s<W>(p); }
    // twiddle tables of the node's lengths, copied to LDS by the kernel
    const cpx<T>* twH;
    const cpx<T>* twW;
    KYM_HD KYM_INL void inv2d(cpx<T>* b, int tid, int) const {
        fft_dif_s<true, W, H, 1, W, false, NT, true>(b, twW, tid);
        fft_dif_s<true, H, W, W, 1, true, NT, true>(b, twH, tid);
    }
    KYM_HD KYM_INL void fwd2d(cpx<T>* b, int tid, int) const {
        fft_dit_s<false, W, H, 1, W, false, NT, true>(b, twW, tid);
        fft_dit_s<false, H, W, W, 1, true, NT, true>(b, twH, tid);
    }
    // Inverse 2D FFT with the modulus in its last pass (columns, radix r_1):
    // R[y (W + 1) + x] = |X[y][x]| scale in natural order over the buffer;
    // sv (if set) receives X before the modulus at its storage index.
    KYM_HD KYM_INL void inv2d_modulus(cpx<T>* b, T* R, T scale, cpx<T>* sv, int tid) const {
        constexpr int R0 = kym_factor_list(H).rad[0];
        fft_dif_s<true, W, H, 1, W, false, NT, true>(b, twW, tid);
        fft_dif_s<true, H, W, W, 1, true, NT, true, T, -1, H, 1>(b, twH, tid);
        dif_last_pass_s<true, H, W, W, 1, true, NT, T>(
            b, tid,
            [&](int t, int bb, int m, cpx<T> v) -> T {
                if (sv) sv[(bb * R0 + m) * W + t] = v;
                return kym_sqrt(v.re * v.re + v.im * v.im) * scale;
            },
            [&](int t, int bb, int m, T val) {
                R[(kym_rev_s<H>(bb * R0) + (H / R0) * m) * (W + 1) + kym_rev_s<W>(t)] = val;
            });
    }
    // Forward 2D FFT of the real field R[y (W + 1) + x] (natural order, over
    // the buffer): the first row pass reads R directly.
    KYM_HD KYM_INL void fwd2d_real(cpx<T>* b, const T* R, int tid) const {
        constexpr int R0 = kym_factor_list(W).rad[0];
        dit_first_pass_real_s<false, W, H, 1, W, false, NT, T>(b, tid, [&](int t, int bb, int m) -> T {
            return R[kym_rev_s<H>(t) * (W + 1) + kym_rev_s<W>(bb * R0) + (W / R0) * m];
        });
        fft_dit_s<false, W, H, 1, W, false, NT, true, T, 1, R0>(b, twW, tid);
        fft_dit_s<false, H, W, W, 1, true, NT, true>(b, twH, tid);
    }
    KYM_HD KYM_INL void leaf_inv2d(cpx<T>* b, int tid, int) const {
        fft_dif_s<true, WJ, HJ, 1, WJ, false, NT>(b, fWJ->tw, tid);
        fft_dif_s<true, HJ, WJ, WJ, 1, true, NT>(b, fHJ->tw, tid);
    }
    KYM_HD KYM_INL void leaf_fwd2d(cpx<T>* b, int tid, int) const {
        fft_dit_s<false, WJ, HJ, 1, WJ, false, NT>(b, fWJ->tw, tid);
        fft_dit_s<false, HJ, WJ, WJ, 1, true, NT>(b, fHJ->tw, tid);
    }
};

// Leaf: periodise U (H x W, natural, row stride W) times phi into `B`
// (HJ x WJ), inverse FFT, real part, unpad, store into the output channel.
template <class T, class G>
KYM_HD void leaf_body(const Scat2DArgs<T>& a, const G& g, const cpx<T>* U, const T* phi, int kk, int ch, int img,
                      cpx<T>* B, int tid, int nth) {
    const T inv_k2 = T(1) / T(kk * kk);
    for (int i = tid; i < g.HJ * g.WJ; i += nth) {
        const int y = i / g.WJ, x = i - (i / g.WJ) * g.WJ;
        B[i] = gather_periodized(U, phi, g.W, g.HJ, g.WJ, kk, y, x) * inv_k2;
    }
    wg_barrier();
    g.leaf_inv2d(B, tid, nth);
    const T inv_n = T(1) / T(g.HJ * g.WJ);
    T* o = a.out + ((long long)img * a.C + ch) * a.oh * a.ow;
    const int* rH = g.revHJ();
    const int* rW = g.revWJ();
    for (int p = tid; p < g.HJ * g.WJ; p += nth) {
        const int py = p / g.WJ, px = p - py * g.WJ;
        const int y = rH[py], x = rW[px];
        if (y >= 1 && y <= g.HJ - 2 && x >= 1 && x <= g.WJ - 2) o[(y - 1) * a.ow + (x - 1)] = B[p].re * inv_n;
    }
    wg_barrier();
}

// ------------------------------------------------------------------ leaf in space
// The real field of a node lives in LDS as R[y * (W + 1) + x] (natural order,
// odd row stride), overlaying the node's complex buffer; the leaf's tables
// follow it (fir_lds_reals; the plan sizes the LDS for the larger of the
// two layouts).
//
// Moving between the digit-reversed complex layout and R permutes the whole
// buffer in place, so every thread first reads its elements into registers,
// then all write after a barrier.  The CPU harness runs one thread, which
// stages through a host array instead.
template <class T, class G>
KYM_HD void stage_modulus(const G& g, cpx<T>* lds, cpx<T>* sv, T inv_n, int tid, int nth) {
    const int n = g.H * g.W, rs = g.W + 1;
    T* R = reinterpret_cast<T*>(lds);
#if !defined(__HIP_DEVICE_COMPILE__)
    if (nth == 1) {
        std::vector<T> m(n);
        for (int s = 0; s < n; ++s) {
            const cpx<T> v = lds[s];
            if (sv) sv[s] = v;
            m[s] = kym_sqrt(v.re * v.re + v.im * v.im) * inv_n;
        }
        for (int s = 0; s < n; ++s) R[g.rh(s / g.W) * rs + g.rw(s % g.W)] = m[s];
        return;
    }
#endif
    constexpr int S = G::STAGE;
    T m[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int s = tid + k * nth;
        if (s < n) {
            const cpx<T> v = lds[s];
            if (sv) sv[s] = v;
            m[k] = kym_sqrt(v.re * v.re + v.im * v.im) * inv_n;
        }
    }
    wg_barrier();
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int s = tid + k * nth;
        if (s < n) {
            const int py = s / g.W, px = s - py * g.W;
            R[g.rh(py) * rs + g.rw(px)] = m[k];
        }
    }
    wg_barrier();
}

// R (natural reals) -> digit-reversed complex buffer, ready for a forward DIT.
template <class T, class G>
KYM_HD void restage_complex(const G& g, cpx<T>* lds, int tid, int nth) {
    const int n = g.H * g.W, rs = g.W + 1;
    const T* R = reinterpret_cast<const T*>(lds);
#if !defined(__HIP_DEVICE_COMPILE__)
    if (nth == 1) {
        std::vector<T> m(n);
        for (int s = 0; s < n; ++s) m[s] = R[g.rh(s / g.W) * rs + g.rw(s % g.W)];
        for (int s = 0; s < n; ++s) lds[s] = {m[s], T(0)};
        return;
    }
#endif
    constexpr int S = G::STAGE;
    T m[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int s = tid + k * nth;
        if (s < n) {
            const int py = s / g.W, px = s - py * g.W;
            m[k] = R[g.rh(py) * rs + g.rw(px)];
        }
    }
    wg_barrier();
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int s = tid + k * nth;
        if (s < n) lds[s] = {m[k], T(0)};
    }
    wg_barrier();
}

// Row stride of Tb: 16 mod 32 floats (column reads of the column pass hit
// distinct banks) when that still fits in the complex buffer, else ow + 1.
KYM_HD KYM_INL int fir_tstride(int W, int ow) {
    const int ts = (ow + 15) / 32 * 32 + 16;
    return (W + 1 + ts <= 2 * W) ? ts : ow + 1;
}

// Reals the in-LDS leaf uses: R, Tb and the two band matrices Bx (Kx x 16)
// and Ay (Ky x 16) with K = min(n, 15 kl + cnt).
KYM_HD KYM_INL long long fir_lds_reals(int H, int W, int ow, int kl, int cnt_y, int cnt_x) {
    const int kx = (15 * kl + cnt_x < W) ? 15 * kl + cnt_x : W;
    const int ky = (15 * kl + cnt_y < H) ? 15 * kl + cnt_y : H;
    return (long long)H * (W + 1) + (long long)H * fir_tstride(W, ow) + 16ll * (kx + ky);
}

// Band length of a leaf pass: the samples one 16-output tile touches.
KYM_HD KYM_INL int fir_kmax(int n, int kl, int cnt) { return (15 * kl + cnt < n) ? 15 * kl + cnt : n; }

// Row passes copy a band table of up to KYM_FIR_BAND_LDS bytes to LDS and
// read a larger one (J >= 5 at 320 and 640 points: 20 and 40 KB, the whole
// row for every output) from global memory, where every workgroup of the
// launch shares it through L2: the rows then fit two workgroups per CU.
#define KYM_FIR_BAND_LDS (16 * 1024)
KYM_HD KYM_INL constexpr bool fir_band_in_lds(int kmax, int elem_bytes) {
    return 16ll * kmax * elem_bytes <= KYM_FIR_BAND_LDS;
}

// Row passes with a band in global memory whose block (up to 16 rows) and
// output width (up to 16) make one MFMA tile split its k-steps over four
// waves, with 4 KB of LDS for their partial sums (KYM_FIR_SPLIT_LDS, where
// the band would be); 8 rows of 640 samples take 160 k-steps (J = 6).
#define KYM_FIR_SPLIT_LDS (4 * 64 * 4 * 4)
KYM_HD KYM_INL constexpr bool fir_rows_split(int rows, int ow) { return rows <= 16 && ow <= 16; }

// band[k][d] = c[(k - kl d) mod n] (zero beyond the taps), k < kmax, d < 16:
// inside a tile's window the circulant factor depends only on (k, d).  The
// frontend stores it next to the taps; passes copy it to LDS.
template <class T>
KYM_HD KYM_INL void fir_band(T* band, const T* filt, const FirAxis& f, int tid, int nth) {
    const T* src = filt + f.band_off;
    copy_batched<8>(16 * f.kmax, [&](int i) { return kym_ldg(src + i); }, [&](int i, T v) { band[i] = v; }, tid, nth);
}

// Row pass on the matrix cores: for rows p < nrows of R (row stride rs,
// natural columns q < W) and outputs x < ow,
//   T[p][x] = sum_q R[p][q] c[(q - kl (x+1) - lo) mod W]
// handed to out(p, x, value).  Bx = fir_band(.., W, kl, c, cnt).
// A 16 x 16 tile reads the window q = qs + kk (mod W), kk < len, in k-steps
// of four (lane l takes kk = 4 s + l / 16): the steps before the window
// wraps and after it run as strided segments (plain loads at fixed strides),
// the step that straddles the wrap and a partial last step one at a time.
// A rows past nrows repeat row nrows - 1 and B columns past ow are read as
// they are: they only reach outputs that are not stored.
template <class T, class OUT>
KYM_HD void fir_rows_mma(const T* R, int rs, int nrows, int W, int kl, int ow, const FirAxis& fx, const T* Bx,
                         OUT out, int tid, int nth) {
    const int lane = lane_id(tid), wv = wave_id(tid), nw = wave_count(nth);
    const int nmt = (nrows + 15) >> 4, nnt = (ow + 15) >> 4;
    for (int t = wv; t < nmt * nnt; t += nw) {
        const int p0 = (t / nnt) << 4, x0 = (t % nnt) << 4;
        const int span = kl * ((ow - 1 - x0 < 15) ? ow - 1 - x0 : 15) + fx.cnt;
        const int len = span < W ? span : W;
        int qs = kl * (x0 + 1) + fx.lo;
        if (qs < 0) qs += W;
        const int nfull = len >> 2;                                // steps with every kk < len
        const int s_a = (W - qs) >> 2 < nfull ? (W - qs) >> 2 : nfull;  // no lane wraps before s_a
        const int s_b = (W - qs + 3) >> 2 < nfull ? (W - qs + 3) >> 2 : nfull;  // every lane wrapped from s_b
        auto row = [&](int l) { const int p = p0 + (l & 15); return (p < nrows ? p : nrows - 1) * rs; };
        auto general = [&](Mma16Acc& c, int s) {
            mma16_step(
                c, lane,
                [&](int l) -> float {
                    const int kk = 4 * s + (l >> 4);
                    int q = qs + kk;
                    q = q >= W ? q - W : q;
                    q = q < W ? q : W - 1;
                    const float v = R[row(l) + q];
                    return kk < len ? v : 0.f;
                },
                [&](int l) -> float {
                    const int kk = 4 * s + (l >> 4);
                    const float v = Bx[(kk < len ? kk : len - 1) * 16 + (l & 15)];
                    return kk < len ? v : 0.f;
                });
        };
        Mma16Acc c;
        mma16_zero(c);
        mma16_strided<4, 64>(
            c, lane, [&](int l) { return R + row(l) + qs + (l >> 4); },
            [&](int l) { return Bx + (l >> 4) * 16 + (l & 15); }, s_a);
        for (int s = s_a; s < s_b; ++s) general(c, s);
        mma16_strided<4, 64>(
            c, lane, [&](int l) { return R + row(l) + qs - W + 4 * s_b + (l >> 4); },
            [&](int l) { return Bx + (4 * s_b + (l >> 4)) * 16 + (l & 15); }, nfull - s_b);
        if (len & 3) general(c, nfull);
        mma16_store(c, lane, [&](int l, int v, float val) {
            const int p = p0 + 4 * (l >> 4) + v, x = x0 + (l & 15);
            if (p < nrows && x < ow) out(p, x, val);
        });
    }
}

// fir_rows_mma for one tile (nrows <= 16, ow <= 16) whose band is read from
// global memory: the tile's k-steps (up to 160, each waiting on L2) are split
// over up to four waves, whose partial sums meet in `red` (KYM_FIR_SPLIT_LDS).
template <class T, class OUT>
KYM_HD void fir_rows_mma_split(const T* R, int rs, int nrows, int W, int kl, int ow, const FirAxis& fx, const T* Bx,
                               OUT out, int tid, int nth, T* red) {
    const int lane = lane_id(tid), wv = wave_id(tid), nw = wave_count(nth);
    const int nnt = (ow + 15) >> 4;
    // k-steps [k0, k1) of tile t into c
    auto tile_steps = [&](Mma16Acc& c, int t, int k0, int k1) {
        const int p0 = (t / nnt) << 4, x0 = (t % nnt) << 4;
        const int span = kl * ((ow - 1 - x0 < 15) ? ow - 1 - x0 : 15) + fx.cnt;
        const int len = span < W ? span : W;
        int qs = kl * (x0 + 1) + fx.lo;
        if (qs < 0) qs += W;
        const int nfull = len >> 2;                                // steps with every kk < len
        const int s_a = (W - qs) >> 2 < nfull ? (W - qs) >> 2 : nfull;  // no lane wraps before s_a
        const int s_b = (W - qs + 3) >> 2 < nfull ? (W - qs + 3) >> 2 : nfull;  // every lane wrapped from s_b
        if (k1 < 0) k1 = nfull + ((len & 3) ? 1 : 0);
        auto row = [&](int l) { const int p = p0 + (l & 15); return (p < nrows ? p : nrows - 1) * rs; };
        auto general = [&](int s) {
            mma16_step(
                c, lane,
                [&](int l) -> float {
                    const int kk = 4 * s + (l >> 4);
                    int q = qs + kk;
                    q = q >= W ? q - W : q;
                    q = q < W ? q : W - 1;
                    const float v = R[row(l) + q];
                    return kk < len ? v : 0.f;
                },
                [&](int l) -> float {
                    const int kk = 4 * s + (l >> 4);
                    const float v = Bx[(kk < len ? kk : len - 1) * 16 + (l & 15)];
                    return kk < len ? v : 0.f;
                });
        };
        const int a0 = k0, a1 = s_a < k1 ? s_a : k1;
        if (a1 > a0)
            mma16_strided<4, 64>(
                c, lane, [&](int l) { return R + row(l) + qs + 4 * a0 + (l >> 4); },
                [&](int l) { return Bx + (4 * a0 + (l >> 4)) * 16 + (l & 15); }, a1 - a0);
        for (int st = (s_a > k0 ? s_a : k0); st < (s_b < k1 ? s_b : k1); ++st) general(st);
        const int b0 = s_b > k0 ? s_b : k0, b1 = nfull < k1 ? nfull : k1;
        if (b1 > b0)
            mma16_strided<4, 64>(
                c, lane, [&](int l) { return R + row(l) + qs - W + 4 * b0 + (l >> 4); },
                [&](int l) { return Bx + (4 * b0 + (l >> 4)) * 16 + (l & 15); }, b1 - b0);
        if ((len & 3) && nfull >= k0 && nfull < k1) general(nfull);
    };
    auto tile_out = [&](int t, int l, int v, float val) {
        const int p = ((t / nnt) << 4) + 4 * (l >> 4) + v, x = ((t % nnt) << 4) + (l & 15);
        if (p < nrows && x < ow) out(p, x, val);
    };
    // One tile (nrows <= 16, ow <= 16) and scratch `red` (4 waves x 64 lanes
    // x 4 floats): up to four waves take a quarter of its k-steps each and
    // sum through LDS.
    const int split = nw < 4 ? nw : 4;
    Mma16Acc c;
    mma16_zero(c);
    if (wv < split) {
        const int span = kl * (ow - 1 < 15 ? ow - 1 : 15) + fx.cnt;  // tile 0: x0 = 0
        const int len = span < W ? span : W;
        const int nsteps = (len >> 2) + ((len & 3) ? 1 : 0);
        tile_steps(c, 0, wv * nsteps / split, (wv + 1) * nsteps / split);
        if (split > 1) mma16_store(c, lane, [&](int l, int v, float val) { red[(wv * 64 + l) * 4 + v] = val; });
    }
    if (split > 1) wg_barrier();
    if (wv == 0)
        mma16_store(c, lane, [&](int l, int v, float val) {
            float sum = val;
            for (int w = 1; w < split; ++w) sum += red[(w * 64 + l) * 4 + v];
            tile_out(0, l, v, sum);
        });
}

// Column pass on the matrix cores: S[y][x] = sum_p c[(p - kl (y+1) - lo) mod H] Tb[p][x]
// for y < oh, x < ow (Tb row stride ts), into o[y * ow + x].  Ay = fir_band(.., H, kl, c, cnt).
// Same k-step segments as the row pass, with the window along p.
template <class T>
KYM_HD void fir_cols_mma(const T* Tb, int ts, int H, int kl, int oh, int ow, const FirAxis& fy, const T* Ay, T* o,
                         int tid, int nth) {
    const int lane = lane_id(tid), wv = wave_id(tid), nw = wave_count(nth);
    const int nmt = (oh + 15) >> 4, nnt = (ow + 15) >> 4;
    for (int t = wv; t < nmt * nnt; t += nw) {
        const int y0 = (t / nnt) << 4, x0 = (t % nnt) << 4;
        const int span = kl * ((oh - 1 - y0 < 15) ? oh - 1 - y0 : 15) + fy.cnt;
        const int len = span < H ? span : H;
        int ps = kl * (y0 + 1) + fy.lo;
        if (ps < 0) ps += H;
        const int nfull = len >> 2;
        const int s_a = (H - ps) >> 2 < nfull ? (H - ps) >> 2 : nfull;
        const int s_b = (H - ps + 3) >> 2 < nfull ? (H - ps + 3) >> 2 : nfull;
        auto col = [&](int l) { const int x = x0 + (l & 15); return x < ow ? x : ow - 1; };
        auto general = [&](Mma16Acc& c, int s) {
            mma16_step(
                c, lane,
                [&](int l) -> float {
                    const int kk = 4 * s + (l >> 4);
                    const float v = Ay[(kk < len ? kk : len - 1) * 16 + (l & 15)];
                    return kk < len ? v : 0.f;
                },
                [&](int l) -> float {
                    const int kk = 4 * s + (l >> 4);
                    int p = ps + kk;
                    p = p >= H ? p - H : p;
                    p = p < H ? p : H - 1;
                    const float v = Tb[p * ts + col(l)];
                    return kk < len ? v : 0.f;
                });
        };
        Mma16Acc c;
        mma16_zero(c);
        mma16_strided<64, 0>(
            c, lane, [&](int l) { return Ay + (l >> 4) * 16 + (l & 15); },
            [&](int l) { return Tb + (ps + (l >> 4)) * ts + col(l); }, s_a, 64, 4 * ts);
        for (int s = s_a; s < s_b; ++s) general(c, s);
        mma16_strided<64, 0>(
            c, lane, [&](int l) { return Ay + (4 * s_b + (l >> 4)) * 16 + (l & 15); },
            [&](int l) { return Tb + (ps - H + 4 * s_b + (l >> 4)) * ts + col(l); }, nfull - s_b, 64, 4 * ts);
        if (len & 3) general(c, nfull);
        mma16_store(c, lane, [&](int l, int v, float val) {
            const int y = y0 + 4 * (l >> 4) + v, x = x0 + (l & 15);
            if (y < oh && x < ow) o[y * ow + x] = val;
        });
    }
}

// Leaf of an LDS node from R (FP32, matrix cores).  Both passes are GEMMs
// with a banded circulant factor:
//   row pass     Tb[p][x] = sum_q R[p][q] c_x[(q - kl (x+1) - lo_x) mod W]
//   column pass  S[y][x]  = sum_p c_y[(p - kl (y+1) - lo_y) mod H] Tb[p][x]
// Each 16 x 16 output tile only needs the window of 15 kl + cnt input
// samples its 16 outputs touch (the whole circle when the taps cover it).
// Callers barrier before reusing the buffer.
template <class T, class G>
KYM_HD void fir_leaf_lds(const Scat2DArgs<T>& a, const G& g, const Node2D& nd, T* R, int img, int tid, int nth) {
    const int r = nd.res, ow = a.ow, oh = a.oh, kl = nd.leaf_k;
    const int H = g.H, W = g.W, rs = W + 1, ts = fir_tstride(W, ow);
    const FirAxis fy = a.fir[2 * r], fx = a.fir[2 * r + 1];
    const int kxm = fir_kmax(W, kl, fx.cnt), kym = fir_kmax(H, kl, fy.cnt);
    T* Tb = R + H * rs;
    T* Bx = Tb + H * ts;
    T* Ay = Bx + 16 * kxm;
    fir_band(Bx, a.filt, fx, tid, nth);
    fir_band(Ay, a.filt, fy, tid, nth);
    wg_barrier();
    fir_rows_mma(R, rs, H, W, kl, ow, fx, Bx, [&](int p, int x, float v) { Tb[p * ts + x] = v; }, tid, nth);
    wg_barrier();
    fir_cols_mma(Tb, ts, H, kl, oh, ow, fy, Ay, a.out + ((long long)img * a.C + nd.leaf_ch) * oh * ow, tid, nth);
}

// ------------------------------------------------------------------ bodies
// A whole node in LDS: root (pad + forward FFT) or order 1/2 (gather +
// periodise, inverse FFT, modulus, forward FFT); then the optional store of
// the spectrum for the children and the optional in-kernel leaf.
// A node of order >= 1 from its gathered spectrum in `lds` (natural
// order): inverse FFT, modulus, leaf, forward FFT and slot if needed.
template <class T, class G>
KYM_HD KYM_INL void node_after_gather(const Scat2DArgs<T>& a, const G& g, const Node2D& nd, int il, cpx<T>* lds, int tid,
                              int nth) {
    const int img = a.img0 + il;
    cpx<T>* wsi = a.ws + il * a.ws_img;
    {
        const T inv_n = T(1) / T(g.H * g.W);
        cpx<T>* sv = (a.save && nd.save >= 0) ? wsi + nd.save : nullptr;
        if constexpr (kym_fir<T>) {
            T* R = reinterpret_cast<T*>(lds);
            if constexpr (G::FUSED_INV) {
                g.inv2d_modulus(lds, R, inv_n, sv, tid);
            } else {
                g.inv2d(lds, tid, nth);
                stage_modulus<T>(g, lds, sv, inv_n, tid, nth);
            }
            fir_leaf_lds(a, g, nd, R, img, tid, nth);
            if (nd.fwd) {
                if constexpr (G::FUSED_FWD) {
                    g.fwd2d_real(lds, R, tid);
                } else {
                    restage_complex<T>(g, lds, tid, nth);
                    g.fwd2d(lds, tid, nth);
                }
            }
        } else {
            g.inv2d(lds, tid, nth);
            for (int i = tid; i < g.H * g.W; i += nth) {
                const cpx<T> v = lds[i];
                if (sv) sv[i] = v;
                lds[i] = {kym_sqrt(v.re * v.re + v.im * v.im) * inv_n, T(0)};
            }
            wg_barrier();
            g.fwd2d(lds, tid, nth);
        }
    }
    if (nd.slot >= 0 && nd.fwd) {
        cpx<T>* dst = wsi + nd.slot;
        for (int i = tid; i < g.H * g.W; i += nth) dst[i] = lds[i];
    }
    if (!kym_fir<T> && nd.leaf_mode == 0)
        leaf_body(a, g, lds, a.filt + a.filt_off[nd.leaf_filt], nd.leaf_k, nd.leaf_ch, img, lds + g.H * g.W, tid,
                  nth);
}

template <class T, class G>
KYM_HD void node_body(const Scat2DArgs<T>& a, const G& g, const Node2D& nd, int il, cpx<T>* lds, int tid, int nth) {
    const int img = a.img0 + il;
    cpx<T>* wsi = a.ws + il * a.ws_img;
    if (nd.order == 0) {
        const T* xin = a.x + (long long)img * a.M * a.N;
        if constexpr (kym_fir<T>) {
            T* R = reinterpret_cast<T*>(lds);
            const int rs = g.W + 1;
            copy_batched<4>(
                g.H * g.W,
                [&](int i) {
                    const int y = i / g.W, x = i - y * g.W;
                    return input_pixel(a, xin, y, x);
                },
                [&](int i, T v) {
                    const int y = i / g.W, x = i - y * g.W;
                    R[y * rs + x] = v;
                },
                tid, nth);
            wg_barrier();
            fir_leaf_lds(a, g, nd, R, img, tid, nth);
            if constexpr (G::FUSED_FWD) {
                g.fwd2d_real(lds, R, tid);
            } else {
                restage_complex<T>(g, lds, tid, nth);
                g.fwd2d(lds, tid, nth);
            }
        } else {
            const int* rH = g.revH();
            const int* rW = g.revW();
            for (int i = tid; i < g.H * g.W; i += nth) {
                const int py = i / g.W, px = i - py * g.W;
                lds[i] = {input_pixel(a, xin, rH[py], rW[px]), T(0)};
            }
            wg_barrier();
            g.fwd2d(lds, tid, nth);
        }
    } else {
        const Node2D& pn = a.nodes[nd.parent];
        gather_whole(lds, wsi + pn.slot, a.filt + a.filt_off[nd.filt], g.H, g.W, nd.k, tid, nth);
        wg_barrier();
        node_after_gather(a, g, nd, il, lds, tid, nth);
        return;
    }
    if (nd.slot >= 0 && nd.fwd) {
        cpx<T>* dst = wsi + nd.slot;
        for (int i = tid; i < g.H * g.W; i += nth) dst[i] = lds[i];
    }
    if (!kym_fir<T> && nd.leaf_mode == 0)
        leaf_body(a, g, lds, a.filt + a.filt_off[nd.leaf_filt], nd.leaf_k, nd.leaf_ch, img, lds + g.H * g.W, tid,
                  nth);
}

template <class T, int MAXP>
KYM_HD DynGeom<T, MAXP> dyn_geom(const Scat2DArgs<T>& a, int r) {
    const int J = a.J;
    return DynGeom<T, MAXP>{a.H[r], a.W[r], a.H[J], a.W[J], &a.lens[a.lenH[r]], &a.lens[a.lenW[r]],
                            &a.lens[a.lenH[J]], &a.lens[a.lenW[J]]};
}

// A whole node in LDS.  task -> (image of chunk, node of group).
template <class T, int MAXP>
KYM_HD void body_lds_node(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int il = task / a.ngroup, gi = task - il * a.ngroup;
    const Node2D& nd = a.nodes[a.group[gi]];
    node_body(a, dyn_geom<T, MAXP>(a, nd.res), nd, il, lds, tid, nth);
}

// The same with compile-time sizes (the plan only selects it when the
// launch's node and leaf sizes equal the template's).
// Complex elements at the start of a compile-time kernel's LDS that hold
// twiddle tables (lengths h and w; one table when equal), kept 16-byte aligned.
KYM_HD constexpr int kym_tw_elems(int h, int w) { return ((w + (h != w ? h : 0)) + 1) / 2 * 2; }

// Copies the twiddles of lengths fh->n and fw->n to tw (LDS); barrier.
template <class T>
KYM_HD KYM_INL void load_twiddles(cpx<T>* tw, const FftLen<T>* fh, const FftLen<T>* fw, int h, int w, int tid,
                                  int nth) {
    for (int i = tid; i < w; i += nth) tw[i] = kym_ldg(fw->tw + i);
    if (h != w)
        for (int i = tid; i < h; i += nth) tw[w + i] = kym_ldg(fh->tw + i);
    wg_barrier();
}

template <class T, int H, int W, int HJ, int WJ, int NT>
KYM_HD void body_lds_node_s(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
    const int il = task / a.ngroup, gi = task - il * a.ngroup;
    const Node2D& nd = a.nodes[a.group[gi]];
    const int r = nd.res, J = a.J;
    constexpr int TWE = kym_tw_elems(H, W);
    const FftLen<T>* fH = &a.lens[a.lenH[r]];
    const FftLen<T>* fW = &a.lens[a.lenW[r]];
    load_twiddles(lds, fH, fW, H, W, tid, NT);
    const StatGeom<T, H, W, HJ, WJ, NT> g{fH, fW, &a.lens[a.lenH[J]], &a.lens[a.lenW[J]], lds + (H != W ? W : 0), lds};
    node_body(a, g, nd, il, lds + TWE, tid, NT);
}

// Gather of two siblings: da[i] = P_K(U fa)[i] scale and db[i] = P_K(U fb)[i]
// scale for i < H W (both in LDS).  Each alias value of the parent spectrum
// U is loaded once for both filters and, by mirror pairs (gather_herm_k),
// for both points i and -i.
template <int K, int H, int W, int NT, class T>
KYM_HD KYM_INL void gather_pair_s(cpx<T>* da, cpx<T>* db, const cpx<T>* U, const T* fa, const T* fb, T scale,
                                  int tid) {
    static_assert(H % 2 == 0 && W % 2 == 0, "mirror pairs need even sizes");
    constexpr int np = herm_pairs(H, W), Wp = W * K;
    for (int p = tid; p < np; p += NT) {
        int y, x;
        const bool h = herm_point(p, H, W, &y, &x);
        const int ym = y == 0 ? 0 : H - y, xm = x == 0 ? 0 : W - x;
        cpx<T> d[2] = {{T(0), T(0)}, {T(0), T(0)}}, m[2] = {{T(0), T(0)}, {T(0), T(0)}};
        if (h) {
            const T* fs[2] = {fa, fb};
            herm_gather<K, 2>(d, m, U, fs, H, W, y, x);
        } else {
            for (int s = 0; s < 2; ++s) {
                const int yy = s ? ym : y, xx = s ? xm : x;
                cpx<T> va = {T(0), T(0)}, vb = {T(0), T(0)};
                auto term = [&](int al, int b) {
                    const int idx = (yy + al * H) * Wp + xx + b * W;
                    const cpx<T> uv = kym_ldg(U + idx);
                    const T wa = kym_ldg(fa + idx), wb = kym_ldg(fb + idx);
                    va.re += uv.re * wa;
                    va.im += uv.im * wa;
                    vb.re += uv.re * wb;
                    vb.im += uv.im * wb;
                };
                if constexpr (K <= 4) {
#pragma unroll
                    for (int al = 0; al < K; ++al)
#pragma unroll
                        for (int b = 0; b < K; ++b) term(al, b);
                } else {  // K^2 terms: rows rolled, four columns per step
                    for (int al = 0; al < K; ++al)
#pragma unroll 4
                        for (int b = 0; b < K; ++b) term(al, b);
                }
                if (s) { m[0] = va; m[1] = vb; } else { d[0] = va; d[1] = vb; }
            }
        }
        da[y * W + x] = d[0] * scale;
        db[y * W + x] = d[1] * scale;
        if (ym != y || xm != x) {
            da[ym * W + xm] = m[0] * scale;
            db[ym * W + xm] = m[1] * scale;
        }
    }
}

// Complex elements of the LDS region one FP32 node uses (its complex buffer
// or, if larger, the leaf's FIR fields; the plan's node_lds).
template <int H, int W, class T>
KYM_HD KYM_INL long long node_region_cpx(const Scat2DArgs<T>& a, const Node2D& nd) {
    const long long reals = fir_lds_reals(H, W, a.ow, nd.leaf_k, a.fir[2 * nd.res].cnt, a.fir[2 * nd.res + 1].cnt);
    const long long c = (reals + 1) / 2;
    return c > (long long)H * W ? c : (long long)H * W;
}

// Two sibling nodes (same parent and resolution; the plan lists them at
// group[2 gi] and group[2 gi + 1]) in one workgroup.  The gather reads the
// parent once for both: the first node's spectrum goes to the node region,
// the second's to a stash after it; the first node runs, then the stash
// moves into the node region and the second runs.  LDS: twiddles, node
// region (node_region_cpx), stash (H W).  task -> (image, pair).
template <class T, int H, int W, int HJ, int WJ, int NT>
KYM_HD KYM_INL void body_lds_pair_s(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
    const int il = task / a.ngroup, gi = task - il * a.ngroup;
    const Node2D& na = a.nodes[a.group[2 * gi]];
    const Node2D& nb = a.nodes[a.group[2 * gi + 1]];
    const int r = na.res, J = a.J;
    constexpr int TWE = kym_tw_elems(H, W);
    const FftLen<T>* fH = &a.lens[a.lenH[r]];
    const FftLen<T>* fW = &a.lens[a.lenW[r]];
    load_twiddles(lds, fH, fW, H, W, tid, NT);
    const StatGeom<T, H, W, HJ, WJ, NT> g{fH, fW, &a.lens[a.lenH[J]], &a.lens[a.lenW[J]], lds + (H != W ? W : 0), lds};
    cpx<T>* buf = lds + TWE;
    cpx<T>* stash = buf + node_region_cpx<H, W>(a, na);
    const cpx<T>* U = a.ws + il * a.ws_img + a.nodes[na.parent].slot;
    const T* fa = a.filt + a.filt_off[na.filt];
    const T* fb = a.filt + a.filt_off[nb.filt];
    const T scale = T(1) / T(na.k * na.k);
    switch (na.k) {  // the plan pairs k = 2^i <= 32
        case 1: gather_pair_s<1, H, W, NT>(buf, stash, U, fa, fb, scale, tid); break;
        case 2: gather_pair_s<2, H, W, NT>(buf, stash, U, fa, fb, scale, tid); break;
        case 4: gather_pair_s<4, H, W, NT>(buf, stash, U, fa, fb, scale, tid); break;
        case 8: gather_pair_s<8, H, W, NT>(buf, stash, U, fa, fb, scale, tid); break;
        case 16: gather_pair_s<16, H, W, NT>(buf, stash, U, fa, fb, scale, tid); break;
        default: gather_pair_s<32, H, W, NT>(buf, stash, U, fa, fb, scale, tid); break;
    }
    wg_barrier();
    node_after_gather(a, g, na, il, buf, tid, NT);
    wg_barrier();
    for (int i = tid; i < H * W; i += NT) buf[i] = stash[i];
    wg_barrier();
    node_after_gather(a, g, nb, il, buf, tid, NT);
}

// Leaf from a spectrum kept in the workspace.  task -> (image, node).
template <class T, int MAXP>
KYM_HD void body_leaf(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int il = task / a.ngroup, gi = task - il * a.ngroup;
    const Node2D& nd = a.nodes[a.group[gi]];
    const int r = nd.res;
    const cpx<T>* U = a.ws + il * a.ws_img + nd.slot;
    const T* phi = a.filt + a.filt_off[nd.leaf_filt];
    leaf_body(a, dyn_geom<T, MAXP>(a, r), U, phi, nd.leaf_k, nd.leaf_ch, a.img0 + il, lds, tid, nth);
}

// ------------------------------------------------------------------ pass FFTs
// Transforms of the big-node passes over a block of nc columns (stride nc,
// column pass) or nr rows (row pass).  DynPass takes the lengths at run time;
// StatPass fixes lengths, block sizes and thread count at compile time (the
// plan selects it only when every block of the launch is full).
template <class T, int MAXP>
struct DynPass {
    static constexpr bool FUSED = false;
    KYM_HD KYM_INL void col_inv(cpx<T>* b, const FftLen<T>& f, int nc, int tid, int nth) const {
        fft_dif<true, MAXP>(b, f, Batch{nc, nc, 1, true}, tid, nth);
    }
    KYM_HD KYM_INL void col_fwd(cpx<T>* b, const FftLen<T>& f, int nc, int tid, int nth) const {
        fft_dit<false, MAXP>(b, f, Batch{nc, nc, 1, true}, tid, nth);
    }
    KYM_HD KYM_INL void row_inv(cpx<T>* b, const FftLen<T>& f, int nr, int tid, int nth) const {
        fft_dif<true, MAXP>(b, f, Batch{nr, 1, f.n, false}, tid, nth);
    }
    KYM_HD KYM_INL void row_fwd(cpx<T>* b, const FftLen<T>& f, int nr, int tid, int nth) const {
        fft_dit<false, MAXP>(b, f, Batch{nr, 1, f.n, false}, tid, nth);
    }
};

// twH / twW: the twiddle tables copied to LDS by the kernel.
template <class T, int H, int W, int CB, int RB, int NT>
struct StatPass {
    const cpx<T>* twH;
    const cpx<T>* twW;
    KYM_HD KYM_INL void col_inv(cpx<T>* b, const FftLen<T>&, int, int tid, int) const {
        fft_dif_s<true, H, CB, CB, 1, true, NT, true>(b, twH, tid);
    }
    KYM_HD KYM_INL void col_fwd(cpx<T>* b, const FftLen<T>&, int, int tid, int) const {
        fft_dit_s<false, H, CB, CB, 1, true, NT, true>(b, twH, tid);
    }
    KYM_HD KYM_INL void row_inv(cpx<T>* b, const FftLen<T>&, int, int tid, int) const {
        fft_dif_s<true, W, RB, 1, W, false, NT, true>(b, twW, tid);
    }
    KYM_HD KYM_INL void row_fwd(cpx<T>* b, const FftLen<T>&, int, int tid, int) const {
        fft_dit_s<false, W, RB, 1, W, false, NT, true>(b, twW, tid);
    }
    // Row inverse whose last pass writes the modulus R[t rs + x] = |X| scale
    // (natural x; R does not overlap the rows) and, if sv is set, X at its
    // storage index; and the row forward whose first pass reads R.
    static constexpr bool FUSED = true;
    KYM_HD KYM_INL void row_inv_modulus(cpx<T>* b, T* R, int rs, T scale, cpx<T>* sv, int tid) const {
        constexpr int R0 = kym_factor_list(W).rad[0];
        fft_dif_s<true, W, RB, 1, W, false, NT, true, T, -1, W, 1>(b, twW, tid);
        dif_last_pass_s<true, W, RB, 1, W, false, NT, T, false>(
            b, tid,
            [&](int t, int bb, int m, cpx<T> v) -> T {
                if (sv) sv[t * W + bb * R0 + m] = v;
                return kym_sqrt(v.re * v.re + v.im * v.im) * scale;
            },
            [&](int t, int bb, int m, T val) { R[t * rs + kym_rev_s<W>(bb * R0) + (W / R0) * m] = val; });
    }
    KYM_HD KYM_INL void row_fwd_real(cpx<T>* b, const T* R, int rs, int tid) const {
        constexpr int R0 = kym_factor_list(W).rad[0];
        dit_first_pass_real_s<false, W, RB, 1, W, false, NT, T, false>(b, tid, [&](int t, int bb, int m) -> T {
            return R[t * rs + kym_rev_s<W>(bb * R0) + (W / R0) * m];
        });
        fft_dit_s<false, W, RB, 1, W, false, NT, true, T, 1, R0>(b, twW, tid);
    }
};

// Row pass of a big node's FIR leaf: rows rr < nr of R (natural columns,
// row stride rs) are the node's rows rev[r0 + rr]; T[y][x] goes to the
// tslot.  Bx: band table space after R (fir_kmax(W, ..) x 16 reals) where
// the band is kept in LDS (fir_band_in_lds); else, where the rows hold one
// MFMA tile (fir_rows_split), the 4 KB of partial sums of the waves that
// share the tile's k-steps.
template <class T>
KYM_HD void fir_rows_to_slot(const Scat2DArgs<T>& a, const Node2D& nd, const T* R, int rs, int nr, int r0,
                             const int* rev, cpx<T>* wsi, T* Bx, int tid, int nth) {
    const int r = nd.res, W = a.W[r], ow = a.ow, kl = nd.leaf_k;
    const FirAxis fx = a.fir[2 * r + 1];
    const bool in_lds = fir_band_in_lds(fx.kmax, (int)sizeof(T));
    if (in_lds) fir_band(Bx, a.filt, fx, tid, nth);
    T* red = (!in_lds && fir_rows_split(a.blk, ow)) ? Bx : nullptr;
    wg_barrier();  // also: R complete
    T* Tg = reinterpret_cast<T*>(wsi + nd.tslot);
    auto out = [&](int p, int x, float v) { Tg[(long long)rev[r0 + p] * ow + x] = v; };
    // separate calls, so that each reads its band through a known address space
    if (in_lds) fir_rows_mma(R, rs, nr, W, kl, ow, fx, Bx, out, tid, nth);
    else if (red) fir_rows_mma_split(R, rs, nr, W, kl, ow, fx, a.filt + fx.band_off, out, tid, nth, red);
    else fir_rows_mma(R, rs, nr, W, kl, ow, fx, a.filt + fx.band_off, out, tid, nth);
}

// The row spectra of a real field are Hermitian along x (row[W - x] =
// conj(row[x])), so the row passes store columns x <= W / 2 only and the
// forward column pass (body_fwd_cols) writes the other half by symmetry.
template <class T>
KYM_HD KYM_INL void store_half_rows(cpx<T>* dst, const cpx<T>* src, int nr, int W, int tid, int nth) {
    const int Wh = W / 2 + 1;
    for (int i = tid; i < nr * Wh; i += nth) {
        const int rr = i / Wh, x = i - rr * Wh;
        dst[(long long)rr * W + x] = src[rr * W + x];
    }
}

// Root row pass: rows of the padded image (row storage positions
// digit-reversed), forward DIT along x.  task -> (image, row block).  With a
// FIR leaf the block's rows are first loaded in natural order as reals (after
// the complex rows in LDS); the leaf's row pass goes to the node's tslot.
template <class T, class P>
KYM_HD void body_root_rows(const P& fp, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int il = task / a.nblk, rb = task - il * a.nblk;
    const int img = a.img0 + il;
    const Node2D& nd = a.nodes[0];
    const int H = a.H[0], W = a.W[0];
    const FftLen<T>& fH = a.lens[a.lenH[0]];
    const FftLen<T>& fW = a.lens[a.lenW[0]];
    const int r0 = rb * a.blk, nr = (H - r0 < a.blk) ? H - r0 : a.blk;
    const T* xin = a.x + (long long)img * a.M * a.N;
    if constexpr (kym_fir<T>) {
        T* R = reinterpret_cast<T*>(lds + nr * W);
        const int rs = W + 1;
        copy_batched<8>(
            nr * W,
            [&](int i) {
                const int rr = i / W, x = i - rr * W;
                return input_pixel(a, xin, fH.rev[r0 + rr], x);
            },
            [&](int i, T v) {
                const int rr = i / W, x = i - rr * W;
                R[rr * rs + x] = v;
            },
            tid, nth);
        wg_barrier();
        for (int i = tid; i < nr * W; i += nth) {
            const int rr = i / W, px = i - rr * W;
            lds[i] = {R[rr * rs + fW.rev[px]], T(0)};
        }
        fir_rows_to_slot(a, nd, R, rs, nr, r0, fH.rev, a.ws + il * a.ws_img, R + nr * rs, tid, nth);
    } else {
        copy_batched<8>(
            nr * W,
            [&](int i) {
                const int rr = i / W, px = i - rr * W;
                return input_pixel(a, xin, fH.rev[r0 + rr], fW.rev[px]);
            },
            [&](int i, T v) { lds[i] = {v, T(0)}; }, tid, nth);
    }
    wg_barrier();
    fp.row_fwd(lds, fW, nr, tid, nth);
    store_half_rows(a.ws + il * a.ws_img + nd.slot + (long long)r0 * W, lds, nr, W, tid, nth);
}

// Gather + inverse column DIF.  Shared by big nodes (filter psi, slot) and
// big leaves (filter phi, leaf_slot).  task -> (image, node, column block).
template <class T, class P>
KYM_HD KYM_INL void icol_common(const P& fp, const Scat2DArgs<T>& a, const cpx<T>* U, const T* f, int Wp, int H, int W, int k,
                        const FftLen<T>& fH, cpx<T>* dst, int c0, int nc, cpx<T>* lds, int tid, int nth) {
    gather_block(lds, nc, c0, U, f, Wp, H, W, k, tid, nth);
    wg_barrier();
    fp.col_inv(lds, fH, nc, tid, nth);
    for (int i = tid; i < H * nc; i += nth) {
        const int p = i / nc, c = i - (i / nc) * nc;
        dst[(long long)p * W + c0 + c] = lds[i];
    }
}

template <class T, class P>
KYM_HD void body_icol(const P& fp, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int per = a.ngroup * a.nblk;
    const int il = task / per, rem = task - il * per;
    const int gi = rem / a.nblk, cb = rem - gi * a.nblk;
    const Node2D& nd = a.nodes[a.group[gi]];
    const Node2D& pn = a.nodes[nd.parent];
    const int r = nd.res, H = a.H[r], W = a.W[r];
    const int c0 = cb * a.blk, nc = (W - c0 < a.blk) ? W - c0 : a.blk;
    cpx<T>* wsi = a.ws + il * a.ws_img;
    icol_common<T>(fp, a, wsi + pn.slot, a.filt + a.filt_off[nd.filt], a.W[pn.res], H, W, nd.k, a.lens[a.lenH[r]],
                wsi + nd.slot, c0, nc, lds, tid, nth);
}

// Gather + inverse column DIF of two sibling big nodes (same parent and
// periodisation; the plan lists them at group[2 gi] and group[2 gi + 1]):
// one load of each parent value serves both filters.  Full column blocks
// only (compile-time passes).  LDS: the two blocks, H x blk each.
// task -> (image, pair, column block).
template <class T, class P>
KYM_HD void body_icol_pair(const P& fp, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int per = a.ngroup * a.nblk;
    const int il = task / per, rem = task - il * per;
    const int gi = rem / a.nblk, cb = rem - gi * a.nblk;
    const Node2D& na = a.nodes[a.group[2 * gi]];
    const Node2D& nb = a.nodes[a.group[2 * gi + 1]];
    const Node2D& pn = a.nodes[na.parent];
    const int r = na.res, H = a.H[r], W = a.W[r];
    const int c0 = cb * a.blk, nc = a.blk;
    cpx<T>* wsi = a.ws + il * a.ws_img;
    cpx<T>* la = lds;
    cpx<T>* lb = lds + H * nc;
    gather_block_pair(la, lb, nc, c0, wsi + pn.slot, a.filt + a.filt_off[na.filt], a.filt + a.filt_off[nb.filt], H, W,
                      na.k, tid, nth);
    wg_barrier();
    const FftLen<T>& fH = a.lens[a.lenH[r]];
    fp.col_inv(la, fH, nc, tid, nth);
    fp.col_inv(lb, fH, nc, tid, nth);
    cpx<T>* da = wsi + na.slot;
    cpx<T>* db = wsi + nb.slot;
    for (int i = tid; i < H * nc; i += nth) {
        const int p = i / nc, c = i - (i / nc) * nc;
        da[(long long)p * W + c0 + c] = la[i];
        db[(long long)p * W + c0 + c] = lb[i];
    }
}

template <class T, int MAXP>
KYM_HD void body_leaf_icol(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const DynPass<T, MAXP> fp;
    const int per = a.ngroup * a.nblk;
    const int il = task / per, rem = task - il * per;
    const int gi = rem / a.nblk, cb = rem - gi * a.nblk;
    const Node2D& nd = a.nodes[a.group[gi]];
    const int J = a.J, HJ = a.H[J], WJ = a.W[J];
    const int c0 = cb * a.blk, nc = (WJ - c0 < a.blk) ? WJ - c0 : a.blk;
    cpx<T>* wsi = a.ws + il * a.ws_img;
    icol_common<T>(fp, a, wsi + nd.slot, a.filt + a.filt_off[nd.leaf_filt], a.W[nd.res], HJ, WJ, nd.leaf_k,
                a.lens[a.lenH[J]], wsi + nd.leaf_slot, c0, nc, lds, tid, nth);
}

// Rows: inverse DIF along x, modulus, forward DIT along x (in place).  With
// a FIR leaf the modulus rows are also kept as natural reals after the complex
// rows, for the leaf's row pass; nodes without children stop there.
template <class T, class P>
KYM_HD void body_rows_mod(const P& fp, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int per = a.ngroup * a.nblk;
    const int il = task / per, rem = task - il * per;
    const int gi = rem / a.nblk, rb = rem - gi * a.nblk;
    const Node2D& nd = a.nodes[a.group[gi]];
    const int r = nd.res, H = a.H[r], W = a.W[r];
    const FftLen<T>& fW = a.lens[a.lenW[r]];
    const int r0 = rb * a.blk, nr = (H - r0 < a.blk) ? H - r0 : a.blk;
    cpx<T>* wsi = a.ws + il * a.ws_img;
    cpx<T>* buf = wsi + nd.slot + (long long)r0 * W;
    copy_batched<8>(nr * W, [&](int i) { return kym_ldg(buf + i); }, [&](int i, cpx<T> v) { lds[i] = v; }, tid, nth);
    wg_barrier();
    const T inv_n = T(1) / T(H * W);
    cpx<T>* sv = (a.save && nd.save >= 0) ? wsi + nd.save + (long long)r0 * W : nullptr;
    T* R = reinterpret_cast<T*>(lds + nr * W);
    const int rs = W + 1;
    if constexpr (P::FUSED && kym_fir<T>) {
        fp.row_inv_modulus(lds, R, rs, inv_n, sv, tid);
        fir_rows_to_slot(a, nd, R, rs, nr, r0, a.lens[a.lenH[r]].rev, wsi, R + nr * rs, tid, nth);
        if (!nd.fwd) return;
        fp.row_fwd_real(lds, R, rs, tid);
        store_half_rows(buf, lds, nr, W, tid, nth);
        return;
    }
    fp.row_inv(lds, fW, nr, tid, nth);
    for (int i = tid; i < nr * W; i += nth) {
        const cpx<T> v = lds[i];
        if (sv) sv[i] = v;
        const T m = kym_sqrt(v.re * v.re + v.im * v.im) * inv_n;
        lds[i] = {m, T(0)};
        if constexpr (kym_fir<T>) {
            const int rr = i / W, px = i - rr * W;
            R[rr * rs + fW.rev[px]] = m;
        }
    }
    wg_barrier();
    if constexpr (kym_fir<T>) fir_rows_to_slot(a, nd, R, rs, nr, r0, a.lens[a.lenH[r]].rev, wsi, R + nr * rs, tid, nth);
    if (!nd.fwd) return;
    fp.row_fwd(lds, fW, nr, tid, nth);
    store_half_rows(buf, lds, nr, W, tid, nth);
}

// Column pass of a big node's FIR leaf: tslot -> LDS (row stride
// fir_tstride) -> output channel.  task -> (image, node).
template <class T, int MAXP>
KYM_HD void body_fir_cols(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    if constexpr (kym_fir<T>) {
        const int il = task / a.ngroup, gi = task - il * a.ngroup;
        const Node2D& nd = a.nodes[a.group[gi]];
        const int r = nd.res, H = a.H[r], W = a.W[r], ow = a.ow, oh = a.oh, kl = nd.leaf_k;
        const FirAxis fy = a.fir[2 * r];
        const int ts = fir_tstride(W, ow), kym = fir_kmax(H, kl, fy.cnt);
        T* Tb = reinterpret_cast<T*>(lds);
        T* Ay = Tb + H * ts;
        const T* Tg = reinterpret_cast<const T*>(a.ws + il * a.ws_img + nd.tslot);
        copy_batched<8>(
            H * ow, [&](int i) { return kym_ldg(Tg + i); },
            [&](int i, T v) {
                const int p = i / ow, x = i - p * ow;
                Tb[p * ts + x] = v;
            },
            tid, nth);
        fir_band(Ay, a.filt, fy, tid, nth);
        wg_barrier();
        fir_cols_mma(Tb, ts, H, kl, oh, ow, fy, Ay, a.out + ((long long)(a.img0 + il) * a.C + nd.leaf_ch) * oh * ow,
                     tid, nth);
    } else {
        (void)a; (void)task; (void)lds; (void)tid; (void)nth;
    }
}

// Rows of a big leaf: inverse DIF along x, real part, unpad, store.
template <class T, int MAXP>
KYM_HD void body_leaf_rows(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int per = a.ngroup * a.nblk;
    const int il = task / per, rem = task - il * per;
    const int gi = rem / a.nblk, rb = rem - gi * a.nblk;
    const Node2D& nd = a.nodes[a.group[gi]];
    const int J = a.J, HJ = a.H[J], WJ = a.W[J];
    const FftLen<T>& fH = a.lens[a.lenH[J]];
    const FftLen<T>& fW = a.lens[a.lenW[J]];
    const int r0 = rb * a.blk, nr = (HJ - r0 < a.blk) ? HJ - r0 : a.blk;
    const cpx<T>* buf = a.ws + il * a.ws_img + nd.leaf_slot + (long long)r0 * WJ;
    for (int i = tid; i < nr * WJ; i += nth) lds[i] = buf[i];
    wg_barrier();
    fft_dif<true, MAXP>(lds, fW, Batch{nr, 1, WJ, false}, tid, nth);
    const T inv_n = T(1) / T(HJ * WJ);
    T* o = a.out + ((long long)(a.img0 + il) * a.C + nd.leaf_ch) * a.oh * a.ow;
    for (int i = tid; i < nr * WJ; i += nth) {
        const int rr = i / WJ, px = i - rr * WJ;
        const int y = fH.rev[r0 + rr], x = fW.rev[px];
        if (y >= 1 && y <= HJ - 2 && x >= 1 && x <= WJ - 2) o[(y - 1) * a.ow + (x - 1)] = lds[i].re * inv_n;
    }
}

// Forward DIT along y over column blocks (rows stored digit-reversed) -> natural.
template <class T, class P>
KYM_HD void body_fwd_cols(const P& fp, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int per = a.ngroup * a.nblk;
    const int il = task / per, rem = task - il * per;
    const int gi = rem / a.nblk, cb = rem - gi * a.nblk;
    const Node2D& nd = a.nodes[a.group[gi]];
    const int r = nd.res, H = a.H[r], W = a.W[r];
    const FftLen<T>& fH = a.lens[a.lenH[r]];
    const int c0 = cb * a.blk, nc = (W - c0 < a.blk) ? W - c0 : a.blk;
    cpx<T>* spec = a.ws + il * a.ws_img + nd.slot;
    const cpx<T>* buf = spec + c0;
    copy_batched<8>(
        H * nc,
        [&](int i) {
            const int y = i / nc, c = i - y * nc;
            return kym_ldg(buf + (long long)y * W + c);
        },
        [&](int i, cpx<T> v) { lds[i] = v; }, tid, nth);
    wg_barrier();
    fp.col_fwd(lds, fH, nc, tid, nth);
    // the launch covers columns 0 .. W / 2 (store_half_rows; columns of its
    // last block beyond that are transformed, not stored); column W - x of the
    // spectrum is X[v][W - x] = conj(X[(H - v) mod H][x])
    const int Wh = W / 2;
    for (int i = tid; i < H * nc; i += nth) {
        const int v = i / nc, x = c0 + i - v * nc;
        if (x > Wh) continue;
        const cpx<T> X = lds[i];
        spec[(long long)v * W + x] = X;
        if (x > 0 && W - x != x) spec[(long long)(v == 0 ? 0 : H - v) * W + (W - x)] = {X.re, -X.im};
    }
}

// Compile-time specialisations X(id, H, W, HJ, WJ, threads, pair) of the LDS
// node kernel (FP32).  A launch uses one when its sizes match; every other
// size runs the run-time-sized bodies.  pair = 1: sibling pairs per
// workgroup (body_lds_pair_s; sizes whose LDS holds the second node's
// spectrum beside the first node); pair = 2: the same, used by default only
// where a launch's deepest periodisation is k >= 16, or when the plan asks
// for every pair kernel (tuning field pair = 2).
#define KYM_NODE_SPECS(X)              \
    X(0, 136, 136, 34, 34, 1024, 0)    \
    X(1, 68, 68, 34, 34, 512, 0)       \
    X(2, 80, 80, 10, 10, 512, 0)       \
    X(3, 40, 40, 10, 10, 256, 0)       \
    X(4, 20, 20, 10, 10, 256, 0)       \
    X(5, 68, 68, 34, 34, 256, 0)       \
    X(6, 80, 80, 10, 10, 256, 0)       \
    X(7, 136, 136, 34, 34, 512, 0)     \
    X(8, 68, 68, 34, 34, 512, 1)       \
    X(9, 68, 68, 34, 34, 256, 1)       \
    X(10, 20, 20, 10, 10, 256, 2)      \
    X(11, 40, 40, 10, 10, 256, 2)      \
    X(12, 80, 80, 10, 10, 512, 2)

// want_threads = 0: the first (default) specialisation of these sizes.
inline int node_spec_id(int H, int W, int HJ, int WJ, int want_threads, int* threads, int want_pair = 0) {
#define KYM_NS_MATCH(id, h, w, hj, wj, nt, pr)                                                  \
    if (H == h && W == w && HJ == hj && WJ == wj && (want_threads == 0 || want_threads == nt) && \
        pr == want_pair) {                                                                      \
        *threads = nt;                                                                          \
        return id;                                                                              \
    }
    KYM_NODE_SPECS(KYM_NS_MATCH)
#undef KYM_NS_MATCH
    return -1;
}

// The body of specialisation SPEC (selected at compile time, so a kernel
// carries only its own body).  EMU = true: one thread per workgroup (CPU
// harness), otherwise the spec's count.
template <class T, int SPEC, bool EMU>
KYM_HD KYM_INL void node_spec_body(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
#define KYM_NS_BODY(sid, h, w, hj, wj, nt, pr)                                                    \
    if constexpr (SPEC == sid) {                                                                 \
        if constexpr (pr) body_lds_pair_s<T, h, w, hj, wj, (EMU ? 1 : nt)>(a, task, lds, tid);    \
        else body_lds_node_s<T, h, w, hj, wj, (EMU ? 1 : nt)>(a, task, lds, tid);                 \
    }
    KYM_NODE_SPECS(KYM_NS_BODY)
#undef KYM_NS_BODY
}

// Run-time selection (the CPU harness).
template <class T, bool EMU>
KYM_HD void dispatch_node_spec(int id, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
    switch (id) {
#define KYM_NS_CASE(sid, h, w, hj, wj, nt, pr) \
    case sid: node_spec_body<T, sid, EMU>(a, task, lds, tid); break;
        KYM_NODE_SPECS(KYM_NS_CASE)
#undef KYM_NS_CASE
        default: break;
    }
}

// Compile-time specialisations X(id, H, W, col block, row block, threads) of
// the big-node passes (FP32): root rows, inverse columns, rows + modulus,
// forward columns, and the backward's passes.  272: 256x256 at J=3; 160 /
// 320 / 640: 128x128 at J=4, 256x256 at J=5 and 512x512 at J=6 (the
// mixed-size batch).  An untuned forward takes the first entry of its sizes
// and rows, so each size lists its forward block first; 160 x 160 by 16 x 8
// serves the backward (its gradient gathers read 16 consecutive columns).
#define KYM_PASS_SPECS(X)         \
    X(0, 272, 272, 16, 8, 256)    \
    X(1, 272, 272, 16, 4, 256)    \
    X(2, 272, 272, 8, 8, 256)     \
    X(3, 272, 272, 32, 8, 256)    \
    X(4, 272, 272, 16, 16, 256)   \
    X(5, 272, 272, 16, 16, 320)   \
    X(6, 272, 272, 17, 17, 320)   \
    X(7, 272, 272, 8, 8, 192)     \
    X(8, 160, 160, 8, 8, 256)     \
    X(9, 320, 320, 8, 8, 256)     \
    X(10, 640, 640, 8, 8, 256)    \
    X(11, 160, 160, 16, 8, 256)   \
    X(12, 160, 160, 32, 8, 256)

// Columns of the first specialisation of these sizes and rows (0: none).
inline int pass_spec_first_cols(int H, int W, int rb) {
#define KYM_PS_FIRST(id, h, w, c, r, nt) \
    if (H == h && W == w && rb == r) return c;
    KYM_PASS_SPECS(KYM_PS_FIRST)
#undef KYM_PS_FIRST
    return 0;
}

// want_threads = 0: the first specialisation of these sizes and blocks.
inline int pass_spec_id(int H, int W, int cb, int rb, int want_threads, int* threads) {
#define KYM_PS_MATCH(id, h, w, c, r, nt)                                                        \
    if (H == h && W == w && cb == c && rb == r && (want_threads == 0 || want_threads == nt)) { \
        *threads = nt;                                                                        \
        return id;                                                                            \
    }
    KYM_PASS_SPECS(KYM_PS_MATCH)
#undef KYM_PS_MATCH
    return -1;
}

enum Kind2D {
    K2_LDS_NODE = 0, K2_LEAF, K2_ROOT_ROWS, K2_ICOL, K2_ROWS_MOD, K2_FWD_COLS, K2_LEAF_ICOL, K2_LEAF_ROWS, K2_FIR_COLS,
    K2_ICOL_PAIR
};

template <class T, int KIND, class P>
KYM_HD KYM_INL void dispatch_pass(const P& fp, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    if constexpr (KIND == K2_ROOT_ROWS) body_root_rows<T>(fp, a, task, lds, tid, nth);
    else if constexpr (KIND == K2_ICOL) body_icol<T>(fp, a, task, lds, tid, nth);
    else if constexpr (KIND == K2_ROWS_MOD) body_rows_mod<T>(fp, a, task, lds, tid, nth);
    else if constexpr (KIND == K2_FWD_COLS) body_fwd_cols<T>(fp, a, task, lds, tid, nth);
    else if constexpr (KIND == K2_ICOL_PAIR) body_icol_pair<T>(fp, a, task, lds, tid, nth);
}

// The plan's FFT length record of length n (one of its H_r or W_r).
template <class T>
KYM_HD KYM_INL const FftLen<T>* len_of(const Scat2DArgs<T>& a, int n) {
    for (int r = 0; r <= a.J; ++r) {
        if (a.H[r] == n) return &a.lens[a.lenH[r]];
        if (a.W[r] == n) return &a.lens[a.lenW[r]];
    }
    return &a.lens[a.lenH[0]];
}

// Static pass body of specialisation SPEC, selected at compile time so a
// kernel carries only its own body.  EMU = true runs it with one thread (CPU
// harness).
template <class T, int KIND, int SPEC, bool EMU>
KYM_HD KYM_INL void pass_spec_body(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
#define KYM_PS_BODY(sid, h, w, c, r, nt)                                                       \
    if constexpr (SPEC == sid) {                                                               \
        const FftLen<T>* fh = len_of(a, h);                                                    \
        const FftLen<T>* fw = len_of(a, w);                                                    \
        load_twiddles(lds, fh, fw, h, w, tid, EMU ? 1 : nt);                                   \
        const StatPass<T, h, w, c, r, (EMU ? 1 : nt)> fp{lds + (h != w ? w : 0), lds};         \
        dispatch_pass<T, KIND>(fp, a, task, lds + kym_tw_elems(h, w), tid, EMU ? 1 : nt);      \
    }
    KYM_PASS_SPECS(KYM_PS_BODY)
#undef KYM_PS_BODY
}

// Run-time selection (the CPU harness).
template <class T, int KIND, bool EMU>
KYM_HD void dispatch_pass_spec(int id, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
    switch (id) {
#define KYM_PS_CASE(sid, h, w, c, r, nt) \
    case sid: pass_spec_body<T, KIND, sid, EMU>(a, task, lds, tid); break;
        KYM_PASS_SPECS(KYM_PS_CASE)
#undef KYM_PS_CASE
        default: break;
    }
}

template <class T, int MAXP>
KYM_HD KYM_INL void dispatch_body2d(int kind, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    switch (kind) {
        case K2_LDS_NODE: body_lds_node<T, MAXP>(a, task, lds, tid, nth); break;
        case K2_LEAF: body_leaf<T, MAXP>(a, task, lds, tid, nth); break;
        case K2_ROOT_ROWS: body_root_rows<T>(DynPass<T, MAXP>(), a, task, lds, tid, nth); break;
        case K2_ICOL: body_icol<T>(DynPass<T, MAXP>(), a, task, lds, tid, nth); break;
        case K2_ROWS_MOD: body_rows_mod<T>(DynPass<T, MAXP>(), a, task, lds, tid, nth); break;
        case K2_FWD_COLS: body_fwd_cols<T>(DynPass<T, MAXP>(), a, task, lds, tid, nth); break;
        case K2_LEAF_ICOL: body_leaf_icol<T, MAXP>(a, task, lds, tid, nth); break;
        case K2_LEAF_ROWS: body_leaf_rows<T, MAXP>(a, task, lds, tid, nth); break;
        case K2_FIR_COLS: body_fir_cols<T, MAXP>(a, task, lds, tid, nth); break;
        default: break;
    }
}
