// Scattering2D backward: hand-written adjoint of every stage, node by node
// in reverse order (order 2, order 1, root), each node pulling the input
// gradients of its children.  With L real and complex gradients written as
// dL/dRe + i dL/dIm (the convention of torch autograd):
//   leaf   S = unpad(Re(ifft2(Y)))         dL/dY = fft2(unpad^T dS) / n_J
//          Y = P_kl(U * phi)               dL/dU += phi * tile(dL/dY) / kl^2
//   child  X_c = P_kc(U * f_c)             dL/dU += f_c * tile(dL/dX_c) / kc^2
//   U = fft2(m), m real                    dL/dm = Re(n ifft2(dL/dU))  (unnormalised inverse)
//   m = |u|                                dL/du = dL/dm * u / |u|, 0 where u = 0
//                                          (kymatio/backend/torch_backend.py:63-96)
//   u = ifft2(X)                           dL/dX = fft2(dL/du) / n
//   root U_0 = fft2(pad(x))                dL/dpad(x) = Re(n ifft2(dL/dU_0)), folded by the
//                                          reflect-pad adjoint into dL/dx.
// The forward run in training mode stores u (unnormalised, in the
// digit-reversed order the inverse DIF leaves it in) in the node's save slot;
// the backward only reads it and writes dL/dX (natural order) to the node's
// gsave slot, so it can run any number of times on one forward.
#pragma once
#include "scat2d_body.h"

// dL/dU at E points (ys[e], xs[e]) of a node of H x W: the leaf term (GY:
// HJ x WJ, already scaled by 1 / (n_J kl^2)) plus the children's dL/dX
// (scaled here by 1 / kc^2), summed in that order.  Each child's loads for
// all E points are issued together, four children at a time.  A child of
// resolution r_p + l reads its dL/dX at (y mod H / 2^l, x mod W / 2^l); the
// plan lists children by ascending resolution, so these residues are kept
// running (one conditional subtraction per level) instead of one integer
// division per child and point; a child out of that order takes the division.
template <int E, class T>
KYM_HD KYM_INL void grad_u_elems(const Scat2DArgs<T>& a, const Node2D& nd, const cpx<T>* wsi, const cpx<T>* GY, int W,
                                 int HJ, int WJ, const int* ys, const int* xs, cpx<T>* acc) {
    const int H = a.H[nd.res];
    long long idx[E];
    int yr[E], xr[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        idx[e] = (long long)ys[e] * W + xs[e];
        yr[e] = ys[e];
        xr[e] = xs[e];
        const T ph = kym_ldg(a.filt + a.filt_off[nd.leaf_filt] + idx[e]);
        const cpx<T> gy = GY[(ys[e] % HJ) * WJ + (xs[e] % WJ)];
        acc[e] = {gy.re * ph, gy.im * ph};
    }
    int lv = 0;  // level of the running residues yr, xr
    auto child_at = [&](const Node2D& cn, int e, int Hc, int Wc) -> long long {
        const int l = cn.res - nd.res;
        if (l < lv) return (long long)(ys[e] % Hc) * Wc + (xs[e] % Wc);
        return (long long)yr[e] * Wc + xr[e];
    };
    auto advance = [&](const Node2D& cn) {
        const int l = cn.res - nd.res;
        while (lv < l) {
            ++lv;
            const int hh = H >> lv, hw = W >> lv;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                yr[e] -= yr[e] >= hh ? hh : 0;
                xr[e] -= xr[e] >= hw ? hw : 0;
            }
        }
    };
    int c = nd.child_begin;
    for (; c + 4 <= nd.child_end; c += 4) {  // four children x E elements of loads in flight
        T fv[4][E];
        cpx<T> gx[4][E];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const Node2D& cn = a.nodes[c + u];
            advance(cn);
            const int Hc = a.H[cn.res], Wc = a.W[cn.res];
            const T* f = a.filt + a.filt_off[cn.filt];
            const T inv = T(1) / T(cn.k * cn.k);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                fv[u][e] = kym_ldg(f + idx[e]) * inv;
                gx[u][e] = kym_ldg(wsi + cn.gsave + child_at(cn, e, Hc, Wc));
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc[e].re += gx[u][e].re * fv[u][e];
                acc[e].im += gx[u][e].im * fv[u][e];
            }
    }
    for (; c < nd.child_end; ++c) {
        const Node2D& cn = a.nodes[c];
        advance(cn);
        const int Hc = a.H[cn.res], Wc = a.W[cn.res];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const T fv = kym_ldg(a.filt + a.filt_off[cn.filt] + idx[e]) / T(cn.k * cn.k);
            const cpx<T> gx = kym_ldg(wsi + cn.gsave + child_at(cn, e, Hc, Wc));
            acc[e].re += gx.re * fv;
            acc[e].im += gx.im * fv;
        }
    }
}

// dL/dU at the A x B alias points (ys[e] + i H/2, xs[e] + j W/2), i < A,
// j < B, of E base points (acc[e][i B + j]).  A child at half the node's
// resolution or coarser, and the leaf at 2^-J of the padded size, have sizes
// dividing H/2 and W/2 (the padded sizes are multiples of 2^J): all alias
// points of a base point read the same child gradient and leaf value, loaded
// once per base point.  Children at the node's own resolution (the root's
// j1 = 0 children, listed first) are read per point, and the filters are.
// Per point the terms are summed in grad_u_elems' order, so results are
// identical to it.
template <int A, int B, int E, class T>
KYM_HD KYM_INL void grad_u_alias(const Scat2DArgs<T>& a, const Node2D& nd, const cpx<T>* wsi, const cpx<T>* GY, int W,
                                 int HJ, int WJ, const int* ys, const int* xs, cpx<T> (*acc)[A * B]) {
    constexpr int P = A * B;
    const int H = a.H[nd.res], hh = H / 2, hw = W / 2;
    int idx[E][P];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const cpx<T> gy = GY[(ys[e] % HJ) * WJ + (xs[e] % WJ)];
        const T* ph = a.filt + a.filt_off[nd.leaf_filt];
#pragma unroll
        for (int i = 0; i < A; ++i)
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const int q = i * B + j;
                idx[e][q] = (ys[e] + i * hh) * W + xs[e] + j * hw;
                const T f = kym_ldg(ph + idx[e][q]);
                acc[e][q] = {gy.re * f, gy.im * f};
            }
    }
    int yr[E], xr[E];
#pragma unroll
    for (int e = 0; e < E; ++e) { yr[e] = ys[e]; xr[e] = xs[e]; }
    int lv = 0;  // level of the running residues yr, xr (as in grad_u_elems)
    auto child_at = [&](const Node2D& cn, int e, int Hc, int Wc) -> long long {
        const int l = cn.res - nd.res;
        if (l < lv) return (long long)(ys[e] % Hc) * Wc + (xs[e] % Wc);
        return (long long)yr[e] * Wc + xr[e];
    };
    auto advance = [&](const Node2D& cn) {
        const int l = cn.res - nd.res;
        while (lv < l) {
            ++lv;
            const int h2 = H >> lv, w2 = W >> lv;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                yr[e] -= yr[e] >= h2 ? h2 : 0;
                xr[e] -= xr[e] >= w2 ? w2 : 0;
            }
        }
    };
    int c = nd.child_begin;
    // leading children at the node's own resolution (the root's j1 = 0
    // children): each point reads its own gradient, at the filter's index
    for (; c < nd.child_end && a.nodes[c].res == nd.res; ++c) {
        const Node2D& cn = a.nodes[c];
        const T* f = a.filt + a.filt_off[cn.filt];
        const T inv = T(1) / T(cn.k * cn.k);
#pragma unroll
        for (int e = 0; e < E; ++e)
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const T fv = kym_ldg(f + idx[e][q]) * inv;
                const cpx<T> gx = kym_ldg(wsi + cn.gsave + idx[e][q]);
                acc[e][q].re += gx.re * fv;
                acc[e][q].im += gx.im * fv;
            }
    }
    for (; c + 4 <= nd.child_end; c += 4) {
        T fv[4][E][P];
        cpx<T> gx[4][E];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const Node2D& cn = a.nodes[c + u];
            advance(cn);
            const int Hc = a.H[cn.res], Wc = a.W[cn.res];
            const T* f = a.filt + a.filt_off[cn.filt];
            const T inv = T(1) / T(cn.k * cn.k);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                gx[u][e] = kym_ldg(wsi + cn.gsave + child_at(cn, e, Hc, Wc));
#pragma unroll
                for (int q = 0; q < P; ++q) fv[u][e][q] = kym_ldg(f + idx[e][q]) * inv;
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e)
#pragma unroll
            for (int q = 0; q < P; ++q)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    acc[e][q].re += gx[u][e].re * fv[u][e][q];
                    acc[e][q].im += gx[u][e].im * fv[u][e][q];
                }
    }
    for (; c < nd.child_end; ++c) {
        const Node2D& cn = a.nodes[c];
        advance(cn);
        const int Hc = a.H[cn.res], Wc = a.W[cn.res];
        const T* f = a.filt + a.filt_off[cn.filt];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const cpx<T> gx = kym_ldg(wsi + cn.gsave + child_at(cn, e, Hc, Wc));
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const T fv = kym_ldg(f + idx[e][q]) / T(cn.k * cn.k);
                acc[e][q].re += gx.re * fv;
                acc[e][q].im += gx.im * fv;
            }
        }
    }
}

// Whether a node's alias points share their child and leaf reads: sizes
// even, the leaf's a divisor of the half sizes (false at J = 0), and, after
// a leading run of children at the node's own resolution (the root's j1 = 0
// children, read per point), every child coarser than the node.
template <class T>
KYM_HD KYM_INL bool grad_alias_ok(const Scat2DArgs<T>& a, const Node2D& nd) {
    const int H = a.H[nd.res], W = a.W[nd.res], HJ = a.H[a.J], WJ = a.W[a.J];
    if (H % 2 != 0 || W % 2 != 0 || (H / 2) % HJ != 0 || (W / 2) % WJ != 0) return false;
    int c = nd.child_begin;
    while (c < nd.child_end && a.nodes[c].res == nd.res) ++c;
    for (; c < nd.child_end; ++c)
        if (a.nodes[c].res <= nd.res) return false;
    return true;
}

// Leaf gradient into B (HJ x WJ, natural), scaled by 1 / (n_J kl^2).
template <class T, class G>
KYM_HD void leaf_grad(const Scat2DArgs<T>& a, const G& g, const Node2D& nd, int img, cpx<T>* B, int tid, int nth) {
    const T* d = a.dS + ((long long)img * a.C + nd.leaf_ch) * a.oh * a.ow;
    const int* rH = g.revHJ();
    const int* rW = g.revWJ();
    for (int p = tid; p < g.HJ * g.WJ; p += nth) {
        const int py = p / g.WJ, px = p - py * g.WJ;
        const int y = rH[py], x = rW[px];
        T v = T(0);
        if (y >= 1 && y <= g.HJ - 2 && x >= 1 && x <= g.WJ - 2) v = d[(y - 1) * a.ow + (x - 1)];
        B[p] = {v, T(0)};
    }
    wg_barrier();
    g.leaf_fwd2d(B, tid, nth);
    const T sc = T(1) / (T(g.HJ * g.WJ) * T(nd.leaf_k * nd.leaf_k));
    for (int p = tid; p < g.HJ * g.WJ; p += nth) B[p] = B[p] * sc;
    wg_barrier();
}

// A whole node's backward in LDS.
template <class T, class G>
KYM_HD void bwd_node_body(const Scat2DArgs<T>& a, const G& g, const Node2D& nd, int il, cpx<T>* lds, int tid,
                          int nth) {
    const int img = a.img0 + il;
    cpx<T>* wsi = a.ws + il * a.ws_img;
    const int n = g.H * g.W;
    const cpx<T>* GY;
    if (nd.leaf_mode == 0) {
        leaf_grad(a, g, nd, img, lds + n, tid, nth);
        GY = lds + n;
    } else {
        GY = wsi + nd.gslot;
    }
    constexpr int E = 2;  // elements per thread and round
    if (nd.child_end > nd.child_begin && grad_alias_ok(a, nd)) {
        // the points (y, x), (y, x + W/2), (y + H/2, x), (y + H/2, x + W/2)
        // share their child and leaf reads (one base point per thread and
        // round; nodes without children keep the plain loop)
        const int hh = g.H / 2, hw = g.W / 2, nb = hh * hw;
        for (int i = tid; i < nb; i += nth) {
            const int y = i / hw, x = i - y * hw;
            cpx<T> gu[1][4];
            grad_u_alias<2, 2, 1>(a, nd, wsi, GY, g.W, g.HJ, g.WJ, &y, &x, gu);
            const int p = y * g.W + x;
            lds[p] = gu[0][0];
            lds[p + hw] = gu[0][1];
            lds[p + hh * g.W] = gu[0][2];
            lds[p + hh * g.W + hw] = gu[0][3];
        }
    } else {
        for (int i0 = tid; i0 < n; i0 += E * nth) {
            int ys[E], xs[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int i = i0 + e * nth < n ? i0 + e * nth : i0;
                ys[e] = i / g.W;
                xs[e] = i - ys[e] * g.W;
            }
            cpx<T> gu[E];
            grad_u_elems<E>(a, nd, wsi, GY, g.W, g.HJ, g.WJ, ys, xs, gu);
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (i0 + e * nth < n) lds[i0 + e * nth] = gu[e];
        }
    }
    wg_barrier();
    g.inv2d(lds, tid, nth);  // unnormalised: Re() is dL/dm (or dL/dpad(x) at the root)
    if (nd.order == 0) {
        T* gp = reinterpret_cast<T*>(wsi + nd.gpad);
        const int* rH = g.revH();
        const int* rW = g.revW();
        for (int p = tid; p < n; p += nth) {
            const int py = p / g.W, px = p - py * g.W;
            gp[rH[py] * g.W + rW[px]] = lds[p].re;
        }
        return;
    }
    const cpx<T>* sv = wsi + nd.save;
    copy_batched_pred<8>(
        n, [&](int p) { return kym_ldg(sv + p); },
        [&](int p, cpx<T> u) {
            const T mag = kym_sqrt(u.re * u.re + u.im * u.im);
            const T gm = lds[p].re;
            lds[p] = mag > T(0) ? cpx<T>{u.re * (gm / mag), u.im * (gm / mag)} : cpx<T>{T(0), T(0)};
        },
        tid, nth);
    wg_barrier();
    g.fwd2d(lds, tid, nth);
    const T inv_n = T(1) / T(n);
    cpx<T>* gs = wsi + nd.gsave;
    for (int i = tid; i < n; i += nth) gs[i] = lds[i] * inv_n;
}

template <class T, int MAXP>
KYM_HD void body_bwd_node(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int il = task / a.ngroup, gi = task - il * a.ngroup;
    const Node2D& nd = a.nodes[a.group[gi]];
    bwd_node_body(a, dyn_geom<T, MAXP>(a, nd.res), nd, il, lds, tid, nth);
}

// A node's backward with compile-time sizes (FP32): twiddles of both lengths
// in LDS first, the node's buffers after them.  task -> (image, node).
template <class T, int H, int W, int HJ, int WJ, int NT>
KYM_HD void body_bwd_node_s(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
    const int il = task / a.ngroup, gi = task - il * a.ngroup;
    const Node2D& nd = a.nodes[a.group[gi]];
    const int r = nd.res, J = a.J;
    constexpr int TWE = kym_tw_elems(H, W);
    const FftLen<T>* fH = &a.lens[a.lenH[r]];
    const FftLen<T>* fW = &a.lens[a.lenW[r]];
    load_twiddles(lds, fH, fW, H, W, tid, NT);
    const StatGeom<T, H, W, HJ, WJ, NT> g{fH, fW, &a.lens[a.lenH[J]], &a.lens[a.lenW[J]], lds + (H != W ? W : 0), lds};
    bwd_node_body(a, g, nd, il, lds + TWE, tid, NT);
}

// The backward body of node specialisation SPEC (its sizes and threads; the
// pair flag does not apply: the backward runs one node per workgroup).
template <class T, int SPEC, bool EMU>
KYM_HD KYM_INL void bwd_node_spec_body(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
#define KYM_BNS_BODY(sid, h, w, hj, wj, nt, pr) \
    if constexpr (SPEC == sid) body_bwd_node_s<T, h, w, hj, wj, (EMU ? 1 : nt)>(a, task, lds, tid);
    KYM_NODE_SPECS(KYM_BNS_BODY)
#undef KYM_BNS_BODY
}

// Run-time selection (the CPU harness).
template <class T, bool EMU>
KYM_HD void dispatch_bwd_node_spec(int id, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
    switch (id) {
#define KYM_BNS_CASE(sid, h, w, hj, wj, nt, pr) \
    case sid: bwd_node_spec_body<T, sid, EMU>(a, task, lds, tid); break;
        KYM_NODE_SPECS(KYM_BNS_CASE)
#undef KYM_BNS_CASE
        default: break;
    }
}

// Leaf gradient of a node whose leaf is not formed in its node kernel.
template <class T, int MAXP>
KYM_HD void body_bwd_leafg(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int il = task / a.ngroup, gi = task - il * a.ngroup;
    const Node2D& nd = a.nodes[a.group[gi]];
    const auto g = dyn_geom<T, MAXP>(a, nd.res);
    leaf_grad(a, g, nd, a.img0 + il, lds, tid, nth);
    cpx<T>* dst = a.ws + il * a.ws_img + nd.gslot;
    for (int i = tid; i < g.HJ * g.WJ; i += nth) dst[i] = lds[i];
}

// Big nodes: dL/dU by column blocks, inverse DIF along y -> U slot (rows digit-reversed).
template <class T, class P>
KYM_HD void body_bwd_gcols(const P& fp, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int per = a.ngroup * a.nblk;
    const int il = task / per, rem = task - il * per;
    const int gi = rem / a.nblk, cb = rem - gi * a.nblk;
    const Node2D& nd = a.nodes[a.group[gi]];
    const int r = nd.res, H = a.H[r], W = a.W[r], J = a.J;
    const int c0 = cb * a.blk, nc = (W - c0 < a.blk) ? W - c0 : a.blk;
    cpx<T>* wsi = a.ws + il * a.ws_img;
    constexpr int E = 2;  // elements (base points) per thread and round
    if (nd.child_end > nd.child_begin && grad_alias_ok(a, nd)) {
        // rows y and y + H/2 of a column share their child and leaf reads
        const int nb = (H / 2) * nc;
        for (int i0 = tid; i0 < nb; i0 += E * nth) {
            int ys[E], xs[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int i = i0 + e * nth < nb ? i0 + e * nth : i0;
                ys[e] = i / nc;
                xs[e] = c0 + i - ys[e] * nc;
            }
            cpx<T> g[E][2];
            grad_u_alias<2, 1, E>(a, nd, wsi, wsi + nd.gslot, W, a.H[J], a.W[J], ys, xs, g);
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (i0 + e * nth < nb) {
                    lds[i0 + e * nth] = g[e][0];
                    lds[i0 + e * nth + nb] = g[e][1];
                }
        }
    } else {
        for (int i0 = tid; i0 < H * nc; i0 += E * nth) {
            int ys[E], xs[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int i = i0 + e * nth < H * nc ? i0 + e * nth : i0;
                ys[e] = i / nc;
                xs[e] = c0 + i - ys[e] * nc;
            }
            cpx<T> g[E];
            grad_u_elems<E>(a, nd, wsi, wsi + nd.gslot, W, a.H[J], a.W[J], ys, xs, g);
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (i0 + e * nth < H * nc) lds[i0 + e * nth] = g[e];
        }
    }
    wg_barrier();
    fp.col_inv(lds, a.lens[a.lenH[r]], nc, tid, nth);
    cpx<T>* dst = wsi + nd.slot;
    for (int i = tid; i < H * nc; i += nth) {
        const int p = i / nc, c = i - (i / nc) * nc;
        dst[(long long)p * W + c0 + c] = lds[i];
    }
}

// Big nodes: rows of the U slot, inverse DIF along x, then modulus adjoint and
// forward DIT along x into the gsave slot (root: dL/dpad(x) into gpad).
template <class T, class P>
KYM_HD void body_bwd_grows(const P& fp, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int per = a.ngroup * a.nblk;
    const int il = task / per, rem = task - il * per;
    const int gi = rem / a.nblk, rb = rem - gi * a.nblk;
    const Node2D& nd = a.nodes[a.group[gi]];
    const int r = nd.res, H = a.H[r], W = a.W[r];
    const FftLen<T>& fW = a.lens[a.lenW[r]];
    const FftLen<T>& fH = a.lens[a.lenH[r]];
    const int r0 = rb * a.blk, nr = (H - r0 < a.blk) ? H - r0 : a.blk;
    cpx<T>* wsi = a.ws + il * a.ws_img;
    const cpx<T>* src = wsi + nd.slot + (long long)r0 * W;
    copy_batched_pred<8>(nr * W, [&](int i) { return kym_ldg(src + i); }, [&](int i, cpx<T> v) { lds[i] = v; }, tid, nth);
    wg_barrier();
    fp.row_inv(lds, fW, nr, tid, nth);
    if (nd.order == 0) {
        T* gp = reinterpret_cast<T*>(wsi + nd.gpad);
        for (int i = tid; i < nr * W; i += nth) {
            const int rr = i / W, px = i - rr * W;
            gp[(long long)fH.rev[r0 + rr] * W + fW.rev[px]] = lds[i].re;
        }
        return;
    }
    const cpx<T>* sv = wsi + nd.save + (long long)r0 * W;
    copy_batched_pred<8>(
        nr * W, [&](int i) { return kym_ldg(sv + i); },
        [&](int i, cpx<T> u) {
            const T mag = kym_sqrt(u.re * u.re + u.im * u.im);
            const T gm = lds[i].re;
            lds[i] = mag > T(0) ? cpx<T>{u.re * (gm / mag), u.im * (gm / mag)} : cpx<T>{T(0), T(0)};
        },
        tid, nth);
    wg_barrier();
    fp.row_fwd(lds, fW, nr, tid, nth);
    cpx<T>* gs = wsi + nd.gsave + (long long)r0 * W;
    for (int i = tid; i < nr * W; i += nth) gs[i] = lds[i];
}

// Big nodes: forward DIT along y of the gsave slot -> natural dL/dX (scaled 1/n).
template <class T, class P>
KYM_HD void body_bwd_gfcols(const P& fp, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    const int per = a.ngroup * a.nblk;
    const int il = task / per, rem = task - il * per;
    const int gi = rem / a.nblk, cb = rem - gi * a.nblk;
    const Node2D& nd = a.nodes[a.group[gi]];
    const int r = nd.res, H = a.H[r], W = a.W[r];
    const int c0 = cb * a.blk, nc = (W - c0 < a.blk) ? W - c0 : a.blk;
    cpx<T>* buf = a.ws + il * a.ws_img + nd.gsave + c0;
    copy_batched_pred<8>(
        H * nc,
        [&](int i) {
            const int y = i / nc, c = i - y * nc;
            return kym_ldg(buf + (long long)y * W + c);
        },
        [&](int i, cpx<T> v) { lds[i] = v; }, tid, nth);
    wg_barrier();
    fp.col_fwd(lds, a.lens[a.lenH[r]], nc, tid, nth);
    const T inv_n = T(1) / T(H * W);
    for (int i = tid; i < H * nc; i += nth) {
        const int y = i / nc, c = i - y * nc;
        buf[(long long)y * W + c] = lds[i] * inv_n;
    }
}

// Reflect-pad adjoint: dL/dx[i][j] = sum of dL/dpad over every padded
// position that reads input pixel (i, j).  task -> (image, row block).
KYM_HD KYM_INL int refl_preimages(int i, int n, int pad, int P, int* out) {
    // positions y in [0, P) with reflect_index(y - pad, n) == i
    int c = 0;
    if (n == 1) {
        for (int y = 0; y < P && c < 64; ++y) out[c++] = y;
        return c;
    }
    const int T_ = 2 * (n - 1);
    int y0 = (pad + i) % T_;
    for (int y = y0; y < P && c < 64; y += T_) out[c++] = y;
    if (i != 0 && i != n - 1) {
        int y1 = ((pad - i) % T_ + T_) % T_;
        for (int y = y1; y < P && c < 64; y += T_) out[c++] = y;
    }
    return c;
}

template <class T, int MAXP>
KYM_HD void body_bwd_fold(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    (void)lds;
    const int il = task / a.nblk, rb = task - il * a.nblk;
    const int img = a.img0 + il;
    const int P = a.H[0], Q = a.W[0];
    const T* gp = reinterpret_cast<const T*>(a.ws + il * a.ws_img + a.nodes[0].gpad);
    T* gx = a.gx + (long long)img * a.M * a.N;
    const int r0 = rb * a.blk, nr = (a.M - r0 < a.blk) ? a.M - r0 : a.blk;
    for (int e = tid; e < nr * a.N; e += nth) {
        const int i = r0 + e / a.N, j = e - (e / a.N) * a.N;
        if (a.pre_pad) {
            gx[(long long)i * a.N + j] = gp[(long long)i * Q + j];
            continue;
        }
        int ys[64], xs[64];
        const int ny = refl_preimages(i, a.M, a.top, P, ys);
        const int nx = refl_preimages(j, a.N, a.left, Q, xs);
        T acc = T(0);
        for (int u = 0; u < ny; ++u)
            for (int v = 0; v < nx; ++v) acc += gp[(long long)ys[u] * Q + xs[v]];
        gx[(long long)i * a.N + j] = acc;
    }
}

enum Kind2DBwd { K2B_NODE = 16, K2B_LEAFG, K2B_GCOLS, K2B_GROWS, K2B_GFCOLS, K2B_FOLD };

template <class T, int MAXP>
KYM_HD KYM_INL void dispatch_bwd2d(int kind, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid, int nth) {
    switch (kind) {
        case K2B_NODE: body_bwd_node<T, MAXP>(a, task, lds, tid, nth); break;
        case K2B_LEAFG: body_bwd_leafg<T, MAXP>(a, task, lds, tid, nth); break;
        case K2B_GCOLS: body_bwd_gcols<T>(DynPass<T, MAXP>(), a, task, lds, tid, nth); break;
        case K2B_GROWS: body_bwd_grows<T>(DynPass<T, MAXP>(), a, task, lds, tid, nth); break;
        case K2B_GFCOLS: body_bwd_gfcols<T>(DynPass<T, MAXP>(), a, task, lds, tid, nth); break;
        case K2B_FOLD: body_bwd_fold<T, MAXP>(a, task, lds, tid, nth); break;
        default: break;
    }
}

// Backward big-node passes with compile-time sizes, blocks and threads (the
// forward's pass specialisations), selected at compile time; EMU = true runs
// one thread (CPU harness).
template <class T, int KIND, int SPEC, bool EMU>
KYM_HD KYM_INL void bwd_pass_spec_body(const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
#define KYM_BPS_BODY(sid, h, w, c, r, nt)                                                            \
    if constexpr (SPEC == sid) {                                                                     \
        constexpr int NTE = EMU ? 1 : nt;                                                            \
        const FftLen<T>* fh = len_of(a, h);                                                          \
        const FftLen<T>* fw = len_of(a, w);                                                          \
        load_twiddles(lds, fh, fw, h, w, tid, NTE);                                                  \
        const StatPass<T, h, w, c, r, NTE> fp{lds + (h != w ? w : 0), lds};                          \
        cpx<T>* b = lds + kym_tw_elems(h, w);                                                        \
        if constexpr (KIND == K2B_GCOLS) body_bwd_gcols<T>(fp, a, task, b, tid, NTE);                \
        else if constexpr (KIND == K2B_GROWS) body_bwd_grows<T>(fp, a, task, b, tid, NTE);           \
        else if constexpr (KIND == K2B_GFCOLS) body_bwd_gfcols<T>(fp, a, task, b, tid, NTE);         \
    }
    KYM_PASS_SPECS(KYM_BPS_BODY)
#undef KYM_BPS_BODY
}

// Run-time selection (the CPU harness).
template <class T, int KIND, bool EMU>
KYM_HD void dispatch_bwd_pass_spec(int id, const Scat2DArgs<T>& a, int task, cpx<T>* lds, int tid) {
    switch (id) {
#define KYM_BPS_CASE(sid, h, w, c, r, nt) \
    case sid: bwd_pass_spec_body<T, KIND, sid, EMU>(a, task, lds, tid); break;
        KYM_PASS_SPECS(KYM_BPS_CASE)
#undef KYM_BPS_CASE
        default: break;
    }
}
