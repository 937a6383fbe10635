// Scattering2D plan: node tree, packed-filter layout, workspace layout and
// the launch list, built once on the host (kymatio/scattering2d/frontend/
// base_frontend.py:19-36 for the padding, filter_bank.py:5-53 for the
// filter order, core/scattering2d.py:1-88 for the path order).
#pragma once
#include <algorithm>
#include <cmath>

#include "plan_common.h"
#include "scat2d_bwd.h"

struct Launch2D {
    int kind;
    size_t group_off;  // byte offset of the node-id list in the blob
    int ngroup, nblk, blk;
    int threads;
    size_t lds_bytes;
    int res;                // resolution of the launch's nodes
    long long model_bytes;  // algorithmic HBM bytes per image (stage-boundary model)
    int spec;               // compile-time specialisation id, -1 = run-time sizes
    int pair = 0;           // 1: a task is a sibling pair (group lists 2 ngroup node ids)
};

// Tuning choices of a plan (0 = the built-in heuristic).  Any non-default
// choice turns the compile-time specialisations off for what it affects.
struct Scat2DTune {
    int max_pow2 = 16;      // largest power-of-two FFT radix: 16, 8, 4 or 2
    int node_threads = 0;   // threads of LDS-node workgroups: 256, 512 or 1024
    int col_block = 0;      // columns per workgroup of the big-node column passes
    int row_block = 0;      // rows per workgroup of the big-node row passes
    int pass_threads = 0;   // threads of the big-node pass workgroups (compiled choices only)
    int streams = 0;        // chunks in flight (inference and training): 1 to 4 (0 = default 2)
    int pair = 0;           // LDS nodes: 1 one node per workgroup, 2 sibling pairs wherever a pair kernel is
                            // compiled (0 = heuristic: the default pair kernels, the others from k = 16);
                            // 0 and 2 also pair the big-node column passes with k = 2, 4 (1 does not)
};

template <class T>
struct Scat2DPlan {
    // configuration
    int M = 0, N = 0, J = 0, L = 0, max_order = 2, pre_pad = 0;
    int Mp = 0, Np = 0, top = 0, left = 0;
    int H[KYM_MAX_RES] = {0}, W[KYM_MAX_RES] = {0};
    int C = 0, oh = 0, ow = 0;
    size_t lds_limit = 0;
    // filters (reference buffer order: phi levels, then psi (j, theta) levels)
    std::vector<long long> filt_off;  // nfilt + 1 entries
    std::vector<int> filt_h, filt_w;
    // tree
    std::vector<Node2D> nodes;
    long long ws_img = 0;  // complex elements of workspace per image
    std::vector<Launch2D> launches;
    // training: forward with saving, then the backward launches
    long long ws_img_train = 0;  // complex elements per image (layout extends the inference one)
    std::vector<Launch2D> bwd_launches;
    bool trainable = true;
    std::string train_error;
    LenTable<T> lens;
    int lenH[KYM_MAX_RES] = {0}, lenW[KYM_MAX_RES] = {0};
    // constants blob (relocated to its device or host home)
    Blob blob;
    size_t off_nodes = 0, off_lens = 0, off_filt = 0;
    size_t max_lds = 0;
    int max_threads = 256;
    int pclass = 7;  // prime class of the kernels (largest odd radix family)

    int phi_index(int r) const { return r; }
    int psi_index(int n1, int level) const { return psi_base[n1] + level; }
    std::vector<int> psi_base, psi_j, psi_nlev;

    static int ilog2(int v) {
        int r = 0;
        while ((1 << r) < v) ++r;
        return r;
    }

    bool use_specs = true;  // select compile-time specialised kernels when sizes match (FP32 only)
    Scat2DTune tune;
    // leaf taps: 2 J axes ([2 r] along y, [2 r + 1] along x) in the packed
    // filter buffer after the filters, at taps_off, taps_len reals in all
    std::vector<FirAxis> fir_axes;
    long long taps_off = 0, taps_len = 0;
    size_t off_fir = 0;

    // Taps of the low-pass filter phi.levels[r] along an axis of n samples.
    // phi.levels[0] is a periodised Gaussian of sigma0 = 0.8 2^(J-1)
    // (kymatio/scattering2d/filter_bank.py: phi = gabor_2d(M, N, 0.8 * 2 ** (J - 1), 0, 0));
    // its taps beyond 7 sigma0 are below 1e-10 of the filter's mass.  Level
    // r >= 1 is its crop to the band |k| < n / 2, which cuts the Gaussian
    // spectrum at c = pi sigma0 / 2^r of its standard deviations and leaves
    // a ringing tail of relative size ~exp(-c^2 / 2): a 7 sigma band is kept
    // only when c >= 7, otherwise every sample of the circle is a tap.
    FirAxis fir_axis(int r, int n, long long off) const {
        const double sigma0 = 0.8 * std::ldexp(1.0, J - 1);
        const double c = r == 0 ? 1e30 : 3.14159265358979 * sigma0 / std::ldexp(1.0, r);
        FirAxis f = {off, -(n / 2), n, 0, 0, 0};
        if (c >= 7.0) {
            const int R = (int)std::ceil(7.0 * sigma0 / std::ldexp(1.0, r));
            if (2 * R + 1 < n) { f.lo = -R; f.cnt = 2 * R + 1; }
        }
        return f;
    }

    bool build(int M_, int N_, int J_, int L_, int max_order_, int pre_pad_, size_t lds_limit_, KymError* e) {
        M = M_; N = N_; J = J_; L = L_; max_order = max_order_; pre_pad = pre_pad_;
        lds_limit = lds_limit_;
        if (!(tune.max_pow2 == 16 || tune.max_pow2 == 8 || tune.max_pow2 == 4 || tune.max_pow2 == 2) ||
            !(tune.node_threads == 0 || tune.node_threads == 256 || tune.node_threads == 512 ||
              tune.node_threads == 1024) ||
            tune.col_block < 0 || tune.col_block > 64 || tune.row_block < 0 || tune.row_block > 64 ||
            !(tune.pass_threads == 0 || tune.pass_threads == 256 || (tune.pass_threads % 64 == 0 &&
                                                                      tune.pass_threads <= 1024)) ||
            !(tune.streams >= 0 && tune.streams <= 4) || !(tune.pair >= 0 && tune.pair <= 2)) {
            *e = {KYM_ERR_VALUE, "invalid tuning choice"};
            return false;
        }
        lens = LenTable<T>();
        lens.max_pow2 = tune.max_pow2;
        const bool specs = use_specs && tune.max_pow2 == 16;
        if (J < 0 || L < 1 || M < 1 || N < 1) { *e = {KYM_ERR_VALUE, "invalid J, L or shape"}; return false; }
        if (J + 1 > KYM_MAX_RES) { *e = {KYM_ERR_UNSUPPORTED, "J too large"}; return false; }
        if ((1 << J) > M || (1 << J) > N) {
            *e = {KYM_ERR_RUNTIME, "The smallest dimension should be larger than 2^J."};
            return false;
        }
        if (max_order != 1 && max_order != 2) { *e = {KYM_ERR_VALUE, "max_order must be 1 or 2"}; return false; }
        // kymatio/scattering2d/utils.py:4-24 and base_frontend.py:19-32
        Mp = ((M + (1 << J)) / (1 << J) + 1) * (1 << J);
        Np = ((N + (1 << J)) / (1 << J) + 1) * (1 << J);
        top = (Mp - M) / 2;
        left = (Np - N) / 2;
        for (int r = 0; r <= J; ++r) {
            H[r] = Mp >> r;
            W[r] = Np >> r;
            lenH[r] = lens.add(H[r]);
            lenW[r] = lens.add(W[r]);
        }
        oh = H[J] - 2;
        ow = W[J] - 2;
        {
            int mp = 1;
            for (int r = 0; r <= J; ++r) mp = std::max(mp, std::max(kym_max_odd_prime(H[r]), kym_max_odd_prime(W[r])));
            pclass = kym_prime_class(mp);
        }
        // filters
        filt_off.clear(); filt_h.clear(); filt_w.clear();
        long long off = 0;
        for (int r = 0; r < J; ++r) {
            filt_off.push_back(off); filt_h.push_back(H[r]); filt_w.push_back(W[r]);
            off += (long long)H[r] * W[r];
        }
        psi_base.clear(); psi_j.clear(); psi_nlev.clear();
        for (int j = 0; j < J; ++j)
            for (int th = 0; th < L; ++th) {
                const int nlev = std::min(j + 1, std::max(J - 1, 1));
                psi_base.push_back((int)filt_off.size());
                psi_j.push_back(j);
                psi_nlev.push_back(nlev);
                for (int lv = 0; lv < nlev; ++lv) {
                    filt_off.push_back(off); filt_h.push_back(H[lv]); filt_w.push_back(W[lv]);
                    off += (long long)H[lv] * W[lv];
                }
            }
        filt_off.push_back(off);
        // leaf taps (FP32 plans)
        const bool fir = kym_fir<T>;
        fir_axes.assign(2 * (size_t)std::max(J, 1), FirAxis{0, 0, 0, 0, 0, 0});
        taps_off = off;
        taps_len = 0;
        if (fir)
            for (int r = 0; r < J; ++r)
                for (int ax = 0; ax < 2; ++ax) {
                    const int n = ax == 0 ? H[r] : W[r];
                    FirAxis& f = fir_axes[2 * r + ax];
                    f = fir_axis(r, n, taps_off + taps_len);
                    taps_len += f.cnt;
                    f.kmax = fir_kmax(n, 1 << (J - r), f.cnt);
                    f.band_off = taps_off + taps_len;
                    taps_len += 16ll * f.kmax;
                }
        // nodes (path order of kymatio/scattering2d/core/scattering2d.py:25-78)
        nodes.clear();
        const int npsi = J * L;
        Node2D root = {};
        root.order = 0; root.res = 0; root.parent = -1; root.filt = -1; root.k = 1;
        root.leaf_ch = 0; root.leaf_filt = phi_index(0); root.leaf_k = 1 << J;
        nodes.push_back(root);
        for (int n1 = 0; n1 < npsi; ++n1) {
            Node2D nd = {};
            const int j1 = psi_j[n1];
            nd.order = 1; nd.res = j1; nd.parent = 0; nd.filt = psi_index(n1, 0); nd.k = 1 << j1;
            nd.leaf_ch = 1 + n1; nd.leaf_filt = phi_index(j1); nd.leaf_k = 1 << (J - j1);
            nodes.push_back(nd);
        }
        int ch = 1 + npsi;
        if (max_order >= 2)
            for (int n1 = 0; n1 < npsi; ++n1)
                for (int n2 = 0; n2 < npsi; ++n2) {
                    const int j1 = psi_j[n1], j2 = psi_j[n2];
                    if (j2 <= j1) continue;
                    Node2D nd = {};
                    nd.order = 2; nd.res = j2; nd.parent = 1 + n1; nd.filt = psi_index(n2, j1);
                    nd.k = 1 << (j2 - j1);
                    nd.leaf_ch = ch++; nd.leaf_filt = phi_index(j2); nd.leaf_k = 1 << (J - j2);
                    nodes.push_back(nd);
                }
        C = ch;
        // placement: LDS-resident or column/row passes; leaf in-kernel or not
        const size_t cs = sizeof(cpx<T>);
        const size_t leaf_elems = (size_t)H[J] * W[J];
        std::vector<char> has_children(nodes.size(), 0);
        for (auto& nd : nodes)
            if (nd.parent >= 0) has_children[nd.parent] = 1;
        auto node_threads = [&](size_t elems) {
            if (tune.node_threads) return tune.node_threads;
            return elems >= 8192 ? 1024 : elems >= 2048 ? 512 : 256;
        };
        // LDS bytes of a node run in one workgroup (leaf not counted for the spectral route)
        auto node_lds = [&](const Node2D& nd) {
            const int r = nd.res;
            size_t b = (size_t)H[r] * W[r] * cs;
            if (fir) {
                const long long reals = fir_lds_reals(H[r], W[r], ow, nd.leaf_k, fir_axes[2 * r].cnt,
                                                      fir_axes[2 * r + 1].cnt);
                b = std::max(b, (size_t)reals * sizeof(T));
            }
            return b;
        };
        ws_img = 0;
        for (size_t i = 0; i < nodes.size(); ++i) {
            Node2D& nd = nodes[i];
            const size_t elems = (size_t)H[nd.res] * W[nd.res];
            // LDS-resident if the node (and, FP32, its in-LDS leaf) fits; leaf_mode
            // places the spectral leaf and, in training, the leaf's gradient
            // (run-time-sized kernels stage at most KYM_STAGE_MAX values per thread;
            // a compile-time kernel for these sizes and threads has no such limit)
            int spec_threads = 0;
            const bool spec_fit =
                specs && sizeof(T) == 4 &&
                node_spec_id(H[nd.res], W[nd.res], H[J], W[J], node_threads(elems), &spec_threads, 0) >= 0 &&
                node_lds(nd) + cs * kym_tw_elems(H[nd.res], W[nd.res]) <= lds_limit;
            const bool fits = node_lds(nd) <= lds_limit &&
                              (!fir || elems <= (size_t)KYM_STAGE_MAX * node_threads(elems) || spec_fit);
            if (fits) { nd.big = 0; nd.leaf_mode = ((elems + leaf_elems) * cs <= lds_limit) ? 0 : 1; }
            else { nd.big = 1; nd.leaf_mode = (leaf_elems * cs <= lds_limit) ? 1 : 2; }
            nd.fir = fir ? (nd.big ? 2 : 1) : 0;
            nd.fwd = (has_children[i] || !fir) ? 1 : 0;
            const bool need_slot = has_children[i] || nd.big || (!fir && nd.leaf_mode != 0);
            nd.slot = need_slot ? ws_img : -1;
            if (need_slot) ws_img += (long long)elems;
            nd.leaf_slot = -1;
            if (!fir && nd.leaf_mode == 2) { nd.leaf_slot = ws_img; ws_img += (long long)leaf_elems; }
            nd.tslot = -1;
            if (nd.fir == 2) {
                nd.tslot = ws_img;
                ws_img += ((long long)H[nd.res] * ow * (long long)sizeof(T) + cs - 1) / cs;
            }
            // keep slots 256-byte aligned
            const long long align = 256 / (long long)cs;
            ws_img = (ws_img + align - 1) / align * align;
        }
        // training layout: save slot (u) and gsave slot (dL/dX) for every non-root node,
        // leaf dL/dY slot where the leaf is not formed in the node kernel, and
        // the root's dL/dpad(x); children as node-index ranges.
        {
            for (auto& nd : nodes) { nd.child_begin = nd.child_end = 0; nd.save = nd.gsave = nd.gslot = nd.gpad = -1; }
            for (size_t i = 1; i < nodes.size(); ++i) {
                Node2D& pn = nodes[nodes[i].parent];
                if (pn.child_end == 0) pn.child_begin = (int)i;
                pn.child_end = (int)i + 1;
            }
            const long long align = 256 / (long long)cs;
            long long w = ws_img;
            for (auto& nd : nodes) {
                if (nd.order > 0) {
                    nd.save = w; w += (long long)H[nd.res] * W[nd.res];
                    w = (w + align - 1) / align * align;
                    nd.gsave = w; w += (long long)H[nd.res] * W[nd.res];
                }
                if (nd.leaf_mode != 0) { nd.gslot = w; w += (long long)leaf_elems; }
                if (nd.order == 0) { nd.gpad = w; w += ((long long)H[0] * W[0] * (long long)sizeof(T) + cs - 1) / cs; }
                w = (w + align - 1) / align * align;
                if (nd.leaf_mode == 2) { trainable = false; train_error = "leaf does not fit LDS: backward unsupported"; }
            }
            ws_img_train = w;
        }
        // blob: nodes, lens, filter offsets, group lists
        blob = Blob();
        std::string err;
        if (!lens.emit(blob, &off_lens, &err)) { *e = {KYM_ERR_UNSUPPORTED, err}; return false; }
        off_nodes = blob.add(sizeof(Node2D) * nodes.size());
        std::memcpy(blob.at<char>(off_nodes), nodes.data(), sizeof(Node2D) * nodes.size());
        off_filt = blob.add(sizeof(long long) * filt_off.size());
        std::memcpy(blob.at<char>(off_filt), filt_off.data(), sizeof(long long) * filt_off.size());
        off_fir = blob.add(sizeof(FirAxis) * fir_axes.size());
        std::memcpy(blob.at<char>(off_fir), fir_axes.data(), sizeof(FirAxis) * fir_axes.size());
        // launches
        launches.clear();
        max_lds = 0;
        max_threads = 256;
        auto add_group = [&](const std::vector<int>& ids) {
            size_t o = blob.add(sizeof(int) * ids.size());
            std::memcpy(blob.at<char>(o), ids.data(), sizeof(int) * ids.size());
            return o;
        };
        // Stage-boundary byte model per node (per image): every tensor that
        // crosses a stage boundary is written once and read once; complex 8 B,
        // real 4 B.  Root: 4 M N input + 16 P_M P_N (U_0 spectrum); other
        // nodes: 40 H_r W_r (periodised input, modulus, re-FFT'd spectrum,
        // each written and read); every leaf: 16 H_J W_J + 4 (H_J-2)(W_J-2).
        const long long leaf_model = 16ll * H[J] * W[J] + 4ll * (H[J] - 2) * (W[J] - 2);
        auto node_model = [&](int id) {
            const Node2D& nd = nodes[id];
            if (nd.order == 0) return 4ll * M * N + 16ll * H[0] * W[0];
            return 40ll * H[nd.res] * W[nd.res];
        };
        auto push = [&](int kind, const std::vector<int>& ids, int nblk, int blk, int threads, size_t lds) {
            Launch2D l;
            l.kind = kind; l.group_off = add_group(ids); l.ngroup = (int)ids.size();
            l.nblk = nblk; l.blk = blk; l.threads = threads; l.lds_bytes = lds;
            l.res = nodes[ids[0]].res;
            long long mb = 0;
            for (int id : ids) {
                const Node2D& nd = nodes[id];
                const bool leaf_here = (kind == K2_LDS_NODE && (nd.fir == 1 || nd.leaf_mode == 0)) ||
                                       kind == K2_LEAF || kind == K2_LEAF_ROWS || kind == K2_FIR_COLS;
                const int npass = nd.order == 0 ? 2 : nd.fwd ? 3 : 2;  // big-node passes sharing the node's bytes
                if (kind == K2_LDS_NODE) mb += node_model(id);
                if (kind == K2_ROOT_ROWS || kind == K2_FWD_COLS || kind == K2_ICOL || kind == K2_ICOL_PAIR ||
                    kind == K2_ROWS_MOD)
                    mb += node_model(id) / npass;
                if (leaf_here) mb += leaf_model;
            }
            l.model_bytes = mb;
            l.spec = -1;
            launches.push_back(l);
            max_lds = std::max(max_lds, lds);
            max_threads = std::max(max_threads, threads);
        };
        const size_t pass_budget = std::min(lds_limit, (size_t)64 * 1024);
        bool no_pass_spec = false;  // a tuned pass thread count with no compiled kernel
        std::vector<int> pending_fir;  // big nodes whose leaf column pass is still to be launched
        // pass blocks: the heuristic keeps a block within 64 KB of LDS; a tuned
        // block may use up to the LDS limit
        auto col_blk = [&](int Hh) {
            if (tune.col_block) return std::max(1, std::min(tune.col_block, (int)(lds_limit / (cs * Hh))));
            return std::max(1, std::min(16, (int)(pass_budget / (cs * Hh))));
        };
        // row passes of FIR nodes also keep the block's rows as reals (stride W + 1)
        // and the leaf's band table along x
        auto row_bytes = [&](int Ww) { return cs * Ww + (fir ? sizeof(T) * (Ww + 1) : 0); };
        auto band_bytes = [&](int r) {
            const int km = fir_axes[2 * r + 1].kmax;
            return fir && fir_band_in_lds(km, (int)sizeof(T)) ? sizeof(T) * 16 * (size_t)km : 0;
        };
        // FIR row passes hold 8 rows at most, so that four of their workgroups share a CU
        auto row_blk = [&](int r) {
            if (tune.row_block) {
                const size_t room = lds_limit > band_bytes(r) ? lds_limit - band_bytes(r) : 0;
                return std::max(1, std::min(tune.row_block, (int)(room / row_bytes(W[r]))));
            }
            int rb = (int)((pass_budget - std::min(pass_budget / 2, band_bytes(r))) / row_bytes(W[r]));
            return std::max(1, std::min(fir ? 8 : 16, rb));
        };
        auto row_lds = [&](int r, int rb) {
            const bool glob = fir && !fir_band_in_lds(fir_axes[2 * r + 1].kmax, (int)sizeof(T));
            return (size_t)rb * row_bytes(W[r]) + band_bytes(r) +
                   (glob && fir_rows_split(rb, ow) ? KYM_FIR_SPLIT_LDS : 0);
        };
        // column pass of a big node's leaf: its row-filtered field plus the band table along y
        auto fir_cols_lds = [&](int r) {
            return sizeof(T) * ((size_t)H[r] * fir_tstride(W[r], ow) +
                                16 * (size_t)fir_axes[2 * r].kmax);
        };
        auto emit_leaves = [&](const std::vector<int>& ids) {
            std::vector<int> small, big, firc;
            for (int id : ids) {
                if (nodes[id].fir == 2) firc.push_back(id);
                else if (nodes[id].fir == 0 && nodes[id].leaf_mode == 1) small.push_back(id);
                else if (nodes[id].fir == 0 && nodes[id].leaf_mode == 2) big.push_back(id);
            }
            for (int id : firc) pending_fir.push_back(id);  // flushed per resolution (one launch)
            if (!small.empty()) push(K2_LEAF, small, 1, 0, node_threads(leaf_elems), leaf_elems * cs);
            if (!big.empty()) {
                const int cb = col_blk(H[J]), rb = std::max(1, std::min(16, (int)(pass_budget / (cs * W[J]))));
                push(K2_LEAF_ICOL, big, (W[J] + cb - 1) / cb, cb, 256, (size_t)H[J] * cb * cs);
                push(K2_LEAF_ROWS, big, (H[J] + rb - 1) / rb, rb, 256, (size_t)rb * W[J] * cs);
            }
        };
        auto emit_nodes = [&](const std::vector<int>& ids, bool is_root) {
            if (ids.empty()) return;
            const Node2D& n0 = nodes[ids[0]];
            const int r = n0.res;
            const size_t elems = (size_t)H[r] * W[r];
            if (!n0.big) {
                const size_t lds = fir ? node_lds(n0) : (elems + (n0.leaf_mode == 0 ? leaf_elems : 0)) * cs;
                // sibling pairs: the nodes sorted by parent, every parent with an even
                // number of children here, periodisation k <= 32, a compiled pair kernel
                if (tune.pair != 1 && specs && sizeof(T) == 4 && n0.fir == 1) {
                    std::vector<int> sib = ids;
                    std::stable_sort(sib.begin(), sib.end(),
                                     [&](int u, int v) { return nodes[u].parent < nodes[v].parent; });
                    bool ok = sib.size() % 2 == 0;
                    for (size_t i = 0; ok && i < sib.size(); i += 2) {
                        const Node2D& u = nodes[sib[i]];
                        const Node2D& v = nodes[sib[i + 1]];
                        ok = u.order >= 1 && u.parent == v.parent && u.k == v.k && u.k <= 32 && v.fir == 1;
                    }
                    int pnt = 0, kmax = 1;
                    for (int id : sib) kmax = std::max(kmax, nodes[id].k);
                    // the other pair kernels by default where the deepest periodisation
                    // of the launch is 16 or more (measured: they gain at 256² J=5 and
                    // 512² J=6, and lose at 128² J=4, k <= 8)
                    int psid = ok ? node_spec_id(H[r], W[r], H[J], W[J], tune.node_threads, &pnt, 1) : -1;
                    if (ok && psid < 0 && (tune.pair == 2 || kmax >= 16))
                        psid = node_spec_id(H[r], W[r], H[J], W[J], tune.node_threads, &pnt, 2);
                    // twiddles, the node region, the second node's stash
                    const size_t need = cs * kym_tw_elems(H[r], W[r]) + (lds + cs - 1) / cs * cs + elems * cs;
                    if (psid >= 0 && need <= lds_limit) {
                        push(K2_LDS_NODE, sib, 1, 0, pnt, need);
                        Launch2D& lb = launches.back();
                        lb.ngroup /= 2;
                        lb.pair = 1;
                        lb.spec = psid;
                        emit_leaves(ids);
                        return;
                    }
                }
                push(K2_LDS_NODE, ids, 1, 0, node_threads(elems), lds);
                int nt = 0;
                const int sid = node_spec_id(H[r], W[r], H[J], W[J], tune.node_threads, &nt, 0);
                if (specs && sizeof(T) == 4 && sid >= 0 && (n0.fir == 1 || n0.leaf_mode == 0) &&
                    lds + cs * kym_tw_elems(H[r], W[r]) <= lds_limit) {
                    launches.back().spec = sid;
                    launches.back().threads = nt;
                    launches.back().lds_bytes = lds + cs * kym_tw_elems(H[r], W[r]);  // + twiddles
                    max_lds = std::max(max_lds, launches.back().lds_bytes);
                }
            } else {
                int cb = col_blk(H[r]);
                const int rb = row_blk(r);
                // an untuned column block is the first compiled one for these sizes
                // and rows (the spec table lists each size's forward choice first:
                // 16 columns at 272, 8 at 160, 320 and 640)
                if (!tune.col_block && specs && sizeof(T) == 4 && H[r] % rb == 0) {
                    const int c0 = pass_spec_first_cols(H[r], W[r], rb);
                    if (c0 > 0 && W[r] % c0 == 0) cb = c0;
                }
                // compile-time passes when the sizes match and every block is full
                int pnt = 0;
                const int want = tune.pass_threads == 256 ? 0 : tune.pass_threads;
                const int psid = (specs && sizeof(T) == 4 && H[r] % rb == 0 && W[r] % cb == 0)
                                     ? pass_spec_id(H[r], W[r], cb, rb, want, &pnt) : -1;
                if (tune.pass_threads && tune.pass_threads != 256 && psid < 0) no_pass_spec = true;
                auto pass = [&](int kind, int nblk, int blk, size_t lds) {
                    push(kind, ids, nblk, blk, 256, lds);
                    const size_t tw = cs * kym_tw_elems(H[r], W[r]);
                    if (psid >= 0 && lds + tw <= lds_limit) {
                        launches.back().spec = psid;
                        launches.back().threads = pnt;
                        launches.back().lds_bytes = lds + tw;  // + twiddles
                        max_lds = std::max(max_lds, launches.back().lds_bytes);
                    }
                };
                if (is_root) pass(K2_ROOT_ROWS, H[r] / rb + (H[r] % rb != 0), rb, row_lds(r, rb));
                else {
                    // sibling pairs in the gather + inverse column pass (one read of
                    // the parent's columns for two filters): compiled passes, k = 2 or
                    // 4 (the mixed batch's 320 and 160 nodes under 640 and 320
                    // parents: 7.34k -> 7.73k images/s; with k = 1, the root's
                    // children at 272, the headline lost 0.9%), the two blocks within
                    // 64 KB; pair = 1 keeps one node per task
                    const size_t two = 2 * (size_t)H[r] * cb * cs, tw = cs * kym_tw_elems(H[r], W[r]);
                    bool paired = false;
                    if (tune.pair != 1 && psid >= 0 && two + tw <= std::min(lds_limit, (size_t)64 * 1024)) {
                        std::vector<int> sib = ids;
                        std::stable_sort(sib.begin(), sib.end(),
                                         [&](int u, int v) { return nodes[u].parent < nodes[v].parent; });
                        bool ok = sib.size() % 2 == 0;
                        for (size_t i = 0; ok && i < sib.size(); i += 2) {
                            const Node2D& u = nodes[sib[i]];
                            const Node2D& v = nodes[sib[i + 1]];
                            ok = u.parent == v.parent && u.k == v.k && (u.k == 2 || u.k == 4);
                        }
                        if (ok) {
                            push(K2_ICOL_PAIR, sib, W[r] / cb, cb, 256, two);
                            Launch2D& lb = launches.back();
                            lb.ngroup /= 2;
                            lb.pair = 1;
                            lb.spec = psid;
                            lb.threads = pnt;
                            lb.lds_bytes = two + tw;
                            max_lds = std::max(max_lds, lb.lds_bytes);
                            paired = true;
                        }
                    }
                    if (!paired) pass(K2_ICOL, W[r] / cb + (W[r] % cb != 0), cb, (size_t)H[r] * cb * cs);
                    pass(K2_ROWS_MOD, H[r] / rb + (H[r] % rb != 0), rb, row_lds(r, rb));
                }
                // the row passes keep columns 0 .. W / 2 of the spectra of real
                // fields; the forward column pass covers them and writes the rest
                // by Hermitian symmetry (body_fwd_cols)
                if (n0.fwd) pass(K2_FWD_COLS, (W[r] / 2 + 1 + cb - 1) / cb, cb, (size_t)H[r] * cb * cs);
            }
            emit_leaves(ids);
        };
        // By resolution: the big nodes of resolution r (order 1, then order 2),
        // then ONE launch for every LDS node of resolution r, first and second
        // order together.  A node's parent always has a lower resolution, so it
        // was written by an earlier launch; merging the orders fills the GPU
        // with one launch instead of a small one plus a large one.
        emit_nodes({0}, true);
        for (int r = 0; r < J; ++r) {
            for (int order = 1; order <= max_order; ++order) {
                std::vector<int> ids;
                for (size_t i = 0; i < nodes.size(); ++i)
                    if (nodes[i].order == order && nodes[i].res == r && nodes[i].big) ids.push_back((int)i);
                emit_nodes(ids, false);
            }
            std::vector<int> ids;
            for (size_t i = 0; i < nodes.size(); ++i)
                if (nodes[i].order >= 1 && nodes[i].res == r && !nodes[i].big) ids.push_back((int)i);
            emit_nodes(ids, false);
            // leaf column passes of this resolution's big nodes (the root's included): one launch
            if (!pending_fir.empty()) {
                push(K2_FIR_COLS, pending_fir, 1, 0, 256, fir_cols_lds(r));
                pending_fir.clear();
            }
        }
        if (!pending_fir.empty()) {  // J = 0: the root alone
            push(K2_FIR_COLS, pending_fir, 1, 0, 256, fir_cols_lds(0));
            pending_fir.clear();
        }
        if (no_pass_spec) {
            *e = {KYM_ERR_VALUE, "invalid tuning choice: no compiled pass for these threads and blocks"};
            return false;
        }
        // backward: order 2, order 1, root; then the pad fold
        std::vector<Launch2D> fwd = launches;
        launches.clear();
        auto emit_bwd = [&](const std::vector<int>& ids) {
            if (ids.empty()) return;
            const Node2D& n0 = nodes[ids[0]];
            const int r = n0.res;
            const size_t elems = (size_t)H[r] * W[r];
            std::vector<int> leafg;
            for (int id : ids)
                if (nodes[id].leaf_mode != 0) leafg.push_back(id);
            if (!leafg.empty()) push(K2B_LEAFG, leafg, 1, 0, node_threads(leaf_elems), leaf_elems * cs);
            if (!n0.big) {
                const size_t lds = (elems + (n0.leaf_mode == 0 ? leaf_elems : 0)) * cs;
                push(K2B_NODE, ids, 1, 0, node_threads(elems), lds);
                // compile-time sizes as in the forward (one node per workgroup)
                int nt = 0;
                const int sid = node_spec_id(H[r], W[r], H[J], W[J], tune.node_threads, &nt, 0);
                const size_t tw = cs * kym_tw_elems(H[r], W[r]);
                if (specs && sizeof(T) == 4 && sid >= 0 && lds + tw <= lds_limit) {
                    launches.back().spec = sid;
                    launches.back().threads = nt;
                    launches.back().lds_bytes = lds + tw;
                    max_lds = std::max(max_lds, launches.back().lds_bytes);
                }
            } else {
                int cb = col_blk(H[r]);
                int rb = std::max(1, std::min(tune.row_block ? tune.row_block : 16,
                                              (int)((tune.row_block ? lds_limit : pass_budget) / (cs * W[r]))));
                // compile-time passes for these sizes and blocks; untuned blocks
                // take 8 x 8 where only that shape is compiled
                int pnt = 0, psid = -1;
                if (specs && sizeof(T) == 4) {
                    const int want = tune.pass_threads == 256 ? 0 : tune.pass_threads;
                    auto spec_for = [&](int c, int rr) {
                        return (H[r] % rr == 0 && W[r] % c == 0) ? pass_spec_id(H[r], W[r], c, rr, want, &pnt) : -1;
                    };
                    // untuned root: 8 x 8 blocks (one node per image: twice the
                    // workgroups of 16 columns)
                    if (n0.order == 0 && !tune.col_block && !tune.row_block && (psid = spec_for(8, 8)) >= 0)
                        cb = rb = 8;
                    if (psid < 0) psid = spec_for(cb, rb);
                    // untuned: the heuristic's columns by 8 rows (160: 16 x 8, 5.7% faster
                    // than 8 x 8 on forward + backward at 128x128 J=4), else 8 x 8
                    if (psid < 0 && !tune.col_block && !tune.row_block && (psid = spec_for(cb, 8)) >= 0) rb = 8;
                    if (psid < 0 && !tune.col_block && !tune.row_block && (psid = spec_for(8, 8)) >= 0) cb = rb = 8;
                }
                auto bpass = [&](int kind, int nblk, int blk, size_t lds) {
                    push(kind, ids, nblk, blk, 256, lds);
                    const size_t tw = cs * kym_tw_elems(H[r], W[r]);
                    if (psid >= 0 && lds + tw <= lds_limit) {
                        launches.back().spec = psid;
                        launches.back().threads = pnt;
                        launches.back().lds_bytes = lds + tw;
                        max_lds = std::max(max_lds, launches.back().lds_bytes);
                    }
                };
                bpass(K2B_GCOLS, (W[r] + cb - 1) / cb, cb, (size_t)H[r] * cb * cs);
                bpass(K2B_GROWS, (H[r] + rb - 1) / rb, rb, (size_t)rb * W[r] * cs);
                if (n0.order > 0) bpass(K2B_GFCOLS, (W[r] + cb - 1) / cb, cb, (size_t)H[r] * cb * cs);
            }
        };
        for (int order = max_order; order >= 1; --order)
            for (int r = 0; r < J; ++r)
                for (int big = 0; big <= 1; ++big) {
                    std::vector<int> ids;
                    for (size_t i = 0; i < nodes.size(); ++i)
                        if (nodes[i].order == order && nodes[i].res == r && nodes[i].big == big) ids.push_back((int)i);
                    emit_bwd(ids);
                }
        emit_bwd({0});
        {
            const int rb = 8;
            push(K2B_FOLD, {0}, ((pre_pad ? Mp : M) + rb - 1) / rb, rb, 256, 0);
        }
        bwd_launches = launches;
        launches = fwd;
        return true;
    }

    // Kernel arguments for one launch; `base` is the relocated blob address.
    Scat2DArgs<T> args(const Launch2D& l, uintptr_t base) const {
        Scat2DArgs<T> a = {};
        a.nodes = reinterpret_cast<const Node2D*>(base + off_nodes);
        a.lens = reinterpret_cast<const FftLen<T>*>(base + off_lens);
        a.filt_off = reinterpret_cast<const long long*>(base + off_filt);
        a.fir = reinterpret_cast<const FirAxis*>(base + off_fir);
        a.group = reinterpret_cast<const int*>(base + l.group_off);
        a.ngroup = l.ngroup; a.nblk = l.nblk; a.blk = l.blk;
        for (int r = 0; r <= J; ++r) { a.lenH[r] = lenH[r]; a.lenW[r] = lenW[r]; a.H[r] = H[r]; a.W[r] = W[r]; }
        a.J = J;
        a.M = pre_pad ? Mp : M; a.N = pre_pad ? Np : N;
        a.top = top; a.left = left; a.pre_pad = pre_pad;
        a.ws_img = ws_img;
        a.C = C; a.oh = oh; a.ow = ow;
        return a;
    }
};
