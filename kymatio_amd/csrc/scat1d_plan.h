// Scattering1D plan: node tree and launch list from the frontend's setup
// (kymatio/scattering1d/frontend/base_frontend.py:37-155 and the generator
// kymatio/scattering1d/core/scattering1d.py:2-107).
#pragma once
#include <algorithm>

#include "plan_common.h"
#include "scat1d_body.h"

struct Launch1D {
    size_t group_off;
    int kind;  // 0 forward node, 1 backward node, 2 backward pad fold
    int ngroup, threads, res;
    size_t lds_bytes;
    long long model_bytes;  // algorithmic bytes per signal (stage-boundary model)
    int spec = -1;          // compile-time node specialisation (FP32 forward), or -1
};

struct Scat1DDesc {
    int N, N_pad, pad_left, log2_stride, mode, oversampling, max_order;
    int n_psi1, n_psi2, phi_nlev;
    const int* psi1_j;
    const int* psi1_nlev;
    const int* psi2_j;
    const int* psi2_nlev;
    const int* ind_start;
    const int* ind_end;
    int n_ind;
};

template <class T>
struct Scat1DPlan {
    int N = 0, N_pad = 0, pad_left = 0, log2_stride = 0, mode = 0, max_order = 2, nres = 0;
    int n[KYM_MAX_RES1] = {0}, len[KYM_MAX_RES1] = {0};
    int ind_start[KYM_MAX_RES1] = {0}, ind_end[KYM_MAX_RES1] = {0};
    std::vector<long long> filt_off;
    std::vector<Node1D> nodes;
    std::vector<int> out_len;  // per output path, in output order
    long long out_row = 0, ws_sig = 0;
    int C = 0, pclass = 7;
    LenTable<T> lens;
    std::vector<Launch1D> launches, bwd_launches;
    long long ws_sig_train = 0;
    Blob blob;
    size_t off_nodes = 0, off_lens = 0, off_filt = 0, max_lds = 0;

    static int ilog2(int v) {
        int r = 0;
        while ((1 << r) < v) ++r;
        return r;
    }

    bool build(const Scat1DDesc& d, size_t lds_limit, KymError* e) {
        N = d.N; N_pad = d.N_pad; pad_left = d.pad_left; log2_stride = d.log2_stride; mode = d.mode;
        max_order = d.max_order;
        if (N < 1 || N_pad < N || (N_pad & (N_pad - 1))) { *e = {KYM_ERR_VALUE, "invalid padded length"}; return false; }
        if (max_order != 1 && max_order != 2) { *e = {KYM_ERR_VALUE, "max_order must be 1 or 2"}; return false; }
        const int lp = ilog2(N_pad);
        nres = lp + 1;
        if (nres > KYM_MAX_RES1 || d.n_ind > KYM_MAX_RES1) { *e = {KYM_ERR_UNSUPPORTED, "signal too long"}; return false; }
        for (int r = 0; r < nres; ++r) { n[r] = N_pad >> r; len[r] = lens.add(n[r]); }
        for (int i = 0; i < d.n_ind; ++i) { ind_start[i] = d.ind_start[i]; ind_end[i] = d.ind_end[i]; }
        pclass = 7;  // power-of-two lengths only
        // packed filters: phi levels, psi1 levels, psi2 levels (torch_frontend.py:30-56 order)
        filt_off.clear();
        long long off = 0;
        std::vector<int> phi_idx, psi1_base, psi2_base;
        for (int l = 0; l < d.phi_nlev; ++l) { phi_idx.push_back((int)filt_off.size()); filt_off.push_back(off); off += n[l]; }
        for (int i = 0; i < d.n_psi1; ++i) {
            psi1_base.push_back((int)filt_off.size());
            for (int l = 0; l < d.psi1_nlev[i]; ++l) { filt_off.push_back(off); off += n[l]; }
        }
        for (int i = 0; i < d.n_psi2; ++i) {
            psi2_base.push_back((int)filt_off.size());
            for (int l = 0; l < d.psi2_nlev[i]; ++l) { filt_off.push_back(off); off += n[l]; }
        }
        filt_off.push_back(off);
        const bool local = mode == AVG1_LOCAL;
        const int ls = log2_stride;
        const int Tlocal = ind_end[ls] - ind_start[ls];
        auto path_len = [&](int unpad_res) {
            if (mode == AVG1_GLOBAL) return 1;
            if (local) return Tlocal;
            return ind_end[unpad_res] - ind_start[unpad_res];
        };
        auto check_level = [&](int have, int want, const char* what) {
            if (want >= have) {
                *e = {KYM_ERR_VALUE, std::string("missing filter level for ") + what};
                return false;
            }
            return true;
        };
        nodes.clear(); out_len.clear();
        out_row = 0;
        auto add = [&](Node1D nd) {
            nd.out_off = (int)out_row;
            const int L = path_len(nd.unpad_res);
            out_len.push_back(L);
            out_row += L;
            nodes.push_back(nd);
        };
        Node1D root = {};
        root.order = 0; root.res = 0; root.parent = -1; root.filt = -1; root.k = 1;
        root.leaf_filt = phi_idx.empty() ? -1 : phi_idx[0];
        root.leaf_k = 1 << ls; root.unpad_res = 0; root.need_fwd = 1;
        if (local && !check_level(d.phi_nlev, 0, "phi")) return false;
        add(root);
        std::vector<int> k1s(d.n_psi1);
        for (int i = 0; i < d.n_psi1; ++i) {
            const int j1 = d.psi1_j[i];
            const int k1 = local ? std::min(j1, ls) : j1;
            k1s[i] = k1;
            Node1D nd = {};
            nd.order = 1; nd.res = k1; nd.parent = 0; nd.filt = psi1_base[i]; nd.k = 1 << k1;
            if (local) {
                if (!check_level(d.phi_nlev, k1, "phi")) return false;
                nd.leaf_filt = phi_idx[k1];
                nd.leaf_k = 1 << std::max(ls - k1, 0);
            }
            nd.unpad_res = std::max(j1 - d.oversampling, 0);
            nd.need_fwd = (local || mode == AVG1_GLOBAL || max_order >= 2) ? 1 : 0;
            add(nd);
        }
        if (max_order >= 2)
            for (int i = 0; i < d.n_psi1; ++i)
                for (int q = 0; q < d.n_psi2; ++q) {
                    const int j1 = d.psi1_j[i], j2 = d.psi2_j[q];
                    if (j2 <= j1) continue;
                    const int k1 = k1s[i];
                    const int sub2 = local ? std::min(j2, ls) : j2;
                    const int k2 = std::max(sub2 - k1, 0);
                    Node1D nd = {};
                    nd.order = 2; nd.res = k1 + k2; nd.parent = 1 + i;
                    if (!check_level(d.psi2_nlev[q], k1, "psi2")) return false;
                    nd.filt = psi2_base[q] + k1; nd.k = 1 << k2;
                    if (local) {
                        if (!check_level(d.phi_nlev, k1 + k2, "phi")) return false;
                        nd.leaf_filt = phi_idx[k1 + k2];
                        nd.leaf_k = 1 << std::max(ls - sub2, 0);
                    }
                    nd.unpad_res = std::max(j2 - d.oversampling, 0);
                    nd.need_fwd = (local || mode == AVG1_GLOBAL) ? 1 : 0;
                    add(nd);
                }
        C = (int)nodes.size();
        for (auto& nd : nodes)
            if (nd.res >= nres) { *e = {KYM_ERR_VALUE, "subsampling beyond the padded length"}; return false; }
        // slots for nodes with children
        std::vector<char> has_children(nodes.size(), 0);
        for (auto& nd : nodes)
            if (nd.parent >= 0) has_children[nd.parent] = 1;
        ws_sig = 0;
        const long long align = 256 / (long long)sizeof(cpx<T>);
        for (size_t i = 0; i < nodes.size(); ++i) {
            nodes[i].slot = has_children[i] ? ws_sig : -1;
            if (has_children[i]) ws_sig += n[nodes[i].res];
            ws_sig = (ws_sig + align - 1) / align * align;
        }
        if (ws_sig == 0) ws_sig = align;
        // training layout (extends the inference one): saved u per non-root
        // node, the root's dL/dpad(x); children as node-index ranges
        {
            for (auto& nd : nodes) { nd.child_begin = nd.child_end = 0; nd.save = nd.gsave = nd.gpad = -1; }
            for (size_t i = 1; i < nodes.size(); ++i) {
                Node1D& pn = nodes[nodes[i].parent];
                if (pn.child_end == 0) pn.child_begin = (int)i;
                pn.child_end = (int)i + 1;
            }
            long long w = ws_sig;
            for (auto& nd : nodes) {
                if (nd.order > 0) {
                    nd.save = w; w += n[nd.res];
                    w = (w + align - 1) / align * align;
                    nd.gsave = w; w += n[nd.res];
                }
                else { nd.gpad = w; w += ((long long)N_pad * (long long)sizeof(T) + sizeof(cpx<T>) - 1) / sizeof(cpx<T>); }
                w = (w + align - 1) / align * align;
            }
            ws_sig_train = w;
        }
        // LDS: node + leaf
        const size_t cs = sizeof(cpx<T>);
        for (auto& nd : nodes) {
            const size_t need = (size_t)(n[nd.res] + (local ? (n[nd.res] / std::max(nd.leaf_k, 1)) : 0)) * cs;
            if (need > lds_limit) {
                *e = {KYM_ERR_UNSUPPORTED, "padded signal length " + std::to_string(N_pad) +
                                               " does not fit one workgroup's LDS for this dtype"};
                return false;
            }
        }
        // blob
        blob = Blob();
        std::string err;
        if (!lens.emit(blob, &off_lens, &err)) { *e = {KYM_ERR_UNSUPPORTED, err}; return false; }
        off_nodes = blob.add(sizeof(Node1D) * nodes.size());
        std::memcpy(blob.at<char>(off_nodes), nodes.data(), sizeof(Node1D) * nodes.size());
        off_filt = blob.add(sizeof(long long) * filt_off.size());
        std::memcpy(blob.at<char>(off_filt), filt_off.data(), sizeof(long long) * filt_off.size());
        // launches: root, then order-1 nodes by resolution, then order-2 nodes by resolution
        launches.clear();
        max_lds = 0;
        const long long leaf_model = local ? 16ll * (N_pad >> ls) + 4ll * Tlocal : 4;
        auto push = [&](const std::vector<int>& ids, int kind) {
            if (ids.empty()) return;
            Launch1D l;
            l.kind = kind;
            l.group_off = blob.add(sizeof(int) * ids.size());
            std::memcpy(blob.at<char>(l.group_off), ids.data(), sizeof(int) * ids.size());
            l.ngroup = (int)ids.size();
            const Node1D& n0 = nodes[ids[0]];
            l.res = n0.res;
            const int len0 = n[n0.res];
            l.threads = std::max(64, std::min(1024, len0 / 16));
            l.lds_bytes = (size_t)(len0 + (local ? len0 / std::max(n0.leaf_k, 1) : 0)) * cs;
            // lds must cover every node of the group (same res => same length, leaf <= node)
            for (int id : ids) {
                const Node1D& nd = nodes[id];
                l.lds_bytes = std::max(l.lds_bytes, (size_t)(n[nd.res] + (local ? n[nd.res] / nd.leaf_k : 0)) * cs);
            }
            long long mb = 0;
            for (int id : ids) {
                const Node1D& nd = nodes[id];
                mb += (nd.order == 0 ? 4ll * N + 16ll * N_pad : 40ll * n[nd.res]) + leaf_model;
            }
            l.model_bytes = mb;
            if (kind == 0 && sizeof(T) == 4) l.spec = node1d_spec_id(len0, l.threads);
            launches.push_back(l);
            max_lds = std::max(max_lds, l.lds_bytes);
        };
        push({0}, 0);
        for (int order = 1; order <= max_order; ++order)
            for (int r = 0; r < nres; ++r) {
                std::vector<int> ids;
                for (size_t i = 0; i < nodes.size(); ++i)
                    if (nodes[i].order == order && nodes[i].res == r) ids.push_back((int)i);
                push(ids, 0);
            }
        std::vector<Launch1D> fwd = launches;
        launches.clear();
        for (int order = max_order; order >= 1; --order)
            for (int r = 0; r < nres; ++r) {
                std::vector<int> ids;
                for (size_t i = 0; i < nodes.size(); ++i)
                    if (nodes[i].order == order && nodes[i].res == r) ids.push_back((int)i);
                push(ids, 1);
            }
        push({0}, 1);
        push({0}, 2);
        launches.back().threads = 256;
        launches.back().lds_bytes = 0;
        bwd_launches = launches;
        launches = fwd;
        return true;
    }

    Scat1DArgs<T> args(const Launch1D& l, uintptr_t base) const {
        Scat1DArgs<T> a = {};
        a.nodes = reinterpret_cast<const Node1D*>(base + off_nodes);
        a.lens = reinterpret_cast<const FftLen<T>*>(base + off_lens);
        a.filt_off = reinterpret_cast<const long long*>(base + off_filt);
        a.group = reinterpret_cast<const int*>(base + l.group_off);
        a.ngroup = l.ngroup;
        for (int r = 0; r < KYM_MAX_RES1; ++r) {
            a.len[r] = len[r]; a.n[r] = n[r];
            a.ind_start[r] = ind_start[r]; a.ind_end[r] = ind_end[r];
        }
        a.mode = mode; a.log2_stride = log2_stride;
        a.N = N; a.pad_left = pad_left;
        a.ws_sig = ws_sig; a.out_row = out_row;
        return a;
    }
};
