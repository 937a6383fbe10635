// CPU test harness (libkymatio_emu.so, built host-only): runs the GPU
// workgroup bodies on the host with one "thread" per workgroup, launch by
// launch, so that the plan, the index arithmetic, the digit-reversal maps
// and the twiddle tables can be checked against the oracle on a machine
// without a GPU.  It is loaded only by the tests; the product library
// (libkymatio_hip.so) contains no CPU compute path.
#include <cstdlib>
#include <string>
#include <vector>

#include "ops_body.h"
#include "scat1d_plan.h"
#include "scat2d_plan.h"

static thread_local std::string g_emu_error;
void kym_set_error(const std::string& msg) { g_emu_error = msg; }

extern "C" const char* kym_emu_last_error(void) { return g_emu_error.c_str(); }

// A launch with a compile-time specialisation, one thread per workgroup.
template <class T>
static void emu_spec(const Launch2D& l, const Scat2DArgs<T>& a, int t, cpx<T>* lds) {
    switch (l.kind) {
        case K2_LDS_NODE: dispatch_node_spec<T, true>(l.spec, a, t, lds, 0); break;
        case K2_ROOT_ROWS: dispatch_pass_spec<T, K2_ROOT_ROWS, true>(l.spec, a, t, lds, 0); break;
        case K2_ICOL: dispatch_pass_spec<T, K2_ICOL, true>(l.spec, a, t, lds, 0); break;
        case K2_ICOL_PAIR: dispatch_pass_spec<T, K2_ICOL_PAIR, true>(l.spec, a, t, lds, 0); break;
        case K2_ROWS_MOD: dispatch_pass_spec<T, K2_ROWS_MOD, true>(l.spec, a, t, lds, 0); break;
        case K2_FWD_COLS: dispatch_pass_spec<T, K2_FWD_COLS, true>(l.spec, a, t, lds, 0); break;
        default: break;
    }
}

template <class T>
static int emu_forward2d(int M, int N, int J, int L, int max_order, int pre_pad, long long lds_limit, const void* x,
                         const void* filt, void* out, long long batch, long long chunk, long long* info,
                         const int* tune = nullptr, int ntune = 0) {
    Scat2DPlan<T> p;
    int* fields[7] = {&p.tune.max_pow2, &p.tune.node_threads, &p.tune.col_block, &p.tune.row_block,
                      &p.tune.pass_threads, &p.tune.streams, &p.tune.pair};
    for (int i = 0; i < ntune && i < 7; ++i)
        if (tune[i] != 0) *fields[i] = tune[i];
    KymError e = {0, ""};
    if (!p.build(M, N, J, L, max_order, pre_pad, (size_t)lds_limit, &e)) {
        g_emu_error = e.msg;
        return e.code;
    }
    p.blob.relocate((uintptr_t)p.blob.bytes.data());
    const uintptr_t base = (uintptr_t)p.blob.bytes.data();
    if (info) {
        info[0] = p.C; info[1] = p.oh; info[2] = p.ow; info[3] = (long long)p.launches.size();
        int nbig = 0, nspec = 0, npair = 0, ncpair = 0;
        for (auto& nd : p.nodes) nbig += nd.big;
        for (auto& l : p.launches) {
            nspec += l.spec >= 0;
            npair += l.pair && l.kind == K2_LDS_NODE;
            ncpair += l.kind == K2_ICOL_PAIR;
        }
        info[4] = nbig;
        info[5] = nspec;
        info[6] = npair;
        int nbspec = 0;
        for (auto& l : p.bwd_launches) nbspec += l.spec >= 0;
        info[7] = nbspec;
        info[8] = ncpair;
    }
    if (!x) return 0;
    if (chunk <= 0 || chunk > batch) chunk = batch;
    std::vector<cpx<T>> ws((size_t)(p.ws_img * chunk) + 1);
    std::vector<cpx<T>> lds(p.max_lds / sizeof(cpx<T>) + 1);
    for (long long i0 = 0; i0 < batch; i0 += chunk) {
        const int nimg = (int)std::min(chunk, batch - i0);
        for (const Launch2D& l : p.launches) {
            Scat2DArgs<T> a = p.args(l, base);
            a.x = static_cast<const T*>(x);
            a.filt = static_cast<const T*>(filt);
            a.ws = ws.data();
            a.out = static_cast<T*>(out);
            a.img0 = (int)i0;
            a.ntask = nimg * l.ngroup * l.nblk;
            for (int t = 0; t < a.ntask; ++t) {
                if (l.spec >= 0) emu_spec(l, a, t, lds.data());
                else dispatch_body2d<T, 31>(l.kind, a, t, lds.data(), 0, 1);
            }
        }
    }
    return 0;
}

template <class T>
static int emu_train2d(int M, int N, int J, int L, int max_order, int pre_pad, long long lds_limit, const void* x,
                       const void* filt, void* out, const void* dS, void* gx, long long batch, int nbwd,
                       const int* tune = nullptr, int ntune = 0) {
    Scat2DPlan<T> p;
    int* fields[7] = {&p.tune.max_pow2, &p.tune.node_threads, &p.tune.col_block, &p.tune.row_block,
                      &p.tune.pass_threads, &p.tune.streams, &p.tune.pair};
    for (int i = 0; i < ntune && i < 7; ++i)
        if (tune[i] != 0) *fields[i] = tune[i];
    KymError e = {0, ""};
    if (!p.build(M, N, J, L, max_order, pre_pad, (size_t)lds_limit, &e)) {
        g_emu_error = e.msg;
        return e.code;
    }
    if (!p.trainable) {
        g_emu_error = p.train_error;
        return KYM_ERR_UNSUPPORTED;
    }
    p.blob.relocate((uintptr_t)p.blob.bytes.data());
    const uintptr_t base = (uintptr_t)p.blob.bytes.data();
    std::vector<cpx<T>> ws((size_t)(p.ws_img_train * batch) + 1);
    std::vector<cpx<T>> lds(p.max_lds / sizeof(cpx<T>) + 1);
    for (int phase = 0; phase < 1 + nbwd; ++phase) {
        const std::vector<Launch2D>& ls = phase == 0 ? p.launches : p.bwd_launches;
        for (const Launch2D& l : ls) {
            Scat2DArgs<T> a = p.args(l, base);
            a.x = static_cast<const T*>(x);
            a.filt = static_cast<const T*>(filt);
            a.ws = ws.data();
            a.ws_img = p.ws_img_train;
            a.out = static_cast<T*>(out);
            a.dS = static_cast<const T*>(dS);
            a.gx = static_cast<T*>(gx);
            a.save = 1;
            a.img0 = 0;
            a.ntask = (int)(batch * l.ngroup * l.nblk);
            for (int t = 0; t < a.ntask; ++t) {
                if (l.kind == K2B_NODE && l.spec >= 0) dispatch_bwd_node_spec<T, true>(l.spec, a, t, lds.data(), 0);
                else if (l.kind == K2B_GCOLS && l.spec >= 0)
                    dispatch_bwd_pass_spec<T, K2B_GCOLS, true>(l.spec, a, t, lds.data(), 0);
                else if (l.kind == K2B_GROWS && l.spec >= 0)
                    dispatch_bwd_pass_spec<T, K2B_GROWS, true>(l.spec, a, t, lds.data(), 0);
                else if (l.kind == K2B_GFCOLS && l.spec >= 0)
                    dispatch_bwd_pass_spec<T, K2B_GFCOLS, true>(l.spec, a, t, lds.data(), 0);
                else if (l.kind >= K2B_NODE) dispatch_bwd2d<T, 31>(l.kind, a, t, lds.data(), 0, 1);
                else if (l.spec >= 0) emu_spec(l, a, t, lds.data());
                else dispatch_body2d<T, 31>(l.kind, a, t, lds.data(), 0, 1);
            }
        }
    }
    return 0;
}

// nbwd: number of backward passes run on one training forward (the backward
// must leave the saved state intact, so every pass gives the same dL/dx).
extern "C" int kym_emu_scat2d_train(int M, int N, int J, int L, int max_order, int pre_pad, int dtype,
                                    long long lds_limit, const void* x, const void* filt, void* out, const void* dS,
                                    void* gx, long long batch, int nbwd) {
    if (dtype == 0) return emu_train2d<float>(M, N, J, L, max_order, pre_pad, lds_limit, x, filt, out, dS, gx, batch, nbwd);
    return emu_train2d<double>(M, N, J, L, max_order, pre_pad, lds_limit, x, filt, out, dS, gx, batch, nbwd);
}

// As kym_emu_scat2d_train with plan tuning choices (kym_scat2d_create_tuned).
extern "C" int kym_emu_scat2d_train_tuned(int M, int N, int J, int L, int max_order, int pre_pad, int dtype,
                                          long long lds_limit, const void* x, const void* filt, void* out,
                                          const void* dS, void* gx, long long batch, int nbwd, const int* tune,
                                          int ntune) {
    if (dtype == 0)
        return emu_train2d<float>(M, N, J, L, max_order, pre_pad, lds_limit, x, filt, out, dS, gx, batch, nbwd, tune,
                                  ntune);
    return emu_train2d<double>(M, N, J, L, max_order, pre_pad, lds_limit, x, filt, out, dS, gx, batch, nbwd, tune,
                               ntune);
}

// Leaf-tap descriptor of the plan (kym_scat2d_leaf_taps).
template <class T>
static int emu_taps2d(int M, int N, int J, int L, int max_order, int pre_pad, long long lds_limit, long long* desc,
                      int n) {
    Scat2DPlan<T> p;
    KymError e = {0, ""};
    if (!p.build(M, N, J, L, max_order, pre_pad, (size_t)lds_limit, &e)) {
        g_emu_error = e.msg;
        return e.code;
    }
    std::vector<long long> d = {p.taps_off, p.taps_len, p.Mp, p.Np};
    for (int r = 0; r < J; ++r)
        for (int ax = 0; ax < 2; ++ax) {
            const FirAxis& f = p.fir_axes[2 * r + ax];
            d.push_back(f.off);
            d.push_back(f.lo);
            d.push_back(f.cnt);
            d.push_back(f.band_off);
            d.push_back(f.kmax);
        }
    if (n < (int)d.size()) {
        g_emu_error = "descriptor too small";
        return KYM_ERR_VALUE;
    }
    for (size_t i = 0; i < d.size(); ++i) desc[i] = d[i];
    return 0;
}

extern "C" int kym_emu_scat2d_taps(int M, int N, int J, int L, int max_order, int pre_pad, int dtype,
                                   long long lds_limit, long long* desc, int n) {
    if (dtype == 0) return emu_taps2d<float>(M, N, J, L, max_order, pre_pad, lds_limit, desc, n);
    return emu_taps2d<double>(M, N, J, L, max_order, pre_pad, lds_limit, desc, n);
}

// As kym_emu_scat2d_forward with plan tuning choices (kym_scat2d_create_tuned).
extern "C" int kym_emu_scat2d_forward_tuned(int M, int N, int J, int L, int max_order, int pre_pad, int dtype,
                                            long long lds_limit, const void* x, const void* filt, void* out,
                                            long long batch, long long chunk, long long* info, const int* tune,
                                            int ntune) {
    if (dtype == 0)
        return emu_forward2d<float>(M, N, J, L, max_order, pre_pad, lds_limit, x, filt, out, batch, chunk, info,
                                    tune, ntune);
    return emu_forward2d<double>(M, N, J, L, max_order, pre_pad, lds_limit, x, filt, out, batch, chunk, info, tune,
                                 ntune);
}

extern "C" int kym_emu_scat2d_forward(int M, int N, int J, int L, int max_order, int pre_pad, int dtype,
                                      long long lds_limit, const void* x, const void* filt, void* out,
                                      long long batch, long long chunk, long long* info) {
    if (dtype == 0)
        return emu_forward2d<float>(M, N, J, L, max_order, pre_pad, lds_limit, x, filt, out, batch, chunk, info);
    return emu_forward2d<double>(M, N, J, L, max_order, pre_pad, lds_limit, x, filt, out, batch, chunk, info);
}

extern "C" int kym_emu_scat1d_forward(int N, int N_pad, int pad_left, int log2_stride, int average_mode,
                                      int oversampling, int max_order, int n_psi1, const int* psi1_j,
                                      const int* psi1_nlev, int n_psi2, const int* psi2_j, const int* psi2_nlev,
                                      int phi_nlev, const int* ind_start, const int* ind_end, int n_ind, int dtype,
                                      const void* x, const void* filt, void* out, long long batch,
                                      long long* info, const void* dS, void* gx, int nbwd) {
    Scat1DDesc d = {N, N_pad, pad_left, log2_stride, average_mode, oversampling, max_order, n_psi1, n_psi2, phi_nlev,
                    psi1_j, psi1_nlev, psi2_j, psi2_nlev, ind_start, ind_end, n_ind};
    auto run = [&](auto tag) -> int {
        using T = decltype(tag);
        Scat1DPlan<T> p;
        KymError e = {0, ""};
        if (!p.build(d, 160 * 1024, &e)) {
            g_emu_error = e.msg;
            return e.code;
        }
        p.blob.relocate((uintptr_t)p.blob.bytes.data());
        const uintptr_t base = (uintptr_t)p.blob.bytes.data();
        if (info) {
            info[0] = p.C; info[1] = p.out_row; info[2] = (long long)p.filt_off.size() - 1; info[3] = p.filt_off.back();
            int nspec = 0;
            for (const Launch1D& l : p.launches) nspec += l.spec >= 0;
            info[4] = nspec;
        }
        if (!x) return 0;
        const bool train = dS != nullptr;
        const long long stride = train ? p.ws_sig_train : p.ws_sig;
        std::vector<cpx<T>> ws((size_t)(stride * batch) + 1);
        std::vector<cpx<T>> lds(p.max_lds / sizeof(cpx<T>) + 1);
        for (int phase = 0; phase < (train ? 1 + nbwd : 1); ++phase)
            for (const Launch1D& l : (phase ? p.bwd_launches : p.launches)) {
                Scat1DArgs<T> a = p.args(l, base);
                a.x = static_cast<const T*>(x);
                a.filt = static_cast<const T*>(filt);
                a.ws = ws.data();
                a.ws_sig = stride;
                a.out = static_cast<T*>(out);
                a.dS = static_cast<const T*>(dS);
                a.gx = static_cast<T*>(gx);
                a.save = train ? 1 : 0;
                a.img0 = 0;
                a.ntask = (int)(batch * l.ngroup);
                for (int t = 0; t < a.ntask; ++t) {
                    if (l.kind == 0 && l.spec >= 0) dispatch_node1d_spec<T, true>(l.spec, a, t, lds.data(), 0);
                    else if (l.kind == 0) body_node1d<T, 31>(a, t, lds.data(), 0, 1);
                    else if (l.kind == 1) body_bwd_node1d<T, 31>(a, t, lds.data(), 0, 1);
                    else body_bwd_fold1d<T>(a, t, 0, 1);
                }
            }
        return 0;
    };
    if (dtype == 0) return run(float(0));
    return run(double(0));
}

// Per-op FFT (kym_op_fft) on the host: same launch sequence and bodies.
template <class T>
static int emu_op_fft(const void* x, void* y, void* work, long long batch, int H, int W, int kind, long long lds_cap) {
    LenTable<T> lt;
    const int lH = lt.add(H), lW = lt.add(W);
    Blob b;
    size_t lens_off = 0;
    std::string err;
    if (!lt.emit(b, &lens_off, &err)) { g_emu_error = err; return KYM_ERR_UNSUPPORTED; }
    b.relocate((uintptr_t)b.bytes.data());
    const FftLen<T>* lens = reinterpret_cast<const FftLen<T>*>(b.bytes.data() + lens_off);
    if (const long long m = fftop_multi_count(H, W, batch)) {
        const long long nblk = (batch + m - 1) / m;
        FftOpArgs o = {x, y, batch, H, W, kind, 0, (int)m, (int)nblk};
        std::vector<cpx<T>> lds((size_t)(m * H) + 1);
        for (long long blk = 0; blk < nblk; ++blk) fftop_multi_body<T>(o, lens, lH, blk, lds.data(), 0, 1);
        return 0;
    }
    FftOpStep steps[2];
    int nsteps = 0;
    const char* e = nullptr;
    const int code = fftop_steps(x, y, work, H, W, kind, (size_t)lds_cap, sizeof(cpx<T>), steps, &nsteps, &e);
    if (code) { g_emu_error = e; return code; }
    for (int q = 0; q < nsteps; ++q) {
        const FftOpStep& st = steps[q];
        FftOpArgs o = {x, y, batch, H, W, kind, 0, st.nrow, st.nblk};
        std::vector<cpx<T>> lds(st.lds / sizeof(cpx<T>) + 1);
        if (st.route == 0)
            for (long long bb = 0; bb < batch; ++bb) fftop_lds_body<T>(o, lens, lH, lW, bb, lds.data(), 0, 1);
        else
            for (long long task = 0; task < batch * st.nblk; ++task)
                fftop_pass_body<T>(o, st.io, lens, lH, lW, st.axis, task, lds.data(), 0, 1);
    }
    return 0;
}

template <class T>
static int emu_filter_ifft(const void* x, const void* f, void* y, long long batch, long long N, int k, int mode,
                           int tlen = 0, int nf = 1, int grp = 0, int gtot = 0, int goff = 0) {
    if (k < 1 || N % k) { g_emu_error = "subsampling factor must divide the size"; return KYM_ERR_RUNTIME; }
    const int n = (int)(N / k);
    LenTable<T> lt;
    const int l = lt.add(n);
    Blob b;
    size_t lens_off = 0;
    std::string err;
    if (!lt.emit(b, &lens_off, &err)) { g_emu_error = err; return KYM_ERR_UNSUPPORTED; }
    b.relocate((uintptr_t)b.bytes.data());
    const FftLen<T>* lens = reinterpret_cast<const FftLen<T>*>(b.bytes.data() + lens_off);
    {  // the compile-time body, as the GPU entry selects it
        FilterIfftArgs o = {x, f, y, batch, (int)N, k, n, mode, 0, 1, tlen, nf, grp, gtot, goff};
        int M = 0;
        const int sid = sizeof(T) == 4 ? filter_ifft_spec_id(n, batch, o, &M) : -1;
        if (sid >= 0) {
            o.nrow = M;
            std::vector<cpx<T>> lds((size_t)M * n + 1);
            for (long long blk = 0; blk < (batch + M - 1) / M; ++blk) {
                switch (sid) {
#define KYM_FI_E(id, nn, mm) \
    case id: filter_ifft_body_s<T, nn, mm, 1>(o, lens[l], blk, lds.data(), 0); break;
                    KYM_FILTER_IFFT_SPECS(KYM_FI_E)
#undef KYM_FI_E
                    default: break;
                }
            }
            return 0;
        }
    }
    int nt = 256, m = 1, split = 1;
    const long long elems = filter_ifft_shape(n, k, batch, &nt, &m, &split, tlen);
    FilterIfftArgs o = {x, f, y, batch, (int)N, k, n, mode, m, split, tlen, nf, grp, gtot, goff};
    std::vector<cpx<T>> lds((size_t)elems + 1);
    for (long long blk = 0; blk < (batch + m - 1) / m; ++blk) filter_ifft_body<T>(o, lens[l], blk, lds.data(), 0, 1);
    return 0;
}

extern "C" int kym_emu_filter_ifft(const void* x, const void* f, void* y, long long batch, long long N, int k,
                                   int mode, int dtype) {
    if (dtype == 0) return emu_filter_ifft<float>(x, f, y, batch, N, k, mode);
    return emu_filter_ifft<double>(x, f, y, batch, N, k, mode);
}

// kym_op_filter_ifft_t on the CPU (one thread per workgroup).
extern "C" int kym_emu_filter_ifft_t(const void* x, const void* f, void* y, long long batch, long long N, int k,
                                     int mode, long long tlen, int dtype) {
    if (tlen < 0 || (tlen > 0 && batch % tlen)) { g_emu_error = "the rows must be a whole number of lines"; return 3; }
    if (dtype == 0) return emu_filter_ifft<float>(x, f, y, batch, N, k, mode, (int)tlen);
    return emu_filter_ifft<double>(x, f, y, batch, N, k, mode, (int)tlen);
}

// kym_op_filter_ifft_g on the CPU (one thread per workgroup).
extern "C" int kym_emu_filter_ifft_g(const void* x, const void* f, void* y, long long rows, long long N, int k,
                                     int mode, int nf, int grp, int gtot, int goff, int dtype) {
    if (nf < 1 || rows % nf || grp < 0 || (grp > 0 && (rows % grp || gtot < goff + grp))) {
        g_emu_error = "invalid rows";
        return 3;
    }
    if (dtype == 0) return emu_filter_ifft<float>(x, f, y, rows, N, k, mode, 0, nf, grp, gtot, goff);
    return emu_filter_ifft<double>(x, f, y, rows, N, k, mode, 0, nf, grp, gtot, goff);
}

template <class T>
static int emu_freq_scatter(const void* x, void* const* ys, const void* const* fs, const int* ks, int nf, long long B,
                            int A, int TT, int c, int N) {
    LenTable<T> lt;
    const int lN = lt.add(N);
    int ln[KYM_FS_MAX];
    for (int i = 0; i < nf; ++i) ln[i] = lt.add(N / ks[i]);
    Blob b;
    size_t lens_off = 0;
    std::string err;
    if (!lt.emit(b, &lens_off, &err)) { g_emu_error = err; return KYM_ERR_UNSUPPORTED; }
    b.relocate((uintptr_t)b.bytes.data());
    const FftLen<T>* lens = reinterpret_cast<const FftLen<T>*>(b.bytes.data() + lens_off);
    FreqScatterArgs o = {};
    o.x = x; o.B = B; o.A = A; o.T = TT; o.c = c; o.N = N; o.nf = nf;
    o.m = freq_scatter_rows(N, TT);
    o.nt_blk = (TT + o.m - 1) / o.m;
    o.lenN = lens + lN;
    for (int i = 0; i < nf; ++i) {
        o.f[i] = fs[i]; o.y[i] = ys[i]; o.k[i] = ks[i]; o.lenn[i] = lens + ln[i];
    }
    std::vector<cpx<T>> lds((size_t)2 * o.m * N + 1);
    const bool spec = sizeof(T) == 4 && freq_scatter_spec_ok(o);  // as the GPU entry selects it
    for (long long blk = 0; blk < B * o.nt_blk; ++blk) {
        if (spec) freq_scatter_body_s<T, 88, 32, 1>(o, blk, lds.data(), 0);
        else freq_scatter_body<T, 31>(o, blk, lds.data(), 0, 1);
    }
    return 0;
}

// kym_op_freq_scatter on the CPU (one thread per workgroup).
extern "C" int kym_emu_freq_scatter(const void* x, void* const* ys, const void* const* fs, const int* ks, int nf,
                                    long long B, int A, int T, int c, int N, int dtype) {
    if (nf < 1 || nf > KYM_FS_MAX || A < 1 || A > N) { g_emu_error = "invalid arguments"; return 3; }
    for (int i = 0; i < nf; ++i)
        if (ks[i] < 1 || N % ks[i]) { g_emu_error = "subsampling factor must divide the size"; return 2; }
    if (dtype == 0) return emu_freq_scatter<float>(x, ys, fs, ks, nf, B, A, T, c, N);
    return emu_freq_scatter<double>(x, ys, fs, ks, nf, B, A, T, c, N);
}

template <class T>
static int emu_lowpass(const void* x, const void* f, void* y, long long outer, int N, int S, int k) {
    if (k < 1 || N % k) { g_emu_error = "subsampling factor must divide the size"; return KYM_ERR_RUNTIME; }
    const int n = N / k;
    LenTable<T> lt;
    const int lN = lt.add(N), ln = lt.add(n);
    Blob b;
    size_t lens_off = 0;
    std::string err;
    if (!lt.emit(b, &lens_off, &err)) { g_emu_error = err; return KYM_ERR_UNSUPPORTED; }
    b.relocate((uintptr_t)b.bytes.data());
    const FftLen<T>* lens = reinterpret_cast<const FftLen<T>*>(b.bytes.data() + lens_off);
    int m = 1, nsb = 1;
    long long nblk = 0;
    const long long elems = lowpass_shape(N, n, S, outer, &m, &nsb, &nblk);
    int M = 0;
    const int sid = sizeof(T) == 4 ? lowpass_spec_id(N, n, S, &M) : -1;
    if (sid >= 0) {  // the compile-time body, as the GPU entry selects it
        LowpassArgs o = {x, f, y, outer, N, S, k, n, M, S == 1 ? 1 : S / M};
        const long long nb = S == 1 ? (outer + M - 1) / M : outer * (S / M);
        std::vector<cpx<T>> lds((size_t)lowpass_spec_lds(N, n, S, M) + 1);
        for (long long blk = 0; blk < nb; ++blk) {
            switch (sid) {
#define KYM_LP_E(id, nn, n2, mm, sv) \
    case id: lowpass_spec_body<T, id, true>(o, lens[lN], lens[ln], blk, lds.data(), 0); break;
                KYM_LOWPASS_SPECS(KYM_LP_E)
#undef KYM_LP_E
                default: break;
            }
        }
        return 0;
    }
    LowpassArgs o = {x, f, y, outer, N, S, k, n, m, nsb};
    std::vector<cpx<T>> lds((size_t)elems + 1);
    for (long long blk = 0; blk < nblk; ++blk) lowpass_body<T>(o, lens[lN], lens[ln], blk, lds.data(), 0, 1);
    return 0;
}

// kym_op_lowpass on the CPU (one thread per workgroup).
extern "C" int kym_emu_lowpass(const void* x, const void* f, void* y, long long outer, long long N, long long S,
                               int k, int dtype) {
    if (dtype == 0) return emu_lowpass<float>(x, f, y, outer, (int)N, (int)S, k);
    return emu_lowpass<double>(x, f, y, outer, (int)N, (int)S, k);
}

extern "C" int kym_emu_op_fft(const void* x, void* y, void* work, long long batch, int H, int W, int kind, int dtype,
                              long long lds_cap) {
    if (dtype == 0) return emu_op_fft<float>(x, y, work, batch, H, W, kind, lds_cap);
    return emu_op_fft<double>(x, y, work, batch, H, W, kind, lds_cap);
}
