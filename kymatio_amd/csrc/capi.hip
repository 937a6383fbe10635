// C ABI of the scattering HIP library (declarations: include/kymatio_hip.h).
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/kymatio_hip.h"
#include "picks.h"
#include "scat1d_plan.h"
#include "scat2d_plan.h"

static thread_local std::string g_last_error;

void kym_set_error(const std::string& msg) { g_last_error = msg; }

static int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define KYM_HIP_TRY(expr)                                                                    \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            return fail(KYM_ERR_HIP, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
    } while (0)

extern "C" const char* kym_last_error(void) { return g_last_error.c_str(); }

int kym_ops_fail(int code, const std::string& msg) { return fail(code, msg); }

extern "C" int kym_version(void) { return KYM_ABI_VERSION; }

static size_t device_lds_limit() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess && v > 0)
        return (size_t)v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
        return (size_t)v;
    return 65536;
}

// ----------------------------------------------------------------- 2D plans
// Optional per-launch timing with HIP events (bench / profiling only).
struct LaunchProfile {
    bool enabled = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> recs;  // launch index, events
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    hipEvent_t get() {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
    void reset() { recs.clear(); used = 0; }
    ~LaunchProfile() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Extra streams of a plan's inference forward: chunks rotate over the
// caller's stream and up to KYM_AUX_STREAMS streams of the plan, forked from
// and joined back to the caller's stream by events, so one chunk's launches
// fill the GPU while another's drain.  The mutex makes each record + wait
// sequence atomic when several host threads share a plan.
#define KYM_AUX_STREAMS 3
struct ChunkStreams {
    hipStream_t aux[KYM_AUX_STREAMS] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[KYM_AUX_STREAMS] = {};
    std::mutex mu;
    int create() {
        KYM_HIP_TRY(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        for (int i = 0; i < KYM_AUX_STREAMS; ++i) {
            KYM_HIP_TRY(hipStreamCreateWithFlags(&aux[i], hipStreamNonBlocking));
            KYM_HIP_TRY(hipEventCreateWithFlags(&join[i], hipEventDisableTiming));
        }
        return KYM_OK;
    }
    ~ChunkStreams() {
        if (fork) (void)hipEventDestroy(fork);
        for (int i = 0; i < KYM_AUX_STREAMS; ++i) {
            if (join[i]) (void)hipEventDestroy(join[i]);
            if (aux[i]) (void)hipStreamDestroy(aux[i]);
        }
    }
};

// Streams a chunked inference call uses: `want` (0 = 2) reduced until every
// stream gets a chunk.  The chunk then fits one stream's share of the
// workspace, with the chunk count a multiple of the stream count and the
// chunks as even as that allows (256 over 3 streams: 6 chunks of <= 43, not
// 85, 85, 85, 1).
static int chunk_streams(const ChunkStreams* cs, int want, int mode, long long batch, long long* chunk) {
    int ns = want == 0 ? 2 : want;
    if (!cs || !cs->aux[0]) ns = 1;
    (void)mode;
    while (ns > 1 && (*chunk < ns || batch <= *chunk / ns * (ns - 1))) --ns;
    if (ns > 1) {
        const long long part = *chunk / ns;
        long long n = (batch + part - 1) / part;
        n = (n + ns - 1) / ns * ns;
        *chunk = (batch + n - 1) / n;
    }
    return ns;
}

// Fork the extra streams of a chunked call from the caller's stream (begin)
// or join them back into it (end).
static int chunk_fork(ChunkStreams* cs, int ns, hipStream_t stream) {
    if (ns <= 1) return KYM_OK;
    std::lock_guard<std::mutex> lock(cs->mu);
    KYM_HIP_TRY(hipEventRecord(cs->fork, stream));
    for (int i = 0; i < ns - 1; ++i) KYM_HIP_TRY(hipStreamWaitEvent(cs->aux[i], cs->fork, 0));
    return KYM_OK;
}

static int chunk_join(ChunkStreams* cs, int ns, hipStream_t stream) {
    if (ns <= 1) return KYM_OK;
    std::lock_guard<std::mutex> lock(cs->mu);
    for (int i = 0; i < ns - 1; ++i) {
        KYM_HIP_TRY(hipEventRecord(cs->join[i], cs->aux[i]));
        KYM_HIP_TRY(hipStreamWaitEvent(stream, cs->join[i], 0));
    }
    return KYM_OK;
}

struct KymScat2D {
    int dtype = KYM_F32;
    int device = 0;
    void* dblob = nullptr;
    std::unique_ptr<Scat2DPlan<float>> pf;
    std::unique_ptr<Scat2DPlan<double>> pd;
    LaunchProfile prof;
    ChunkStreams cs;
};

template <class T>
static int upload_and_prepare(Scat2DPlan<T>& p, void** dblob) {
    const size_t bytes = p.blob.bytes.size();
    KYM_HIP_TRY(hipMalloc(dblob, bytes));
    p.blob.relocate((uintptr_t)*dblob);
    KYM_HIP_TRY(hipMemcpy(*dblob, p.blob.bytes.data(), bytes, hipMemcpyHostToDevice));
    for (const std::vector<Launch2D>* ls : {&p.launches, &p.bwd_launches})
        for (const Launch2D& l : *ls) {
            Scat2DKernel<T> k = pick_scat2d<T>(l.kind, l.threads, p.pclass, l.spec);
            if (!k) return fail(KYM_ERR_UNSUPPORTED, "no kernel for launch kind");
            if (l.lds_bytes > 65536)
                KYM_HIP_TRY(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                (int)l.lds_bytes));
        }
    return KYM_OK;
}

extern "C" int kym_scat2d_create_tuned(int M, int N, int J, int L, int max_order, int pre_pad, int dtype,
                                       long long lds_limit, const int* tune, int ntune, void** out) {
    if (!out) return fail(KYM_ERR_VALUE, "null plan pointer");
    *out = nullptr;
    if (dtype != KYM_F32 && dtype != KYM_F64) return fail(KYM_ERR_TYPE, "dtype must be float32 or float64");
    if (ntune < 0 || ntune > 7 || (ntune > 0 && !tune)) return fail(KYM_ERR_VALUE, "tuning array must hold 0-7 values");
    Scat2DTune t;
    int* fields[7] = {&t.max_pow2, &t.node_threads, &t.col_block, &t.row_block, &t.pass_threads, &t.streams, &t.pair};
    for (int i = 0; i < ntune; ++i)
        if (tune[i] != 0) *fields[i] = tune[i];
    auto h = std::make_unique<KymScat2D>();
    h->dtype = dtype;
    KYM_HIP_TRY(hipGetDevice(&h->device));
    size_t lim = device_lds_limit();
    if (lds_limit > 0 && (size_t)lds_limit < lim) lim = (size_t)lds_limit;
    KymError e = {KYM_OK, ""};
    int rc;
    if (dtype == KYM_F32) {
        h->pf = std::make_unique<Scat2DPlan<float>>();
        h->pf->tune = t;
        if (!h->pf->build(M, N, J, L, max_order, pre_pad, lim, &e)) return fail(e.code, e.msg);
        rc = upload_and_prepare(*h->pf, &h->dblob);
    } else {
        h->pd = std::make_unique<Scat2DPlan<double>>();
        h->pd->tune = t;
        if (!h->pd->build(M, N, J, L, max_order, pre_pad, lim, &e)) return fail(e.code, e.msg);
        rc = upload_and_prepare(*h->pd, &h->dblob);
    }
    if (rc == KYM_OK) rc = h->cs.create();
    if (rc != KYM_OK) {
        if (h->dblob) (void)hipFree(h->dblob);
        return rc;
    }
    *out = h.release();
    return KYM_OK;
}

extern "C" int kym_scat2d_create(int M, int N, int J, int L, int max_order, int pre_pad, int dtype,
                                 long long lds_limit, void** out) {
    return kym_scat2d_create_tuned(M, N, J, L, max_order, pre_pad, dtype, lds_limit, nullptr, 0, out);
}

extern "C" int kym_scat2d_destroy(void* plan) {
    if (!plan) return KYM_OK;
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(h->device);
    if (h->dblob) (void)hipFree(h->dblob);
    delete h;  // the plan's stream and events, on its device
    (void)hipSetDevice(cur);
    return KYM_OK;
}

template <class T>
static void fill_info(const Scat2DPlan<T>& p, long long* info) {
    info[0] = p.C;
    info[1] = p.oh;
    info[2] = p.ow;
    info[3] = p.Mp;
    info[4] = p.Np;
    info[5] = (long long)p.filt_off.size() - 1;
    info[6] = p.taps_off + p.taps_len;
    info[7] = p.ws_img * (long long)sizeof(cpx<T>);
    info[8] = (long long)p.launches.size();
    info[9] = p.top;
    info[10] = p.left;
    info[11] = p.trainable ? p.ws_img_train * (long long)sizeof(cpx<T>) : -1;
}

extern "C" int kym_scat2d_info(void* plan, long long* info) {
    if (!plan || !info) return fail(KYM_ERR_VALUE, "null argument");
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    if (h->pf) fill_info(*h->pf, info);
    else fill_info(*h->pd, info);
    return KYM_OK;
}

extern "C" int kym_scat2d_filter_offsets(void* plan, long long* offs) {
    if (!plan || !offs) return fail(KYM_ERR_VALUE, "null argument");
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    const std::vector<long long>& v = h->pf ? h->pf->filt_off : h->pd->filt_off;
    for (size_t i = 0; i < v.size(); ++i) offs[i] = v[i];
    return KYM_OK;
}

template <class T>
static int leaf_taps(const Scat2DPlan<T>& p, long long* desc, int n) {
    std::vector<long long> d = {p.taps_off, p.taps_len, p.Mp, p.Np};
    for (int r = 0; r < p.J; ++r)
        for (int ax = 0; ax < 2; ++ax) {
            const FirAxis& f = p.fir_axes[2 * r + ax];
            d.push_back(f.off);
            d.push_back(f.lo);
            d.push_back(f.cnt);
            d.push_back(f.band_off);
            d.push_back(f.kmax);
        }
    if (n < (int)d.size()) return fail(KYM_ERR_VALUE, "leaf taps descriptor needs 4 + 10 J entries");
    for (size_t i = 0; i < d.size(); ++i) desc[i] = d[i];
    return KYM_OK;
}

extern "C" int kym_scat2d_leaf_taps(void* plan, long long* desc, int n) {
    if (!plan || !desc) return fail(KYM_ERR_VALUE, "null argument");
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    return h->pf ? leaf_taps(*h->pf, desc, n) : leaf_taps(*h->pd, desc, n);
}

// mode 0: inference forward (batch in chunks); 1: training forward (saves
// u, whole batch in the workspace); 2: backward from the saved workspace.
// The kernel, grid and arguments of one launch of a forward chunk.
template <class T>
struct Launch2DCall {
    Scat2DKernel<T> k;
    unsigned grid, threads;
    size_t lds;
    Scat2DArgs<T> a;
};

template <class T>
static int launch_call(const Scat2DPlan<T>& p, const Launch2D& l, int mode, const void* x, const void* filt,
                       void* out, const void* dS, void* gx, long long img_elems, void* ws, long long i0, int nimg,
                       uintptr_t base, Launch2DCall<T>* c) {
    Scat2DArgs<T> a = p.args(l, base);
    a.x = static_cast<const T*>(x);
    a.filt = static_cast<const T*>(filt);
    a.ws = static_cast<cpx<T>*>(ws);
    a.ws_img = img_elems;
    a.out = static_cast<T*>(out);
    a.dS = static_cast<const T*>(dS);
    a.gx = static_cast<T*>(gx);
    a.save = mode == 1 ? 1 : 0;
    a.img0 = (int)i0;
    const long long ntask = (long long)nimg * l.ngroup * l.nblk;
    if (ntask > (1ll << 30)) return fail(KYM_ERR_VALUE, "too many tasks in one launch");
    a.ntask = (int)ntask;
    c->grid = (unsigned)((ntask + 7) / 8 * 8);
    c->threads = (unsigned)l.threads;
    c->lds = l.lds_bytes;
    c->k = pick_scat2d<T>(l.kind, l.threads, p.pclass, l.spec);
    if (!c->k) return fail(KYM_ERR_UNSUPPORTED, "no kernel for launch kind");
    c->a = a;
    return KYM_OK;
}

// Images per chunk of a forward over `batch` images with ws_bytes of workspace.
template <class T>
static int chunk_of(const Scat2DPlan<T>& p, int mode, long long batch, long long ws_bytes, long long* chunk) {
    const long long img_elems = mode == 0 ? p.ws_img : p.ws_img_train;
    const long long per = img_elems * (long long)sizeof(cpx<T>);
    if (mode != 0 && !p.trainable) return fail(KYM_ERR_UNSUPPORTED, p.train_error);
    long long c = per > 0 ? ws_bytes / per : batch;
    if (c < 1) return fail(KYM_ERR_VALUE, "workspace too small for one image");
    if (c > batch) c = batch;
    if (mode != 0 && c < batch) return fail(KYM_ERR_VALUE, "training workspace must hold the whole batch");
    *chunk = c;
    return KYM_OK;
}

// mode 0: inference forward (batch in chunks); 1: training forward (saves
// u, whole batch in the workspace); 2: backward from the saved workspace.
// With S streams (tune.streams, 0 = 2; at least S images) chunk k runs on
// stream k mod S.  Inference splits the workspace in S parts and chunk k
// uses part k mod S, so chunk k + S reuses a part only after chunk k on the
// same stream; training keeps image i's state at image offset i of the
// workspace whatever the chunks, so the backward finds what the forward
// saved.  The extra streams wait for the caller's stream before their first
// chunk, and the caller's stream waits for them at the end.
template <class T>
static int run2d(const Scat2DPlan<T>& p, int mode, const void* x, const void* filt, void* out, const void* dS,
                 void* gx, long long batch, void* ws, long long ws_bytes, hipStream_t stream, uintptr_t base,
                 LaunchProfile* prof, ChunkStreams* cs = nullptr) {
    const long long img_elems = mode == 0 ? p.ws_img : p.ws_img_train;
    long long chunk = 0;
    int rc = chunk_of(p, mode, batch, ws_bytes, &chunk);
    if (rc != KYM_OK) return rc;
    const std::vector<Launch2D>& launches = mode == 2 ? p.bwd_launches : p.launches;
    const int ns = chunk_streams(cs, p.tune.streams, mode, batch, &chunk);
    rc = chunk_fork(cs, ns, stream);
    if (rc != KYM_OK) return rc;
    int k = 0;
    for (long long i0 = 0; i0 < batch; i0 += chunk, ++k) {
        const int nimg = (int)std::min(chunk, batch - i0);
        const int si = k % ns;
        const hipStream_t st = si ? cs->aux[si - 1] : stream;
        void* wsk = static_cast<char*>(ws) + (mode == 0 ? si * chunk : i0) * img_elems * (long long)sizeof(cpx<T>);
        for (const Launch2D& l : launches) {
            Launch2DCall<T> c;
            rc = launch_call(p, l, mode, x, filt, out, dS, gx, img_elems, wsk, i0, nimg, base, &c);
            if (rc != KYM_OK) return rc;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (prof && prof->enabled) {
                e0 = prof->get();
                e1 = prof->get();
                if (e0) KYM_HIP_TRY(hipEventRecord(e0, st));
            }
            hipLaunchKernelGGL(c.k, dim3(c.grid), dim3(c.threads), c.lds, st, c.a);
            KYM_HIP_TRY(hipGetLastError());
            if (e0 && e1) {
                KYM_HIP_TRY(hipEventRecord(e1, st));
                prof->recs.push_back({(int)(&l - launches.data()), {e0, e1}});
            }
        }
    }
    return chunk_join(cs, ns, stream);
}

extern "C" int kym_scat2d_forward(void* plan, const void* x, const void* filt, void* out, long long batch,
                                  void* ws, long long ws_bytes, void* stream) {
    if (!plan) return fail(KYM_ERR_VALUE, "null plan");
    if (batch < 0) return fail(KYM_ERR_VALUE, "negative batch");
    if (batch == 0) return KYM_OK;
    if (!x || !filt || !out || !ws) return fail(KYM_ERR_VALUE, "null device pointer");
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    int cur = 0;
    KYM_HIP_TRY(hipGetDevice(&cur));
    if (cur != h->device) return fail(KYM_ERR_TYPE, "plan was created for another device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->pf)
        return run2d(*h->pf, 0, x, filt, out, nullptr, nullptr, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, &h->prof,
                     &h->cs);
    return run2d(*h->pd, 0, x, filt, out, nullptr, nullptr, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, &h->prof,
                 &h->cs);
}

// ------------------------------------------------- 2D forward, no allocation
// Workspace of kym_scat2d_forward_noalloc: one file-scope device array, a
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol):
// the device address of each GPU's copy comes from hipGetSymbolAddress on the
// first call with that GPU current, and that address is what the kernels get.
#define KYM_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[KYM_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[KYM_NOALLOC_MAX_DEVICES] = {};

static int scratch_of_device(int dev, void** ptr) {
    if (dev < 0 || dev >= KYM_NOALLOC_MAX_DEVICES) return fail(KYM_ERR_VALUE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        KYM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(KYM_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return KYM_OK;
}

extern "C" int kym_scat2d_forward_noalloc(void* plan, const void* x, const void* filt, void* out, long long batch,
                                          void* stream) {
    if (!plan) return fail(KYM_ERR_VALUE, "null plan");
    if (batch < 0) return fail(KYM_ERR_VALUE, "negative batch");
    if (!x || !filt || !out) return fail(KYM_ERR_VALUE, "null device pointer");
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    int cur = 0;
    KYM_HIP_TRY(hipGetDevice(&cur));
    if (cur != h->device) return fail(KYM_ERR_TYPE, "plan was created for another device");
    const long long per_image = h->pf ? h->pf->ws_img * (long long)sizeof(cpx<float>)
                                      : h->pd->ws_img * (long long)sizeof(cpx<double>);
    if (per_image > KYM_NOALLOC_SCRATCH_BYTES)
        return fail(KYM_ERR_VALUE, "kym_scat2d_forward_noalloc: one image needs " + std::to_string(per_image) +
                                       " bytes of workspace, more than the " +
                                       std::to_string((long long)KYM_NOALLOC_SCRATCH_BYTES) +
                                       "-byte static scratch; use kym_scat2d_forward with a workspace");
    void* ws = nullptr;
    int rc = scratch_of_device(cur, &ws);
    if (rc != KYM_OK) return rc;
    if (batch == 0) return KYM_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // no per-launch profiling here: its events would be created on demand
    if (h->pf)
        return run2d(*h->pf, 0, x, filt, out, nullptr, nullptr, batch, ws, KYM_NOALLOC_SCRATCH_BYTES, s,
                     (uintptr_t)h->dblob, nullptr, &h->cs);
    return run2d(*h->pd, 0, x, filt, out, nullptr, nullptr, batch, ws, KYM_NOALLOC_SCRATCH_BYTES, s,
                 (uintptr_t)h->dblob, nullptr, &h->cs);
}

extern "C" int kym_scat2d_forward_train(void* plan, const void* x, const void* filt, void* out, long long batch,
                                        void* ws, long long ws_bytes, void* stream) {
    if (!plan) return fail(KYM_ERR_VALUE, "null plan");
    if (batch < 0) return fail(KYM_ERR_VALUE, "negative batch");
    if (batch == 0) return KYM_OK;
    if (!x || !filt || !out || !ws) return fail(KYM_ERR_VALUE, "null device pointer");
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    int cur = 0;
    KYM_HIP_TRY(hipGetDevice(&cur));
    if (cur != h->device) return fail(KYM_ERR_TYPE, "plan was created for another device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->pf)
        return run2d(*h->pf, 1, x, filt, out, nullptr, nullptr, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, &h->prof,
                     &h->cs);
    return run2d(*h->pd, 1, x, filt, out, nullptr, nullptr, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, &h->prof,
                 &h->cs);
}

extern "C" int kym_scat2d_backward(void* plan, const void* dS, const void* filt, void* gx, long long batch,
                                   void* ws, long long ws_bytes, void* stream) {
    if (!plan) return fail(KYM_ERR_VALUE, "null plan");
    if (batch < 0) return fail(KYM_ERR_VALUE, "negative batch");
    if (batch == 0) return KYM_OK;
    if (!dS || !filt || !gx || !ws) return fail(KYM_ERR_VALUE, "null device pointer");
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    int cur = 0;
    KYM_HIP_TRY(hipGetDevice(&cur));
    if (cur != h->device) return fail(KYM_ERR_TYPE, "plan was created for another device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->pf)
        return run2d(*h->pf, 2, nullptr, filt, nullptr, dS, gx, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, nullptr,
                     &h->cs);
    return run2d(*h->pd, 2, nullptr, filt, nullptr, dS, gx, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, nullptr,
                 &h->cs);
}

template <class T>
static void launch_rec(const Scat2DPlan<T>& p, int i, long long* rec) {
    const Launch2D& l = p.launches[i];
    rec[0] = l.kind; rec[1] = l.pair ? 2 * l.ngroup : l.ngroup; rec[2] = l.nblk; rec[3] = l.threads;
    rec[4] = (long long)l.lds_bytes; rec[5] = l.model_bytes; rec[6] = l.res;
}

extern "C" int kym_scat2d_launch_info(void* plan, int i, long long* rec) {
    if (!plan || !rec) return fail(KYM_ERR_VALUE, "null argument");
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    const int n = h->pf ? (int)h->pf->launches.size() : (int)h->pd->launches.size();
    if (i < 0 || i >= n) return fail(KYM_ERR_VALUE, "launch index out of range");
    if (h->pf) launch_rec(*h->pf, i, rec);
    else launch_rec(*h->pd, i, rec);
    return KYM_OK;
}

extern "C" int kym_scat2d_profile(void* plan, int enable) {
    if (!plan) return fail(KYM_ERR_VALUE, "null plan");
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    h->prof.enabled = enable != 0;
    h->prof.reset();
    return KYM_OK;
}

extern "C" int kym_scat2d_profile_read(void* plan, double* ms, long long* count, int n) {
    if (!plan || !ms || !count) return fail(KYM_ERR_VALUE, "null argument");
    KymScat2D* h = static_cast<KymScat2D*>(plan);
    for (int i = 0; i < n; ++i) { ms[i] = 0.0; count[i] = 0; }
    for (auto& r : h->prof.recs) {
        KYM_HIP_TRY(hipEventSynchronize(r.second.second));
        float t = 0.f;
        KYM_HIP_TRY(hipEventElapsedTime(&t, r.second.first, r.second.second));
        if (r.first >= 0 && r.first < n) { ms[r.first] += t; count[r.first] += 1; }
    }
    h->prof.reset();
    return KYM_OK;
}

// ----------------------------------------------------------------- 2D graphs
// One executable HIP graph over the inference forwards of several plans (one
// per image size), bound to fixed buffers: each plan's launches form one
// dependency chain, chains are independent branches.
struct KymScat2DGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
    ~KymScat2DGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (g) (void)hipGraphDestroy(g);
    }
};

// dev_count (counted graphs): the chain is laid out for `batch` images (the
// maximum) and its kernels read the actual count from *dev_count at replay.
template <class T>
static int add_plan_chain(hipGraph_t g, const Scat2DPlan<T>& p, const void* x, const void* filt, void* out,
                          long long batch, void* ws, long long ws_bytes, uintptr_t base,
                          const int* dev_count = nullptr) {
    long long chunk = 0;
    int rc = chunk_of(p, 0, batch, ws_bytes, &chunk);
    if (rc != KYM_OK) return rc;
    if (dev_count && batch > 0x7fffffffll) return fail(KYM_ERR_VALUE, "maximum batch too large");
    hipGraphNode_t prev = nullptr;
    for (long long i0 = 0; i0 < batch; i0 += chunk) {
        const int nimg = (int)std::min(chunk, batch - i0);
        for (const Launch2D& l : p.launches) {
            Launch2DCall<T> c;
            rc = launch_call(p, l, 0, x, filt, out, nullptr, nullptr, p.ws_img, ws, i0, nimg, base, &c);
            if (rc != KYM_OK) return rc;
            c.a.dev_count = dev_count;
            c.a.max_count = (int)batch;
            void* args[1] = {&c.a};
            hipKernelNodeParams kp = {};
            kp.func = reinterpret_cast<void*>(c.k);
            kp.gridDim = dim3(c.grid);
            kp.blockDim = dim3(c.threads);
            kp.sharedMemBytes = (unsigned)c.lds;
            kp.kernelParams = args;
            kp.extra = nullptr;
            hipGraphNode_t node;
            KYM_HIP_TRY(hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &kp));
            prev = node;
        }
    }
    return KYM_OK;
}

extern "C" int kym_scat2d_graph_create(void* const* plans, int nplans, const void* const* x, const void* const* filt,
                                       void* const* out, const long long* batch, void* const* ws,
                                       const long long* ws_bytes, void** graph) {
    if (!graph) return fail(KYM_ERR_VALUE, "null graph pointer");
    *graph = nullptr;
    if (nplans < 1 || !plans || !x || !filt || !out || !batch || !ws || !ws_bytes)
        return fail(KYM_ERR_VALUE, "null argument");
    auto G = std::make_unique<KymScat2DGraph>();
    KYM_HIP_TRY(hipGetDevice(&G->device));
    KYM_HIP_TRY(hipGraphCreate(&G->g, 0));
    for (int i = 0; i < nplans; ++i) {
        KymScat2D* h = static_cast<KymScat2D*>(plans[i]);
        if (!h) return fail(KYM_ERR_VALUE, "null plan");
        if (h->device != G->device) return fail(KYM_ERR_TYPE, "plan was created for another device");
        if (batch[i] < 0) return fail(KYM_ERR_VALUE, "negative batch");
        if (batch[i] == 0) continue;
        if (!x[i] || !filt[i] || !out[i] || !ws[i]) return fail(KYM_ERR_VALUE, "null device pointer");
        const int rc = h->pf ? add_plan_chain(G->g, *h->pf, x[i], filt[i], out[i], batch[i], ws[i], ws_bytes[i],
                                              (uintptr_t)h->dblob)
                             : add_plan_chain(G->g, *h->pd, x[i], filt[i], out[i], batch[i], ws[i], ws_bytes[i],
                                              (uintptr_t)h->dblob);
        if (rc != KYM_OK) return rc;
    }
    KYM_HIP_TRY(hipGraphInstantiate(&G->exec, G->g, nullptr, nullptr, 0));
    *graph = G.release();
    return KYM_OK;
}

extern "C" int kym_scat2d_graph_create_counted(void* const* plans, int nplans, const void* const* x,
                                               const void* const* filt, void* const* out, const long long* max_batch,
                                               void* const* ws, const long long* ws_bytes, const int* dev_counts,
                                               void** graph) {
    if (!graph) return fail(KYM_ERR_VALUE, "null graph pointer");
    *graph = nullptr;
    if (nplans < 1 || !plans || !x || !filt || !out || !max_batch || !ws || !ws_bytes || !dev_counts)
        return fail(KYM_ERR_VALUE, "null argument");
    auto G = std::make_unique<KymScat2DGraph>();
    KYM_HIP_TRY(hipGetDevice(&G->device));
    KYM_HIP_TRY(hipGraphCreate(&G->g, 0));
    for (int i = 0; i < nplans; ++i) {
        KymScat2D* h = static_cast<KymScat2D*>(plans[i]);
        if (!h) return fail(KYM_ERR_VALUE, "null plan");
        if (h->device != G->device) return fail(KYM_ERR_TYPE, "plan was created for another device");
        if (max_batch[i] < 0) return fail(KYM_ERR_VALUE, "negative batch");
        if (max_batch[i] == 0) continue;
        if (!x[i] || !filt[i] || !out[i] || !ws[i]) return fail(KYM_ERR_VALUE, "null device pointer");
        const int rc = h->pf ? add_plan_chain(G->g, *h->pf, x[i], filt[i], out[i], max_batch[i], ws[i], ws_bytes[i],
                                              (uintptr_t)h->dblob, dev_counts + i)
                             : add_plan_chain(G->g, *h->pd, x[i], filt[i], out[i], max_batch[i], ws[i], ws_bytes[i],
                                              (uintptr_t)h->dblob, dev_counts + i);
        if (rc != KYM_OK) return rc;
    }
    KYM_HIP_TRY(hipGraphInstantiate(&G->exec, G->g, nullptr, nullptr, 0));
    *graph = G.release();
    return KYM_OK;
}

extern "C" int kym_scat2d_graph_launch(void* graph, void* stream) {
    if (!graph) return fail(KYM_ERR_VALUE, "null graph");
    KymScat2DGraph* G = static_cast<KymScat2DGraph*>(graph);
    int cur = 0;
    KYM_HIP_TRY(hipGetDevice(&cur));
    if (cur != G->device) return fail(KYM_ERR_TYPE, "graph was created for another device");
    KYM_HIP_TRY(hipGraphLaunch(G->exec, static_cast<hipStream_t>(stream)));
    return KYM_OK;
}

extern "C" int kym_scat2d_graph_destroy(void* graph) {
    if (!graph) return KYM_OK;
    KymScat2DGraph* G = static_cast<KymScat2DGraph*>(graph);
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(G->device);
    delete G;
    (void)hipSetDevice(cur);
    return KYM_OK;
}

// ----------------------------------------------------------------- 1D plans
struct KymScat1D {
    int dtype = KYM_F32;
    int device = 0;
    void* dblob = nullptr;
    std::unique_ptr<Scat1DPlan<float>> pf;
    std::unique_ptr<Scat1DPlan<double>> pd;
    LaunchProfile prof;
    ChunkStreams cs;
    int streams = 0;  // inference chunks in flight (kym_scat1d_set_streams; 0 = 2)
};

template <class T>
static int upload1d(Scat1DPlan<T>& p, void** dblob) {
    const size_t bytes = p.blob.bytes.size();
    KYM_HIP_TRY(hipMalloc(dblob, bytes));
    p.blob.relocate((uintptr_t)*dblob);
    KYM_HIP_TRY(hipMemcpy(*dblob, p.blob.bytes.data(), bytes, hipMemcpyHostToDevice));
    for (const std::vector<Launch1D>* ls : {&p.launches, &p.bwd_launches})
        for (const Launch1D& l : *ls)
            if (l.lds_bytes > 65536)
                KYM_HIP_TRY(hipFuncSetAttribute((const void*)pick_scat1d<T>(l.threads, l.kind, l.spec),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)l.lds_bytes));
    return KYM_OK;
}

extern "C" int kym_scat1d_create(int N, int N_pad, int pad_left, int log2_stride, int average_mode,
                                 int oversampling, int max_order, int n_psi1, const int* psi1_j,
                                 const int* psi1_nlev, int n_psi2, const int* psi2_j, const int* psi2_nlev,
                                 int phi_nlev, const int* ind_start, const int* ind_end, int n_ind, int dtype,
                                 long long lds_limit, void** out) {
    if (!out) return fail(KYM_ERR_VALUE, "null plan pointer");
    *out = nullptr;
    if (dtype != KYM_F32 && dtype != KYM_F64) return fail(KYM_ERR_TYPE, "dtype must be float32 or float64");
    if (average_mode < 0 || average_mode > 2) return fail(KYM_ERR_VALUE, "average_mode must be 0, 1 or 2");
    Scat1DDesc d = {N, N_pad, pad_left, log2_stride, average_mode, oversampling, max_order, n_psi1, n_psi2, phi_nlev,
                    psi1_j, psi1_nlev, psi2_j, psi2_nlev, ind_start, ind_end, n_ind};
    auto h = std::make_unique<KymScat1D>();
    h->dtype = dtype;
    KYM_HIP_TRY(hipGetDevice(&h->device));
    size_t lim = device_lds_limit();
    if (lds_limit > 0 && (size_t)lds_limit < lim) lim = (size_t)lds_limit;
    KymError e = {KYM_OK, ""};
    int rc;
    if (dtype == KYM_F32) {
        h->pf = std::make_unique<Scat1DPlan<float>>();
        if (!h->pf->build(d, lim, &e)) return fail(e.code, e.msg);
        rc = upload1d(*h->pf, &h->dblob);
    } else {
        h->pd = std::make_unique<Scat1DPlan<double>>();
        if (!h->pd->build(d, lim, &e)) return fail(e.code, e.msg);
        rc = upload1d(*h->pd, &h->dblob);
    }
    if (rc == KYM_OK) rc = h->cs.create();
    if (rc != KYM_OK) {
        if (h->dblob) (void)hipFree(h->dblob);
        return rc;
    }
    *out = h.release();
    return KYM_OK;
}

extern "C" int kym_scat1d_destroy(void* plan) {
    if (!plan) return KYM_OK;
    KymScat1D* h = static_cast<KymScat1D*>(plan);
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(h->device);
    if (h->dblob) (void)hipFree(h->dblob);
    delete h;  // the plan's streams and events, on its device
    (void)hipSetDevice(cur);
    return KYM_OK;
}

extern "C" int kym_scat1d_set_streams(void* plan, int streams) {
    if (!plan) return fail(KYM_ERR_VALUE, "null plan");
    if (streams < 0 || streams > KYM_AUX_STREAMS + 1) return fail(KYM_ERR_VALUE, "streams must be 0-4");
    static_cast<KymScat1D*>(plan)->streams = streams;
    return KYM_OK;
}

template <class T>
static void fill_info1d(const Scat1DPlan<T>& p, long long* info) {
    info[0] = p.C;
    info[1] = p.out_row;
    info[2] = (long long)p.filt_off.size() - 1;
    info[3] = p.filt_off.back();
    info[4] = p.ws_sig * (long long)sizeof(cpx<T>);
    info[5] = (long long)p.launches.size();
    info[6] = p.ws_sig_train * (long long)sizeof(cpx<T>);
}

extern "C" int kym_scat1d_info(void* plan, long long* info) {
    if (!plan || !info) return fail(KYM_ERR_VALUE, "null argument");
    KymScat1D* h = static_cast<KymScat1D*>(plan);
    if (h->pf) fill_info1d(*h->pf, info);
    else fill_info1d(*h->pd, info);
    return KYM_OK;
}

extern "C" int kym_scat1d_layout(void* plan, long long* filt_offs, long long* path_len) {
    if (!plan) return fail(KYM_ERR_VALUE, "null plan");
    KymScat1D* h = static_cast<KymScat1D*>(plan);
    const std::vector<long long>& fo = h->pf ? h->pf->filt_off : h->pd->filt_off;
    const std::vector<int>& pl = h->pf ? h->pf->out_len : h->pd->out_len;
    if (filt_offs)
        for (size_t i = 0; i < fo.size(); ++i) filt_offs[i] = fo[i];
    if (path_len)
        for (size_t i = 0; i < pl.size(); ++i) path_len[i] = pl[i];
    return KYM_OK;
}

// mode 0 inference, 1 training forward (saves u, whole batch), 2 backward
template <class T>
static int run1d(const Scat1DPlan<T>& p, int mode, const void* x, const void* filt, void* out, const void* dS,
                 void* gx, long long batch, void* ws, long long ws_bytes, hipStream_t stream, uintptr_t base,
                 LaunchProfile* prof, ChunkStreams* cs = nullptr, int streams = 1) {
    const long long sig_elems = mode == 0 ? p.ws_sig : p.ws_sig_train;
    const long long per = sig_elems * (long long)sizeof(cpx<T>);
    long long chunk = ws_bytes / per;
    if (chunk < 1) return fail(KYM_ERR_VALUE, "workspace too small for one signal");
    if (chunk > batch) chunk = batch;
    if (mode != 0 && chunk < batch) return fail(KYM_ERR_VALUE, "training workspace must hold the whole batch");
    const std::vector<Launch1D>& launches = mode == 2 ? p.bwd_launches : p.launches;
    // inference chunks rotate over the plan's streams as in run2d
    const int ns = chunk_streams(cs, streams, mode, batch, &chunk);
    int rc = chunk_fork(cs, ns, stream);
    if (rc != KYM_OK) return rc;
    int k = 0;
    for (long long i0 = 0; i0 < batch; i0 += chunk, ++k) {
        const int nimg = (int)std::min(chunk, batch - i0);
        const int si = k % ns;
        const hipStream_t st = si ? cs->aux[si - 1] : stream;
        for (const Launch1D& l : launches) {
            Scat1DArgs<T> a = p.args(l, base);
            a.x = static_cast<const T*>(x);
            a.filt = static_cast<const T*>(filt);
            a.ws = reinterpret_cast<cpx<T>*>(static_cast<char*>(ws) + si * chunk * per);
            a.ws_sig = sig_elems;
            a.out = static_cast<T*>(out);
            a.dS = static_cast<const T*>(dS);
            a.gx = static_cast<T*>(gx);
            a.save = mode == 1 ? 1 : 0;
            a.img0 = (int)i0;
            const long long ntask = (long long)nimg * l.ngroup;
            if (ntask > (1ll << 30)) return fail(KYM_ERR_VALUE, "too many tasks in one launch");
            a.ntask = (int)ntask;
            const unsigned grid = (unsigned)((ntask + 7) / 8 * 8);
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (prof && prof->enabled) {
                e0 = prof->get();
                e1 = prof->get();
                if (e0) KYM_HIP_TRY(hipEventRecord(e0, st));
            }
            hipLaunchKernelGGL(pick_scat1d<T>(l.threads, l.kind, l.spec), dim3(grid), dim3(l.threads), l.lds_bytes, st, a);
            KYM_HIP_TRY(hipGetLastError());
            if (e0 && e1) {
                KYM_HIP_TRY(hipEventRecord(e1, st));
                prof->recs.push_back({(int)(&l - launches.data()), {e0, e1}});
            }
        }
    }
    return chunk_join(cs, ns, stream);
}

extern "C" int kym_scat1d_forward(void* plan, const void* x, const void* filt, void* out, long long batch, void* ws,
                                  long long ws_bytes, void* stream) {
    if (!plan) return fail(KYM_ERR_VALUE, "null plan");
    if (batch < 0) return fail(KYM_ERR_VALUE, "negative batch");
    if (batch == 0) return KYM_OK;
    if (!x || !filt || !out || !ws) return fail(KYM_ERR_VALUE, "null device pointer");
    KymScat1D* h = static_cast<KymScat1D*>(plan);
    int cur = 0;
    KYM_HIP_TRY(hipGetDevice(&cur));
    if (cur != h->device) return fail(KYM_ERR_TYPE, "plan was created for another device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->pf)
        return run1d(*h->pf, 0, x, filt, out, nullptr, nullptr, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, &h->prof,
                     &h->cs, h->streams);
    return run1d(*h->pd, 0, x, filt, out, nullptr, nullptr, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, &h->prof,
                 &h->cs, h->streams);
}

extern "C" int kym_scat1d_forward_train(void* plan, const void* x, const void* filt, void* out, long long batch,
                                        void* ws, long long ws_bytes, void* stream) {
    if (!plan) return fail(KYM_ERR_VALUE, "null plan");
    if (batch < 0) return fail(KYM_ERR_VALUE, "negative batch");
    if (batch == 0) return KYM_OK;
    if (!x || !filt || !out || !ws) return fail(KYM_ERR_VALUE, "null device pointer");
    KymScat1D* h = static_cast<KymScat1D*>(plan);
    int cur = 0;
    KYM_HIP_TRY(hipGetDevice(&cur));
    if (cur != h->device) return fail(KYM_ERR_TYPE, "plan was created for another device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->pf) return run1d(*h->pf, 1, x, filt, out, nullptr, nullptr, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, &h->prof);
    return run1d(*h->pd, 1, x, filt, out, nullptr, nullptr, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, &h->prof);
}

extern "C" int kym_scat1d_backward(void* plan, const void* dS, const void* filt, void* gx, long long batch, void* ws,
                                   long long ws_bytes, void* stream) {
    if (!plan) return fail(KYM_ERR_VALUE, "null plan");
    if (batch < 0) return fail(KYM_ERR_VALUE, "negative batch");
    if (batch == 0) return KYM_OK;
    if (!dS || !filt || !gx || !ws) return fail(KYM_ERR_VALUE, "null device pointer");
    KymScat1D* h = static_cast<KymScat1D*>(plan);
    int cur = 0;
    KYM_HIP_TRY(hipGetDevice(&cur));
    if (cur != h->device) return fail(KYM_ERR_TYPE, "plan was created for another device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->pf) return run1d(*h->pf, 2, nullptr, filt, nullptr, dS, gx, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, nullptr);
    return run1d(*h->pd, 2, nullptr, filt, nullptr, dS, gx, batch, ws, ws_bytes, s, (uintptr_t)h->dblob, nullptr);
}
