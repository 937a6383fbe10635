// CPU test harness (libkymatio_emu.so, built host-only): runs the GPU
// workgroup bodies on the host with one "thread" per workgroup, launch by
// launch, so that the plan, the index arithmetic, the digit-reversal maps
// and the twiddle tables can be checked against the oracle on a machine
// without a GPU.  It is loaded only by the tests; the product library
// (libkymatio_hip.so) contains no CPU compute path.
#include <cstdlib>
#include <string>
#include <vector>

#include "scat2d_plan.h"

static thread_local std::string g_emu_error;
void kym_set_error(const std::string& msg) { g_emu_error = msg; }

extern "C" const char* kym_emu_last_error(void) { return g_emu_error.c_str(); }

template <class T>
static int emu_forward2d(int M, int N, int J, int L, int max_order, int pre_pad, long long lds_limit, const void* x,
                         const void* filt, void* out, long long batch, long long chunk, long long* info) {
    Scat2DPlan<T> p;
    KymError e = {0, ""};
    if (!p.build(M, N, J, L, max_order, pre_pad, (size_t)lds_limit, &e)) {
        g_emu_error = e.msg;
        return e.code;
    }
    p.blob.relocate((uintptr_t)p.blob.bytes.data());
    const uintptr_t base = (uintptr_t)p.blob.bytes.data();
    if (info) {
        info[0] = p.C; info[1] = p.oh; info[2] = p.ow; info[3] = (long long)p.launches.size();
        int nbig = 0;
        for (auto& nd : p.nodes) nbig += nd.big;
        info[4] = nbig;
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
            for (int t = 0; t < a.ntask; ++t) dispatch_body2d<T, 31>(l.kind, a, t, lds.data(), 0, 1);
        }
    }
    return 0;
}

extern "C" int kym_emu_scat2d_forward(int M, int N, int J, int L, int max_order, int pre_pad, int dtype,
                                      long long lds_limit, const void* x, const void* filt, void* out,
                                      long long batch, long long chunk, long long* info) {
    if (dtype == 0)
        return emu_forward2d<float>(M, N, J, L, max_order, pre_pad, lds_limit, x, filt, out, batch, chunk, info);
    return emu_forward2d<double>(M, N, J, L, max_order, pre_pad, lds_limit, x, filt, out, batch, chunk, info);
}
