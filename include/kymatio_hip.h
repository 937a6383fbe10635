/* kymatio_hip.h -- C ABI of the MI355X scattering library (libkymatio_hip.so).
 *
 * Plain pointers and sizes only.  Device pointers are HIP device addresses
 * (e.g. torch.Tensor.data_ptr() of a tensor on the current device); `stream`
 * is a hipStream_t (torch.cuda.current_stream().cuda_stream).  Every call is
 * stream-ordered and never synchronises the host.  All functions return a
 * KymStatus code; kym_last_error() gives the message of the last failure on
 * the calling thread.  Status codes map to the reference's exception types:
 *   1 -> TypeError, 2 -> RuntimeError, 3 -> ValueError, 4 HIP failure,
 *   5 unsupported configuration.
 *
 * Ownership: the caller owns inputs, outputs, filters and workspace (torch
 * tensors); a plan owns its constant tables (twiddles, digit-reversal maps,
 * node tree) on the device that was current when it was created.  A plan is
 * immutable after creation, so calls on different streams may share it.
 */
#ifndef KYMATIO_HIP_H
#define KYMATIO_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define KYM_ABI_VERSION 1

/* dtype codes */
#define KYM_DTYPE_F32 0
#define KYM_DTYPE_F64 1

const char* kym_last_error(void);
int kym_version(void);

/* ---------------------------------------------------------------- 2D ----
 * Replaces the core call scattering2d(x, pad, unpad, backend, J, L, phi, psi,
 * max_order, out_type) made by ScatteringTorch2D.scattering
 * (kymatio/scattering2d/frontend/torch_frontend.py:98-99; core in
 * kymatio/scattering2d/core/scattering2d.py:1-88) with one fused pipeline:
 * reflect pad, fft2, filter multiply + periodisation, ifft2, modulus, fft2,
 * low-pass, ifft2, unpad, written directly into the stacked output.
 *
 * kym_scat2d_create: plan for input shape (M, N), scale J, L angles,
 *   max_order 1|2, pre_pad 0|1 (kymatio/scattering2d/frontend/
 *   torch_frontend.py:9-21; padding from base_frontend.py:19-32).
 *   lds_limit caps the LDS bytes a workgroup may use (0 = device maximum);
 *   smaller values force the multi-pass route and exist for testing.
 *   Fails with status 2 and the reference message 'The smallest dimension
 *   should be larger than 2^J.' (base_frontend.py:22-23).
 */
int kym_scat2d_create(int M, int N, int J, int L, int max_order, int pre_pad, int dtype,
                      long long lds_limit, void** plan);

/* kym_scat2d_create with tuning choices (the autotuner's knobs; no
 * counterpart in the reference): tune[0..ntune) = {largest
 * power-of-two FFT radix (16/8/4/2), LDS-node threads (256/512/1024),
 * big-node column block, big-node row block, big-node pass threads (a
 * compiled choice for these sizes and blocks), streams (1: the inference
 * forward runs its chunks one after the other on the caller's stream; S in
 * 2..4, default 2: the workspace is split in S parts and chunks rotate over
 * the caller's stream and S - 1 streams of the plan, joined back to the
 * caller's stream before kym_scat2d_forward's work ends; forward_train and
 * backward split the batch in S chunks the same way, each image's saved
 * state at its own offset of the training workspace), pair (1:
 * one LDS node per workgroup; 2: sibling nodes of one parent share a
 * workgroup and one read of the parent's spectrum wherever a pair kernel
 * for the size is compiled; 0, the default: where that is the default
 * kernel for the size, or the launch's deepest periodisation is 16 or
 * more; 0 and 2 also pair siblings with k = 2 or 4 in the big nodes'
 * gather + inverse column pass where it is compiled and two blocks fit
 * 64 KB of LDS, 1 does not)}; ntune <= 7; 0 = built-in heuristic.  Results
 * are the same for every choice up to FP32 rounding. */
int kym_scat2d_create_tuned(int M, int N, int J, int L, int max_order, int pre_pad, int dtype,
                            long long lds_limit, const int* tune, int ntune, void** plan);
int kym_scat2d_destroy(void* plan);

/* info[12] = {C, out_h, out_w, M_padded, N_padded, n_filters, filter_elems,
 *             workspace_bytes_per_image, n_launches, pad_top, pad_left,
 *             training_workspace_bytes_per_image (-1 if not differentiable)} */
int kym_scat2d_info(void* plan, long long* info);

/* Packed filter layout: filter i (the reference's buffer tensor{i}: phi
 * levels first, then each psi's levels in (j, theta) order,
 * kymatio/scattering2d/frontend/torch_frontend.py:28-43) occupies elements
 * [offs[i], offs[i+1]) of one contiguous real array of the plan's dtype,
 * of filter_elems elements in all (info[6]); the low-pass taps follow. */
int kym_scat2d_filter_offsets(void* plan, long long* offs);

/* Low-pass taps (FP32 plans).  The leaf of resolution r,
 * unpad(Re(ifft2(P_kl(U * phi.levels[r])))) of the reference
 * (kymatio/scattering2d/core/scattering2d.py:19-23, 42-46, 69-73), equals a
 * separable real filter-and-subsample of the field behind U, because
 * phi.levels[r] is separable: Re(ifft2(phi.levels[r]))[a][b] = g_r[a] h_r[b].
 * desc (4 + 10 J entries, n its capacity) = {taps_off, taps_len, M_padded,
 * N_padded}, then per r < J and axis (y, then x):
 * {off, lo, cnt, band_off, kmax}.  For the axis of n = M_padded / 2^r (y) or
 * N_padded / 2^r (x) samples, with c = g_r (y) or h_r (x), the caller fills
 * packed[off + i] = c[(lo + i) mod n] for i < cnt, and the band table
 * packed[band_off + 16 k + d] = packed[off + t] with t = (k - 2^(J-r) d) mod n
 * when t < cnt, else 0, for k < kmax, d < 16.  taps_len = 0 (FP64 plans)
 * means no taps.  No counterpart in the reference (derived filter data). */
int kym_scat2d_leaf_taps(void* plan, long long* desc, int n);

/* x: (batch, M, N) real (or padded size with pre_pad); filt: packed filters;
 * out: (batch, C, out_h, out_w); ws: workspace of ws_bytes (any multiple of
 * workspace_bytes_per_image >= 1 image; the batch is processed in chunks of
 * ws_bytes / workspace_bytes_per_image images). */
int kym_scat2d_forward(void* plan, const void* x, const void* filt, void* out, long long batch,
                       void* ws, long long ws_bytes, void* stream);

/* kym_scat2d_forward without a workspace argument and without allocating:
 * no counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The workspace is a static device array of
 * KYM_NOALLOC_SCRATCH_BYTES, one copy per GPU, held until the process ends;
 * the batch runs in chunks of KYM_NOALLOC_SCRATCH_BYTES /
 * workspace_bytes_per_image images (info[7]).  Status 3 if one image needs
 * more than KYM_NOALLOC_SCRATCH_BYTES.  Every call on a GPU shares that GPU's
 * array, so calls that overlap in time on one GPU (other host threads, other
 * streams) race; calls ordered on one stream do not.  No part of the array
 * needs to start at zero (every element a kernel reads was written earlier in
 * the same call).  Results equal kym_scat2d_forward's with a workspace of
 * KYM_NOALLOC_SCRATCH_BYTES. */
#define KYM_NOALLOC_SCRATCH_BYTES (64ll << 20)
int kym_scat2d_forward_noalloc(void* plan, const void* x, const void* filt, void* out, long long batch,
                               void* stream);

/* Differentiation (replaces torch autograd through the reference's ops:
 * ModulusStable.backward, kymatio/backend/torch_backend.py:63-96, and the
 * autograd of fft2/ifft2, mean, the real filter product, ReflectionPad2d and
 * the unpad slice).  forward_train is kym_scat2d_forward that also keeps
 * every node's pre-modulus field in `ws`, which must hold the whole batch
 * (training_workspace_bytes_per_image * batch).  backward then reads dS
 * (= dL/dS, layout of out) and that workspace and writes gx = dL/dx
 * (layout of x); it must be given the same ws, untouched in between. */
int kym_scat2d_forward_train(void* plan, const void* x, const void* filt, void* out, long long batch,
                             void* ws, long long ws_bytes, void* stream);
int kym_scat2d_backward(void* plan, const void* dS, const void* filt, void* gx, long long batch, void* ws,
                        long long ws_bytes, void* stream);

/* Launch list introspection: rec[7] = {kind, nodes, tasks_per_node, threads,
 * lds_bytes, model_bytes_per_image, resolution}.  model_bytes is the
 * algorithmic HBM traffic the launch accounts for per image under the
 * stage-boundary byte model (each stage-crossing tensor written and read
 * once); it feeds the roofline figure of the benchmark.  No counterpart in
 * the reference (measurement support). */
int kym_scat2d_launch_info(void* plan, int index, long long* rec);

/* Mixed-size batches (no counterpart in the reference, whose modules take one
 * image size each): one executable HIP graph over the inference forwards of
 * nplans plans (one per image size) bound to fixed device buffers: x[i],
 * filt[i], out[i], batch[i] images and a workspace ws[i] of ws_bytes[i] per
 * plan, as kym_scat2d_forward takes them.  Each plan's launches form one
 * dependency chain; the chains are independent branches of the graph, so
 * sizes run concurrently.  graph_launch replays it on a stream; the plans
 * and buffers must outlive the graph. */
int kym_scat2d_graph_create(void* const* plans, int nplans, const void* const* x, const void* const* filt,
                            void* const* out, const long long* batch, void* const* ws, const long long* ws_bytes,
                            void** graph);
/* The same graph laid out for at most max_batch[i] images per plan, whose
 * kernels read the actual image count of plan i from dev_counts[i] (a device
 * array of nplans ints) at every replay: counts 0..max_batch[i] replay the
 * one graph without rebuilding it (write the counts, e.g. with a copy on the
 * replay's stream, before graph_launch; a count above max_batch[i] is
 * clamped to it).  Images i < count of plan i are read from x[i] and written
 * to out[i]; chunks of the layout past the count exit at once. */
int kym_scat2d_graph_create_counted(void* const* plans, int nplans, const void* const* x, const void* const* filt,
                                    void* const* out, const long long* max_batch, void* const* ws,
                                    const long long* ws_bytes, const int* dev_counts, void** graph);
int kym_scat2d_graph_launch(void* graph, void* stream);
int kym_scat2d_graph_destroy(void* graph);

/* Per-launch HIP-event timing of subsequent kym_scat2d_forward calls on this
 * plan (enable 1/0; enabling clears earlier records).  profile_read waits
 * for the recorded events and returns, per launch index i < n, the summed
 * milliseconds and the number of launches.  Not thread-safe; measurement
 * support with no counterpart in the reference. */
int kym_scat2d_profile(void* plan, int enable);
int kym_scat2d_profile_read(void* plan, double* ms, long long* count, int n);

/* ---------------------------------------------------------------- 1D ----
 * Replaces the generator call scattering1d(U_0, backend, filters,
 * log2_stride, average_local) and the per-path unpad / stack of
 * ScatteringBase1D.scattering (kymatio/scattering1d/frontend/
 * base_frontend.py:106-155; core in kymatio/scattering1d/core/
 * scattering1d.py:2-107) with one fused pipeline.
 *
 * kym_scat1d_create: the frontend's setup (base_frontend.py:37-98) passed
 * in: input length N, padded length N_pad (power of two), pad_left,
 * log2_stride, average_mode (0 local 'T', 1 none 'T=0', 2 'global'),
 * oversampling, max_order 1|2, the first/second order filters' dyadic
 * subsampling j and level counts, phi's level count and the border indices
 * ind_start/ind_end[0..n_ind).  Status 5 if N_pad does not fit one
 * workgroup's LDS for the dtype.
 */
int kym_scat1d_create(int N, int N_pad, int pad_left, int log2_stride, int average_mode, int oversampling,
                      int max_order, int n_psi1, const int* psi1_j, const int* psi1_nlev, int n_psi2,
                      const int* psi2_j, const int* psi2_nlev, int phi_nlev, const int* ind_start,
                      const int* ind_end, int n_ind, int dtype, long long lds_limit, void** plan);
int kym_scat1d_destroy(void* plan);

/* info[7] = {n_paths, out_row_elems, n_filters, filter_elems,
 *            workspace_bytes_per_signal, n_launches,
 *            training_workspace_bytes_per_signal} */
int kym_scat1d_info(void* plan, long long* info);

/* filt_offs[n_filters + 1]: packed filter layout in the reference's buffer
 * order (phi levels, psi1 levels, psi2 levels; kymatio/scattering1d/frontend/
 * torch_frontend.py:30-56).  path_len[n_paths]: output length of each path in
 * output order (sorted by (order, n) as base_frontend.py:147-150). */
int kym_scat1d_layout(void* plan, long long* filt_offs, long long* path_len);

/* x: (batch, N) real; out: (batch, out_row) with path c at
 * [sum(path_len[:c]), ...); ws: workspace (>= 1 signal).  Chunks of the
 * batch rotate over the caller's stream and the plan's streams as in
 * kym_scat2d_forward (kym_scat1d_set_streams); the work ends on `stream`. */
int kym_scat1d_forward(void* plan, const void* x, const void* filt, void* out, long long batch, void* ws,
                       long long ws_bytes, void* stream);

/* Inference chunks in flight for kym_scat1d_forward (no counterpart in the
 * reference; a scheduling choice): 1 = one after the other on the caller's
 * stream, S in 2..4 = the workspace split in S parts and chunks rotating over
 * the caller's stream and S - 1 streams of the plan; 0 = default 2.  Results
 * do not depend on it. */
int kym_scat1d_set_streams(void* plan, int streams);

/* Differentiation of the 1D path, as kym_scat2d_forward_train /
 * kym_scat2d_backward: the training forward keeps every node's pre-modulus
 * field in ws (training_workspace_bytes_per_signal * batch); backward reads
 * dS (layout of out) and writes gx (layout of x).  Covers all three
 * averaging modes (the un-averaged and global outputs feed the modulus
 * gradient directly). */
int kym_scat1d_forward_train(void* plan, const void* x, const void* filt, void* out, long long batch, void* ws,
                             long long ws_bytes, void* stream);
int kym_scat1d_backward(void* plan, const void* dS, const void* filt, void* gx, long long batch, void* ws,
                        long long ws_bytes, void* stream);

/* ---------------------------------------------------------- per-op ----
 * The backend namespace as single kernels, for the backend plugin route and
 * op-level tests (forward only).  Complex arrays are interleaved (re, im)
 * pairs, i.e. the reference's (..., 2) layout; real arrays its (..., 1).
 */
/* Pad (kymatio/scattering2d/backend/torch_backend.py:19-86): reflect pad of
 * (batch, M, N) to (batch, M+top+bottom, N+left+right), numpy 'reflect'
 * semantics (so pad == size matches the reference's special case). */
int kym_op_pad2d(const void* x, void* y, long long batch, int M, int N, int top, int bottom, int left, int right,
                 int dtype, void* stream);
/* pad (kymatio/scattering1d/backend/torch_backend.py:50-82); status 3
 * 'Indefinite padding size (larger than tensor).' when a pad >= N. */
int kym_op_pad1d(const void* x, void* y, long long batch, int N, int left, int right, int dtype, void* stream);
/* cdgmm (kymatio/backend/torch_backend.py:147-219): C = A * B, A complex
 * (batch, n), B real (n) or complex (n) when b_complex. */
int kym_op_cdgmm(const void* A, const void* B, void* C, long long batch, long long n, int b_complex, int dtype,
                 void* stream);
/* modulus (kymatio/backend/torch_backend.py:137-141): y = |x|, n complex in, n real out. */
int kym_op_modulus(const void* x, void* y, long long n, int dtype, void* stream);
/* subsample_fourier (kymatio/scattering2d/backend/torch_backend.py:92-129;
 * 1D: kymatio/scattering1d/backend/torch_backend.py:18-48 with W = 1, kw = 1):
 * y[b, i, j] = mean over the kh x kw aliases of x (batch, H, W) complex. */
int kym_op_subsample(const void* x, void* y, long long batch, int H, int W, int kh, int kw, int dtype, void* stream);
/* unpad (kymatio/scattering2d/backend/torch_backend.py:157-176): y = x[..., 1:-1, 1:-1]. */
int kym_op_unpad2d(const void* x, void* y, long long batch, int H, int W, int dtype, void* stream);
/* fft2 / ifft2 / rfft / irfft (kymatio/scattering2d/backend/torch_backend.py:
 * 133-155; 1D with W = 1, kymatio/scattering1d/backend/torch_backend.py:
 * 108-141) over (batch, H, W): kind 0 forward c2c, 1 inverse c2c (1/(HW)),
 * 2 forward from real input, 3 inverse keeping the real part.  lds_limit as
 * in kym_scat2d_create (0 = device maximum).  work: batch * H * W complex
 * scratch, needed only by kind 3 when a transform does not fit in LDS
 * (status 3 if it is then null); may be null otherwise. */
int kym_op_fft(const void* x, void* y, void* work, long long batch, int H, int W, int kind, int dtype,
               long long lds_limit, void* stream);
/* The filter-periodise-invert step of the 1D and time-frequency paths in one
 * kernel: y[b] = ifft_n(subsample_fourier(cdgmm(x[b], f), k)), n = N / k,
 * for x (batch, N) complex and a real filter f (N) -- the reference's
 * ifft(subsample_fourier(cdgmm(...))) and irfft(...) of
 * kymatio/scattering1d/backend/torch_backend.py:18-48, 108-141 with
 * kymatio/backend/torch_backend.py:147-219, as used by
 * kymatio/scattering1d/core/timefrequency_scattering.py.  mode 0: y (batch,
 * n) complex; 1: its real part (irfft); 2: its modulus, y (batch, n) real.
 * Status 2 if k does not divide N. */
int kym_op_filter_ifft(const void* x, const void* f, void* y, long long batch, long long N, int k, int mode, int dtype,
                       void* stream);
/* kym_op_filter_ifft with the output transposed: the batch rows are (b, t),
 * t < tlen, and y is laid out (batch / tlen, n, tlen) -- the ifft and
 * swap_time_frequency of kymatio/scattering1d/core/
 * timefrequency_scattering.py:266-267 (backend/torch_backend.py:172-186)
 * in one kernel; with mode 2 also the modulus that
 * timefrequency_scattering.py:409 applies next.  tlen = 0 is
 * kym_op_filter_ifft.  Status 3 if tlen does not divide batch. */
int kym_op_filter_ifft_t(const void* x, const void* f, void* y, long long batch, long long N, int k, int mode,
                         long long tlen, int dtype, void* stream);
/* kym_op_filter_ifft over grouped rows: nf > 1 filters the x rows with a
 * bank of nf filters f (nf, N) -- row r is x row r / nf with filter r mod nf,
 * y (rows, n) -- the per-n1 filters of the first order of
 * kymatio/scattering1d/core/timefrequency_scattering.py:133-153 for all n1 of
 * one j1 at once; grp > 0 writes row r to output row (r / grp) gtot + goff +
 * r mod grp, so a group's (B, grp) rows land in the (B, gtot) stack that
 * timefrequency_scattering.py:157 and :184 build with backend.stack.
 * Status 3 if nf does not divide rows or the row map does not fit. */
int kym_op_filter_ifft_g(const void* x, const void* f, void* y, long long rows, long long N, int k, int mode, int nf,
                         int grp, int gtot, int goff, int dtype, void* stream);
/* Frequency scattering of one input in one kernel: x (B, A, T) real (c = 1)
 * or complex (c = 2) along the log-frequency axis a < A, zero-padded to N;
 * for each of nf <= 16 wavelets f[i] (N reals, host array of device
 * pointers) y[i] (B, N / ks[i], T) real = |ifft(subsample_fourier(cdgmm(
 * fft(x_padded), f[i]), ks[i]))| along that axis -- the swap_time_frequency,
 * pad_frequency, cfft / rfft, and per wavelet ifft and swap_time_frequency
 * of kymatio/scattering1d/core/timefrequency_scattering.py:244-267, with the
 * modulus that timefrequency_scattering.py:409 applies next, reading x once.
 * Status 2 if a ks[i] does not divide N. */
int kym_op_freq_scatter(const void* x, void* const* ys, const void* const* fs, const int* ks, int nf, long long B,
                        int A, int T, int c, int N, int dtype, void* stream);
/* Low-pass of real lines in one kernel: y = Re ifft_n(subsample_fourier(
 * cdgmm(fft_N(x), f), k)), n = N / k -- rfft, cdgmm, subsample_fourier and
 * irfft of kymatio/scattering1d/backend/torch_backend.py:118-136 with
 * kymatio/backend/torch_backend.py:148 as time_averaging and
 * frequency_averaging apply them (kymatio/scattering1d/core/
 * timefrequency_scattering.py:278-373).  x is
 * (outer, N, S) real, y (outer, n, S) real: the lines run along the middle
 * axis, so frequency_averaging's two swap_time_frequency are not needed
 * (S = 1: contiguous lines).  Status 2 if k does not divide N. */
int kym_op_lowpass(const void* x, const void* f, void* y, long long outer, long long N, long long S, int k, int dtype,
                   void* stream);
/* modulus_rotation (kymatio/scattering3d/backend/torch_backend.py:97-124):
 * y = sqrt(prev^2 + |x|^2) elementwise over n complex x; prev may be null
 * (the first m of a solid-harmonic order l). */
int kym_op_modulus_rotation(const void* x, const void* prev, void* y, long long n, int dtype, void* stream);
/* All m of one (l, j) of the 3D solid-harmonic loop at once
 * (kymatio/scattering3d/core/scattering3d.py:31-35, 51-55, whose m loop calls
 * cdgmm3d, ifft and modulus_rotation once per m).
 * harmonic_products: out (batch, nm, n) complex = U (batch, n) complex times
 *   F (nm, n) complex, the cdgmm3d(U, F[m]) of every m.
 * modulus_rotation_sum: y[b * ystride + i] for (batch, nm, n) complex x, the
 *   result of the reference's nm chained modulus_rotation calls
 *   (kymatio/scattering3d/backend/torch_backend.py:97-124):
 *   p = sqrt(p^2 + |x_m|^2) from p = 0; nm = 1 is the plain modulus
 *   (rotation_covariant=False, scattering3d.py:36-39). */
int kym_op_harmonic_products(const void* U, const void* F, void* out, long long batch, int nm, long long n, int dtype,
                             void* stream);
int kym_op_modulus_rotation_sum(const void* x, void* y, long long batch, int nm, long long n, long long ystride,
                                int dtype, void* stream);
/* compute_integrals (kymatio/scattering3d/backend/torch_backend.py:126-151):
 * out[b, q] = sum over the n elements of x[b] of x^powers[q], q < npow <= 8
 * (accumulated in FP64, stored in the dtype). */
int kym_op_integrals(const void* x, void* out, long long batch, long long n, const double* powers, int npow, int dtype,
                     void* stream);

/* ------------------------------------------------------ learned part ----
 * Kernels of the warm-start decoder (kymatio_amd/learned).  They have no
 * counterpart in the reference, whose closest precedents are the learned
 * inverse and the descent reconstruction examples
 * (examples/2d/regularized_inverse_scattering_MNIST_torch.py:38-131,
 * examples/1d/reconstruct_torch.py:77-129).  FP32 tensors unless noted.
 *
 * conv_fwd: out (B, Cout, T) = conv1d_k9_same(scale * upsample_up(in)) + bias,
 *   optional ReLU; in (B, Cin, T/up); w bf16 (Cout_pad, 9, Cin_pad);
 *   T % 64 == 0, Cin_pad % 16 == 0, Cout_pad % 16 == 0 <= 128.  Also the
 *   backward-data pass when given the transposed, flipped weights.
 * conv_bwd_weight: dw (Cout, 9, Cin) += dW of that conv (atomics; zero it
 *   first); Cin_pad % 64 == 0; ktiles = 64-sample time tiles per workgroup.
 * up_relu_adjoint: din (B, C, Tin) = sum of up fine positions of dx,
 *   masked by act > 0 when mask.
 * bias_grad: db (C) += sum over (b, t) of dout (B, C, T) (atomics).
 * mse: loss += mean((y - x)^2), dy = 2 (y - x) / n (dy may be null).
 * adam: Adam step (bias-corrected, step >= 1) on n parameters.
 * pack_weights: FP32 (Cout, 9, Cin) -> bf16 forward (Cout_pad, 9, Cin_pad)
 *   and backward-data (Cin_pad, 9, Cout_pad, taps reversed) layouts (wt may
 *   be null).
 * chirps: B signals of N samples, 1-8 linear chirps + noise * N(0,1), a
 *   deterministic function of (seed, index0 + signal index).
 * mfma_layout_check: writes the 16x16 products of fixed integer matrices
 *   through the bf16 MFMA builtins (K = 16 and K = 32) for a layout test. */
int kym_dec_conv_fwd(const void* in, const void* w, const void* bias, const void* scale, void* out, int B, int Cin,
                     int Cin_pad, int Cout, int Cout_pad, int T, int up, int relu, void* stream);
int kym_dec_conv_bwd_weight(const void* dout, const void* in, const void* scale, void* dw, int B, int Cin,
                            int Cin_pad, int Cout, int Cout_pad, int T, int up, int ktiles, void* stream);
int kym_dec_up_relu_adjoint(const void* dx, const void* act, void* din, int B, int C, int Tin, int up, int mask,
                            void* stream);
int kym_dec_bias_grad(const void* dout, void* db, int B, int C, int T, void* stream);
int kym_dec_mse(const void* y, const void* x, void* dy, void* loss, long long n, void* stream);
int kym_dec_adam(void* p, const void* g, void* m, void* v, long long n, float lr, float b1, float b2, float eps,
                 int step, void* stream);
int kym_dec_pack_weights(const void* w, void* wf, void* wt, int Cout, int Cout_pad, int Cin, int Cin_pad,
                         void* stream);
int kym_chirps(void* x, int B, int N, unsigned long long seed, long long index0, float noise, void* stream);
int kym_mfma_layout_check(void* out16, void* out32, void* stream);

#ifdef __cplusplus
}
#endif
#endif
