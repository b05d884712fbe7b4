/*
 * klt.h -- C ABI of the MI355X imaging library (libklt.so).
 *
 * The library is a drop-in for K-Litho's imaging hot path.  K-Litho has no
 * library API of its own: its two tools call internal C++ functions that take
 * std::vector buffers.  Each compute entry point below replaces one of those
 * functions and is cited against it; the klitho_tcc / klitho_socs tools in
 * this package call these entry points in place of the reference's ones.
 *
 * Conventions
 *   - Every function returns KLT_OK (0) or a negative KLT_ERR_* code;
 *     klt_last_error() gives the message of the calling thread's last error.
 *   - Arrays are row-major.  Complex arrays are interleaved (re, im).
 *   - Array arguments are device pointers unless KLT_HOST_PTRS is passed in
 *     `flags`; then they are host pointers and the library stages them.
 *   - A context owns one HIP stream (or a caller-provided one), device
 *     buffers, twiddle tables and plans.  One context per host thread.
 *   - Frequency index of a kernel / TCC entry: ID(nx,ny) =
 *     (ny+Ny)*(2Nx+1) + (nx+Nx), |nx|<=Nx, |ny|<=Ny
 *     (K-Litho-TCC/klitho_tcc.cpp:221,245).
 */
#ifndef KLT_H
#define KLT_H

#ifdef __cplusplus
extern "C" {
#endif

#define KLT_OK 0
#define KLT_ERR_ARG -1      /* invalid argument (reference: exit 1 on bad argv) */
#define KLT_ERR_SIZE -2     /* 4N+1 > L (reference: "ERROR in FI()", klitho_socs.cpp:166-169) */
#define KLT_ERR_HIP -3      /* HIP runtime failure */
#define KLT_ERR_NODEV -4    /* no usable GPU */
#define KLT_ERR_IO -5       /* file could not be opened / read / written */
#define KLT_ERR_NOMEM -6    /* allocation failure (reference: bad_alloc, klitho_tcc.cpp:368-371) */
#define KLT_ERR_EIG -7      /* eigensolver did not converge (reference: zheevr info != 0) */

#define KLT_HOST_PTRS 1u    /* array arguments are host memory */
#define KLT_EIG_PSD 2u      /* eigensolver input is positive semidefinite (see klt_eig_topk) */
#define KLT_EIG_WARM 4u     /* with KLT_EIG_PSD: start from the previous solve's block (see klt_eig_topk) */

typedef struct klt_ctx klt_ctx;
typedef struct klt_kernels klt_kernels;

/* ---- context --------------------------------------------------------- */
const char* klt_last_error(void);
int klt_version(void);
/* Number of visible GPUs (0 when none). */
int klt_device_count(void);
int klt_ctx_create(int device, klt_ctx** out);
void klt_ctx_destroy(klt_ctx* ctx);
/* Use `stream` (a hipStream_t) for all work of this context; NULL restores
 * the context's own stream.  The caller keeps ownership.  Work already queued
 * by the context on the previous stream comes first: the new stream waits for
 * it (an event, no host wait), since both use the context's buffers.  The
 * previous stream must still exist; if it is being captured, no such wait is
 * added (the capture orders its own work). */
int klt_ctx_set_stream(klt_ctx* ctx, void* stream);
void* klt_ctx_stream(klt_ctx* ctx);
int klt_ctx_sync(klt_ctx* ctx);
/* Stream ordering with a caller's stream (a hipStream_t; NULL is the null
 * stream), without a host wait (no reference counterpart: the reference is
 * single-threaded host code).  klt_ctx_wait_stream: later work of the context
 * starts after everything queued so far on `stream`, e.g. the kernels that
 * produced its device inputs.  klt_ctx_join_stream: later work on `stream`
 * starts after everything queued so far on the context's stream, e.g. the
 * consumers of its device outputs. */
int klt_ctx_wait_stream(klt_ctx* ctx, void* stream);
int klt_ctx_join_stream(klt_ctx* ctx, void* stream);
/* Per-kernel timing (no reference counterpart; the reference prints only a
 * whole-run wall clock, klitho_socs.cpp:355-361).  While on, every kernel
 * launch is bracketed by HIP events on the context stream. */
int klt_ctx_set_profiling(klt_ctx* ctx, int on);
/* Writes one "name calls total_ms" line per kernel launched since the last
 * report and clears the record; returns the byte count or an error. */
int klt_ctx_profile_report(klt_ctx* ctx, char* buf, int len);

/* Launch-shape tuning of the batched SOCS passes (no reference counterpart).
 * klt_socs_forward and klt_socs_backward split their kernels over more blocks
 * when a batch is too small to fill the GPU: column ranges of the band pass,
 * row ranges of the synthesis pass, groups of kernels in the coarse-grid
 * stage.  With tuning on (the default) the first call of each (batch, size,
 * kernel set) runs the candidate splits once each, timed with HIP events,
 * and keeps the fastest for later calls; off, a fixed rule decides. */
int klt_ctx_set_autotune(klt_ctx* ctx, int on);
/* Pins the split factors of later klt_socs_forward / klt_socs_backward calls
 * (no reference counterpart): ks column ranges of the band pass (at most 8),
 * rs row ranges of the synthesis pass (at most 16), kg kernel groups of the
 * coarse-grid stage; 0 leaves a factor to the tuner or the fixed rule.  A
 * factor that does not fit a launch shape is lowered to the largest that
 * does.  Pinned shapes are not timed and not recorded in the tune report. */
int klt_ctx_set_splits(klt_ctx* ctx, int ks, int rs, int kg);
/* Pins the tile height of the mask band pass of later klt_socs_forward /
 * klt_socs_backward calls: 64 or 128 rows per block (0: tuned, or 128 under
 * the fixed rule).  Like klt_ctx_set_splits, it has no reference counterpart:
 * it exposes a choice of the autotuner. */
int klt_ctx_set_band_rows(klt_ctx* ctx, int rows);
/* One line per tuned shape: "socs_forward|socs_backward|socs_vjp B Lx Ly Nx
 * Ny nk column_splits row_splits kernel_groups band_rows ms"; returns the
 * byte count or an error. */
int klt_ctx_tune_report(klt_ctx* ctx, char* buf, int len);

/* ---- host boundary: sources, masks, sizing, file formats ---------------
 * Host-only; bit-identical to the reference's float arithmetic. */
/* K-Litho-TCC/source.cpp:105-142 (createSource): srcType is "Annular",
 * "Dipole", "CrossQuadrupole", "Point" or "Import" (sourceFile: ".bin" raw
 * float32, otherwise text); the result is normalised to sum 1. */
int klt_source_create(float* src, int srcSize, const char* srcType, float innerR,
                      float outerR, float radius, float offset, int onXAxis, float ptX,
                      float ptY, const char* sourceFile);
/* common/mask.cpp:31-60 (createMask): "LineSpace" or "Import", centred in the
 * Lx x Ly period, times dose. */
int klt_mask_create(float* msk, int Lx, int Ly, int maskSizeX, int maskSizeY,
                    const char* maskType, int lineWidth, int spaceWidth, int isHorizontal,
                    const char* maskFile, float dose);
/* K-Litho-TCC/klitho_tcc.cpp:52-67 (getOuterSigma), not clamped. */
float klt_outer_sigma(const float* src, int srcSize);
/* K-Litho-TCC/klitho_tcc.cpp:479-486: clamps sigma to 1 and sets
 * N = floor(f32(NA) * L * (1 + sigma) / lambda). */
int klt_tcc_sizing(float NA, int Lx, int Ly, float outerSigma, float lambda, int* Nx, int* Ny);
/* common/file_io.cpp:232-260 / 204-230 (readFloatArrayFrom{Binary,Txt}):
 * reader chosen by extension; a size mismatch warns and zero-pads/truncates. */
int klt_read_floats(const char* path, float* out, int size);
/* common/file_io.cpp:137-150 (writeFloatArrayToBinary). */
int klt_write_floats(const char* path, const float* data, int size);
/* common/file_io.cpp:262-280 (writeKernelInfo). */
int klt_write_kernel_info(const char* path, int sizeX, int sizeY, int nk, const float* scales);
/* common/file_io.cpp:282-307 (readKernelInfo): returns nk (>= 0) or an error. */
int klt_read_kernel_info(const char* path, int* sizeX, int* sizeY, float* scales, int cap);

/* ---- PNG views (common/file_io.cpp:7-135) --------------------------------
 * The reference's three PNG writers, pixel for pixel: 8-bit grayscale of
 * clamp((v - min) / (max - min), 0, 1) * 255 (:7-23, source.png and mask.png),
 * the 7-colour partition of that value into [0, 1/6), ..., [5/6, 1), {1}
 * (:25-80, image_s.png with min 0 and max TARGET_INTENSITY), and the
 * continuous blue-aqua-green-yellow-orange-red map (:82-135, image.png,
 * tcc_{r,i}.png, the kernel PNGs); min == max gives uniform green in the colour
 * maps, as in the reference.  The files use stored (uncompressed) deflate
 * blocks.  data: sy rows of sx values.  KLT_ERR_IO ("Error: Failed to write
 * PNG: ...") when the file cannot be written. */
int klt_write_png_gray(const char* path, int sx, int sy, const float* data, float min, float max);
int klt_write_png_7colors(const char* path, int sx, int sy, const float* data, float min,
                          float max);
int klt_write_png_continuous(const char* path, int sx, int sy, const float* data, float min,
                             float max);

/* ---- TCC build (K-Litho-TCC/klitho_tcc.cpp:187-267, calcTCC) -----------
 * TCC[i][j] = sum_s S_s P(i,s) conj(P(j,s)) over the source points the
 * reference uses (S != 0, inside the unit disc, |p|,|q| <= ceil(sh*sigma)).
 * Pupil membership and phase follow the reference's float32 arithmetic.
 * nmed is the image-side immersion index: 1 reproduces the reference's
 * sqrt(1 - rho^2 NA^2) defocus phase; NA may not exceed nmed.  Zero defocus
 * gives a unit pupil (the reference's NaN for NA*rho > 1 is avoided).
 * tcc: T x T complex double (T = (2Nx+1)(2Ny+1)), both triangles. */
#define KLT_TCC_FP64 0       /* FP64 Gram, accumulated like the reference */
#define KLT_TCC_SPLIT_F16 1  /* split hi+lo FP16 MFMA Gram, FP32 accumulation */
int klt_tcc_build(klt_ctx* ctx, const float* src, int srcSize, float outerSigma, float lambda,
                  float NA, float defocus, float nmed, int Lx, int Ly, int Nx, int Ny,
                  int precision, double* tcc, unsigned flags);

/* ---- eigensolver (klitho_tcc.cpp:301-377, calcKernels / zheevr_) --------
 * FP64 on the GPU: Householder tridiagonalisation, bisection for the k
 * largest eigenvalues, inverse iteration with re-orthogonalisation inside
 * clusters, back-transformation.  a: n x n complex double Hermitian (full
 * storage; like zheevr_ with uplo='U' the result depends only on the upper
 * triangle for an exactly Hermitian input).  w: k eigenvalues, descending;
 * vecs: k x n complex, row q the unit eigenvector of w[q] (phase arbitrary,
 * as in LAPACK).  With KLT_EIG_PSD in flags the caller states that the matrix
 * is positive semidefinite, as every TCC of a non-negative source is
 * (TCC = P^H diag(S) P); the solver then takes the k <= 48 largest eigenpairs
 * by block subspace iteration (64 columns, rounds of X <- orth(A^2 X) with
 * CholQR2, Rayleigh-Ritz) to residuals of 5e-14 of the largest eigenvalue, and
 * uses the Householder path for matrices it does not reach (n < 128, k > 48,
 * no convergence in 24 rounds).  The flag applies to klt_eig_topk_batch and
 * klt_calc_kernels as well.  With KLT_EIG_WARM besides (klt_eig_topk and
 * klt_calc_kernels), the iteration starts from the 64-column block that the
 * previous positive-semidefinite solve of a matrix of the same order left on
 * this context, and tests convergence from its fifth round on: for a
 * sequence of slowly changing matrices, such as the TCCs of a source
 * optimisation.  The result meets the same residual test; without such a
 * previous solve the flag is ignored. */
int klt_eig_topk(klt_ctx* ctx, const double* a, int n, int k, double* w, double* vecs,
                 unsigned flags);
/* A batch of `count` Hermitian matrices of the same size in one call (no
 * reference counterpart; the reference runs zheevr once per TCC,
 * klitho_tcc.cpp:301-377): a [count][n][n] complex, w [count][k] and vecs
 * [count][k][n] complex, each matrix's results equal to klt_eig_topk's.  Every
 * kernel launch of the solve covers the whole batch, so the per-launch latency
 * that bounds a single solve is shared by the kernel sets of several defocus
 * or dose conditions. */
int klt_eig_topk_batch(klt_ctx* ctx, const double* a, int count, int n, int k, double* w,
                       double* vecs, unsigned flags);
/* calcKernels: the nk (clamped to T) kernels of the largest eigenvalues in the
 * krn_i_{r,i}.bin layout, krn[q][j] = z_q[T-1-j] as float (klitho_tcc.cpp:
 * 358-362), and the eigenvalues as float (:344-347).  krn_r, krn_i: nk x T. */
int klt_calc_kernels(klt_ctx* ctx, const double* tcc, int T, int nk, float* krn_r,
                     float* krn_i, float* scales, unsigned flags);

/* ---- Hopkins image (klitho_tcc.cpp:534-544: calcImage + FT_c2r) ----------
 * The image klitho_tcc writes to image.bin, from the full TCC:
 * H[g] = sum_n1 M^[n1+g] conj(M^[n1]) TCC[ID(n1)][ID(n1+g)] for g_x >= 0, then
 * the c2r synthesis.  tcc: T x T complex double; mask, image: Ly x Lx. */
int klt_hopkins_image(klt_ctx* ctx, const float* mask, int Lx, int Ly, const double* tcc,
                      int Nx, int Ny, float* image, unsigned flags);

/* ---- SOCS kernel bank --------------------------------------------------
 * Replaces the kernel read loop of klitho_socs.cpp:297-318: nk kernels of
 * sizeX x sizeY (odd) in the krn_i_{r,i}.bin layout, and their eigenvalues
 * (kernel_info.txt).  The bank is packed on the device once. */
int klt_kernels_create(klt_ctx* ctx, const float* krn_r, const float* krn_i,
                       const float* scales, int nk, int sizeX, int sizeY,
                       unsigned flags, klt_kernels** out);
void klt_kernels_destroy(klt_kernels* kb);

/* ---- SOCS imaging (klitho_socs.cpp:338-343: FT_r2c + calcSOCS) ---------
 * image(x) = sum_k scales[k] |sum_n K_k[n] M^[n] e^{+2 pi i n.(x-c)/L}|^2 with
 * M^ the mask spectrum of FT_r2c (klitho_socs.cpp:71-121).  mask and image
 * are Ly x Lx float.  Errors with KLT_ERR_SIZE when 4N+1 exceeds Lx or Ly.
 * klt_socs_image is the plain single-image path: FP32 arithmetic throughout. */
int klt_socs_image(klt_ctx* ctx, const klt_kernels* kb, const float* mask, int Lx,
                   int Ly, float* image, unsigned flags);
/* nmask masks of one size, stored back to back.  For every N <= 31 the
 * full-field DFT passes run as split hi+lo FP16 GEMMs on the matrix cores with
 * FP32 accumulation (image relative L2 error ~1e-6 against FP64); the
 * coarse-grid FFTs are FP32, on a 64-point grid for N <= 15 and a 128-point
 * grid for N 16 .. 31.  Above 31 the call fails with KLT_ERR_ARG. */
int klt_socs_forward(klt_ctx* ctx, const klt_kernels* kb, const float* masks,
                     int nmask, int Lx, int Ly, float* images, unsigned flags);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the images of klt_socs_forward (same arithmetic and
 * kernels, kernel half-size <= 15, device pointers only) with every table and
 * intermediate in one fixed device array of KLT_NOALLOC_SCRATCH_BYTES per GPU,
 * which the library holds until the process ends.  No allocation and no
 * synchronous copy: the call only launches kernels on the context stream (the
 * twiddle tables are generated on the device at the start of every call).
 * Lx and Ly are limited to KLT_NOALLOC_MAX_SIDE (KLT_ERR_ARG beyond); larger
 * batches run in sub-batches that fit the array.  All calls on one GPU share
 * the array and it keeps the last call's contents: calls that overlap on one
 * GPU (two streams, two host threads) race.  The split factors come from the
 * fixed rule, not the tuner. */
#define KLT_NOALLOC_SCRATCH_BYTES (32u << 20)
#define KLT_NOALLOC_MAX_SIDE 4096
int klt_socs_forward_noalloc(klt_ctx* ctx, const klt_kernels* kb, const float* masks,
                             int nmask, int Lx, int Ly, float* images);

/* Masks of different sizes under several kernel sets in one call (no
 * reference counterpart: klitho_socs.cpp:338-343 images one mask per run).
 * masks[m] (Ly[m] x Lx[m]) under set s gives images[m * nbank + s].  Every
 * set must have the same kernel size, half-size <= 31.  Each mask's band is
 * computed once and shared by its sets.  The call's launches are captured in
 * a hipGraph on first use and replayed when the same arguments (pointers,
 * sizes, sets) come again, so a caller that refills the same buffers pays one
 * graph launch per batch.  A stream the caller is itself capturing receives
 * the launches directly, for a batch whose arguments have been seen before
 * (the first call of a batch uploads its job tables with synchronous copies,
 * which a capture does not allow: it returns KLT_ERR_ARG under capture); the
 * library then keeps that batch's buffers until the context is destroyed, as
 * the caller's graph refers to them.  Device pointers only (the pointer
 * arrays are host memory). */
int klt_socs_forward_batch(klt_ctx* ctx, int nmask, const float* const* masks, const int* Lx,
                           const int* Ly, int nbank, const klt_kernels* const* banks,
                           float* const* images, unsigned flags);
/* Completion time of each image of the context's last klt_socs_forward_batch
 * call, in microseconds from the start of that batch on the device (device
 * wall clock).  Waits for the batch; writes min(cap, n) values and returns the
 * image count n. */
int klt_socs_batch_latency(klt_ctx* ctx, float* us, int cap);

/* ---- resist step (klitho_tcc.cpp:442,547; klitho_socs.cpp:275,353;
 * common/file_io.cpp:25-80) ------------------------------------------------
 * The class of each pixel of nimg images (Ly x Lx) in the reference's 7-level
 * view: v = clamp((I - min) / (max - min), 0, 1) computed in float, level 0..5
 * for v in [k/6, (k+1)/6), level 6 for v = 1; min == max gives level 3.  With
 * min = 0 and max = KLT_TARGET_INTENSITY, level 6 is the printed resist
 * (I >= threshold).  levels: nimg * Ly * Lx bytes. */
#define KLT_TARGET_INTENSITY 0.225f
int klt_resist_levels(klt_ctx* ctx, const float* images, int nimg, int Lx, int Ly, float min,
                      float max, unsigned char* levels, unsigned flags);

/* ---- adjoint (no reference counterpart) ---------------------------------
 * Resist loss of each SOCS image, L = sum_x (sigmoid(a (I(x) - thr)) - T(x))^2,
 * and its gradient with respect to the mask, dL/dm (Ly x Lx per mask), through
 * the same band-limited passes as klt_socs_forward (kernel half-size <= 31).
 * masks (continuous transmittance allowed), targets, grads: nmask x Ly x Lx;
 * losses: nmask; images: optional (NULL) output of the forward images. */
int klt_socs_backward(klt_ctx* ctx, const klt_kernels* kb, const float* masks,
                      const float* targets, int nmask, int Lx, int Ly, float a, float thr,
                      float* grads, float* losses, float* images, unsigned flags);

/* No reference counterpart.  The vector-Jacobian product of the SOCS image
 * for any loss: given dL/dI (nmask x Ly x Lx, I the klt_socs_forward images of
 * masks), writes dL/dm = sum_x dL/dI(x) dI(x)/dm (nmask x Ly x Lx) through the
 * adjoint passes of klt_socs_backward (kernel half-size <= 31; dL/dI's band
 * with |g| <= 2N is all that enters).  The forward image is not formed. */
int klt_socs_vjp(klt_ctx* ctx, const klt_kernels* kb, const float* masks, const float* dL_dI,
                 int nmask, int Lx, int Ly, float* grads, unsigned flags);

/* Gradient of the same loss with respect to the source, in the Abbe form:
 * with I_A(x) = sum_s S_s |E_s(x)|^2 the untruncated image of the source src
 * (equal to the Hopkins image of its TCC), dL/dS_s = sum_x G(x) |E_s(x)|^2,
 * G = dL/dI_A, for every grid point inside the unit disc (zero elsewhere),
 * with respect to the weights as given.  The optics are those of
 * klt_tcc_build (FP64 Gram).  Kernel half-size <= 31.  grad_src: nmask x
 * srcSize^2; losses: nmask (loss of I_A); images: optional I_A (device
 * pointers only, else NULL). */
int klt_abbe_source_grad(klt_ctx* ctx, const float* masks, const float* targets, int nmask,
                         int Lx, int Ly, const float* src, int srcSize, float lambda, float NA,
                         float defocus, float nmed, int Nx, int Ny, float a, float thr,
                         float* grad_src, float* losses, float* images, unsigned flags);

/* ---- source-mask optimisation (no reference counterpart) ----------------
 * Projected, normalised gradient descent of the resist loss of
 * klt_socs_backward over nmask masks and their shared source, `steps`
 * iterations of: the TCC of the current source (klt_tcc_build, split-FP16
 * Gram, at the kernel half-size of a filled pupil, N = floor(NA L 2 / lambda),
 * so that the source may grow anywhere inside the unit disc), its
 * K = min(nk, T) kernels (as klt_calc_kernels with KLT_EIG_PSD, and
 * KLT_EIG_WARM from the second iteration on, to residuals of 1e-10 of the
 * largest eigenvalue), the loss and mask gradient g_m through the SOCS
 * model (klt_socs_backward), the source gradient g_S through the Abbe model
 * (klt_abbe_source_grad, summed over the masks), then
 *   m <- clip(m - step_mask g_m / max|g_m|, 0, 1)                    per mask,
 *   S <- S' / sum S',  S' = max(S - step_src max(S) g_S / max|g_S|, 0)
 * (grid points outside the unit disc stay zero).  masks: device, nmask x Ly x
 * Lx, updated in place; targets: device; src: host srcSize^2, normalised,
 * updated in place.  history[steps + 1]: the summed SOCS loss before each step
 * and after the last; mismatch (optional, NULL) [steps + 1]: the count of
 * pixels whose printed resist (I >= thr) differs from the target (T >= 0.5). */
int klt_smo_run(klt_ctx* ctx, float* masks, const float* targets, int nmask, int Lx, int Ly,
                float* src, int srcSize, float lambda, float NA, float defocus, float nmed,
                int nk, float a, float thr, float step_mask, float step_src, int steps,
                double* history, long long* mismatch);

#ifdef __cplusplus
}
#endif
#endif /* KLT_H */
