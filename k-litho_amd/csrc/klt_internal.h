// klt_internal.h -- shared host-side plumbing of libklt.so: error state,
// the context object, device buffer helpers and HIP error checks.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/klt.h"
#include "klt_error.h"

namespace klt {

#define KLT_HIP(expr)                                                         \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::klt::set_error(KLT_ERR_HIP, std::string(#expr) + ": " +        \
                                               hipGetErrorString(e_));        \
  } while (0)

// Growable device buffer owned by a context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t need);
  void release();
  ~DevBuf() { release(); }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
};

// Twiddle tables of one SOCS geometry (Lx, Ly, Nx, Ny).
// Rows of the input tables are padded with zeros to NP = round_up(max(Nx,Ny)+1, 8)
// entries and the output-row table to G = round_up(2Ny+1, 8), so that the
// register-blocked passes can run fixed-size loops.
struct SocsPlan {
  int Lx = 0, Ly = 0, Nx = 0, Ny = 0, M = 0, NP = 0, G = 0;
  DevBuf twy_in;   // [Ly][NP]    e^{-2 pi i ny (y - floor(Ly/2))/Ly}, ny < NP
  DevBuf twx_in;   // [Lx][NP]    e^{-2 pi i nx (x - floor(Lx/2))/Lx}, nx < NP
  DevBuf twx_out;  // [Lx][2Nx+1] e^{+2 pi i gx (x - (Lx+1)/2)/Lx}
  DevBuf twy_out;  // [Ly][G]     e^{+2 pi i gy (y - (Ly+1)/2)/Ly}, gy < G
  DevBuf twM;      // [M]         e^{+2 pi i j/M}
};

// Half-precision operand tables of the fused SOCS path (socs_fused.hip):
// the band kernel's row twiddles Wf (rows n = 2 nx: cos, 2 nx + 1: -sin, 32
// rows, Kpad = round_up(Lx, 128) columns, zero past Lx) as hi / lo fp16 pairs
// in MFMA fragment order, TY[y][j] (j = 2 gy: w cos, 2 gy + 1: -w sin;
// Lypad x 64) as hi / lo pairs, and e^{+2 pi i j/64}.
struct FusedTables {
  int Kpad = 0, Lypad = 0;
  DevBuf Wf, TYh, TYl, tw64;
  DevBuf Ef;  // synthesis column twiddles, hi / lo fp16 in MFMA fragment order
  // adjoint: resist-gradient band rows Wgf (64 rows, centre (Lx+1)/2, fragment
  // order) and the column twiddles e^{-2 pi i gy (y - (Ly+1)/2)/Ly} (Ly x 32)
  DevBuf Wgf, twyg;
  // fused adjoint synthesis (k_synth_resist): the column twiddles of the
  // resist band as A fragments over 16-row tiles (TYgf) and the synthesis
  // column twiddles E as B fragments over 32-column steps (Ebf)
  DevBuf TYgf, Ebf;
  // kernel half-sizes 16..31 (k_synth_wide): Wf then holds 64 twiddle rows;
  // TYw = TY (Lypad x 128) as A fragments over 16-row tiles, Efw the
  // synthesis column twiddles as A fragments over 8 steps of 32
  DevBuf TYw, Efw, tw128;  // (tw128: e^{+2 pi i j/128}, j < 128)
  // fused adjoint synthesis of half-sizes 16..31 (k_synth_resist_wide): the
  // resist band's column twiddles over 16-row tiles (TYgw, 128 j) and the
  // synthesis column twiddles as B fragments over 32-column steps (Ebw, 256 k)
  DevBuf TYgw, Ebw;
};

// Device pointers of one fused forward launch (socs_fused_launch): the tables
// (layouts of FusedTables / SocsPlan: Wf fragment-ordered row twiddles with
// Kpad columns, twy the [Ly][ldtw] column twiddles, TYh / TYl [Lypad][64], Ef
// the synthesis column fragments, tw64 e^{+2 pi i j/64}) and the intermediates
// of nb images: qpart [nb][nblk * ks][2Ny+1][16], spec [nb][32][4Nx+1], and
// for kg > 1 ipart [nb][kg][64 * 64].
struct FusedView {
  int Kpad, Lypad, ldtw;
  const float4* Wf;
  const float2* twy;
  const _Float16* TYh;
  const _Float16* TYl;
  const float4* Ef;
  const float2* tw64;
};
struct FusedScratch {
  float2* qpart;
  float2* spec;
  float* ipart;
};

// Work descriptors of the mixed-size batch (klt_socs_forward_batch): one per
// block of each stage, so that one launch covers masks of different sizes
// and several kernel sets.
struct BandJob {        // row block `blk` of one mask
  const float* in;      // the mask, Ly x Lx
  const float4* Wf;     // its row twiddles (fragment order)
  const float2* twy;    // its column twiddles, [Ly][ldtw]
  float2* out;          // its band partials, [nblk][2Ny+1][16] (32 for N > 15)
  int Lx, Ly, Kpad, ldtw, Nx, Ny, blk, pad_;
};
struct MidJob {         // one mask x one kernel set
  const float2* mpart;  // the mask's band partials
  const float2* kern;   // the set's kernels, [nk][2Ny+1][2Nx+1]
  const float* scales;  // and eigenvalues
  float2* spec;         // the image's intensity spectrum, [32][4Nx+1] (64 rows for N > 15)
  float* ipart;         // kg > 1: the image's partial coarse intensities, [kg][64 * 64]
  int nblk, Nx, Ny, nk;
  float inv_area;
  int k0, k1, grp, kg;  // this block's kernels [k0, k1), group grp of kg
  int img;              // the image's index in the batch
};
struct SynthJob {       // one 128-column strip of one image
  const float2* spec;
  const float4* Ef;     // synthesis column twiddles (fragment order; Efw for N > 15)
  const _Float16* TYh;  // (N > 15: the TYw fragment table)
  const _Float16* TYl;
  float* img;
  int Nx, Lx, Ly, Lypad, x0, image, nstrips, Ny;
};

// One captured mixed batch: its job tables, intermediates and the graph that
// replays its three stages; `key` is the call's arguments.  `held`: its
// launches went into a graph that the caller was capturing, which keeps
// pointers to these buffers, so the entry stays until the context is
// destroyed instead of taking part in the cache's eviction.
struct BatchGraph {
  std::vector<uint64_t> key;
  int nimg = 0, nband = 0, nmid = 0, nsynth = 0;
  int kg = 1;         // kernel groups per image in the narrow coarse stage
  bool wide = false;  // kernel half-size above 15 (the wide passes)
  int lds_mid = 0;    // dynamic LDS of the wide coarse stage
  bool full = true;
  bool held = false;
  DevBuf band_jobs, mid_jobs, synth_jobs, mpart, spec, strips_done, done_at, t0, ipart;
  const float2* spec_tw = nullptr;  // e^{+2 pi i j/64} (j/128 for N > 15) of the coarse FFTs
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~BatchGraph();
};

}  // namespace klt

struct klt_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  // per-kernel timing (klt_ctx_set_profiling): events bracketing launches
  bool profiling = false;
  struct TimedLaunch {
    const char* name;
    hipEvent_t start, stop;
  };
  std::vector<TimedLaunch> timed;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t take_event();
  // Pinned host staging for the small host-to-device uploads of a call (source
  // points, tile and job lists): upload() copies `bytes` from host memory into
  // the next free piece of the pinned area and queues an asynchronous copy to
  // `dev` on the context stream, so the call need not wait for the copy before
  // its host data goes out of scope.  The area is reused from its start once
  // full, after the last queued copy has completed.
  char* up_base = nullptr;
  size_t up_cap = 0, up_used = 0;
  hipEvent_t up_last = nullptr;
  int upload(void* dev, const void* host, size_t bytes);
  // pinned host area for small device-to-host reads that the host waits for
  // with an event (so that kernels queued after the copy run meanwhile)
  void* dl_base = nullptr;
  size_t dl_cap = 0;
  int download_area(size_t bytes, void** out);
  // scratch reused across calls (grown on demand)
  klt::DevBuf stage_in, stage_out, qpart, band, spec, ccols, misc, eig_ws;
  klt::DevBuf gpart, gspec, lpart, bwd_img;  // adjoint scratch
  klt::DevBuf vjp_sum;                        // klt_socs_vjp: the (unused) sums of G^2
  klt::DevBuf abbe_tcc, abbe_band, abbe_in;   // source-gradient scratch / staging
  klt::DevBuf abbe_pupil;                     // pupil table of the source-gradient points
  klt::DevBuf tcc_pts;                        // source points of the TCC build
  klt::DevBuf tcc_part;                       // split-K partial sums of the TCC Gram
  klt::DevBuf tcc_jobs;                       // per-block jobs of the TCC Gram
  klt::DevBuf coarse_ws;                      // coarse-grid scratch of the N > 18 SOCS stage
  klt::DevBuf tcc_tab;                        // split16 TCC: points, chunk masks, jobs
  std::vector<char> tcc_tab_host;             // its last uploaded contents
  // the split16 TCC plan of the last call: source values and the geometry
  // that fixes pupil membership (defocus changes only the phase)
  struct TccPlan {
    std::vector<float> src;
    std::array<float, 4> fp{};   // outerSigma, lambda, NA, nmed
    std::array<int, 5> ip{};     // srcSize, Lx, Ly, Nx, Ny
    int ns = 0, Kpad = 0, Tpad = 0, nc = 0, ntl = 0, njobs = 0, nblocks = 0;
    float scale = 0.f;
    size_t oP = 0, oL = 0, oT = 0, oTJ = 0, oJ = 0, oO = 0, oC = 0;
    bool valid = false;
  } tcc_plan;
  klt::DevBuf tw64;                           // e^{+2 pi i j/64}, j < 64
  klt::DevBuf ipart, gam_part;                // kernel-group partials of the coarse stages
  klt::DevBuf mhat, gfull, gcoarse;           // adjoint: per-image M^, G^ and coarse-grid g_c
  klt::DevBuf hspec;                          // Hopkins spectra of the source gradient
  klt::DevBuf sub_ws;                         // blocks of the subspace eigensolver
  int sub_warm_n = 0;  // order of the last single-matrix PSD solve whose block sub_ws holds
  std::map<std::tuple<int, int, int, int>, std::unique_ptr<klt::SocsPlan>> socs_plans;
  std::map<std::tuple<int, int, int, int>, std::unique_ptr<klt::FusedTables>> fused_tables;
  // split factors of the fused passes per launch shape (kind, batch, Lx, Ly,
  // Nx, Ny, nk), chosen by timing on first use while `autotune` is on
  struct Tuned {
    int ks, rs, kg;  // band column ranges, synthesis row ranges, kernel groups
    int rb;          // rows per band block (64 or 128)
    float ms;
  };
  bool autotune = true;
  std::map<std::array<int, 7>, Tuned> fused_tune;
  // split factors pinned by klt_ctx_set_splits (0: tuned or fixed rule)
  int pin_ks = 0, pin_rs = 0, pin_kg = 0, pin_rb = 0;
  // captured mixed batches, most recent last (a few kept), and the last one run
  std::vector<std::unique_ptr<klt::BatchGraph>> batches;
  klt::BatchGraph* last_batch = nullptr;
  int activate() const;  // hipSetDevice(device)
};

struct klt_kernels {
  klt_ctx* ctx = nullptr;
  int nk = 0, Nx = 0, Ny = 0;
  klt::DevBuf kern;    // [nk][2Ny+1][2Nx+1] float2
  klt::DevBuf scales;  // [nk] float
  std::vector<float> h_scales;
};

// Launches `kernel` on the context stream; with profiling on, the launch is
// bracketed by HIP events and recorded under `name`.
#define KLT_LAUNCH(ctx, name, kernel, grid, block, lds, ...)                        \
  do {                                                                              \
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;                                      \
    if ((ctx)->profiling) {                                                         \
      ev0_ = (ctx)->take_event();                                                   \
      ev1_ = (ctx)->take_event();                                                   \
      (void)hipEventRecord(ev0_, (ctx)->stream);                                    \
    }                                                                               \
    hipLaunchKernelGGL(kernel, grid, block, lds, (ctx)->stream, __VA_ARGS__);       \
    if ((ctx)->profiling) {                                                         \
      (void)hipEventRecord(ev1_, (ctx)->stream);                                    \
      (ctx)->timed.push_back({name, ev0_, ev1_});                                   \
    }                                                                               \
  } while (0)

namespace klt {
// SOCS geometry plan (cached in the context).
int get_socs_plan(klt_ctx* ctx, int Lx, int Ly, int Nx, int Ny, SocsPlan** out);
// SOCS forward passes: the FP32 path (any N <= 31) and the fused
// split-precision path (N <= 15).
int socs_forward_device(klt_ctx* ctx, const klt_kernels* kb, const float* masks, int B,
                        int Lx, int Ly, float* images);
bool socs_fused_supported(const klt_kernels* kb);
// Coarse-grid stage (64-point FFTs, FP32) of the single-image path, kernel
// half-size <= 15: band [B][2Ny+1][2Nx+1] (normalised) to the intensity
// spectra spec [B][G][4Nx+1].
int socs_mid_from_band(klt_ctx* ctx, const klt_kernels* kb, const float2* band, int B,
                       float2* spec, int G);
// Coarse-grid stage of any half-size N <= 31 (socs.hip): band
// [B][2Ny+1][2Nx+1] (normalised) to spec [B][G][4Nx+1]; N <= 15 through
// socs_mid_from_band, above that by direct transforms on the 128-point grid.
int socs_coarse_spec(klt_ctx* ctx, const klt_kernels* kb, const SocsPlan* p,
                     const float2* band, int B, float2* spec, int G);
// Batched SOCS forward for kernel half-sizes 16..31 (socs_fused.hip): the
// band pass on the matrix cores (64 twiddle rows), the coarse stage as
// 128-point FFTs, and the matrix-core synthesis at K = 128.
bool socs_wide_supported(const klt_kernels* kb);
int socs_forward_wide(klt_ctx* ctx, const klt_kernels* kb, const float* masks, int B, int Lx,
                      int Ly, float* images);
int socs_forward_fused(klt_ctx* ctx, const klt_kernels* kb, const float* masks, int B, int Lx,
                       int Ly, float* images);
// The launches of one fused forward pass over nb masks (kernel half-size
// <= 15) with explicit tables and intermediates, split factors ks (band column
// ranges, <= 8), rs (synthesis row ranges, <= 16) and kg (kernel groups), and
// rb rows per band block (64 or 128; qpart then holds ceil(Ly/rb) ks partials
// per image).
int socs_fused_launch(klt_ctx* ctx, const klt_kernels* kb, const float* masks, int nb, int Lx,
                      int Ly, float* images, const FusedView& v, const FusedScratch& w,
                      int ks, int rs, int kg, int rb = 128);
// Intermediate bytes per image of socs_fused_launch, and the fixed split rule
// (two blocks per CU) for nb images.
size_t fused_scratch_bytes(int Lx, int Ly, int Nx, int Ny, int ks, int kg);
void fused_fixed_splits(int nb, int Lx, int Ly, int nk, int* ks, int* rs, int* kg);
// Band partial sums of the fused passes (socs_fused.hip), nblk = ceil(Ly/128)
// row blocks split into ks column ranges (band_split: the fixed rule of the
// batched passes, parts = nblk * ks):
// mask band  [B][parts][2Ny+1][16]  (nx in [0, Nx], centre floor(L/2));
// resist band [B][parts][4Ny+1][32] (gx in [0, 2Nx], centre (L+1)/2), with the
// loss partials lpart [B][parts].  Kernel half-size <= 15.
int band_split(int Lx, int Ly, int B);
int band_mask_partials(klt_ctx* ctx, int Lx, int Ly, int Nx, int Ny, const float* masks, int B,
                       int ks, float2* out);
int band_resist_partials(klt_ctx* ctx, int Lx, int Ly, int Nx, int Ny, const float* images,
                         const float* targets, int B, int ks, float a_sig, float thr,
                         float2* out, float* lpart);
// TCC on the device (tcc.hip) and the Hopkins image (socs.hip).
int tcc_build_device(klt_ctx* ctx, const float* src_host, int srcSize, float outerSigma,
                     float lambda, float NA, float defocus, float nmed, int Lx, int Ly,
                     int Nx, int Ny, int precision, double* tcc_dev);
int hopkins_image_device(klt_ctx* ctx, const float* mask, int Lx, int Ly, const double* tcc,
                         int Nx, int Ny, float* image);
// Hopkins intensity spectra of B images from their bands [B][(2Ny+1)(2Nx+1)]
// (normalised, centred): spec [B][rows][4Nx+1] (socs.hip).
int hopkins_spec_device(klt_ctx* ctx, const float2* band, int B, const double* tcc, int Nx,
                        int Ny, int rows, float2* spec);
// The source gradient's stages for kernel half-sizes 16..31 (socs_fused.hip):
// the normalised bands Mfull [B][(2Ny+1)(2Nx+1)] of B masks; from the
// Hopkins spectra of a device TCC, the resist loss per mask, the 128-point
// coarse-grid g_c [B][256][64] (thread-major) and, with images != NULL, the
// Hopkins images; the gradient per source point on the 128-point grid.
int band_mask_full_wide(klt_ctx* ctx, int Lx, int Ly, int Nx, int Ny, const float* masks, int B,
                        float2* Mfull);
int abbe_resist_wide(klt_ctx* ctx, const float2* Mfull, int B, int Lx, int Ly, const double* tcc,
                     int Nx, int Ny, const float* targets, float a_sig, float thr, float* images,
                     float* loss, float* gc);
struct PupilGeom;
int abbe_grad_coarse128(klt_ctx* ctx, const float2* Mfull, const float* gc, int B, int Lx,
                        int Ly, const PupilGeom& geom, const int2* pts, int npts, int srcSize,
                        float* grad);
// Hopkins images of B masks of Lx x Ly from their bands, through the batched
// matrix-core synthesis (kernel half-size <= 15; socs_fused.hip).
int hopkins_images_fused(klt_ctx* ctx, const float2* band, int B, int Lx, int Ly,
                         const double* tcc, int Nx, int Ny, float* images);
struct PupilGeom;  // klt_pupil.h
// From the resist-band partials gpart and loss partials lpart
// (band_resist_partials, `parts` per image) of B images: the losses, and
// g_c(u) = sum_g G^[g] e^{+2 pi i g.u/64} on the 64 x 64 coarse grid, thread-major
// [B][256][16] floats (socs_fused.hip).  Kernel half-size <= 15.
int resist_coarse_grid(klt_ctx* ctx, const float2* gpart, const float* lpart, int parts, int B,
                       int Nx, int Ny, float inv_area, float* loss, float* gc);
// Source gradient of the resist loss for B masks (socs_fused.hip): from each
// mask's normalised band M^ [B][(2Ny+1)(2Nx+1)] and its coarse-grid g_c
// (resist_coarse_grid), dL/dS at the npts source points pts (grid offsets from
// the centre) into grad [B][srcSize^2].  Kernel half-size <= 15.
int abbe_grad_coarse(klt_ctx* ctx, const float2* Mfull, const float* gc, int B, int Lx, int Ly,
                     const PupilGeom& geom, const int2* pts, int npts, int srcSize,
                     float* grad);
// Top-k eigenpairs of a Hermitian matrix on the device (eig.hip).
int eig_topk_device(klt_ctx* ctx, const double* a_dev, int n, int k, double* w_dev,
                    double* vec_dev);
// The same for `count` matrices at once (a_dev [count][n][n] complex, w_dev
// [count][k], vec_dev [count][k][n] complex): every launch covers the batch.
int eig_topk_batch_device(klt_ctx* ctx, const double* a_dev, int count, int n, int k,
                          double* w_dev, double* vec_dev);
// The same for positive semidefinite matrices by block subspace iteration
// (eig_subspace.hip); KLT_ERR_EIG when a matrix is outside its
// reach (k + 16 > 64, n < 128, no convergence), for the caller to fall back.
// warm (count 1): start from the block of the previous solve of order n;
// tol: the residual target relative to the largest eigenvalue (0: 5e-14).
int eig_topk_psd_batch_device(klt_ctx* ctx, const double* a_dev, int count, int n, int k,
                              double* w_dev, double* vec_dev, bool warm = false,
                              double tol = 0.0);
// klt_calc_kernels on device pointers with the subspace solver's residual
// target tol (0: the public 5e-14), for callers that need less, such as the
// source-mask optimisation loop.
int calc_kernels_device(klt_ctx* ctx, const double* tcc, int T, int nk, float* krn_r,
                        float* krn_i, float* scales, unsigned flags, double tol,
                        bool sync = true);
// New kernels and scales (device pointers, the bank's count and size) into an
// existing bank without allocating (socs.hip).
int kernels_refill(klt_kernels* kb, const float* krn_r, const float* krn_i,
                   const float* scales);
// Mixed-size, multi-set batch (socs_fused.hip): images[m * nbank + s] is mask
// m under kernel set s.  Captured into a graph on first use, replayed after.
int socs_forward_batch(klt_ctx* ctx, int nmask, const float* const* masks, const int* Lx,
                       const int* Ly, int nbank, const klt_kernels* const* banks,
                       float* const* images);
// Device address of socs_noalloc.hip's scratch array on `device` (the current
// device), looked up once per device; called by klt_ctx_create.
int noalloc_setup(int device);
// Copy helpers honouring KLT_HOST_PTRS.
int to_device(klt_ctx* ctx, const void* src, size_t bytes, unsigned flags, DevBuf& stage,
              const void** dptr);
}  // namespace klt
