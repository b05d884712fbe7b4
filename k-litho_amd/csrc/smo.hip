// smo.hip -- source-mask optimisation over the adjoint (no counterpart in the
// reference, which images one mask under one source; it turns the resist
// step of klitho_tcc.cpp:442 / the sigmoid resist loss into an optimisation
// loop).
//
// Each iteration, for the current source S (host, srcSize^2) and masks m
// (device):
//   1. TCC of S (klt_tcc_build, split hi+lo FP16 Gram on the matrix cores) at
//      the kernel half-size of a fully filled pupil,
//      N = floor(NA L (1 + 1) / lambda), so that the source may grow anywhere
//      inside the unit disc without changing the kernel size;
//   2. its nk kernels (klt_calc_kernels, positive-semidefinite path; from the
//      second iteration on the subspace iteration starts from the previous
//      iteration's block, as the source changes little between steps);
//   3. the resist loss and dL/dm through the SOCS model (klt_socs_backward),
//      with the images for the printed-resist count;
//   4. dL/dS through the Abbe model (klt_abbe_source_grad), summed over masks;
//   5. m <- clip(m - step_mask g_m / max|g_m|, 0, 1) per mask on the device,
//      S <- S' / sum S', S' = max(S - step_src max(S) g_S / max|g_S|, 0) on
//      the host (points outside the unit disc stay zero).
// The loss and the count of pixels whose printed resist (I >= thr) differs
// from the target (T >= 0.5) are recorded before every step and after the
// last.
#include <algorithm>
#include <cmath>
#include <vector>

#include "klt_internal.h"

namespace {

constexpr int kRedBlocks = 256;  // blocks a mask of the reductions (partials)

// Block max over 256 threads, valid in every thread.
__device__ __forceinline__ float block_max(float m, float* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// grid (kRedBlocks, B): part[b][x] = max |g| over block x's share of mask b
// (float4 loads where the mask is aligned).
__global__ __launch_bounds__(256) void k_smo_absmax(const float* __restrict__ g, size_t n,
                                                    float* __restrict__ part) {
  const float* gb = g + (size_t)blockIdx.y * n;
  const size_t stride = (size_t)gridDim.x * 256;
  float m = 0.f;
  if ((n & 3) == 0 && (reinterpret_cast<uintptr_t>(gb) & 15) == 0) {
    const float4* g4 = reinterpret_cast<const float4*>(gb);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n / 4; i += stride) {
      const float4 v = g4[i];
      m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
    }
  } else {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
      m = fmaxf(m, fabsf(gb[i]));
  }
  __shared__ float red[4];
  m = block_max(m, red);
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = m;
}

// m <- clip(m - step g / max|g|, 0, 1), max|g| of mask b from its kRedBlocks
// partials; a zero gradient leaves the mask
__global__ __launch_bounds__(256) void k_smo_mask_step(float* __restrict__ m,
                                                       const float* __restrict__ g, size_t n,
                                                       const float* __restrict__ part,
                                                       float step) {
  __shared__ float red[4];
  const float gmax = block_max(part[(size_t)blockIdx.y * kRedBlocks + threadIdx.x], red);
  if (!(gmax > 0.f)) return;
  const float s = step / gmax;
  float* mb = m + (size_t)blockIdx.y * n;
  const float* gb = g + (size_t)blockIdx.y * n;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    mb[i] = fminf(1.f, fmaxf(0.f, mb[i] - s * gb[i]));
}

// cnt[x] = the count of pixels with (I >= thr) != (T >= 0.5) in block x's
// share (summed on the host)
__global__ __launch_bounds__(256) void k_smo_mismatch(const float* __restrict__ img,
                                                      const float* __restrict__ tgt, size_t n,
                                                      float thr, unsigned* __restrict__ cnt) {
  unsigned c = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    c += ((img[i] >= thr) != (tgt[i] >= 0.5f)) ? 1u : 0u;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  __shared__ unsigned red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) cnt[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace

extern "C" {

int klt_smo_run(klt_ctx* ctx, float* masks, const float* targets, int nmask, int Lx, int Ly,
                float* src, int srcSize, float lambda, float NA, float defocus, float nmed,
                int nk, float a_sig, float thr, float step_mask, float step_src, int steps,
                double* history, long long* mismatch) {
  if (!ctx || !masks || !targets || !src || !history || nmask <= 0 || Lx <= 0 || Ly <= 0 ||
      srcSize <= 0 || nk <= 0 || steps < 0)
    return klt::set_error(KLT_ERR_ARG, "klt_smo_run: bad argument");
  KLT_TRY(ctx->activate());
  int Nx = 0, Ny = 0;
  KLT_TRY(klt_tcc_sizing(NA, Lx, Ly, 1.0f, lambda, &Nx, &Ny));
  if (4 * Nx + 1 > Lx || 4 * Ny + 1 > Ly) return klt::set_error(KLT_ERR_SIZE, "ERROR in FI()");
  const int T = (2 * Nx + 1) * (2 * Ny + 1), K = std::min(nk, T), n2 = srcSize * srcSize;
  const size_t px = (size_t)Lx * Ly, tot = px * nmask;
  klt::DevBuf tcc, kr, ki, sc, grad, gsrc, loss, lossA, img, dsrc, mx, cnt;
  KLT_TRY(tcc.ensure((size_t)T * T * 16));
  KLT_TRY(kr.ensure((size_t)K * T * 4));
  KLT_TRY(ki.ensure((size_t)K * T * 4));
  KLT_TRY(sc.ensure((size_t)K * 4));
  KLT_TRY(grad.ensure(tot * 4));
  KLT_TRY(gsrc.ensure((size_t)nmask * n2 * 4));
  KLT_TRY(loss.ensure((size_t)nmask * 4));
  KLT_TRY(lossA.ensure((size_t)nmask * 4));
  KLT_TRY(img.ensure(tot * 4));
  KLT_TRY(dsrc.ensure((size_t)n2 * 4));
  KLT_TRY(mx.ensure((size_t)nmask * kRedBlocks * 4));
  const unsigned cblocks = 1024;
  KLT_TRY(cnt.ensure(cblocks * 4));
  const int sh = (srcSize - 1) / 2;
  std::vector<float> h_loss(nmask), h_g((size_t)nmask * n2), gs(n2);
  std::vector<unsigned> h_cnt(cblocks);
  const unsigned blocks = (unsigned)std::min<size_t>((px + 255) / 256, 512);
  struct Bank {  // destroyed on every way out of the loop
    klt_kernels* kb = nullptr;
    explicit operator bool() const { return kb != nullptr; }
    ~Bank() { klt_kernels_destroy(kb); }
  } bank;
  for (int it = 0; it <= steps; ++it) {
    KLT_HIP(hipMemcpyAsync(dsrc.ptr, src, (size_t)n2 * 4, hipMemcpyHostToDevice, ctx->stream));
    KLT_TRY(klt_tcc_build(ctx, dsrc.as<float>(), srcSize, 1.0f, lambda, NA, defocus, nmed, Lx,
                          Ly, Nx, Ny, KLT_TCC_SPLIT_F16, tcc.as<double>(), 0));
    // the source stays non-negative, so the TCC is positive semidefinite; from
    // the second iteration on the previous eigensolve's block starts this one
    // (residuals to 1e-10 of the largest eigenvalue: the loop's gradients are
    // good to 1e-4, and the public calls' 5e-14 costs about three more rounds)
    KLT_TRY(klt::calc_kernels_device(ctx, tcc.as<double>(), T, K, kr.as<float>(),
                                     ki.as<float>(), sc.as<float>(),
                                     KLT_EIG_PSD | (it > 0 ? KLT_EIG_WARM : 0u), 1e-10, false));
    // one kernel bank for the loop, refilled in place after the first step
    if (!bank)
      KLT_TRY(klt_kernels_create(ctx, kr.as<float>(), ki.as<float>(), sc.as<float>(), K,
                                 2 * Nx + 1, 2 * Ny + 1, 0, &bank.kb));
    else
      KLT_TRY(klt::kernels_refill(bank.kb, kr.as<float>(), ki.as<float>(), sc.as<float>()));
    KLT_TRY(klt_socs_backward(ctx, bank.kb, masks, targets, nmask, Lx, Ly, a_sig, thr,
                              grad.as<float>(), loss.as<float>(), img.as<float>(), 0));
    KLT_LAUNCH(ctx, "k_smo_mismatch", k_smo_mismatch, dim3(cblocks), dim3(256), 0,
               img.as<float>(), targets, tot, thr, cnt.as<unsigned>());
    KLT_HIP(hipMemcpyAsync(h_loss.data(), loss.ptr, (size_t)nmask * 4, hipMemcpyDeviceToHost,
                           ctx->stream));
    KLT_HIP(hipMemcpyAsync(h_cnt.data(), cnt.ptr, cblocks * 4, hipMemcpyDeviceToHost,
                           ctx->stream));
    if (it < steps) {
      KLT_TRY(klt_abbe_source_grad(ctx, masks, targets, nmask, Lx, Ly, dsrc.as<float>(),
                                   srcSize, lambda, NA, defocus, nmed, Nx, Ny, a_sig, thr,
                                   gsrc.as<float>(), lossA.as<float>(), nullptr, 0));
      KLT_HIP(hipMemcpyAsync(h_g.data(), gsrc.ptr, h_g.size() * 4, hipMemcpyDeviceToHost,
                             ctx->stream));
      KLT_LAUNCH(ctx, "k_smo_absmax", k_smo_absmax, dim3(kRedBlocks, nmask), dim3(256), 0,
                 grad.as<float>(), px, mx.as<float>());
      KLT_LAUNCH(ctx, "k_smo_mask_step", k_smo_mask_step, dim3(blocks, nmask), dim3(256), 0,
                 masks, grad.as<float>(), px, mx.as<float>(), step_mask);
      KLT_HIP(hipGetLastError());
    }
    KLT_HIP(hipStreamSynchronize(ctx->stream));
    double L = 0.0;
    for (float v : h_loss) L += v;
    history[it] = L;
    if (mismatch) {
      long long c = 0;
      for (unsigned v : h_cnt) c += v;
      mismatch[it] = c;
    }
    if (it == steps) break;
    // source step on the host (srcSize^2 values)
    std::fill(gs.begin(), gs.end(), 0.f);
    for (int b = 0; b < nmask; ++b)
      for (int i = 0; i < n2; ++i) gs[i] += h_g[(size_t)b * n2 + i];
    float gmax = 0.f, smax = 0.f;
    for (int i = 0; i < n2; ++i) {
      gmax = std::max(gmax, std::fabs(gs[i]));
      smax = std::max(smax, src[i]);
    }
    if (gmax > 0.f && smax > 0.f) {
      double sum = 0.0;
      for (int q = 0; q < srcSize; ++q)
        for (int p = 0; p < srcSize; ++p) {
          const int i = q * srcSize + p, dp = p - sh, dq = q - sh;
          float v = 0.f;
          if (dp * dp + dq * dq <= sh * sh)
            v = std::max(0.f, src[i] - step_src * smax * gs[i] / gmax);
          src[i] = v;
          sum += v;
        }
      if (sum > 0.0)
        for (int i = 0; i < n2; ++i) src[i] = (float)(src[i] / sum);
    }
  }
  return KLT_OK;
}

}  // extern "C"
