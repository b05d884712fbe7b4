// abbe.hip -- gradient of the resist loss with respect to the source
// (no reference counterpart), in the Abbe form of the partially coherent
// image.
//
// With the kernels stored index-reversed (klitho_tcc.cpp:358-362), the image
// of all SOCS kernels equals I_A(x) = sum_s S_s |E_s(x)|^2 with
//   E_s(x) = sum_n A_s[n] e^{+2 pi i n.(x - c)/L},  A_s[n] = P(-n, s) M^[n],
// which is also the Hopkins image of the TCC of the same source.  For the
// loss L = sum_x (sigmoid(a (I_A - thr)) - T)^2, G = dL/dI_A and
//   dL/dS_s = sum_x G(x) |E_s(x)|^2
//           = Lx Ly sum_{n, n'} A_s[n] conj(A_s[n']) conj(G^[n - n'])
// with G^ the |g| <= 2N band of G (the expression is exact: |E_s|^2 is band
// limited to 2N).  The gradient is taken with respect to the source weights as
// given (after createSource's normalisation), for every grid point inside the
// unit disc (S_s = 0 included).
//
//   1. mask bands by the matrix-core band pass, M^ (2N+1)^2 per mask
//      (k_abbe_assemble), queued before the host reads a device source;
//   2. TCC of the source (split-FP16 matrix-core Gram: within 2e-7 of the
//      FP64 Gram in relative Frobenius norm, far inside the gradient's
//      tolerance), and the Hopkins image I_A of each mask from M^ (one wave
//      per spectrum entry, then the matrix-core synthesis);
//   3. resist bands (G, loss partials) by the band pass, then per mask the
//      loss and the coarse-grid g_c of G^ (resist_coarse_grid);
//   4. the double sum on the 64 x 64 coarse grid (abbe_grad_coarse in
//      socs_fused.hip): Lx Ly sum_{n, n'} A_s[n] conj(A_s[n']) conj(G^[n - n'])
//      = Lx Ly / 64^2 sum_u g_c(u) |e_s(u)|^2 with e_s, g_c the 64-point
//      transforms of A_s and G^, one 2D transform per source point (two
//      points per pass, the pupil values from a table built once per call).
// Kernel half-sizes 16..31 take the wide passes of socs_fused.hip: the band
// with 64 twiddle rows, the Hopkins image formed tile by tile inside the fused
// resist band (k_synth_resist_wide, stored only when asked for), and steps 3-4
// on the 128 x 128 coarse grid (Lx Ly / 128^2 sum_u g_c |e_s|^2).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "klt_internal.h"
#include "klt_pupil.h"

namespace {

constexpr int kThreads = 256;

// Per mask: M^ (2N+1)^2 from the mask-band partials.
__global__ __launch_bounds__(kThreads) void k_abbe_assemble(const float2* __restrict__ mpart,
                                                            int nblk, int Nx, int Ny,
                                                            float inv_area,
                                                            float2* __restrict__ Mfull) {
  const int b = blockIdx.x;
  const int Wx = 2 * Nx + 1, Wy = 2 * Ny + 1;
  for (int i = threadIdx.x; i < Wx * Wy; i += kThreads) {
    const int nyi = i / Wx, nx = i - nyi * Wx - Nx, ny = nyi - Ny;
    const int r = nx >= 0 ? nyi : (-ny + Ny), c = nx >= 0 ? nx : -nx;
    float2 s = make_float2(0.f, 0.f);
#pragma unroll 8
    for (int k = 0; k < nblk; ++k) {
      const float2 p = mpart[(((size_t)b * nblk + k) * Wy + r) * 16 + c];
      s.x += p.x;
      s.y += p.y;
    }
    Mfull[(size_t)b * Wx * Wy + i] = make_float2(s.x * inv_area, (nx < 0 ? -s.y : s.y) * inv_area);
  }
}

__global__ void k_fill_zero(float* __restrict__ x, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.f;
}

}  // namespace

extern "C" {

int klt_abbe_source_grad(klt_ctx* ctx, const float* masks, const float* targets, int nmask,
                         int Lx, int Ly, const float* src, int srcSize, float lambda, float NA,
                         float defocus, float nmed, int Nx, int Ny, float a_sig, float thr,
                         float* grad_src, float* losses, float* images, unsigned flags) {
  if (!ctx || !masks || !targets || !src || !grad_src || !losses)
    return klt::set_error(KLT_ERR_ARG, "klt_abbe_source_grad: NULL argument");
  if (nmask <= 0 || Lx <= 0 || Ly <= 0 || srcSize <= 0 || Nx < 0 || Ny < 0)
    return klt::set_error(KLT_ERR_ARG, "klt_abbe_source_grad: bad size");
  if (Nx > 31 || Ny > 31)
    return klt::set_error(KLT_ERR_ARG, "klt_abbe_source_grad: half-size N above 31");
  // half-sizes 16..31: the 128-point coarse grid and the wide passes
  const bool wide = Nx > 15 || Ny > 15;
  if (4 * Nx + 1 > Lx || 4 * Ny + 1 > Ly) return klt::set_error(KLT_ERR_SIZE, "ERROR in FI()");
  if (NA < 0 || !(nmed >= 1.0f) || NA > nmed)
    return klt::set_error(KLT_ERR_ARG, "klt_abbe_source_grad: NA must not exceed the medium index");
  KLT_TRY(ctx->activate());
  hipStream_t st = ctx->stream;
  const size_t npix = (size_t)Lx * Ly, bytes = (size_t)nmask * npix * sizeof(float);
  const size_t ns2 = (size_t)srcSize * srcSize;
  std::vector<float> hsrc(ns2);
  const float *dm = masks, *dt = targets;
  if (flags & KLT_HOST_PTRS) {
    std::memcpy(hsrc.data(), src, ns2 * sizeof(float));
    // own staging buffer: the TCC build below uses the context's stage_in
    KLT_TRY(ctx->abbe_in.ensure(2 * bytes));
    char* s = ctx->abbe_in.as<char>();
    KLT_HIP(hipMemcpyAsync(s, masks, bytes, hipMemcpyHostToDevice, st));
    KLT_HIP(hipMemcpyAsync(s + bytes, targets, bytes, hipMemcpyHostToDevice, st));
    dm = reinterpret_cast<const float*>(s);
    dt = reinterpret_cast<const float*>(s + bytes);
  }
  // a device source is read into pinned memory and waited for with an event
  // after the mask band pass (which does not need it) has been queued, so that
  // the band pass runs while the host prepares the TCC build
  hipEvent_t src_read = nullptr;
  void* src_pinned = nullptr;
  if (!(flags & KLT_HOST_PTRS)) {
    KLT_TRY(ctx->download_area(ns2 * sizeof(float), &src_pinned));
    KLT_HIP(hipMemcpyAsync(src_pinned, src, ns2 * sizeof(float), hipMemcpyDeviceToHost, st));
    src_read = ctx->take_event();
    KLT_HIP(hipEventRecord(src_read, st));
  }
  const int Wx = 2 * Nx + 1, Wy = 2 * Ny + 1, T = Wx * Wy;
  const int Gy = 4 * Ny + 1;
  // band partials: row blocks of 128 rows, each split into ks column ranges
  const int ks = klt::band_split(Lx, Ly, nmask);
  const int nblk = (Ly + 127) / 128 * ks;
  // outputs (device)
  float *dg = grad_src, *dl = losses, *di = images;
  if (flags & KLT_HOST_PTRS) {
    KLT_TRY(ctx->stage_out.ensure((size_t)nmask * ns2 * 4 + (size_t)nmask * 4 + 512));
    dg = ctx->stage_out.as<float>();
    dl = dg + (size_t)nmask * ns2 + 64;
    di = nullptr;
  }
  // (the wide passes form the image inside the fused resist band and write it
  // only when the caller asks for it)
  float* I = di;
  if (!I && (!wide || images)) {
    KLT_TRY(ctx->bwd_img.ensure(bytes));
    I = ctx->bwd_img.as<float>();
  }
  // 1. mask bands (matrix-core band pass)
  KLT_TRY(ctx->abbe_band.ensure((size_t)nmask * T * sizeof(float2)));
  float2* Mfull = ctx->abbe_band.as<float2>();
  const float inv_area = 1.0f / ((float)Lx * (float)Ly);
  if (wide) {
    KLT_TRY(klt::band_mask_full_wide(ctx, Lx, Ly, Nx, Ny, dm, nmask, Mfull));
  } else {
    KLT_TRY(ctx->qpart.ensure((size_t)nmask * nblk * Wy * 16 * sizeof(float2)));
    KLT_TRY(ctx->gpart.ensure((size_t)nmask * nblk * Gy * 32 * sizeof(float2)));
    KLT_TRY(ctx->lpart.ensure((size_t)nmask * nblk * sizeof(float)));
    KLT_TRY(klt::band_mask_partials(ctx, Lx, Ly, Nx, Ny, dm, nmask, ks, ctx->qpart.as<float2>()));
    KLT_LAUNCH(ctx, "k_abbe_assemble", k_abbe_assemble, dim3(nmask), dim3(kThreads), 0,
               ctx->qpart.as<float2>(), nblk, Nx, Ny, inv_area, Mfull);
  }
  // 2. TCC of the source, then the Hopkins images from the bands
  if (src_read) {
    const hipError_t e = hipEventSynchronize(src_read);
    ctx->event_pool.push_back(src_read);
    KLT_HIP(e);
    std::memcpy(hsrc.data(), src_pinned, ns2 * sizeof(float));
  }
  float sigma = klt_outer_sigma(hsrc.data(), srcSize);
  if (sigma > 1.0f) sigma = 1.0f;
  KLT_TRY(ctx->abbe_tcc.ensure((size_t)T * T * 16));
  double* tcc = ctx->abbe_tcc.as<double>();
  KLT_TRY(klt::tcc_build_device(ctx, hsrc.data(), srcSize, sigma, lambda, NA, defocus, nmed, Lx,
                                Ly, Nx, Ny, KLT_TCC_SPLIT_F16, tcc));
  const int M = wide ? 128 : 64;  // coarse grid
  KLT_TRY(ctx->gcoarse.ensure((size_t)nmask * M * M * sizeof(float)));
  if (wide) {
    // 2b-3. the Hopkins image tile by tile, its resist band, the losses and g_c
    KLT_TRY(klt::abbe_resist_wide(ctx, Mfull, nmask, Lx, Ly, tcc, Nx, Ny, dt, a_sig, thr, I, dl,
                                  ctx->gcoarse.as<float>()));
  } else {
    KLT_TRY(klt::hopkins_images_fused(ctx, Mfull, nmask, Lx, Ly, tcc, Nx, Ny, I));
    // 3. resist bands and loss partials
    KLT_TRY(klt::band_resist_partials(ctx, Lx, Ly, Nx, Ny, I, dt, nmask, ks, a_sig, thr,
                                      ctx->gpart.as<float2>(), ctx->lpart.as<float>()));
  }
  // 4. per source point
  const int sh = (int)((srcSize - 1) / 2.0);
  std::vector<int2> pts;
  for (int q = -sh; q <= sh; ++q)
    for (int p = -sh; p <= sh; ++p)
      if (p * p + q * q <= sh * sh) pts.push_back(make_int2(p, q));
  KLT_TRY(ctx->misc.ensure(pts.size() * sizeof(int2)));
  KLT_TRY(ctx->upload(ctx->misc.ptr, pts.data(), pts.size() * sizeof(int2)));
  const size_t ng = (size_t)nmask * ns2;
  KLT_LAUNCH(ctx, "k_fill_zero", k_fill_zero, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0,
             dg, ng);
  const klt::PupilGeom g = klt::make_geom(srcSize, lambda, NA, defocus, nmed, Lx, Ly, Nx, Ny);
  if (wide) {
    KLT_TRY(klt::abbe_grad_coarse128(ctx, Mfull, ctx->gcoarse.as<float>(), nmask, Lx, Ly, g,
                                     ctx->misc.as<int2>(), (int)pts.size(), srcSize, dg));
  } else {
    KLT_TRY(klt::resist_coarse_grid(ctx, ctx->gpart.as<float2>(), ctx->lpart.as<float>(), nblk,
                                    nmask, Nx, Ny, inv_area, dl, ctx->gcoarse.as<float>()));
    KLT_TRY(klt::abbe_grad_coarse(ctx, Mfull, ctx->gcoarse.as<float>(), nmask, Lx, Ly, g,
                                  ctx->misc.as<int2>(), (int)pts.size(), srcSize, dg));
  }
  if (flags & KLT_HOST_PTRS) {
    KLT_HIP(hipMemcpyAsync(grad_src, dg, ng * 4, hipMemcpyDeviceToHost, st));
    KLT_HIP(hipMemcpyAsync(losses, dl, (size_t)nmask * 4, hipMemcpyDeviceToHost, st));
    if (images && I) KLT_HIP(hipMemcpyAsync(images, I, bytes, hipMemcpyDeviceToHost, st));
  }
  if (flags & KLT_HOST_PTRS) KLT_HIP(hipStreamSynchronize(st));  // the host outputs are complete
  return KLT_OK;
}

}  // extern "C"
