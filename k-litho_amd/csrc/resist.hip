// resist.hip -- the resist step after the aerial image on the GPU.
//
// The reference thresholds its image at TARGET_INTENSITY = 0.225
// (K-Litho-TCC/klitho_tcc.cpp:442, K-Litho-SOCS/klitho_socs.cpp:275) and
// writes the 7-level view image_s.png (klitho_tcc.cpp:547,
// klitho_socs.cpp:353; common/file_io.cpp:25-80): with
// v = clamp((I - min) / (max - min), 0, 1) in float and min = 0,
// max = TARGET_INTENSITY, each pixel falls in one of [0, 1/6), [1/6, 2/6),
// ..., [5/6, 1) or {1}; the last class, I >= TARGET_INTENSITY, is the printed
// resist.  k_resist_levels writes that class (0..6) per pixel as one byte: a
// read of 4 bytes and a write of 1 byte per pixel (HBM bound), 16 pixels per
// thread.  This file is compiled with -ffp-contract=off and divides with the
// correctly rounded FP32 division, so every class equals the reference's.
#include <algorithm>

#include "klt_internal.h"

namespace {

__device__ __forceinline__ unsigned level_of(float d, float mn, float mx, bool flat) {
  if (flat) return 3u;  // min == max: the reference's uniform green
  const float q = (d - mn) / (mx - mn);
  const float t = (q < 1.0f) ? q : 1.0f;  // std::min(1.0f, q)
  const double v = (0.0f < t) ? t : 0.0f;  // std::max(0.0f, t)
  return v < 1.0 / 6.0 ? 0u : v < 2.0 / 6.0 ? 1u : v < 3.0 / 6.0 ? 2u : v < 4.0 / 6.0 ? 3u
       : v < 5.0 / 6.0 ? 4u : v < 6.0 / 6.0 ? 5u : 6u;
}

// n pixels; thread t handles pixels 16 t .. 16 t + 15 (vector loads and one
// 16-byte store when the whole group is in range and aligned)
__global__ __launch_bounds__(256) void k_resist_levels(const float* __restrict__ img, size_t n,
                                                       float mn, float mx,
                                                       unsigned char* __restrict__ lev) {
  const bool flat = mn == mx;
  const size_t stride = (size_t)gridDim.x * blockDim.x * 16;
  const bool vec = (reinterpret_cast<uintptr_t>(img) & 15) == 0 &&
                   (reinterpret_cast<uintptr_t>(lev) & 15) == 0;
  for (size_t i0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 16; i0 < n; i0 += stride) {
    if (vec && i0 + 16 <= n) {
      unsigned w[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 a = reinterpret_cast<const float4*>(img + i0)[q];
        w[q] = level_of(a.x, mn, mx, flat) | (level_of(a.y, mn, mx, flat) << 8) |
               (level_of(a.z, mn, mx, flat) << 16) | (level_of(a.w, mn, mx, flat) << 24);
      }
      *reinterpret_cast<uint4*>(lev + i0) = make_uint4(w[0], w[1], w[2], w[3]);
    } else {
      for (size_t i = i0; i < n && i < i0 + 16; ++i)
        lev[i] = (unsigned char)level_of(img[i], mn, mx, flat);
    }
  }
}

}  // namespace

extern "C" {

int klt_resist_levels(klt_ctx* ctx, const float* images, int nimg, int Lx, int Ly, float min,
                      float max, unsigned char* levels, unsigned flags) {
  if (!ctx || !images || !levels || nimg <= 0 || Lx <= 0 || Ly <= 0)
    return klt::set_error(KLT_ERR_ARG, "klt_resist_levels: bad argument");
  KLT_TRY(ctx->activate());
  const size_t n = (size_t)nimg * Lx * Ly;
  const float* din = images;
  unsigned char* dout = levels;
  if (flags & KLT_HOST_PTRS) {
    KLT_TRY(ctx->stage_in.ensure(n * sizeof(float)));
    KLT_TRY(ctx->stage_out.ensure(n));
    KLT_HIP(hipMemcpyAsync(ctx->stage_in.ptr, images, n * sizeof(float), hipMemcpyHostToDevice,
                           ctx->stream));
    din = ctx->stage_in.as<float>();
    dout = ctx->stage_out.as<unsigned char>();
  }
  const size_t groups = (n + 15) / 16;
  const unsigned blocks = (unsigned)std::min<size_t>((groups + 255) / 256, 4096);
  KLT_LAUNCH(ctx, "k_resist_levels", k_resist_levels, dim3(blocks), dim3(256), 0, din, n, min,
             max, dout);
  KLT_HIP(hipGetLastError());
  if (flags & KLT_HOST_PTRS) {
    KLT_HIP(hipMemcpyAsync(levels, dout, n, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return KLT_OK;
}

}  // extern "C"
