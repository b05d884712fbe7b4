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
// read of 4 bytes and a write of 1 byte per pixel (HBM bound).  This file is
// compiled with -ffp-contract=off and divides with the correctly rounded FP32
// division, so every class equals the reference's.
//
// For max > min every step of the class (subtraction, division by a positive
// number, the clamps, the comparisons) is monotone in the pixel value, so the
// class of a pixel is the number of six thresholds t_1 .. t_6 at or below it,
// t_k the least float whose class is at least k.  The host finds each t_k by
// bisection over the ordered floats with the same float operations, and
// k_resist_thresholds compares (NaN, which the reference's clamps send to
// class 6, separately).  k_resist_levels keeps the per-pixel arithmetic for
// max < min and max == min.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "klt_internal.h"

namespace {

__host__ __device__ __forceinline__ unsigned level_of(float d, float mn, float mx, bool flat) {
  if (flat) return 3u;  // min == max: the reference's uniform green
  const float q = (d - mn) / (mx - mn);
  const float t = (q < 1.0f) ? q : 1.0f;  // std::min(1.0f, q)
  const double v = (0.0f < t) ? t : 0.0f;  // std::max(0.0f, t)
  return v < 1.0 / 6.0 ? 0u : v < 2.0 / 6.0 ? 1u : v < 3.0 / 6.0 ? 2u : v < 4.0 / 6.0 ? 3u
       : v < 5.0 / 6.0 ? 4u : v < 6.0 / 6.0 ? 5u : 6u;
}

// class by the thresholds (max > min)
__device__ __forceinline__ unsigned level_thr(float d, const float (&t)[6]) {
  unsigned c = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) c += d >= t[k] ? 1u : 0u;
  return d != d ? 6u : c;
}

// Chunks of 4096 pixels a block: thread x of the block takes the float4
// groups x, x + 256, x + 512, x + 768 of the chunk (each wave load reads
// 1 KB, each store writes 256 contiguous bytes).
__global__ __launch_bounds__(256) void k_resist_thresholds(const float* __restrict__ img,
                                                           size_t n, float t0, float t1,
                                                           float t2, float t3, float t4,
                                                           float t5,
                                                           unsigned char* __restrict__ lev) {
  const float t[6] = {t0, t1, t2, t3, t4, t5};
  const bool vec = (reinterpret_cast<uintptr_t>(img) & 15) == 0 &&
                   (reinterpret_cast<uintptr_t>(lev) & 3) == 0;
  for (size_t c0 = (size_t)blockIdx.x * 4096; c0 < n; c0 += (size_t)gridDim.x * 4096) {
    if (vec && c0 + 4096 <= n) {
      float4 a[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        a[q] = reinterpret_cast<const float4*>(img + c0)[q * 256 + threadIdx.x];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        reinterpret_cast<unsigned*>(lev + c0)[q * 256 + threadIdx.x] =
            level_thr(a[q].x, t) | (level_thr(a[q].y, t) << 8) | (level_thr(a[q].z, t) << 16) |
            (level_thr(a[q].w, t) << 24);
    } else {
      for (size_t i = c0 + threadIdx.x; i < n && i < c0 + 4096; i += 256)
        lev[i] = (unsigned char)level_thr(img[i], t);
    }
  }
}

// The least float (in the order of the floats, NaNs excluded) whose class is
// at least k, 1 <= k <= 6, for mn < mx: +inf has class 6, so it exists.
float level_threshold(int k, float mn, float mx) {
  auto key = [](float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  };
  auto val = [](uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
  };
  uint32_t lo = key(-INFINITY), hi = key(INFINITY);
  if (level_of(val(lo), mn, mx, false) >= (unsigned)k) return val(lo);
  while (hi - lo > 1) {  // level(lo) < k <= level(hi)
    const uint32_t mid = lo + (hi - lo) / 2;
    if (level_of(val(mid), mn, mx, false) >= (unsigned)k)
      hi = mid;
    else
      lo = mid;
  }
  return val(hi);
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
  if (max > min) {
    float t[6];
    for (int k = 1; k <= 6; ++k) t[k - 1] = level_threshold(k, min, max);
    const unsigned blocks = (unsigned)std::min<size_t>((n + 4095) / 4096, 8192);
    KLT_LAUNCH(ctx, "k_resist_levels", k_resist_thresholds, dim3(blocks), dim3(256), 0, din, n,
               t[0], t[1], t[2], t[3], t[4], t[5], dout);
  } else {
    const size_t groups = (n + 15) / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((groups + 255) / 256, 4096);
    KLT_LAUNCH(ctx, "k_resist_levels", k_resist_levels, dim3(blocks), dim3(256), 0, din, n, min,
               max, dout);
  }
  KLT_HIP(hipGetLastError());
  if (flags & KLT_HOST_PTRS) {
    KLT_HIP(hipMemcpyAsync(levels, dout, n, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return KLT_OK;
}

}  // extern "C"
