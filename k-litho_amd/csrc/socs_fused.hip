// socs_fused.hip -- batched SOCS forward for kernel half-size N <= 15 on
// MI355X.  Same mathematics as socs.hip; the two full-field passes run on the
// matrix cores and the coarse grid runs as register-resident radix-8 x 8 FFTs:
//
//   k_band_mfma   R[y][n] = sum_x m[y][x] W[n][x]: the row DFT of the mask as a
//                 GEMM (rows x 1024 x 32) on v_mfma_f32_16x16x32_f16 with the
//                 mask and the twiddles split into hi + lo halves (three
//                 products, FP32 accumulation); the epilogue contracts each
//                 128-row block with the column twiddles, giving per-block
//                 partial sums of the (2N+1) x (N+1) mask-spectrum band.
//   k_socs_mid    one block per image: sums the partials, forms A_k = K_k M^
//                 on a 64 x 64 grid, inverse-FFTs it (8 threads per 64-point
//                 transform, radix 8 x 8, exchange through LDS), accumulates
//                 lambda_k |E_k|^2 in registers over all kernels, and forward-
//                 transforms the coarse intensity to its |g| <= 2N spectrum.
//                 FP32 throughout.
//   k_synth_mfma  per 128-column strip: P[j][x] = sum_gx J[gy][gx] e^{+i gx x}
//                 in FP32, split to hi + lo halves in LDS, then
//                 I = TY (rows x 64) . P (64 x 128) on the matrix cores.
// 64 >= 4N+1 for N <= 15, so the coarse grid is exact (no aliasing of the
// |g| <= 2N intensity band).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "klt_internal.h"

namespace {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kM = 64;        // coarse grid
constexpr int kRowsBand = 128;  // mask rows per band block
constexpr int kColsSynth = 128; // image columns per synthesis block
constexpr int kPStride = 72;    // LDS row stride (halves) of the P operand

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ void split8(const float* v, half8_t& hi, half8_t& lo) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const _Float16 h = (_Float16)v[t];
    hi[t] = h;
    lo[t] = (_Float16)(v[t] - (float)h);
  }
}

__device__ __forceinline__ float4_t mfma3(half8_t ah, half8_t al, half8_t bh, half8_t bl,
                                          float4_t c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, c, 0, 0, 0);
  return c;
}

// 8 consecutive floats of a mask row starting at x (zero beyond Lx / for
// rows outside the image).
__device__ __forceinline__ void load8(const float* row, int x, int Lx, bool valid, bool vec,
                                      float* v) {
  if (valid && vec && x + 8 <= Lx) {
    const float4 a = *reinterpret_cast<const float4*>(row + x);
    const float4 b = *reinterpret_cast<const float4*>(row + x + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = (valid && x + t < Lx) ? row[x + t] : 0.f;
  }
}

// ---------------------------------------------------------------------------
// Stage 1.  grid (ceil(Ly/128), B), 256 threads.  Wave w owns rows
// blk*128 + 32w .. +32 (two 16-row MFMA blocks) and all 32 output columns
// (n = 2 nx: cos part, n = 2 nx + 1: -sin part, nx <= 15).
// mpart: [B][nblk][2Ny+1][16] float2, entry (ny, nx) for nx in [0, Nx].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_band_mfma(const float* __restrict__ mask, int Lx,
                                                   int Ly, int Kpad,
                                                   const _Float16* __restrict__ Wh,
                                                   const _Float16* __restrict__ Wl,
                                                   const float2* __restrict__ twy, int NP,
                                                   int Nx, int Ny,
                                                   float2* __restrict__ mpart) {
  __shared__ float sR[kRowsBand][33];
  const int blk = blockIdx.x, b = blockIdx.y, nblk = gridDim.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int ybase = blk * kRowsBand + wave * 32;
  const float* img = mask + (size_t)b * Lx * Ly;
  const int ya = ybase + r16, yb = ybase + 16 + r16;
  const bool va = ya < Ly, vb = yb < Ly;
  const float* ra = img + (size_t)(va ? ya : 0) * Lx;
  const float* rb = img + (size_t)(vb ? yb : 0) * Lx;
  const bool vec = (Lx & 3) == 0;
  float4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = float4_t{0.f, 0.f, 0.f, 0.f};
  const _Float16* w0h = Wh + (size_t)r16 * Kpad + kq;
  const _Float16* w0l = Wl + (size_t)r16 * Kpad + kq;
  const _Float16* w1h = Wh + (size_t)(16 + r16) * Kpad + kq;
  const _Float16* w1l = Wl + (size_t)(16 + r16) * Kpad + kq;
  for (int k = 0; k < Kpad; k += 32) {
    float va8[8], vb8[8];
    load8(ra, k + kq, Lx, va, vec, va8);
    load8(rb, k + kq, Lx, vb, vec, vb8);
    half8_t ah, al, bh, bl;
    split8(va8, ah, al);
    split8(vb8, bh, bl);
    const half8_t c0h = *reinterpret_cast<const half8_t*>(w0h + k);
    const half8_t c0l = *reinterpret_cast<const half8_t*>(w0l + k);
    const half8_t c1h = *reinterpret_cast<const half8_t*>(w1h + k);
    const half8_t c1l = *reinterpret_cast<const half8_t*>(w1l + k);
    acc[0][0] = mfma3(ah, al, c0h, c0l, acc[0][0]);
    acc[0][1] = mfma3(ah, al, c1h, c1l, acc[0][1]);
    acc[1][0] = mfma3(bh, bl, c0h, c0l, acc[1][0]);
    acc[1][1] = mfma3(bh, bl, c1h, c1l, acc[1][1]);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        sR[wave * 32 + a * 16 + (lane >> 4) * 4 + r][c * 16 + r16] = acc[a][c][r];
  __syncthreads();
  // partial band: sum over this block's rows of R[y][nx] e^{-2 pi i ny (y-cy)/Ly}
  const int Wy = 2 * Ny + 1, nxp = Nx + 1;
  const int y0 = blk * kRowsBand;
  const int nrow = min(kRowsBand, Ly - y0);
  float2* out = mpart + ((size_t)b * nblk + blk) * Wy * 16;
  for (int o = threadIdx.x; o < Wy * nxp; o += 256) {
    const int nyi = o / nxp, nx = o - nyi * nxp;
    const int ny = nyi - Ny;
    const int any = ny < 0 ? -ny : ny;
    float sr = 0.f, si = 0.f;
    for (int yy = 0; yy < nrow; ++yy) {
      float2 w = twy[(size_t)(y0 + yy) * NP + any];
      if (ny < 0) w.y = -w.y;
      const float rr = sR[yy][2 * nx], ri = sR[yy][2 * nx + 1];
      sr += rr * w.x - ri * w.y;
      si += rr * w.y + ri * w.x;
    }
    out[nyi * 16 + nx] = make_float2(sr, si);
  }
}

// 8-point DFT, X[k] = sum_n x[n] e^{s 2 pi i nk/8}, s = +1 (inverse) or -1.
template <int S>
__device__ __forceinline__ void dft8(float2* x) {
  constexpr float R = 0.70710678118654752440f;
  float2 a[8];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    a[i] = make_float2(x[i].x + x[i + 4].x, x[i].y + x[i + 4].y);
    a[i + 4] = make_float2(x[i].x - x[i + 4].x, x[i].y - x[i + 4].y);
  }
  // a5 *= (R, S R); a6 *= S i; a7 *= (-R, S R)
  {
    const float2 v = a[5];
    a[5] = make_float2(R * v.x - S * R * v.y, S * R * v.x + R * v.y);
    const float2 u = a[6];
    a[6] = make_float2(-S * u.y, S * u.x);
    const float2 q = a[7];
    a[7] = make_float2(-R * q.x - S * R * q.y, S * R * q.x - R * q.y);
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const float2* b = a + 4 * h;
    const float2 c0 = make_float2(b[0].x + b[2].x, b[0].y + b[2].y);
    const float2 c2 = make_float2(b[0].x - b[2].x, b[0].y - b[2].y);
    const float2 c1 = make_float2(b[1].x + b[3].x, b[1].y + b[3].y);
    const float2 d = make_float2(b[1].x - b[3].x, b[1].y - b[3].y);
    const float2 c3 = make_float2(-S * d.y, S * d.x);
    x[0 + h] = make_float2(c0.x + c1.x, c0.y + c1.y);
    x[2 + h] = make_float2(c2.x + c3.x, c2.y + c3.y);
    x[4 + h] = make_float2(c0.x - c1.x, c0.y - c1.y);
    x[6 + h] = make_float2(c2.x - c3.x, c2.y - c3.y);
  }
}

// One 64-point transform per group of 8 threads (64 = 8 x 8).  On entry
// thread t holds x[t + 8 n2] in v[n2]; on exit it holds X[t + 8 k1] in v[k1].
// tw[k2] = e^{S 2 pi i t k2 / 64}.  xs: the group's 64-entry LDS scratch.
// Every thread of the block must call it (it synchronises the block).
template <int S>
__device__ __forceinline__ void fft64_group(float2* v, const float2* tw, float2* xs, int t,
                                            bool active) {
  if (active) {
    dft8<S>(v);
#pragma unroll
    for (int k2 = 0; k2 < 8; ++k2) xs[t * 8 + k2] = cmulf(v[k2], tw[k2]);
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int n1 = 0; n1 < 8; ++n1) v[n1] = xs[n1 * 8 + t];
    dft8<S>(v);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Stage 2.  grid (B), 256 threads = 32 groups of 8.
// spec: [B][32][4Nx+1], row gy in [0, 2Ny] (rows above zero), entry gx + 2Nx.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_socs_mid(const float2* __restrict__ mpart, int nblk,
                                                  int Nx, int Ny, float inv_area,
                                                  const float2* __restrict__ kern,
                                                  const float* __restrict__ scales, int nk,
                                                  const float2* __restrict__ twM,
                                                  float2* __restrict__ spec) {
  __shared__ float2 sM[31 * 31];
  __shared__ float2 sB[31 * kM];
  __shared__ float2 sX[32 * 64];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int g = tid >> 3, t = tid & 7;
  const int Wx = 2 * Nx + 1, Wy = 2 * Ny + 1, T = Wx * Wy;
  const int Gx = 4 * Nx + 1;
  // mask-spectrum band from the per-block partials (Hermitian for nx < 0)
  for (int i = tid; i < T; i += 256) {
    const int nyi = i / Wx, nx = i - nyi * Wx - Nx;
    const int ny = nyi - Ny;
    const int r = nx >= 0 ? nyi : (-ny + Ny), c = nx >= 0 ? nx : -nx;
    float2 s = make_float2(0.f, 0.f);
    for (int k = 0; k < nblk; ++k) {
      const float2 p = mpart[(((size_t)b * nblk + k) * Wy + r) * 16 + c];
      s.x += p.x;
      s.y += p.y;
    }
    if (nx < 0) s.y = -s.y;
    sM[i] = make_float2(s.x * inv_area, s.y * inv_area);
  }
  float2 twi[8], twf[8];
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) {
    twi[k2] = twM[(t * k2) & 63];
    twf[k2] = make_float2(twi[k2].x, -twi[k2].y);
  }
  float acc[2][8];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[q][k] = 0.f;
  float2* xs = sX + g * 64;
  __syncthreads();
  for (int k = 0; k < nk; ++k) {
    const float2* K = kern + (size_t)k * T;
    const float lam = scales[k];
    // rows: B[r][u] = sum_nx A[r][nx] e^{+2 pi i nx u / 64}
    {
      const bool act = g < Wy;
      float2 v[8];
      if (act) {
#pragma unroll
        for (int n2 = 0; n2 < 8; ++n2) {
          const int n = t + 8 * n2;
          const int nx = n < 32 ? n : n - 64;
          float2 val = make_float2(0.f, 0.f);
          if (nx >= -Nx && nx <= Nx) val = cmulf(K[g * Wx + nx + Nx], sM[g * Wx + nx + Nx]);
          v[n2] = val;
        }
      }
      fft64_group<1>(v, twi, xs, t, act);
      if (act) {
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) sB[g * kM + t + 8 * k1] = v[k1];
      }
    }
    __syncthreads();
    // columns: E[uy][ux] = sum_ny B[ny][ux] e^{+2 pi i ny uy / 64}
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int ux = g + 32 * q;
      float2 v[8];
#pragma unroll
      for (int n2 = 0; n2 < 8; ++n2) {
        const int n = t + 8 * n2;
        const int ny = n < 32 ? n : n - 64;
        v[n2] = (ny >= -Ny && ny <= Ny) ? sB[(ny + Ny) * kM + ux] : make_float2(0.f, 0.f);
      }
      fft64_group<1>(v, twi, xs, t, true);
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) acc[q][k1] += lam * (v[k1].x * v[k1].x + v[k1].y * v[k1].y);
    }
  }
  // forward transform of the coarse intensity.  Thread (g, t) holds column
  // ux = g + 32 q at rows uy = t + 8 k1: the input layout of a column FFT.
  const int G2 = 2 * Ny + 1;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int ux = g + 32 * q;
    float2 v[8];
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) v[k1] = make_float2(acc[q][k1], 0.f);
    fft64_group<-1>(v, twf, xs, t, true);
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) {
      const int gy = t + 8 * k1;
      if (gy < G2) sB[gy * kM + ux] = v[k1];
    }
  }
  __syncthreads();
  // rows gy in [0, 2Ny]: J[gy][gx] = 1/64^2 sum_ux C[gy][ux] e^{-2 pi i gx ux/64}
  const float inv = 1.0f / (float)(kM * kM);
  float2* out = spec + (size_t)b * 32 * Gx;
  {
    const bool act = g < G2;
    float2 v[8];
    if (act) {
#pragma unroll
      for (int n2 = 0; n2 < 8; ++n2) v[n2] = sB[g * kM + t + 8 * n2];
    }
    fft64_group<-1>(v, twf, xs, t, act);
    if (act) {
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) {
        const int kk = t + 8 * k1;
        const int gx = kk < 32 ? kk : kk - 64;
        if (gx >= -2 * Nx && gx <= 2 * Nx)
          out[g * Gx + gx + 2 * Nx] = make_float2(v[k1].x * inv, v[k1].y * inv);
      }
    }
  }
  for (int i = G2 * Gx + tid; i < 32 * Gx; i += 256) out[i] = make_float2(0.f, 0.f);
}

// ---------------------------------------------------------------------------
// Stage 3.  grid (ceil(Lx/128), B), 256 threads.
// P[x][j] (j = 2 gy: Re, 2 gy + 1: Im) for the block's 128 columns, scaled by
// a power of two and split to hi/lo halves in LDS; then wave w computes rows
// y = 16 ((w>>1) + 2 i) ..
// for the 64 columns (w & 1) of the strip, K = 64.
// TY: [Lypad][64] hi/lo, TY[y][2gy] = w_gy cos, TY[y][2gy+1] = -w_gy sin.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_synth_mfma(const float2* __restrict__ spec, int Nx,
                                                    int Lx, int Ly, int Lypad,
                                                    const float2* __restrict__ twx,
                                                    const _Float16* __restrict__ TYh,
                                                    const _Float16* __restrict__ TYl,
                                                    float* __restrict__ image) {
  __shared__ float2 sJ[32 * 61];
  __shared__ __attribute__((aligned(16))) _Float16 sPh[kColsSynth * kPStride];
  __shared__ __attribute__((aligned(16))) _Float16 sPl[kColsSynth * kPStride];
  const int b = blockIdx.y, x0 = blockIdx.x * kColsSynth, tid = threadIdx.x;
  const int Gx = 4 * Nx + 1, Hx = 2 * Nx + 1;
  __shared__ float red[4];
  float part = 0.f;
  for (int i = tid; i < 32 * Gx; i += 256) {
    const float2 v = spec[(size_t)b * 32 * Gx + i];
    sJ[i] = v;
    part += fabsf(v.x) + fabsf(v.y);
  }
  // Power-of-two scale putting max |P| <= sum |J| near 2^12, so that the lo
  // half of the split stays a normal fp16 number whatever the image's level.
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
  if ((tid & 63) == 0) red[tid >> 6] = part;
  __syncthreads();
  const float bound = red[0] + red[1] + red[2] + red[3];
  int ex = 0;
  (void)frexpf(bound > 0.f ? bound : 1.f, &ex);
  const float pscale = ldexpf(1.0f, 12 - ex), inv_pscale = ldexpf(1.0f, ex - 12);
  {
    const int c = tid & (kColsSynth - 1), half = tid >> 7;
    const int x = x0 + c;
    float2 acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = make_float2(0.f, 0.f);
    if (x < Lx) {
      const float2* tx = twx + (size_t)x * Hx;
      for (int gx = -2 * Nx; gx <= 2 * Nx; ++gx) {
        float2 w = tx[gx < 0 ? -gx : gx];
        if (gx < 0) w.y = -w.y;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const float2 j = sJ[(half * 16 + q) * Gx + gx + 2 * Nx];
          acc[q].x += j.x * w.x - j.y * w.y;
          acc[q].y += j.x * w.y + j.y * w.x;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int j = 2 * (half * 16 + q);
      acc[q].x *= pscale;
      acc[q].y *= pscale;
      const _Float16 hr = (_Float16)acc[q].x, hi = (_Float16)acc[q].y;
      sPh[c * kPStride + j] = hr;
      sPl[c * kPStride + j] = (_Float16)(acc[q].x - (float)hr);
      sPh[c * kPStride + j + 1] = hi;
      sPl[c * kPStride + j + 1] = (_Float16)(acc[q].y - (float)hi);
    }
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int ch = wave & 1, rp = wave >> 1;
  half8_t bh[4][2], bl[4][2];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int c = ch * 64 + cb * 16 + r16;
      bh[cb][ks] = *reinterpret_cast<const half8_t*>(sPh + c * kPStride + ks * 32 + kq);
      bl[cb][ks] = *reinterpret_cast<const half8_t*>(sPl + c * kPStride + ks * 32 + kq);
    }
  float* img = image + (size_t)b * Lx * Ly;
  for (int yb = rp * 16; yb < Lypad; yb += 32) {
    const _Float16* th = TYh + (size_t)(yb + r16) * 64 + kq;
    const _Float16* tl = TYl + (size_t)(yb + r16) * 64 + kq;
    const half8_t ah0 = *reinterpret_cast<const half8_t*>(th);
    const half8_t ah1 = *reinterpret_cast<const half8_t*>(th + 32);
    const half8_t al0 = *reinterpret_cast<const half8_t*>(tl);
    const half8_t al1 = *reinterpret_cast<const half8_t*>(tl + 32);
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      float4_t d = float4_t{0.f, 0.f, 0.f, 0.f};
      d = mfma3(ah0, al0, bh[cb][0], bl[cb][0], d);
      d = mfma3(ah1, al1, bh[cb][1], bl[cb][1], d);
      const int x = x0 + ch * 64 + cb * 16 + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int y = yb + (lane >> 4) * 4 + r;
        if (y < Ly && x < Lx) img[(size_t)y * Lx + x] = d[r] * inv_pscale;
      }
    }
  }
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

void split_half(double v, _Float16* hi, _Float16* lo) {
  const _Float16 h = (_Float16)(float)v;
  *hi = h;
  *lo = (_Float16)(float)(v - (double)(float)h);
}

}  // namespace

namespace klt {

static int get_fused_tables(klt_ctx* ctx, SocsPlan* p, FusedTables** out) {
  auto key = std::make_tuple(p->Lx, p->Ly, p->Nx, p->Ny);
  auto it = ctx->fused_tables.find(key);
  if (it != ctx->fused_tables.end()) {
    *out = it->second.get();
    return KLT_OK;
  }
  auto t = std::make_unique<FusedTables>();
  const int Lx = p->Lx, Ly = p->Ly, Nx = p->Nx, Ny = p->Ny;
  t->Kpad = round_up(Lx, 32);
  t->Lypad = round_up(Ly, 32);
  const int cix = Lx / 2, coy = (Ly + 1) / 2;
  std::vector<_Float16> h((size_t)32 * t->Kpad), l((size_t)32 * t->Kpad);
  for (int n = 0; n < 32; ++n)
    for (int x = 0; x < t->Kpad; ++x) {
      double v = 0.0;
      const int nx = n / 2;
      if (nx <= Nx && x < Lx) {
        long long r = ((long long)nx * (x - cix)) % Lx;
        if (r < 0) r += Lx;
        const double a = 2.0 * M_PI * (double)r / (double)Lx;
        v = (n & 1) ? -std::sin(a) : std::cos(a);
      }
      split_half(v, &h[(size_t)n * t->Kpad + x], &l[(size_t)n * t->Kpad + x]);
    }
  KLT_TRY(t->Wh.ensure(h.size() * sizeof(_Float16)));
  KLT_TRY(t->Wl.ensure(l.size() * sizeof(_Float16)));
  KLT_HIP(hipMemcpy(t->Wh.ptr, h.data(), h.size() * sizeof(_Float16), hipMemcpyHostToDevice));
  KLT_HIP(hipMemcpy(t->Wl.ptr, l.data(), l.size() * sizeof(_Float16), hipMemcpyHostToDevice));
  h.assign((size_t)t->Lypad * 64, (_Float16)0.f);
  l.assign((size_t)t->Lypad * 64, (_Float16)0.f);
  for (int y = 0; y < Ly; ++y)
    for (int gy = 0; gy <= 2 * Ny; ++gy) {
      long long r = ((long long)gy * (y - coy)) % Ly;
      if (r < 0) r += Ly;
      const double a = 2.0 * M_PI * (double)r / (double)Ly;
      const double w = gy == 0 ? 1.0 : 2.0;
      split_half(w * std::cos(a), &h[(size_t)y * 64 + 2 * gy], &l[(size_t)y * 64 + 2 * gy]);
      split_half(-w * std::sin(a), &h[(size_t)y * 64 + 2 * gy + 1],
                 &l[(size_t)y * 64 + 2 * gy + 1]);
    }
  KLT_TRY(t->TYh.ensure(h.size() * sizeof(_Float16)));
  KLT_TRY(t->TYl.ensure(l.size() * sizeof(_Float16)));
  KLT_HIP(hipMemcpy(t->TYh.ptr, h.data(), h.size() * sizeof(_Float16), hipMemcpyHostToDevice));
  KLT_HIP(hipMemcpy(t->TYl.ptr, l.data(), l.size() * sizeof(_Float16), hipMemcpyHostToDevice));
  std::vector<float2> tw(kM);
  for (int j = 0; j < kM; ++j) {
    const double a = 2.0 * M_PI * j / kM;
    tw[j] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  KLT_TRY(t->tw64.ensure(kM * sizeof(float2)));
  KLT_HIP(hipMemcpy(t->tw64.ptr, tw.data(), kM * sizeof(float2), hipMemcpyHostToDevice));
  *out = t.get();
  ctx->fused_tables[key] = std::move(t);
  return KLT_OK;
}

bool socs_fused_supported(const klt_kernels* kb) { return kb->Nx <= 15 && kb->Ny <= 15; }

int socs_forward_fused(klt_ctx* ctx, const klt_kernels* kb, const float* masks, int B, int Lx,
                       int Ly, float* images) {
  SocsPlan* p = nullptr;
  KLT_TRY(get_socs_plan(ctx, Lx, Ly, kb->Nx, kb->Ny, &p));
  FusedTables* ft = nullptr;
  KLT_TRY(get_fused_tables(ctx, p, &ft));
  const int Nx = kb->Nx, Ny = kb->Ny, Wy = 2 * Ny + 1, Gx = 4 * Nx + 1;
  const int nblk = (Ly + kRowsBand - 1) / kRowsBand;
  const int SB = std::min(B, 1024);
  KLT_TRY(ctx->qpart.ensure((size_t)SB * nblk * Wy * 16 * sizeof(float2)));
  KLT_TRY(ctx->spec.ensure((size_t)SB * 32 * Gx * sizeof(float2)));
  const float inv_area = 1.0f / ((float)Lx * (float)Ly);
  for (int b0 = 0; b0 < B; b0 += SB) {
    const int nb = std::min(SB, B - b0);
    const float* m = masks + (size_t)b0 * Lx * Ly;
    float* out = images + (size_t)b0 * Lx * Ly;
    KLT_LAUNCH(ctx, "k_band_mfma", k_band_mfma, dim3(nblk, nb), dim3(256), 0, m, Lx, Ly,
               ft->Kpad, ft->Wh.as<_Float16>(), ft->Wl.as<_Float16>(),
               p->twy_in.as<float2>(), p->NP, Nx, Ny, ctx->qpart.as<float2>());
    KLT_LAUNCH(ctx, "k_socs_mid", k_socs_mid, dim3(nb), dim3(256), 0, ctx->qpart.as<float2>(),
               nblk, Nx, Ny, inv_area, kb->kern.as<float2>(), kb->scales.as<float>(), kb->nk,
               ft->tw64.as<float2>(), ctx->spec.as<float2>());
    KLT_LAUNCH(ctx, "k_synth_mfma", k_synth_mfma,
               dim3((Lx + kColsSynth - 1) / kColsSynth, nb), dim3(256), 0,
               ctx->spec.as<float2>(), Nx, Lx, Ly, ft->Lypad, p->twx_out.as<float2>(),
               ft->TYh.as<_Float16>(), ft->TYl.as<_Float16>(), out);
    KLT_HIP(hipGetLastError());
  }
  return KLT_OK;
}

}  // namespace klt
