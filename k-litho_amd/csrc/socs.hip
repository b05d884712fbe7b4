// socs.hip -- SOCS aerial image on MI355X (replaces klitho_socs.cpp:71-241).
//
// The reference takes a full L x L FFT of the mask (FT_r2c), multiplies each
// kernel into a (4N+1)^2 grid, inverse-FFTs it, accumulates lambda_k |E_k|^2
// and Fourier-interpolates the coarse intensity back to L x L (FI).  Only the
// (2N+1)^2 band of the mask spectrum is ever used and the intensity is band
// limited to |g| <= 2N, so this implementation:
//   1. k_band_cols / k_band_rows: computes just that band of the mask
//      spectrum with two separable DFT passes (one read of the mask);
//   2. per image, forms A_k = K_k * M^ on an M x M grid (M = power of two
//      >= 4N+1, which is exact: |g| <= 2N never aliases), inverse-transforms
//      it, accumulates lambda_k |E_k|^2, and forward-transforms the coarse
//      intensity to its |g| <= 2N spectrum: for N <= 15 with the 64-point
//      LDS FFTs of socs_fused.hip (k_socs_mid_band), above that with direct
//      transforms, in one block per image while they fit in LDS
//      (k_socs_coarse, N <= 18), else in four launches over device scratch
//      (k_coarse_b / _int / _rx / _gy, N <= 31);
//   3. k_synth_cols / k_synth_rows: evaluates the band-limited intensity at
//      every pixel with two separable passes (one write of the image).
// HBM traffic per image is one mask read and one image write plus small
// per-image intermediates.
#include <algorithm>
#include <cstring>

#include "klt_internal.h"

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---------------------------------------------------------------------------
// 1a. column pass: Q[ny][x] = sum_y m[y][x] e^{-2 pi i ny (y-cy)/Ly}, ny in
// [0, NP) (table rows beyond Ny are zero).  Each thread owns two columns;
// a block covers 512 columns and `rows` rows (blockIdx.y picks the chunk).
// Output: qpart[b][chunk][NP][Lx].
// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(kThreads) void k_band_cols(const float* __restrict__ mask,
                                                        int Lx, int Ly, int rows,
                                                        const float2* __restrict__ twy,
                                                        float2* __restrict__ qpart) {
  const int x0 = 2 * (blockIdx.x * kThreads + threadIdx.x);
  const int chunk = blockIdx.y, chunks = gridDim.y, b = blockIdx.z;
  const int y0 = chunk * rows;
  const int y1 = min(Ly, y0 + rows);
  float ar0[NP], ai0[NP], ar1[NP], ai1[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) ar0[j] = ai0[j] = ar1[j] = ai1[j] = 0.f;
  const float* img = mask + (size_t)b * Lx * Ly;
  const bool v0 = x0 < Lx, v1 = x0 + 1 < Lx;
  const bool even = (Lx & 1) == 0;
  for (int y = y0; y < y1; ++y) {
    const float* row = img + (size_t)y * Lx;
    float m0 = 0.f, m1 = 0.f;
    if (even) {
      if (v0) {
        float2 v = *reinterpret_cast<const float2*>(row + x0);
        m0 = v.x;
        m1 = v.y;
      }
    } else {
      if (v0) m0 = row[x0];
      if (v1) m1 = row[x0 + 1];
    }
    const float2* tw = twy + (size_t)y * NP;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      float2 w = tw[j];
      ar0[j] += m0 * w.x;
      ai0[j] += m0 * w.y;
      ar1[j] += m1 * w.x;
      ai1[j] += m1 * w.y;
    }
  }
  float2* out = qpart + (size_t)(b * chunks + chunk) * NP * Lx;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    if (v0) out[(size_t)j * Lx + x0] = make_float2(ar0[j], ai0[j]);
    if (v1) out[(size_t)j * Lx + x0 + 1] = make_float2(ar1[j], ai1[j]);
  }
}

// Block-wide sum of v over 256 threads (4 waves of 64); result valid in
// thread 0.  red: 4 floats of LDS.
__device__ __forceinline__ float block_sum(float v, float* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// ---------------------------------------------------------------------------
// 1b. row pass: M^[ny][nx] = 1/(Lx Ly) sum_x Q[ny][x] e^{-2 pi i nx (x-cx)/Lx}
// for ny in [0,Ny], nx in [-Nx,Nx]; the ny < 0 half is its conjugate mirror
// (the mask is real).  One block per (ny, image).  band: [b][2Ny+1][2Nx+1].
// ---------------------------------------------------------------------------
template <int NP>
__global__ __launch_bounds__(kThreads) void k_band_rows(const float2* __restrict__ qpart,
                                                        int chunks, int Lx, int Ly,
                                                        int Nx, int Ny,
                                                        const float2* __restrict__ twx,
                                                        float2* __restrict__ band) {
  __shared__ float red[4];
  const int ny = blockIdx.x, b = blockIdx.y;
  float pr[NP], pi[NP], nr[NP], ni[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) pr[j] = pi[j] = nr[j] = ni[j] = 0.f;
  for (int x = threadIdx.x; x < Lx; x += kThreads) {
    float2 q = make_float2(0.f, 0.f);
    for (int c = 0; c < chunks; ++c) {
      float2 t = qpart[((size_t)(b * chunks + c) * NP + ny) * Lx + x];
      q.x += t.x;
      q.y += t.y;
    }
    const float2* tw = twx + (size_t)x * NP;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      float2 w = tw[j];
      // +nx: q*w ; -nx: q*conj(w)
      pr[j] += q.x * w.x - q.y * w.y;
      pi[j] += q.x * w.y + q.y * w.x;
      nr[j] += q.x * w.x + q.y * w.y;
      ni[j] += q.y * w.x - q.x * w.y;
    }
  }
  const int Wx = 2 * Nx + 1;
  const float inv = 1.0f / ((float)Lx * (float)Ly);
  float2* out = band + (size_t)b * Wx * (2 * Ny + 1);
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    if (j > Nx) break;
    float sr = block_sum(pr[j], red) * inv;
    float si = block_sum(pi[j], red) * inv;
    float tr = block_sum(nr[j], red) * inv;
    float ti = block_sum(ni[j], red) * inv;
    if (threadIdx.x == 0) {
      out[(Ny + ny) * Wx + Nx + j] = make_float2(sr, si);
      out[(Ny + ny) * Wx + Nx - j] = make_float2(tr, ti);
      if (ny > 0) {
        out[(Ny - ny) * Wx + Nx - j] = make_float2(sr, -si);
        out[(Ny - ny) * Wx + Nx + j] = make_float2(tr, -ti);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 2. coarse grid: per image, I_c(u) = sum_k l_k |sum_n A_k[n] e^{2 pi i n.u/M}|^2,
// then J[g] = 1/M^2 sum_u I_c(u) e^{-2 pi i g.u/M} for gy in [0,2Ny], gx in
// [-2Nx,2Nx].  spec: [b][G][4Nx+1] (rows gy > 2Ny zero).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_socs_coarse(
    const float2* __restrict__ band, const float2* __restrict__ kern,
    const float* __restrict__ scales, int nk, int Nx, int Ny, int M,
    const float2* __restrict__ twM, float2* __restrict__ spec, int G) {
  extern __shared__ float4 smem4[];
  const int Wx = 2 * Nx + 1, Wy = 2 * Ny + 1, T = Wx * Wy;
  const int Gx = 4 * Nx + 1;
  const int nB = max(Wy * M, M * Gx);
  float2* sband = reinterpret_cast<float2*>(smem4);
  float2* sA = sband + T;
  float2* sB = sA + T;
  float* sI = reinterpret_cast<float*>(sB + nB);
  float2* stw = reinterpret_cast<float2*>(sI + M * M);
  const int b = blockIdx.x, tid = threadIdx.x;
  const int mm = M - 1;
  for (int i = tid; i < T; i += kThreads) sband[i] = band[(size_t)b * T + i];
  for (int i = tid; i < M; i += kThreads) stw[i] = twM[i];
  for (int i = tid; i < M * M; i += kThreads) sI[i] = 0.f;
  __syncthreads();
  for (int k = 0; k < nk; ++k) {
    const float2* K = kern + (size_t)k * T;
    for (int i = tid; i < T; i += kThreads) sA[i] = cmul(K[i], sband[i]);
    __syncthreads();
    // B[r][ux] = sum_nx A[r][nx] e^{2 pi i (nx-Nx) ux / M}
    for (int o = tid; o < Wy * M; o += kThreads) {
      const int r = o / M, ux = o - r * M;
      float2 acc = make_float2(0.f, 0.f);
      for (int nx = 0; nx < Wx; ++nx) {
        float2 a = sA[r * Wx + nx];
        float2 w = stw[((nx - Nx) * ux) & mm];
        acc.x += a.x * w.x - a.y * w.y;
        acc.y += a.x * w.y + a.y * w.x;
      }
      sB[o] = acc;
    }
    __syncthreads();
    const float lam = scales[k];
    for (int o = tid; o < M * M; o += kThreads) {
      const int uy = o / M, ux = o - uy * M;
      float2 acc = make_float2(0.f, 0.f);
      for (int r = 0; r < Wy; ++r) {
        float2 v = sB[r * M + ux];
        float2 w = stw[((r - Ny) * uy) & mm];
        acc.x += v.x * w.x - v.y * w.y;
        acc.y += v.x * w.y + v.y * w.x;
      }
      sI[o] += lam * (acc.x * acc.x + acc.y * acc.y);
    }
    __syncthreads();
  }
  // R[uy][gi] = sum_ux I[uy][ux] e^{-2 pi i gx ux/M}, gx = gi - 2Nx
  for (int o = tid; o < M * Gx; o += kThreads) {
    const int uy = o / Gx, gi = o - uy * Gx, gx = gi - 2 * Nx;
    float2 acc = make_float2(0.f, 0.f);
    for (int ux = 0; ux < M; ++ux) {
      float v = sI[uy * M + ux];
      float2 w = stw[(-gx * ux) & mm];
      acc.x += v * w.x;
      acc.y += v * w.y;
    }
    sB[o] = acc;
  }
  __syncthreads();
  const float inv = 1.0f / ((float)M * (float)M);
  float2* out = spec + (size_t)b * G * Gx;
  for (int o = tid; o < G * Gx; o += kThreads) {
    const int gy = o / Gx, gi = o - gy * Gx;
    float2 acc = make_float2(0.f, 0.f);
    if (gy <= 2 * Ny) {
      for (int uy = 0; uy < M; ++uy) {
        float2 v = sB[uy * Gx + gi];
        float2 w = stw[(-gy * uy) & mm];
        acc.x += v.x * w.x - v.y * w.y;
        acc.y += v.x * w.y + v.y * w.x;
      }
    }
    out[o] = make_float2(acc.x * inv, acc.y * inv);
  }
}

// ---------------------------------------------------------------------------
// 2'. the same coarse-grid stage in four launches over device scratch, for
// half-sizes whose single-block stage does not fit in LDS (N > 18; M = 128):
//   k_coarse_b   B[b][k][r][ux] = sum_nx K_k[r][nx] band[b][r][nx] e^{2 pi i (nx-Nx) ux/M}
//   k_coarse_int I[b][uy][ux]   = sum_k lambda_k |sum_r B[b][k][r][ux] e^{2 pi i (r-Ny) uy/M}|^2
//   k_coarse_rx  R[b][uy][gi]   = sum_ux I[b][uy][ux] e^{-2 pi i gx ux/M}, gx = gi - 2Nx
//   k_coarse_gy  spec[b][gy][gi] = 1/M^2 sum_uy R[b][uy][gi] e^{-2 pi i gy uy/M}
// with the sums of k_socs_coarse in the same order (same FP32 results).  One
// thread per output; ux (or gi) runs fastest, so a wave shares its other
// indices and reads the band, kernel and twiddle entries as broadcasts.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_coarse_b(const float2* __restrict__ band,
                                                       const float2* __restrict__ kern, int nk,
                                                       int Nx, int Ny, int M,
                                                       const float2* __restrict__ twM,
                                                       float2* __restrict__ Bm, size_t n) {
  const size_t o = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (o >= n) return;
  const int Wx = 2 * Nx + 1, Wy = 2 * Ny + 1, T = Wx * Wy, mm = M - 1;
  const int ux = (int)(o % M);
  const size_t q = o / M;
  const int r = (int)(q % Wy);
  const size_t bk = q / Wy;
  const int k = (int)(bk % nk), b = (int)(bk / nk);
  const float2* K = kern + (size_t)k * T + r * Wx;
  const float2* bd = band + (size_t)b * T + r * Wx;
  float2 acc = make_float2(0.f, 0.f);
  for (int nx = 0; nx < Wx; ++nx) {
    const float2 a = cmul(K[nx], bd[nx]);
    const float2 w = twM[((nx - Nx) * ux) & mm];
    acc.x += a.x * w.x - a.y * w.y;
    acc.y += a.x * w.y + a.y * w.x;
  }
  Bm[o] = acc;
}

__global__ __launch_bounds__(kThreads) void k_coarse_int(const float2* __restrict__ Bm,
                                                         const float* __restrict__ scales,
                                                         int nk, int Ny, int M,
                                                         const float2* __restrict__ twM,
                                                         float* __restrict__ I, size_t n) {
  const size_t o = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (o >= n) return;
  const int Wy = 2 * Ny + 1, mm = M - 1;
  const int ux = (int)(o % M), uy = (int)((o / M) % M), b = (int)(o / ((size_t)M * M));
  float s = 0.f;
  for (int k = 0; k < nk; ++k) {
    const float2* Bk = Bm + (((size_t)b * nk + k) * Wy) * M + ux;
    float2 acc = make_float2(0.f, 0.f);
    for (int r = 0; r < Wy; ++r) {
      const float2 v = Bk[(size_t)r * M];
      const float2 w = twM[((r - Ny) * uy) & mm];
      acc.x += v.x * w.x - v.y * w.y;
      acc.y += v.x * w.y + v.y * w.x;
    }
    s += scales[k] * (acc.x * acc.x + acc.y * acc.y);
  }
  I[o] = s;
}

__global__ __launch_bounds__(kThreads) void k_coarse_rx(const float* __restrict__ I, int Nx,
                                                        int M, const float2* __restrict__ twM,
                                                        float2* __restrict__ R, size_t n) {
  const size_t o = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (o >= n) return;
  const int Gx = 4 * Nx + 1, mm = M - 1;
  const int gi = (int)(o % Gx), gx = gi - 2 * Nx;
  const size_t row = o / Gx;  // b M + uy
  const float* Ir = I + row * M;
  float2 acc = make_float2(0.f, 0.f);
  for (int ux = 0; ux < M; ++ux) {
    const float v = Ir[ux];
    const float2 w = twM[(-gx * ux) & mm];
    acc.x += v * w.x;
    acc.y += v * w.y;
  }
  R[o] = acc;
}

__global__ __launch_bounds__(kThreads) void k_coarse_gy(const float2* __restrict__ R, int Nx,
                                                        int Ny, int M,
                                                        const float2* __restrict__ twM,
                                                        float2* __restrict__ spec, int G,
                                                        size_t n) {
  const size_t o = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (o >= n) return;
  const int Gx = 4 * Nx + 1, mm = M - 1;
  const int gi = (int)(o % Gx), gy = (int)((o / Gx) % G), b = (int)(o / ((size_t)G * Gx));
  float2 acc = make_float2(0.f, 0.f);
  if (gy <= 2 * Ny) {
    const float2* Rb = R + (size_t)b * M * Gx + gi;
    for (int uy = 0; uy < M; ++uy) {
      const float2 v = Rb[(size_t)uy * Gx];
      const float2 w = twM[(-gy * uy) & mm];
      acc.x += v.x * w.x - v.y * w.y;
      acc.y += v.x * w.y + v.y * w.x;
    }
  }
  const float inv = 1.0f / ((float)M * (float)M);
  spec[o] = make_float2(acc.x * inv, acc.y * inv);
}

// ---------------------------------------------------------------------------
// 3a. C[gy][x] = w_gy sum_gx J[gy][gx] e^{+2 pi i gx (x-cx)/Lx}, w_0 = 1,
// w_gy = 2 (the gy < 0 half is the conjugate and is folded in by taking real
// parts in 3b).  One thread per column.  ccols: [b][G][Lx].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_synth_cols(const float2* __restrict__ spec,
                                                         int G, int Nx, int Ny, int Lx,
                                                         const float2* __restrict__ twx,
                                                         float2* __restrict__ ccols) {
  extern __shared__ float4 smem4[];
  float2* sJ = reinterpret_cast<float2*>(smem4);
  const int Gx = 4 * Nx + 1, Hx = 2 * Nx + 1;
  const int b = blockIdx.y;
  for (int i = threadIdx.x; i < G * Gx; i += kThreads) sJ[i] = spec[(size_t)b * G * Gx + i];
  __syncthreads();
  const int x = blockIdx.x * kThreads + threadIdx.x;
  if (x >= Lx) return;
  const float2* tw = twx + (size_t)x * Hx;
  for (int gy = 0; gy < G; ++gy) {
    float2 acc = make_float2(0.f, 0.f);
    if (gy <= 2 * Ny) {
      const float2* row = sJ + gy * Gx + 2 * Nx;
      float2 j0 = row[0];
      acc = j0;
      for (int g = 1; g < Hx; ++g) {
        float2 w = tw[g];
        float2 p = row[g], m = row[-g];
        // p*w + m*conj(w)
        acc.x += (p.x + m.x) * w.x - (p.y - m.y) * w.y;
        acc.y += (p.x - m.x) * w.y + (p.y + m.y) * w.x;
      }
      if (gy > 0) {
        acc.x *= 2.f;
        acc.y *= 2.f;
      }
    }
    ccols[((size_t)b * G + gy) * Lx + x] = acc;
  }
}

// ---------------------------------------------------------------------------
// 3b. I[y][x] = sum_gy Re(e^{+2 pi i gy (y-cy)/Ly} C[gy][x]).  Each thread owns
// two columns, a block 512 columns and `rows` rows.
// ---------------------------------------------------------------------------
template <int G>
__global__ __launch_bounds__(kThreads) void k_synth_rows(const float2* __restrict__ ccols,
                                                         int Lx, int Ly, int rows,
                                                         const float2* __restrict__ twy,
                                                         float* __restrict__ image) {
  const int x0 = 2 * (blockIdx.x * kThreads + threadIdx.x);
  const int chunk = blockIdx.y, b = blockIdx.z;
  const int y0 = chunk * rows;
  const int y1 = min(Ly, y0 + rows);
  const bool v0 = x0 < Lx, v1 = x0 + 1 < Lx;
  float c0r[G], c0i[G], c1r[G], c1i[G];
  const float2* cb = ccols + (size_t)b * G * Lx;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    float2 a = v0 ? cb[(size_t)g * Lx + x0] : make_float2(0.f, 0.f);
    float2 c = v1 ? cb[(size_t)g * Lx + x0 + 1] : make_float2(0.f, 0.f);
    c0r[g] = a.x;
    c0i[g] = -a.y;
    c1r[g] = c.x;
    c1i[g] = -c.y;
  }
  float* img = image + (size_t)b * Lx * Ly;
  const bool even = (Lx & 1) == 0;
  for (int y = y0; y < y1; ++y) {
    const float2* tw = twy + (size_t)y * G;
    float s0 = 0.f, s1 = 0.f;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      float2 w = tw[g];
      s0 += w.x * c0r[g];
      s0 += w.y * c0i[g];
      s1 += w.x * c1r[g];
      s1 += w.y * c1i[g];
    }
    float* row = img + (size_t)y * Lx;
    if (even) {
      if (v0) *reinterpret_cast<float2*>(row + x0) = make_float2(s0, s1);
    } else {
      if (v0) row[x0] = s0;
      if (v1) row[x0 + 1] = s1;
    }
  }
}

// Hopkins spectrum (calcImage, K-Litho-TCC/klitho_tcc.cpp:270-290):
//   H[g] = sum_{n1} M^[n1+g] conj(M^[n1]) TCC[ID(n1)][ID(n1+g)], g_x >= 0,
// then FT_c2r (:151-183) keeps the real part of the g_x = 0 column.  That
// equals synthesising the Hermitian completion J (J[g] = H[g] for g_x > 0,
// J[-g] = conj(H[g]), J[gy][0] = (H[gy][0] + conj(H[-gy][0])) / 2), written
// here in the half-plane layout of the synthesis passes.  FP64 sums.
// One wave's share of H[g] = sum_n1 M^[n1 + g] conj(M^[n1]) TCC[ID(n1)][ID(n1 + g)]
// (lanes stride over the n1 rectangle; FP64), summed over the wave.
__device__ double2 hopkins_term(const float2* __restrict__ band,
                                const double* __restrict__ tcc, int Nx, int Ny, int gx,
                                int gy) {
  const int Wx = 2 * Nx + 1, T = Wx * (2 * Ny + 1);
  const int y0 = max(-Ny, -Ny - gy), y1 = min(Ny, Ny - gy);
  const int x0 = max(-Nx, -Nx - gx), x1 = min(Nx, Nx - gx);
  const int nxr = x1 - x0 + 1, total = (y1 - y0 + 1) * nxr;
  double sr = 0.0, si = 0.0;
  for (int j = threadIdx.x & 63; j < total; j += 64) {
    const int ny = y0 + j / nxr, nx = x0 + j % nxr;
    const int i1 = (ny + Ny) * Wx + nx + Nx, i2 = (ny + gy + Ny) * Wx + nx + gx + Nx;
    const float2 a = band[i2], b = band[i1];
    // a conj(b)
    const double pr = (double)a.x * b.x + (double)a.y * b.y;
    const double pi = (double)a.y * b.x - (double)a.x * b.y;
    const double tr = tcc[2 * ((size_t)i1 * T + i2)], ti = tcc[2 * ((size_t)i1 * T + i2) + 1];
    sr += pr * tr - pi * ti;
    si += pr * ti + pi * tr;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sr += __shfl_xor(sr, o, 64);
    si += __shfl_xor(si, o, 64);
  }
  return make_double2(sr, si);
}

// Hopkins intensity spectrum J[gy][gx] (gy in [0, 2Ny], gx in [-2Nx, 2Nx],
// rows up to `rows` zero) of image blockIdx.y from its band [T]; one wave per
// entry.  spec: [B][rows][4Nx+1].
__global__ __launch_bounds__(kThreads) void k_hopkins_spec(const float2* __restrict__ band,
                                                           const double* __restrict__ tcc,
                                                           int Nx, int Ny, int rows,
                                                           float2* __restrict__ spec) {
  const int Gx = 4 * Nx + 1, T = (2 * Nx + 1) * (2 * Ny + 1);
  const int o = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6), b = blockIdx.y;
  if (o >= rows * Gx) return;  // whole waves
  const int gy = o / Gx, gx = o - gy * Gx - 2 * Nx;
  const float2* bd = band + (size_t)b * T;
  float2 out = make_float2(0.f, 0.f);
  if (gy <= 2 * Ny) {
    double2 j;
    if (gx > 0) {
      j = hopkins_term(bd, tcc, Nx, Ny, gx, gy);
    } else if (gx < 0) {
      const double2 h = hopkins_term(bd, tcc, Nx, Ny, -gx, -gy);
      j = make_double2(h.x, -h.y);
    } else {
      const double2 a = hopkins_term(bd, tcc, Nx, Ny, 0, gy);
      const double2 c = hopkins_term(bd, tcc, Nx, Ny, 0, -gy);
      j = make_double2(0.5 * (a.x + c.x), 0.5 * (a.y - c.y));
    }
    out = make_float2((float)j.x, (float)j.y);
  }
  if ((threadIdx.x & 63) == 0) spec[(size_t)b * rows * Gx + o] = out;
}

__global__ void k_pack_kernels(const float* __restrict__ kr, const float* __restrict__ ki,
                               int n, float2* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2(kr[i], ki[i]);
}

}  // namespace

namespace klt {

template <int NP>
static void launch_band(klt_ctx* ctx, const float* masks, int B, int Lx, int Ly, int chunks,
                        int rows, const SocsPlan* p, float2* qpart, float2* band) {
  dim3 g1((Lx + 2 * kThreads - 1) / (2 * kThreads), chunks, B);
  KLT_LAUNCH(ctx, "k_band_cols", k_band_cols<NP>, g1, dim3(kThreads), 0, masks, Lx, Ly, rows,
             p->twy_in.as<float2>(), qpart);
  dim3 g2(p->Ny + 1, B);
  KLT_LAUNCH(ctx, "k_band_rows", k_band_rows<NP>, g2, dim3(kThreads), 0, qpart, chunks, Lx,
             Ly, p->Nx, p->Ny, p->twx_in.as<float2>(), band);
}

template <int G>
static void launch_rows(klt_ctx* ctx, const float2* ccols, int B, int Lx, int Ly, int chunks,
                        int rows, const SocsPlan* p, float* images) {
  dim3 g((Lx + 2 * kThreads - 1) / (2 * kThreads), chunks, B);
  KLT_LAUNCH(ctx, "k_synth_rows", k_synth_rows<G>, g, dim3(kThreads), 0, ccols, Lx, Ly, rows,
             p->twy_out.as<float2>(), images);
}

// Coarse-grid stage of nb images: band [nb][2Ny+1][2Nx+1] (normalised,
// centred) to the intensity spectra spec [nb][G][4Nx+1] (rows 2Ny < gy < G
// zero).  N <= 15: the 64-point FFT stage of socs_fused.hip; above that the
// direct transforms on the M = 128 grid, in one block per image while they
// fit in LDS (N <= 18), else in four launches over device scratch.
int socs_coarse_spec(klt_ctx* ctx, const klt_kernels* kb, const SocsPlan* p,
                     const float2* band, int nb, float2* spec, int G) {
  if (socs_fused_supported(kb)) return socs_mid_from_band(ctx, kb, band, nb, spec, G);
  const int Nx = kb->Nx, Ny = kb->Ny, M = p->M;
  const int T = (2 * Nx + 1) * (2 * Ny + 1), Gx = 4 * Nx + 1;
  const int nB = std::max((2 * Ny + 1) * M, M * Gx);
  const size_t lds_coarse =
      sizeof(float2) * (2 * (size_t)T + nB + M) + sizeof(float) * (size_t)M * M;
  if (lds_coarse <= 160 * 1024) {
    if (lds_coarse > 64 * 1024)
      KLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_socs_coarse),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_coarse));
    KLT_LAUNCH(ctx, "k_socs_coarse", k_socs_coarse, dim3(nb), dim3(kThreads), lds_coarse, band,
               kb->kern.as<float2>(), kb->scales.as<float>(), kb->nk, Nx, Ny, M,
               p->twM.as<float2>(), spec, G);
    KLT_HIP(hipGetLastError());
    return KLT_OK;
  }
  // sub-batches of cb images bound the scratch to 256 MiB
  const size_t szB = (size_t)kb->nk * (2 * Ny + 1) * M * sizeof(float2);
  const size_t szI = (size_t)M * M * sizeof(float), szR = (size_t)M * Gx * sizeof(float2);
  const int cb = (int)std::max<size_t>(1, std::min<size_t>(nb, (256u << 20) / (szB + szI + szR)));
  KLT_TRY(ctx->coarse_ws.ensure((szB + szI + szR) * cb));
  float2* Bm = ctx->coarse_ws.as<float2>();
  float* Ic = reinterpret_cast<float*>(ctx->coarse_ws.as<char>() + szB * cb);
  float2* R = reinterpret_cast<float2*>(ctx->coarse_ws.as<char>() + (szB + szI) * cb);
  const float2* tw = p->twM.as<float2>();
  auto grid = [](size_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); };
  for (int c0 = 0; c0 < nb; c0 += cb) {
    const int nc = std::min(cb, nb - c0);
    const float2* bd = band + (size_t)c0 * T;
    const size_t nbk = (size_t)nc * kb->nk * (2 * Ny + 1) * M, ni = (size_t)nc * M * M,
                 nr = (size_t)nc * M * Gx, ns = (size_t)nc * G * Gx;
    KLT_LAUNCH(ctx, "k_coarse_b", k_coarse_b, grid(nbk), dim3(kThreads), 0, bd,
               kb->kern.as<float2>(), kb->nk, Nx, Ny, M, tw, Bm, nbk);
    KLT_LAUNCH(ctx, "k_coarse_int", k_coarse_int, grid(ni), dim3(kThreads), 0, Bm,
               kb->scales.as<float>(), kb->nk, Ny, M, tw, Ic, ni);
    KLT_LAUNCH(ctx, "k_coarse_rx", k_coarse_rx, grid(nr), dim3(kThreads), 0, Ic, Nx, M, tw, R,
               nr);
    KLT_LAUNCH(ctx, "k_coarse_gy", k_coarse_gy, grid(ns), dim3(kThreads), 0, R, Nx, Ny, M, tw,
               spec + (size_t)c0 * G * Gx, G, ns);
  }
  KLT_HIP(hipGetLastError());
  return KLT_OK;
}

// Runs the SOCS forward for B masks of Lx x Ly already on the device.
int socs_forward_device(klt_ctx* ctx, const klt_kernels* kb, const float* masks, int B,
                        int Lx, int Ly, float* images) {
  SocsPlan* p = nullptr;
  KLT_TRY(get_socs_plan(ctx, Lx, Ly, kb->Nx, kb->Ny, &p));
  const int NP = p->NP, G = p->G;
  const int Nx = kb->Nx, Ny = kb->Ny;
  const int T = (2 * Nx + 1) * (2 * Ny + 1), Gx = 4 * Nx + 1;
  const int xblocks = (Lx + 2 * kThreads - 1) / (2 * kThreads);
  // enough blocks to fill 256 CUs several times over
  int want = std::max(1, 2048 / std::max(1, xblocks * B));
  int chunks = std::min(std::max(1, want), std::max(1, Ly / 8));
  int rows = (Ly + chunks - 1) / chunks;
  chunks = (Ly + rows - 1) / rows;
  if (NP > 32 || G > 64)
    return set_error(KLT_ERR_ARG, "SOCS kernel half-size N above 31 is not supported");
  // sub-batches bound the scratch footprint
  const int SB = std::min(B, 512);
  KLT_TRY(ctx->qpart.ensure((size_t)SB * chunks * NP * Lx * sizeof(float2)));
  KLT_TRY(ctx->band.ensure((size_t)SB * T * sizeof(float2)));
  KLT_TRY(ctx->spec.ensure((size_t)SB * G * Gx * sizeof(float2)));
  KLT_TRY(ctx->ccols.ensure((size_t)SB * G * Lx * sizeof(float2)));
  for (int b0 = 0; b0 < B; b0 += SB) {
    const int nb = std::min(SB, B - b0);
    const float* m = masks + (size_t)b0 * Lx * Ly;
    float* out = images + (size_t)b0 * Lx * Ly;
    float2* qpart = ctx->qpart.as<float2>();
    float2* band = ctx->band.as<float2>();
    switch (NP) {
      case 8: launch_band<8>(ctx, m, nb, Lx, Ly, chunks, rows, p, qpart, band); break;
      case 16: launch_band<16>(ctx, m, nb, Lx, Ly, chunks, rows, p, qpart, band); break;
      case 24: launch_band<24>(ctx, m, nb, Lx, Ly, chunks, rows, p, qpart, band); break;
      default: launch_band<32>(ctx, m, nb, Lx, Ly, chunks, rows, p, qpart, band); break;
    }
    KLT_TRY(socs_coarse_spec(ctx, kb, p, band, nb, ctx->spec.as<float2>(), G));
    KLT_LAUNCH(ctx, "k_synth_cols", k_synth_cols, dim3((Lx + kThreads - 1) / kThreads, nb),
               dim3(kThreads), sizeof(float2) * G * Gx, ctx->spec.as<float2>(), G, Nx, Ny, Lx,
               p->twx_out.as<float2>(), ctx->ccols.as<float2>());
    switch (G) {
      case 8: launch_rows<8>(ctx, ctx->ccols.as<float2>(), nb, Lx, Ly, chunks, rows, p, out); break;
      case 16: launch_rows<16>(ctx, ctx->ccols.as<float2>(), nb, Lx, Ly, chunks, rows, p, out); break;
      case 24: launch_rows<24>(ctx, ctx->ccols.as<float2>(), nb, Lx, Ly, chunks, rows, p, out); break;
      case 32: launch_rows<32>(ctx, ctx->ccols.as<float2>(), nb, Lx, Ly, chunks, rows, p, out); break;
      case 40: launch_rows<40>(ctx, ctx->ccols.as<float2>(), nb, Lx, Ly, chunks, rows, p, out); break;
      case 48: launch_rows<48>(ctx, ctx->ccols.as<float2>(), nb, Lx, Ly, chunks, rows, p, out); break;
      case 56: launch_rows<56>(ctx, ctx->ccols.as<float2>(), nb, Lx, Ly, chunks, rows, p, out); break;
      default: launch_rows<64>(ctx, ctx->ccols.as<float2>(), nb, Lx, Ly, chunks, rows, p, out); break;
    }
    KLT_HIP(hipGetLastError());
  }
  return KLT_OK;
}

// Hopkins image of one mask from a device TCC (T x T complex double).
int hopkins_spec_device(klt_ctx* ctx, const float2* band, int B, const double* tcc, int Nx,
                        int Ny, int rows, float2* spec) {
  const int waves = kThreads / 64, n = rows * (4 * Nx + 1);
  KLT_LAUNCH(ctx, "k_hopkins_spec", k_hopkins_spec, dim3((n + waves - 1) / waves, B),
             dim3(kThreads), 0, band, tcc, Nx, Ny, rows, spec);
  KLT_HIP(hipGetLastError());
  return KLT_OK;
}

int hopkins_image_device(klt_ctx* ctx, const float* mask, int Lx, int Ly, const double* tcc,
                         int Nx, int Ny, float* image) {
  SocsPlan* p = nullptr;
  KLT_TRY(get_socs_plan(ctx, Lx, Ly, Nx, Ny, &p));
  const int NP = p->NP, G = p->G;
  const int T = (2 * Nx + 1) * (2 * Ny + 1), Gx = 4 * Nx + 1;
  if (NP > 32 || G > 64)
    return set_error(KLT_ERR_ARG, "Hopkins image: half-size N above 31 is not supported");
  const int xblocks = (Lx + 2 * kThreads - 1) / (2 * kThreads);
  int chunks = std::min(std::max(1, 2048 / xblocks), std::max(1, Ly / 8));
  int rows = (Ly + chunks - 1) / chunks;
  chunks = (Ly + rows - 1) / rows;
  KLT_TRY(ctx->qpart.ensure((size_t)chunks * NP * Lx * sizeof(float2)));
  KLT_TRY(ctx->band.ensure((size_t)T * sizeof(float2)));
  KLT_TRY(ctx->spec.ensure((size_t)G * Gx * sizeof(float2)));
  KLT_TRY(ctx->ccols.ensure((size_t)G * Lx * sizeof(float2)));
  float2* qpart = ctx->qpart.as<float2>();
  float2* band = ctx->band.as<float2>();
  switch (NP) {
    case 8: launch_band<8>(ctx, mask, 1, Lx, Ly, chunks, rows, p, qpart, band); break;
    case 16: launch_band<16>(ctx, mask, 1, Lx, Ly, chunks, rows, p, qpart, band); break;
    case 24: launch_band<24>(ctx, mask, 1, Lx, Ly, chunks, rows, p, qpart, band); break;
    default: launch_band<32>(ctx, mask, 1, Lx, Ly, chunks, rows, p, qpart, band); break;
  }
  KLT_TRY(hopkins_spec_device(ctx, band, 1, tcc, Nx, Ny, G, ctx->spec.as<float2>()));
  KLT_LAUNCH(ctx, "k_synth_cols", k_synth_cols, dim3((Lx + kThreads - 1) / kThreads, 1),
             dim3(kThreads), sizeof(float2) * G * Gx, ctx->spec.as<float2>(), G, Nx, Ny, Lx,
             p->twx_out.as<float2>(), ctx->ccols.as<float2>());
  float2* cc = ctx->ccols.as<float2>();
  switch (G) {
    case 8: launch_rows<8>(ctx, cc, 1, Lx, Ly, chunks, rows, p, image); break;
    case 16: launch_rows<16>(ctx, cc, 1, Lx, Ly, chunks, rows, p, image); break;
    case 24: launch_rows<24>(ctx, cc, 1, Lx, Ly, chunks, rows, p, image); break;
    case 32: launch_rows<32>(ctx, cc, 1, Lx, Ly, chunks, rows, p, image); break;
    case 40: launch_rows<40>(ctx, cc, 1, Lx, Ly, chunks, rows, p, image); break;
    case 48: launch_rows<48>(ctx, cc, 1, Lx, Ly, chunks, rows, p, image); break;
    case 56: launch_rows<56>(ctx, cc, 1, Lx, Ly, chunks, rows, p, image); break;
    default: launch_rows<64>(ctx, cc, 1, Lx, Ly, chunks, rows, p, image); break;
  }
  KLT_HIP(hipGetLastError());
  return KLT_OK;
}

}  // namespace klt

extern "C" {

int klt_hopkins_image(klt_ctx* ctx, const float* mask, int Lx, int Ly, const double* tcc,
                      int Nx, int Ny, float* image, unsigned flags) {
  if (!ctx || !mask || !tcc || !image)
    return klt::set_error(KLT_ERR_ARG, "klt_hopkins_image: NULL argument");
  if (Lx <= 0 || Ly <= 0 || Nx < 0 || Ny < 0)
    return klt::set_error(KLT_ERR_ARG, "Error: Invalid input parameters. Please check the values.");
  if (4 * Nx + 1 > Lx || 4 * Ny + 1 > Ly)
    return klt::set_error(KLT_ERR_SIZE, "Hopkins image: 4N+1 exceeds the period");
  KLT_TRY(ctx->activate());
  const int T = (2 * Nx + 1) * (2 * Ny + 1);
  const size_t mbytes = (size_t)Lx * Ly * sizeof(float), tbytes = (size_t)T * T * 16;
  const float* dm = mask;
  const double* dt = tcc;
  float* di = image;
  if (flags & KLT_HOST_PTRS) {
    KLT_TRY(ctx->stage_in.ensure(mbytes + tbytes + 256));
    char* b = ctx->stage_in.as<char>();
    KLT_HIP(hipMemcpyAsync(b, tcc, tbytes, hipMemcpyHostToDevice, ctx->stream));
    KLT_HIP(hipMemcpyAsync(b + tbytes, mask, mbytes, hipMemcpyHostToDevice, ctx->stream));
    dt = reinterpret_cast<const double*>(b);
    dm = reinterpret_cast<const float*>(b + tbytes);
    KLT_TRY(ctx->stage_out.ensure(mbytes));
    di = ctx->stage_out.as<float>();
  }
  KLT_TRY(klt::hopkins_image_device(ctx, dm, Lx, Ly, dt, Nx, Ny, di));
  if (flags & KLT_HOST_PTRS) {
    KLT_HIP(hipMemcpyAsync(image, di, mbytes, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return KLT_OK;
}


int klt_kernels_create(klt_ctx* ctx, const float* krn_r, const float* krn_i,
                       const float* scales, int nk, int sizeX, int sizeY, unsigned flags,
                       klt_kernels** out) {
  if (!ctx || !out || !krn_r || !krn_i || !scales)
    return klt::set_error(KLT_ERR_ARG, "klt_kernels_create: NULL argument");
  *out = nullptr;
  if (nk <= 0 || sizeX <= 0 || sizeY <= 0)
    return klt::set_error(KLT_ERR_ARG, "klt_kernels_create: empty kernel bank");
  if (sizeX % 2 == 0 || sizeY % 2 == 0)  // klitho_socs.cpp:298-301
    return klt::set_error(KLT_ERR_ARG, "Error: Kernel size must be odd.");
  KLT_TRY(ctx->activate());
  auto kb = std::make_unique<klt_kernels>();
  kb->ctx = ctx;
  kb->nk = nk;
  kb->Nx = (sizeX - 1) / 2;
  kb->Ny = (sizeY - 1) / 2;
  const int T = sizeX * sizeY;
  const size_t n = (size_t)nk * T;
  KLT_TRY(kb->kern.ensure(n * sizeof(float2)));
  KLT_TRY(kb->scales.ensure(nk * sizeof(float)));
  kb->h_scales.resize(nk);
  if (flags & KLT_HOST_PTRS) {
    std::vector<float2> h(n);
    for (size_t i = 0; i < n; ++i) h[i] = make_float2(krn_r[i], krn_i[i]);
    KLT_HIP(hipMemcpy(kb->kern.ptr, h.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    std::memcpy(kb->h_scales.data(), scales, nk * sizeof(float));
    KLT_HIP(hipMemcpy(kb->scales.ptr, scales, nk * sizeof(float), hipMemcpyHostToDevice));
  } else {
    hipLaunchKernelGGL(k_pack_kernels, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, krn_r, krn_i, (int)n, kb->kern.as<float2>());
    KLT_HIP(hipGetLastError());
    KLT_HIP(hipMemcpyAsync(kb->scales.ptr, scales, nk * sizeof(float),
                           hipMemcpyDeviceToDevice, ctx->stream));
    KLT_HIP(hipMemcpyAsync(kb->h_scales.data(), scales, nk * sizeof(float),
                           hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
  }
  *out = kb.release();
  return KLT_OK;
}

void klt_kernels_destroy(klt_kernels* kb) {
  if (!kb) return;
  if (kb->ctx) (void)hipSetDevice(kb->ctx->device);
  delete kb;
}

static int socs_entry(klt_ctx* ctx, const klt_kernels* kb, const float* masks, int nmask,
                      int Lx, int Ly, float* images, unsigned flags, bool plain) {
  if (!ctx || !kb || !masks || !images)
    return klt::set_error(KLT_ERR_ARG, "klt_socs_forward: NULL argument");
  if (nmask <= 0 || Lx <= 0 || Ly <= 0)
    return klt::set_error(KLT_ERR_ARG, "Error: Invalid input parameters. Please check the values.");
  if (4 * kb->Nx + 1 > Lx || 4 * kb->Ny + 1 > Ly)  // klitho_socs.cpp:164-169
    return klt::set_error(KLT_ERR_SIZE, "ERROR in FI()");
  KLT_TRY(ctx->activate());
  const size_t bytes = (size_t)nmask * Lx * Ly * sizeof(float);
  const void* dm = nullptr;
  KLT_TRY(klt::to_device(ctx, masks, bytes, flags, ctx->stage_in, &dm));
  float* dout = images;
  if (flags & KLT_HOST_PTRS) {
    KLT_TRY(ctx->stage_out.ensure(bytes));
    dout = ctx->stage_out.as<float>();
  }
  if (!plain && klt::socs_fused_supported(kb))
    KLT_TRY(klt::socs_forward_fused(ctx, kb, static_cast<const float*>(dm), nmask, Lx, Ly, dout));
  else if (!plain && klt::socs_wide_supported(kb))
    KLT_TRY(klt::socs_forward_wide(ctx, kb, static_cast<const float*>(dm), nmask, Lx, Ly, dout));
  else
    KLT_TRY(klt::socs_forward_device(ctx, kb, static_cast<const float*>(dm), nmask, Lx, Ly, dout));
  if (flags & KLT_HOST_PTRS) {
    KLT_HIP(hipMemcpyAsync(images, dout, bytes, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return KLT_OK;
}

int klt_socs_forward(klt_ctx* ctx, const klt_kernels* kb, const float* masks, int nmask,
                     int Lx, int Ly, float* images, unsigned flags) {
  return socs_entry(ctx, kb, masks, nmask, Lx, Ly, images, flags, false);
}

int klt_socs_image(klt_ctx* ctx, const klt_kernels* kb, const float* mask, int Lx, int Ly,
                   float* image, unsigned flags) {
  return socs_entry(ctx, kb, mask, 1, Lx, Ly, image, flags, true);
}

}  // extern "C"

namespace klt {

// New kernels and scales (device pointers, same count and size) into an
// existing bank, on the context stream, without allocating (the bank's host
// copy of the scales is left as it was: no path reads it).
int kernels_refill(klt_kernels* kb, const float* krn_r, const float* krn_i,
                   const float* scales) {
  klt_ctx* ctx = kb->ctx;
  const size_t n = (size_t)kb->nk * (2 * kb->Nx + 1) * (2 * kb->Ny + 1);
  hipLaunchKernelGGL(k_pack_kernels, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     ctx->stream, krn_r, krn_i, (int)n, kb->kern.as<float2>());
  KLT_HIP(hipGetLastError());
  KLT_HIP(hipMemcpyAsync(kb->scales.ptr, scales, kb->nk * sizeof(float),
                         hipMemcpyDeviceToDevice, ctx->stream));
  return KLT_OK;
}

}  // namespace klt
