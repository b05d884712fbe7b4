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
// Stage 1 (and the adjoint's resist-gradient band).  grid (ceil(Ly/128), B),
// 256 threads.  Wave w owns rows blk*128 + 32w .. +32 (two 16-row MFMA
// blocks) and all 16*NCB output columns (n = 2 nx: cos part, n = 2 nx + 1:
// -sin part).  The GEMM input is
//   RESIST = false: the mask m[y][x];
//   RESIST = true : G(x) = dL/dI = 2 (s - T) a s (1 - s), s = sigmoid(a (I - thr)),
//                   from the image I and target T (the loss sum (s - T)^2 of
//                   the block goes to lpart[b][blk]).
// Epilogue: out[b][blk][ny + By][nx] = sum_{y in block} R[y][nx] tw(ny, y) for
// |ny| <= By, 0 <= nx <= Bx; tw(ny, y) = twy[y][|ny|] (conjugated for ny < 0).
// ---------------------------------------------------------------------------
template <int NCB, bool RESIST>
__global__ __launch_bounds__(256) void k_band_mfma(
    const float* __restrict__ in0, const float* __restrict__ in1, float a_sig, float thr,
    int Lx, int Ly, int Kpad, const _Float16* __restrict__ Wh, const _Float16* __restrict__ Wl,
    const float2* __restrict__ twy, int ldtw, int Bx, int By, int ldo,
    float2* __restrict__ out_part, float* __restrict__ lpart) {
  __shared__ float sR[kRowsBand][16 * NCB + 1];
  __shared__ float2 sTw[kRowsBand][8 * NCB];
  __shared__ float red[4];
  const int blk = blockIdx.x, b = blockIdx.y, nblk = gridDim.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int ybase = blk * kRowsBand + wave * 32;
  const size_t ib = (size_t)b * Lx * Ly;
  const int ya = ybase + r16, yb = ybase + 16 + r16;
  const bool va = ya < Ly, vb = yb < Ly;
  const size_t ra = ib + (size_t)(va ? ya : 0) * Lx, rb = ib + (size_t)(vb ? yb : 0) * Lx;
  const bool vec = (Lx & 3) == 0;
  float4_t acc[2][NCB];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < NCB; ++c) acc[a][c] = float4_t{0.f, 0.f, 0.f, 0.f};
  float lsum = 0.f;
  // software pipeline: the loads of step k + 32 are issued before the
  // conversions and matrix products of step k
  float na8[8], nb8[8], nta8[8], ntb8[8];
  load8(in0 + ra, kq, Lx, va, vec, na8);
  load8(in0 + rb, kq, Lx, vb, vec, nb8);
  if (RESIST) {
    load8(in1 + ra, kq, Lx, va, vec, nta8);
    load8(in1 + rb, kq, Lx, vb, vec, ntb8);
  }
  for (int k = 0; k < Kpad; k += 32) {
    float va8[8], vb8[8], ta8[8], tb8[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      va8[t] = na8[t];
      vb8[t] = nb8[t];
      if (RESIST) {
        ta8[t] = nta8[t];
        tb8[t] = ntb8[t];
      }
    }
    if (k + 32 < Kpad) {
      load8(in0 + ra, k + 32 + kq, Lx, va, vec, na8);
      load8(in0 + rb, k + 32 + kq, Lx, vb, vec, nb8);
      if (RESIST) {
        load8(in1 + ra, k + 32 + kq, Lx, va, vec, nta8);
        load8(in1 + rb, k + 32 + kq, Lx, vb, vec, ntb8);
      }
    }
    if (RESIST) {
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const bool ina = va && k + kq + t < Lx, inb = vb && k + kq + t < Lx;
        const float sa = 1.f / (1.f + __expf(-a_sig * (va8[t] - thr)));
        const float sb = 1.f / (1.f + __expf(-a_sig * (vb8[t] - thr)));
        const float da = sa - ta8[t], db = sb - tb8[t];
        lsum += (ina ? da * da : 0.f) + (inb ? db * db : 0.f);
        va8[t] = ina ? 2.f * da * a_sig * sa * (1.f - sa) : 0.f;
        vb8[t] = inb ? 2.f * db * a_sig * sb * (1.f - sb) : 0.f;
      }
    }
    half8_t ah, al, bh, bl;
    split8(va8, ah, al);
    split8(vb8, bh, bl);
#pragma unroll
    for (int c = 0; c < NCB; ++c) {
      const size_t wo = (size_t)(c * 16 + r16) * Kpad + k + kq;
      const half8_t ch = *reinterpret_cast<const half8_t*>(Wh + wo);
      const half8_t cl = *reinterpret_cast<const half8_t*>(Wl + wo);
      acc[0][c] = mfma3(ah, al, ch, cl, acc[0][c]);
      acc[1][c] = mfma3(bh, bl, ch, cl, acc[1][c]);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < NCB; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        sR[wave * 32 + a * 16 + (lane >> 4) * 4 + r][c * 16 + r16] = acc[a][c][r];
  if (RESIST) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (lane == 0) red[wave] = lsum;
  }
  __syncthreads();
  if (RESIST && threadIdx.x == 0) lpart[(size_t)b * nblk + blk] = red[0] + red[1] + red[2] + red[3];
  const int Wy = 2 * By + 1, nxp = Bx + 1;
  const int y0 = blk * kRowsBand;
  const int nrow = min(kRowsBand, Ly - y0);
  // this block's column twiddles, staged once (|ny| <= By < 8 NCB)
  for (int e = threadIdx.x; e < kRowsBand * 8 * NCB; e += 256) {
    const int yy = e / (8 * NCB), c = e - yy * (8 * NCB);
    sTw[yy][c] = (yy < nrow && c <= By) ? twy[(size_t)(y0 + yy) * ldtw + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2* out = out_part + ((size_t)b * nblk + blk) * Wy * ldo;
  for (int o = threadIdx.x; o < Wy * nxp; o += 256) {
    const int nyi = o / nxp, nx = o - nyi * nxp;
    const int ny = nyi - By;
    const int any = ny < 0 ? -ny : ny;
    float sr = 0.f, si = 0.f;
    for (int yy = 0; yy < nrow; ++yy) {
      float2 w = sTw[yy][any];
      if (ny < 0) w.y = -w.y;
      const float rr = sR[yy][2 * nx], ri = sR[yy][2 * nx + 1];
      sr += rr * w.x - ri * w.y;
      si += rr * w.y + ri * w.x;
    }
    out[nyi * ldo + nx] = make_float2(sr, si);
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
// Adjoint mid stage.  grid (B), 256 threads.  With G = dL/dI and its band
// G^[g] = 1/(Lx Ly) sum_x G(x) e^{-2 pi i g.(x - c_out)/L}, |g| <= 2N:
//   g_c(u)   = sum_g G^[g] e^{+2 pi i g.u/64}              (coarse grid, registers)
//   e_k(u)   = sum_n A_k[n] e^{+2 pi i n.u/64}, A_k = K_k M^
//   F_k[-q]  = 1/64^2 sum_u conj(e_k(u)) g_c(u) e^{-2 pi i q.u/64}
//            = sum_n' conj(A_k[n']) G^[n' + q]              (|n'+q| <= 3N < 64 - N)
//   Gam[n]   = sum_k lambda_k K_k[n] F_k[n]
// and dL/dm(x) = 2 Re sum_n Gam[n] e^{-2 pi i n.(x - c_in)/L}, written as the
// Hermitian spectrum J[n] = Gam[-n] + conj(Gam[n]) (times e^{2 pi i n.(c_out -
// c_in)/L}) in the synthesis layout [32][4Nx+1] (rows gy in [0, Ny]).
// loss[b] = sum of the per-block loss partials.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_socs_mid_bwd(
    const float2* __restrict__ mpart, const float2* __restrict__ gpart,
    const float* __restrict__ lpart, int nblk, int Nx, int Ny, float inv_area,
    const float2* __restrict__ kern, const float* __restrict__ scales, int nk,
    const float2* __restrict__ tw64, int dcx, int dcy, int Lx, int Ly,
    float2* __restrict__ gspec, float* __restrict__ loss) {
  __shared__ float2 sM[31 * 31];
  __shared__ float2 sB[61 * kM];
  __shared__ float2 sX[32 * 64];
  __shared__ float2 sGam[31 * 31];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int g = tid >> 3, t = tid & 7;
  const int Wx = 2 * Nx + 1, Wy = 2 * Ny + 1, T = Wx * Wy;
  const int Gx = 4 * Nx + 1, Gy = 4 * Ny + 1;
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
    sGam[i] = make_float2(0.f, 0.f);
  }
  if (tid == 0) {
    float l = 0.f;
    for (int k = 0; k < nblk; ++k) l += lpart[(size_t)b * nblk + k];
    loss[b] = l;
  }
  float2 twi[8], twf[8];
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) {
    twi[k2] = tw64[(t * k2) & 63];
    twf[k2] = make_float2(twi[k2].x, -twi[k2].y);
  }
  float2* xs = sX + g * 64;
  __syncthreads();
  // g_c rows: gy in [-2Ny, 2Ny] (row index gy + 2Ny), two rounds of 32 groups
#pragma unroll
  for (int rr = 0; rr < 2; ++rr) {
    const int r = g + 32 * rr;
    const bool act = r < Gy;
    const int gy = r - 2 * Ny;
    float2 v[8];
    if (act) {
#pragma unroll
      for (int n2 = 0; n2 < 8; ++n2) {
        const int n = t + 8 * n2;
        const int gx = n < 32 ? n : n - 64;
        float2 val = make_float2(0.f, 0.f);
        if (gx >= -2 * Nx && gx <= 2 * Nx) {
          const int row = gx >= 0 ? r : (-gy + 2 * Ny), col = gx >= 0 ? gx : -gx;
          for (int k = 0; k < nblk; ++k) {
            const float2 p = gpart[(((size_t)b * nblk + k) * Gy + row) * 32 + col];
            val.x += p.x;
            val.y += p.y;
          }
          val = make_float2(val.x * inv_area, gx >= 0 ? val.y * inv_area : -val.y * inv_area);
        }
        v[n2] = val;
      }
    }
    fft64_group<1>(v, twi, xs, t, act);
    if (act) {
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) sB[r * kM + t + 8 * k1] = v[k1];
    }
  }
  __syncthreads();
  float2 gc[2][8];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int ux = g + 32 * q;
    float2 v[8];
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2) {
      const int n = t + 8 * n2;
      const int gy = n < 32 ? n : n - 64;
      v[n2] = (gy >= -2 * Ny && gy <= 2 * Ny) ? sB[(gy + 2 * Ny) * kM + ux] : make_float2(0.f, 0.f);
    }
    fft64_group<1>(v, twi, xs, t, true);
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) gc[q][k1] = v[k1];
  }
  const float inv = 1.0f / (float)(kM * kM);
  for (int k = 0; k < nk; ++k) {
    const float2* K = kern + (size_t)k * T;
    const float lam = scales[k];
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
      // conj(e) g_c, then forward along uy
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) {
        const float2 e = v[k1], c = gc[q][k1];
        v[k1] = make_float2(e.x * c.x + e.y * c.y, e.x * c.y - e.y * c.x);
      }
      fft64_group<-1>(v, twf, xs, t, true);
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) {
        const int qq = t + 8 * k1;
        const int qy = qq < 32 ? qq : qq - 64;
        if (qy >= -Ny && qy <= Ny) sB[(qy + Ny) * kM + ux] = v[k1];
      }
    }
    __syncthreads();
    {
      const bool act = g < Wy;
      const int qy = g - Ny;
      float2 v[8];
      if (act) {
#pragma unroll
        for (int n2 = 0; n2 < 8; ++n2) v[n2] = sB[g * kM + t + 8 * n2];
      }
      fft64_group<-1>(v, twf, xs, t, act);
      if (act) {
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) {
          const int qq = t + 8 * k1;
          const int qx = qq < 32 ? qq : qq - 64;
          if (qx >= -Nx && qx <= Nx) {
            const int idx = (-qy + Ny) * Wx + (-qx + Nx);
            const float2 f = make_float2(v[k1].x * inv, v[k1].y * inv);
            const float2 kf = cmulf(K[idx], f);
            sGam[idx].x += lam * kf.x;
            sGam[idx].y += lam * kf.y;
          }
        }
      }
    }
  }
  __syncthreads();
  float2* out = gspec + (size_t)b * 32 * Gx;
  for (int i = tid; i < 32 * Gx; i += 256) {
    const int gy = i / Gx, gx = i - gy * Gx - 2 * Nx;
    float2 j = make_float2(0.f, 0.f);
    if (gy <= Ny && gx >= -Nx && gx <= Nx) {
      const float2 gm = sGam[(-gy + Ny) * Wx + (-gx + Nx)];  // Gam[-n]
      const float2 gp = sGam[(gy + Ny) * Wx + (gx + Nx)];    // Gam[n]
      j = make_float2(gm.x + gp.x, gm.y - gp.y);
      if (dcx | dcy) {
        float sn, cs;
        sincospif(2.0f * ((float)(gx * dcx) / (float)Lx + (float)(gy * dcy) / (float)Ly), &sn, &cs);
        j = cmulf(j, make_float2(cs, sn));
      }
    }
    out[i] = j;
  }
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
  // software pipeline over row blocks: the next block's TY operands load
  // while this block's products and stores run
  half8_t nh0, nh1, nl0, nl1;
  {
    const int y1 = rp * 16;
    const _Float16* th = TYh + (size_t)(y1 + r16) * 64 + kq;
    const _Float16* tl = TYl + (size_t)(y1 + r16) * 64 + kq;
    nh0 = *reinterpret_cast<const half8_t*>(th);
    nh1 = *reinterpret_cast<const half8_t*>(th + 32);
    nl0 = *reinterpret_cast<const half8_t*>(tl);
    nl1 = *reinterpret_cast<const half8_t*>(tl + 32);
  }
  for (int yb = rp * 16; yb < Lypad; yb += 32) {
    const half8_t ah0 = nh0, ah1 = nh1, al0 = nl0, al1 = nl1;
    if (yb + 32 < Lypad) {
      const _Float16* th = TYh + (size_t)(yb + 32 + r16) * 64 + kq;
      const _Float16* tl = TYl + (size_t)(yb + 32 + r16) * 64 + kq;
      nh0 = *reinterpret_cast<const half8_t*>(th);
      nh1 = *reinterpret_cast<const half8_t*>(th + 32);
      nl0 = *reinterpret_cast<const half8_t*>(tl);
      nl1 = *reinterpret_cast<const half8_t*>(tl + 32);
    }
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
  // resist-gradient band: rows n = 2 gx (cos), 2 gx + 1 (-sin), gx <= 2Nx,
  // centre (Lx+1)/2; column twiddles e^{-2 pi i gy (y - (Ly+1)/2)/Ly}, gy <= 2Ny
  {
    const int cox = (Lx + 1) / 2;
    std::vector<_Float16> gh((size_t)64 * t->Kpad), gl((size_t)64 * t->Kpad);
    for (int n = 0; n < 64; ++n)
      for (int x = 0; x < t->Kpad; ++x) {
        double v = 0.0;
        const int gx = n / 2;
        if (gx <= 2 * Nx && x < Lx) {
          long long r = ((long long)gx * (x - cox)) % Lx;
          if (r < 0) r += Lx;
          const double a = 2.0 * M_PI * (double)r / (double)Lx;
          v = (n & 1) ? -std::sin(a) : std::cos(a);
        }
        split_half(v, &gh[(size_t)n * t->Kpad + x], &gl[(size_t)n * t->Kpad + x]);
      }
    KLT_TRY(t->Wgh.ensure(gh.size() * sizeof(_Float16)));
    KLT_TRY(t->Wgl.ensure(gl.size() * sizeof(_Float16)));
    KLT_HIP(hipMemcpy(t->Wgh.ptr, gh.data(), gh.size() * sizeof(_Float16), hipMemcpyHostToDevice));
    KLT_HIP(hipMemcpy(t->Wgl.ptr, gl.data(), gl.size() * sizeof(_Float16), hipMemcpyHostToDevice));
    std::vector<float2> ty((size_t)Ly * 32, make_float2(0.f, 0.f));
    for (int y = 0; y < Ly; ++y)
      for (int gy = 0; gy <= 2 * Ny; ++gy) {
        long long r = ((long long)gy * (y - coy)) % Ly;
        if (r < 0) r += Ly;
        const double a = 2.0 * M_PI * (double)r / (double)Ly;
        ty[(size_t)y * 32 + gy] = make_float2((float)std::cos(a), (float)-std::sin(a));
      }
    KLT_TRY(t->twyg.ensure(ty.size() * sizeof(float2)));
    KLT_HIP(hipMemcpy(t->twyg.ptr, ty.data(), ty.size() * sizeof(float2), hipMemcpyHostToDevice));
  }
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
    KLT_LAUNCH(ctx, "k_band_mfma", (k_band_mfma<2, false>), dim3(nblk, nb), dim3(256), 0, m,
               nullptr, 0.f, 0.f, Lx, Ly, ft->Kpad, ft->Wh.as<_Float16>(),
               ft->Wl.as<_Float16>(), p->twy_in.as<float2>(), p->NP, Nx, Ny, 16,
               ctx->qpart.as<float2>(), nullptr);
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

int band_mask_partials(klt_ctx* ctx, int Lx, int Ly, int Nx, int Ny, const float* masks, int B,
                       float2* out) {
  SocsPlan* p = nullptr;
  KLT_TRY(get_socs_plan(ctx, Lx, Ly, Nx, Ny, &p));
  FusedTables* ft = nullptr;
  KLT_TRY(get_fused_tables(ctx, p, &ft));
  const int nblk = (Ly + kRowsBand - 1) / kRowsBand;
  KLT_LAUNCH(ctx, "k_band_mfma", (k_band_mfma<2, false>), dim3(nblk, B), dim3(256), 0, masks,
             nullptr, 0.f, 0.f, Lx, Ly, ft->Kpad, ft->Wh.as<_Float16>(), ft->Wl.as<_Float16>(),
             p->twy_in.as<float2>(), p->NP, Nx, Ny, 16, out, nullptr);
  KLT_HIP(hipGetLastError());
  return KLT_OK;
}

int band_resist_partials(klt_ctx* ctx, int Lx, int Ly, int Nx, int Ny, const float* images,
                         const float* targets, int B, float a_sig, float thr, float2* out,
                         float* lpart) {
  SocsPlan* p = nullptr;
  KLT_TRY(get_socs_plan(ctx, Lx, Ly, Nx, Ny, &p));
  FusedTables* ft = nullptr;
  KLT_TRY(get_fused_tables(ctx, p, &ft));
  const int nblk = (Ly + kRowsBand - 1) / kRowsBand;
  KLT_LAUNCH(ctx, "k_band_resist", (k_band_mfma<4, true>), dim3(nblk, B), dim3(256), 0, images,
             targets, a_sig, thr, Lx, Ly, ft->Kpad, ft->Wgh.as<_Float16>(),
             ft->Wgl.as<_Float16>(), ft->twyg.as<float2>(), 32, 2 * Nx, 2 * Ny, 32, out, lpart);
  KLT_HIP(hipGetLastError());
  return KLT_OK;
}

int socs_backward_fused(klt_ctx* ctx, const klt_kernels* kb, const float* masks,
                        const float* targets, int B, int Lx, int Ly, float a_sig, float thr,
                        float* grads, float* losses, float* images) {
  SocsPlan* p = nullptr;
  KLT_TRY(get_socs_plan(ctx, Lx, Ly, kb->Nx, kb->Ny, &p));
  FusedTables* ft = nullptr;
  KLT_TRY(get_fused_tables(ctx, p, &ft));
  const int Nx = kb->Nx, Ny = kb->Ny, Wy = 2 * Ny + 1, Gx = 4 * Nx + 1, Gy = 4 * Ny + 1;
  const int nblk = (Ly + kRowsBand - 1) / kRowsBand;
  const int SB = std::min(B, 512);
  KLT_TRY(ctx->qpart.ensure((size_t)SB * nblk * Wy * 16 * sizeof(float2)));
  KLT_TRY(ctx->spec.ensure((size_t)SB * 32 * Gx * sizeof(float2)));
  KLT_TRY(ctx->gpart.ensure((size_t)SB * nblk * Gy * 32 * sizeof(float2)));
  KLT_TRY(ctx->gspec.ensure((size_t)SB * 32 * Gx * sizeof(float2)));
  KLT_TRY(ctx->lpart.ensure((size_t)SB * nblk * sizeof(float)));
  float* img = images;
  if (!img) {
    KLT_TRY(ctx->bwd_img.ensure((size_t)SB * Lx * Ly * sizeof(float)));
  }
  const float inv_area = 1.0f / ((float)Lx * (float)Ly);
  const int dcx = (Lx + 1) / 2 - Lx / 2, dcy = (Ly + 1) / 2 - Ly / 2;
  for (int b0 = 0; b0 < B; b0 += SB) {
    const int nb = std::min(SB, B - b0);
    const size_t off = (size_t)b0 * Lx * Ly;
    float* I = images ? images + off : ctx->bwd_img.as<float>();
    KLT_LAUNCH(ctx, "k_band_mfma", (k_band_mfma<2, false>), dim3(nblk, nb), dim3(256), 0,
               masks + off, nullptr, 0.f, 0.f, Lx, Ly, ft->Kpad, ft->Wh.as<_Float16>(),
               ft->Wl.as<_Float16>(), p->twy_in.as<float2>(), p->NP, Nx, Ny, 16,
               ctx->qpart.as<float2>(), nullptr);
    KLT_LAUNCH(ctx, "k_socs_mid", k_socs_mid, dim3(nb), dim3(256), 0, ctx->qpart.as<float2>(),
               nblk, Nx, Ny, inv_area, kb->kern.as<float2>(), kb->scales.as<float>(), kb->nk,
               ft->tw64.as<float2>(), ctx->spec.as<float2>());
    KLT_LAUNCH(ctx, "k_synth_mfma", k_synth_mfma,
               dim3((Lx + kColsSynth - 1) / kColsSynth, nb), dim3(256), 0,
               ctx->spec.as<float2>(), Nx, Lx, Ly, ft->Lypad, p->twx_out.as<float2>(),
               ft->TYh.as<_Float16>(), ft->TYl.as<_Float16>(), I);
    KLT_LAUNCH(ctx, "k_band_resist", (k_band_mfma<4, true>), dim3(nblk, nb), dim3(256), 0, I,
               targets + off, a_sig, thr, Lx, Ly, ft->Kpad, ft->Wgh.as<_Float16>(),
               ft->Wgl.as<_Float16>(), ft->twyg.as<float2>(), 32, 2 * Nx, 2 * Ny, 32,
               ctx->gpart.as<float2>(), ctx->lpart.as<float>());
    KLT_LAUNCH(ctx, "k_socs_mid_bwd", k_socs_mid_bwd, dim3(nb), dim3(256), 0,
               ctx->qpart.as<float2>(), ctx->gpart.as<float2>(), ctx->lpart.as<float>(), nblk,
               Nx, Ny, inv_area, kb->kern.as<float2>(), kb->scales.as<float>(), kb->nk,
               ft->tw64.as<float2>(), dcx, dcy, Lx, Ly, ctx->gspec.as<float2>(), losses + b0);
    KLT_LAUNCH(ctx, "k_synth_mfma", k_synth_mfma,
               dim3((Lx + kColsSynth - 1) / kColsSynth, nb), dim3(256), 0,
               ctx->gspec.as<float2>(), Nx, Lx, Ly, ft->Lypad, p->twx_out.as<float2>(),
               ft->TYh.as<_Float16>(), ft->TYl.as<_Float16>(), grads + off);
    KLT_HIP(hipGetLastError());
  }
  return KLT_OK;
}

}  // namespace klt

extern "C" {

int klt_socs_backward(klt_ctx* ctx, const klt_kernels* kb, const float* masks,
                      const float* targets, int nmask, int Lx, int Ly, float a_sig, float thr,
                      float* grads, float* losses, float* images, unsigned flags) {
  if (!ctx || !kb || !masks || !targets || !grads || !losses)
    return klt::set_error(KLT_ERR_ARG, "klt_socs_backward: NULL argument");
  if (nmask <= 0 || Lx <= 0 || Ly <= 0)
    return klt::set_error(KLT_ERR_ARG, "klt_socs_backward: bad size");
  if (4 * kb->Nx + 1 > Lx || 4 * kb->Ny + 1 > Ly)
    return klt::set_error(KLT_ERR_SIZE, "ERROR in FI()");
  if (!klt::socs_fused_supported(kb))
    return klt::set_error(KLT_ERR_ARG, "klt_socs_backward: kernel half-size above 15");
  KLT_TRY(ctx->activate());
  const size_t bytes = (size_t)nmask * Lx * Ly * sizeof(float);
  const float *dm = masks, *dt = targets;
  float *dg = grads, *dl = losses, *di = images;
  if (flags & KLT_HOST_PTRS) {
    KLT_TRY(ctx->stage_in.ensure(2 * bytes));
    char* s = ctx->stage_in.as<char>();
    KLT_HIP(hipMemcpyAsync(s, masks, bytes, hipMemcpyHostToDevice, ctx->stream));
    KLT_HIP(hipMemcpyAsync(s + bytes, targets, bytes, hipMemcpyHostToDevice, ctx->stream));
    dm = reinterpret_cast<const float*>(s);
    dt = reinterpret_cast<const float*>(s + bytes);
    KLT_TRY(ctx->stage_out.ensure(2 * bytes + (size_t)nmask * 4 + 256));
    char* o = ctx->stage_out.as<char>();
    dg = reinterpret_cast<float*>(o);
    di = images ? reinterpret_cast<float*>(o + bytes) : nullptr;
    dl = reinterpret_cast<float*>(o + 2 * bytes);
  }
  KLT_TRY(klt::socs_backward_fused(ctx, kb, dm, dt, nmask, Lx, Ly, a_sig, thr, dg, dl, di));
  if (flags & KLT_HOST_PTRS) {
    KLT_HIP(hipMemcpyAsync(grads, dg, bytes, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipMemcpyAsync(losses, dl, (size_t)nmask * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (images) KLT_HIP(hipMemcpyAsync(images, di, bytes, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return KLT_OK;
}

}  // extern "C"
