// tcc.hip -- TCC build on MI355X (replaces calcTCC, klitho_tcc.cpp:187-267).
//
// TCC[i][j] = sum_s S_s P(i,s) conj(P(j,s)) over the source points the
// reference uses.  The pupil P(i, s) (pupil_value) takes the reference's
// float32 membership decisions and phase, evaluated with correctly rounded
// operations in the reference's order (this file is compiled with
// -ffp-contract=off).  Two precisions:
//   split16: k_split_pupil evaluates P straight into the operand
//     X = sqrt(2^e S) P (real and imaginary parts interleaved along K, split
//     into hi + lo halves), then k_gram_split16: Re TCC = X X^T and
//     Im TCC = X X'^T, X' the lane-local pair swap (x_re, x_im) -> (-x_im, x_re)
//     of X, each product ~ hh + hl + lh with FP32 accumulation
//     (v_mfma_f32_16x16x32_f16), upper-triangle 128 x 128 tiles over the
//     K-chunks that reach both panels, and k_gram_reduce;
//   FP64: k_pupil writes P[s][i], and k_gram_f64 accumulates the Gram in
//     double like the reference.
// The TCC is written as T x T complex double, row-major, both triangles.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "klt_internal.h"
#include "klt_pupil.h"

namespace {

using klt::PupilGeom;
using klt::SrcPoint;
using klt::pupil_value;

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// P[s][i] for every used source point (blockIdx.y) and frequency i.
__global__ __launch_bounds__(256) void k_pupil(PupilGeom g, const SrcPoint* __restrict__ pts,
                                               float2* __restrict__ P) {
  const int s = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= g.T) return;
  const int W = 2 * g.Nx + 1;
  const int ny = i / W - g.Ny, nx = i % W - g.Nx;
  float2 v = make_float2(0.f, 0.f);
  const SrcPoint sp = pts[s];
  pupil_value(g, sp.p, sp.q, nx, ny, &v);
  P[(size_t)s * g.T + i] = v;
}

// The split operand of the MFMA Gram, [Tpad][Kpad] (k contiguous), straight
// from the pupil function:
//   X[i][2 s] = sqrt(scale S_s) Re P(i, s),  X[i][2 s + 1] = sqrt(scale S_s) Im P(i, s)
// as hi + lo fp16 (zero past T and past 2 ns).  S >= 0, so
//   sum_k X[i][k] X[j][k] = scale Re sum_s S_s P(i,s) conj(P(j,s)).
// grid (Tpad / 64, Kpad / 64), block (x, y) takes rows i0 = 64 x .. + 63 and
// source points 32 y .. + 31 (K columns 64 y .. + 63).  A block whose two
// 16-point chunks are dead for its 128-row panel (live[panel * nc + chunk] ==
// 0: no tile of the Gram reads them) writes nothing.
__global__ __launch_bounds__(256) void k_split_pupil(PupilGeom g, const SrcPoint* __restrict__ pts,
                                                     int ns, int Kpad, float scale,
                                                     const unsigned char* __restrict__ live,
                                                     int nc, _Float16* __restrict__ Xh,
                                                     _Float16* __restrict__ Xl) {
  const int i0 = blockIdx.x * 64, kb = blockIdx.y * 64;
  const int panel = i0 / 128, c0 = kb / 32;
  if (!live[panel * nc + c0] && !live[panel * nc + c0 + 1]) return;
  const int W = 2 * g.Nx + 1;
  for (int e = threadIdx.x; e < 64 * 32; e += 256) {
    const int ii = e >> 5, ss = e & 31;
    const int i = i0 + ii, s = kb / 2 + ss;
    float2 v = make_float2(0.f, 0.f);
    float w = 0.f;
    if (i < g.T && s < ns) {
      const SrcPoint sp = pts[s];
      pupil_value(g, sp.p, sp.q, i % W - g.Nx, i / W - g.Ny, &v);
      w = sqrtf(scale * sp.S);
    }
    const float xr = w * v.x, xi = w * v.y;
    const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
    const _Float16 lr = (_Float16)(xr - (float)hr), li = (_Float16)(xi - (float)hi);
    const size_t o = (size_t)i * Kpad + kb + 2 * ss;
    typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
    *reinterpret_cast<half2_t*>(Xh + o) = half2_t{hr, hi};
    *reinterpret_cast<half2_t*>(Xl + o) = half2_t{lr, li};
  }
}

// (x_0, x_1, ..., x_7) -> (-x_1, x_0, -x_3, x_2, ...): the operand of the
// imaginary part, exact on both split halves.
__device__ __forceinline__ half8_t swap_neg(half8_t b) {
  half8_t r;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    r[2 * t] = -b[2 * t + 1];
    r[2 * t + 1] = b[2 * t];
  }
  return r;
}

// Split-FP16 MFMA Gram over the live K-chunks: one 128 x 128 output tile over
// a run of 32-wide K-chunks per block.  order[blockIdx.x] is the block's job
// (or < 0 for an idle block); jobs[job] = (row panel, column panel, offset,
// count) of its run in chunks[] (chunk indices, each 32 halves of K = 16
// source points).  A tile's runs cover exactly the chunks whose source points
// can reach both of its panels: a chunk that one panel's frequencies lie
// outside of every pupil disc of contributes zero, and is skipped, as the
// reference skips P = 0 (klitho_tcc.cpp:246).  The host deals runs of equal
// length to the blocks and orders them so that the blocks of one XCD (block
// ids equal mod 8) take consecutive jobs.  Waves 2 x 2, each a 64 x 64
// sub-tile of 4 x 4 MFMA blocks with real (X_r . X_c) and imaginary
// (X_r . X'_c) accumulators; partial sums go to part[job][128][128] (re, im).
// v_mfma_f32_16x16x32_f16 operand layout: lane l holds A[m = l%16][k = 8(l/16)
// .. +7] and B[k = 8(l/16) .. +7][n = l%16], so the pair swap of X' stays in
// the lane; the result D[m = 4(l/16) + r][n = l%16], r = 0..3.  Per chunk the
// block stages the chunk's X rows of both panels (hi, lo: 4 x 128 rows x 32
// halves) in LDS and issues the next chunk's global loads before its products.
constexpr int kGT = 128;     // output tile
// Gram blocks: one a CU.  Each writes a 128 x 128 partial (128 KB) that the
// reduction reads back; at T = 1225 the whole build measured 0.078 ms with
// 256, 0.086 ms with 128 (a slower Gram) and with 512 (a slower reduction).
constexpr int kGramJobs = 256;
constexpr int kGLd = 40;     // LDS row stride (halves): 80 bytes
__global__ __launch_bounds__(256, 1) void k_gram_split16(
    const _Float16* __restrict__ Xh, const _Float16* __restrict__ Xl, int Kpad,
    const int* __restrict__ order, const int4* __restrict__ jobs,
    const int* __restrict__ chunks, float2* __restrict__ part) {
  // X rows of the tile's row panel (hi, lo) and of its column panel (hi, lo)
  __shared__ __attribute__((aligned(16))) _Float16 sop[4][kGT * kGLd];
  const int jid = order[blockIdx.x];
  if (jid < 0) return;
  const int4 job = jobs[jid];
  const int2 t = make_int2(job.x, job.y);
  const int* run = chunks + job.z;
  const int nrun = job.w;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wr = wave >> 1, wc = wave & 1;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  // staging: thread tid moves 16-byte chunks tid and tid + 256 of each array
  // (chunk c: row c / 4, halves 8 (c % 4) .. + 8 of the step); offsets of
  // its two chunks in the U rows (ou) and the V rows (ov) at k = 0
  const int row0 = tid >> 2, col0 = (tid & 3) * 8;  // chunk tid + 256: row + 64
  const size_t ou0 = (size_t)(t.x * kGT + row0) * Kpad + col0, ou1 = ou0 + (size_t)64 * Kpad;
  const size_t ov0 = (size_t)(t.y * kGT + row0) * Kpad + col0, ov1 = ov0 + (size_t)64 * Kpad;
  uint4 g0, g1, g2, g3, g4, g5, g6, g7;
#define KLT_GRAM_FETCH(k)                                                   \
  do {                                                                      \
    g0 = *reinterpret_cast<const uint4*>(Xh + ou0 + (k));                   \
    g1 = *reinterpret_cast<const uint4*>(Xh + ou1 + (k));                   \
    g2 = *reinterpret_cast<const uint4*>(Xl + ou0 + (k));                   \
    g3 = *reinterpret_cast<const uint4*>(Xl + ou1 + (k));                   \
    g4 = *reinterpret_cast<const uint4*>(Xh + ov0 + (k));                   \
    g5 = *reinterpret_cast<const uint4*>(Xh + ov1 + (k));                   \
    g6 = *reinterpret_cast<const uint4*>(Xl + ov0 + (k));                   \
    g7 = *reinterpret_cast<const uint4*>(Xl + ov1 + (k));                   \
  } while (0)
  float4_t accR[4][4], accI[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      accR[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
      accI[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
    }
  if (nrun > 0) KLT_GRAM_FETCH(run[0] * 32);
  const int l0 = row0 * kGLd + col0, l1 = l0 + 64 * kGLd;
  for (int u = 0; u < nrun; ++u) {
    __syncthreads();  // the previous chunk's operand reads are done
    *reinterpret_cast<uint4*>(&sop[0][l0]) = g0;
    *reinterpret_cast<uint4*>(&sop[0][l1]) = g1;
    *reinterpret_cast<uint4*>(&sop[1][l0]) = g2;
    *reinterpret_cast<uint4*>(&sop[1][l1]) = g3;
    *reinterpret_cast<uint4*>(&sop[2][l0]) = g4;
    *reinterpret_cast<uint4*>(&sop[2][l1]) = g5;
    *reinterpret_cast<uint4*>(&sop[3][l0]) = g6;
    *reinterpret_cast<uint4*>(&sop[3][l1]) = g7;
    __syncthreads();
    if (u + 1 < nrun) KLT_GRAM_FETCH(run[u + 1] * 32);
    // A fragments of the wave's 4 row blocks (hi, lo), kept over the columns
    half8_t ah[4], al[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int row = wr * 64 + a * 16 + r16;
      ah[a] = *reinterpret_cast<const half8_t*>(&sop[0][row * kGLd + kq]);
      al[a] = *reinterpret_cast<const half8_t*>(&sop[1][row * kGLd + kq]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col = wc * 64 + b * 16 + r16;
      const half8_t b1h = *reinterpret_cast<const half8_t*>(&sop[2][col * kGLd + kq]);
      const half8_t b1l = *reinterpret_cast<const half8_t*>(&sop[3][col * kGLd + kq]);
      const half8_t b2h = swap_neg(b1h), b2l = swap_neg(b1l);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], b1h, accR[a][b], 0, 0, 0);
        accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], b1l, accR[a][b], 0, 0, 0);
        accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], b1h, accR[a][b], 0, 0, 0);
        accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], b2h, accI[a][b], 0, 0, 0);
        accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], b2l, accI[a][b], 0, 0, 0);
        accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], b2h, accI[a][b], 0, 0, 0);
      }
    }
  }
#undef KLT_GRAM_FETCH
  float2* out = part + (size_t)jid * kGT * kGT;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ii = wr * 64 + a * 16 + (lane >> 4) * 4 + r, jj = wc * 64 + b * 16 + r16;
        out[ii * kGT + jj] = make_float2(accR[a][b][r], accI[a][b][r]);
      }
}
// Sums the partials of each 128 x 128 tile (jobs tj.x .. tj.x + tj.y - 1 of
// tile_jobs[tile]; none for a tile that no chunk reaches) in a fixed order,
// scales them and writes both triangles.  grid (ntiles, kGT / 8): block
// (tile, r) takes rows 8 r .. + 7 of the tile, read and written along rows;
// the strip passes through LDS so that the mirrored entries tcc[j][i] go out
// in runs of 8 along rows as well.
__global__ __launch_bounds__(256) void k_gram_reduce(const float2* __restrict__ part,
                                                     const int2* __restrict__ tile_jobs,
                                                     const int2* __restrict__ tiles, int T,
                                                     double inv_scale,
                                                     double* __restrict__ tcc) {
  __shared__ double2 sv[8][kGT + 1];
  const int tile = blockIdx.x, r0 = blockIdx.y * 8;
  const int2 t = tiles[tile], tj = tile_jobs[tile];
  const bool diag_tile = t.x == t.y;
  for (int e = threadIdx.x; e < 8 * kGT; e += 256) {
    const int rr = e >> 7, jj = e & (kGT - 1);
    const int i = t.x * kGT + r0 + rr, j = t.y * kGT + jj;
    double re = 0.0, im = 0.0;
    for (int q = 0; q < tj.y; ++q) {
      const float2 v = part[(size_t)(tj.x + q) * kGT * kGT + (r0 + rr) * kGT + jj];
      re += (double)v.x;
      im += (double)v.y;
    }
    re *= inv_scale;
    im = i == j ? 0.0 : im * inv_scale;
    sv[rr][jj] = make_double2(re, im);
    if (i < T && j < T && !(diag_tile && i > j)) {
      tcc[2 * ((size_t)i * T + j)] = re;
      tcc[2 * ((size_t)i * T + j) + 1] = im;
    }
  }
  __syncthreads();
  // mirror: row j of the TCC gets conj of column j of the strip
  for (int e = threadIdx.x; e < 8 * kGT; e += 256) {
    const int jj = e >> 3, rr = e & 7;
    const int i = t.x * kGT + r0 + rr, j = t.y * kGT + jj;
    if (i >= T || j >= T || i >= j) continue;
    const double2 v = sv[rr][jj];
    tcc[2 * ((size_t)j * T + i)] = v.x;
    tcc[2 * ((size_t)j * T + i) + 1] = -v.y;
  }
}

// FP64 reference Gram: 32 x 32 upper-triangle output tiles, source points
// staged through LDS 16 at a time; thread (ty, tx) owns a 2 x 2 block.
__global__ __launch_bounds__(256) void k_gram_f64(const float2* __restrict__ P,
                                                  const SrcPoint* __restrict__ pts, int ns,
                                                  int T, const int2* __restrict__ tiles,
                                                  double* __restrict__ tcc) {
  __shared__ double2 sx[16][32];
  __shared__ double2 sy[16][32];
  const int2 t = tiles[blockIdx.x];
  const int i0 = t.x * 32, j0 = t.y * 32;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double2 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = make_double2(0.0, 0.0);
  for (int s0 = 0; s0 < ns; s0 += 16) {
    for (int e = threadIdx.x; e < 16 * 32; e += 256) {
      const int ss = e >> 5, c = e & 31;
      const int s = s0 + ss;
      double2 x = make_double2(0.0, 0.0), y = make_double2(0.0, 0.0);
      if (s < ns) {
        const double S = (double)pts[s].S;
        if (i0 + c < T) {
          const float2 p = P[(size_t)s * T + i0 + c];
          x = make_double2((double)p.x * S, (double)p.y * S);
        }
        if (j0 + c < T) {
          const float2 p = P[(size_t)s * T + j0 + c];
          y = make_double2((double)p.x, (double)p.y);
        }
      }
      sx[ss][c] = x;
      sy[ss][c] = y;
    }
    __syncthreads();
#pragma unroll 4
    for (int ss = 0; ss < 16; ++ss) {
      const double2 x0 = sx[ss][ty * 2], x1 = sx[ss][ty * 2 + 1];
      const double2 y0 = sy[ss][tx * 2], y1 = sy[ss][tx * 2 + 1];
      // x * conj(y)
      acc[0][0].x += x0.x * y0.x + x0.y * y0.y;
      acc[0][0].y += x0.y * y0.x - x0.x * y0.y;
      acc[0][1].x += x0.x * y1.x + x0.y * y1.y;
      acc[0][1].y += x0.y * y1.x - x0.x * y1.y;
      acc[1][0].x += x1.x * y0.x + x1.y * y0.y;
      acc[1][0].y += x1.y * y0.x - x1.x * y0.y;
      acc[1][1].x += x1.x * y1.x + x1.y * y1.y;
      acc[1][1].y += x1.y * y1.x - x1.x * y1.y;
    }
    __syncthreads();
  }
  const bool diag_tile = t.x == t.y;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const int i = i0 + ty * 2 + a, j = j0 + tx * 2 + b;
      if (i >= T || j >= T || (diag_tile && i > j)) continue;
      const double im = i == j ? 0.0 : acc[a][b].y;
      tcc[2 * ((size_t)i * T + j)] = acc[a][b].x;
      tcc[2 * ((size_t)i * T + j) + 1] = im;
      if (i != j) {
        tcc[2 * ((size_t)j * T + i)] = acc[a][b].x;
        tcc[2 * ((size_t)j * T + i) + 1] = -im;
      }
    }
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace

namespace klt {

// Source points in the reference's order (q outer, p inner), klitho_tcc.cpp:202-205.
static std::vector<SrcPoint> source_points(const float* src, int srcSize, float outerSigma) {
  const int sh = (int)((srcSize - 1) / 2.0);
  const int oS = (int)std::ceil((float)sh * outerSigma);
  std::vector<SrcPoint> pts;
  for (int q = -oS; q <= oS; ++q)
    for (int p = -oS; p <= oS; ++p) {
      if (q + sh < 0 || q + sh >= srcSize || p + sh < 0 || p + sh >= srcSize) continue;
      const float S = src[(q + sh) * srcSize + p + sh];
      if (S != 0.0f && p * p + q * q <= sh * sh) pts.push_back({p, q, S});
    }
  return pts;
}

// Pupil matrix of every used source point into ctx->misc ([ns][T] float2);
// the points are left in ctx->tcc_pts.
static int build_pupil(klt_ctx* ctx, const float* src_host, int srcSize, float outerSigma,
                       const PupilGeom& g, int* ns_out, std::vector<SrcPoint>& pts) {
  pts = source_points(src_host, srcSize, outerSigma);
  const int ns = (int)pts.size();
  *ns_out = ns;
  if (ns == 0) return KLT_OK;
  KLT_TRY(ctx->tcc_pts.ensure(sizeof(SrcPoint) * ns));
  KLT_TRY(ctx->upload(ctx->tcc_pts.ptr, pts.data(), sizeof(SrcPoint) * ns));
  KLT_TRY(ctx->misc.ensure(sizeof(float2) * (size_t)ns * g.T));
  KLT_LAUNCH(ctx, "k_pupil", k_pupil, dim3((g.T + 255) / 256, ns), dim3(256), 0, g,
             ctx->tcc_pts.as<SrcPoint>(), ctx->misc.as<float2>());
  KLT_HIP(hipGetLastError());
  return KLT_OK;
}

// Conservative reach of each 16-point K-chunk (32 halves) of the split
// operand: live[a * nc + c] = 1 when some frequency of the 128-row panel a lies
// in the bounding box (widened by one cell) that pupil_value checks first for
// some source point of chunk c; outside every box the operand is zero.
static std::vector<unsigned char> chunk_reach(const PupilGeom& g, const std::vector<SrcPoint>& pts,
                                              int np, int nc) {
  const int T = g.T, W = 2 * g.Nx + 1;
  std::vector<unsigned char> live((size_t)np * nc, 0);
  for (size_t si = 0; si < pts.size(); ++si) {
    const int c = (int)(si / 16);
    int bx0 = -g.Nx, bx1 = g.Nx, by0 = -g.Ny, by1 = g.Ny;
    if (g.sh > 0) {
      const float sx = (float)pts[si].p / (float)g.sh, sy = (float)pts[si].q / (float)g.sh;
      bx0 = (int)((-1.0f - sx) * g.Lxn) - 1;
      bx1 = (int)((1.0f - sx) * g.Lxn) + 1;
      by0 = (int)((-1.0f - sy) * g.Lyn) - 1;
      by1 = (int)((1.0f - sy) * g.Lyn) + 1;
    }
    for (int a = 0; a < np; ++a) {
      unsigned char& m = live[(size_t)a * nc + c];
      const int i_lo = a * kGT, i_hi = std::min(T, i_lo + kGT) - 1;
      if (m || i_lo > i_hi) continue;
      for (int r = i_lo / W; r <= i_hi / W && !m; ++r) {
        const int ny = r - g.Ny;
        if (ny < by0 || ny > by1) continue;
        const int x_lo = (r == i_lo / W ? i_lo % W : 0) - g.Nx;
        const int x_hi = (r == i_hi / W ? i_hi % W : W - 1) - g.Nx;
        if (x_hi >= bx0 && x_lo <= bx1) m = 1;
      }
    }
  }
  return live;
}

// The split16 plan of one source and geometry: points, scale, chunk reach,
// tiles, job runs and block order, packed into ctx->tcc_tab (uploaded when it
// changes); sizes and offsets into *pl.
static int plan_split16(klt_ctx* ctx, const PupilGeom& g, const float* src_host, int srcSize,
                        float outerSigma, klt_ctx::TccPlan* pl) {
  const std::vector<SrcPoint> pts = source_points(src_host, srcSize, outerSigma);
  const int ns = (int)pts.size();
  pl->ns = ns;
  if (ns == 0) return KLT_OK;
  // power-of-two scale bringing max |S| into [0.5, 1)
  float smax = 0.f;
  for (auto& p : pts) smax = std::max(smax, std::fabs(p.S));
  int e = 0;
  (void)std::frexp(smax, &e);
  const float scale = std::ldexp(1.0f, -e);
  const int Kpad = round_up(2 * ns, 64), Tpad = round_up(g.T, kGT);
  const int np = Tpad / kGT, nc = Kpad / 32;
  const std::vector<unsigned char> live = chunk_reach(g, pts, np, nc);
  // upper-triangle tiles and the chunks each needs (live for both panels)
  std::vector<int2> tiles;
  std::vector<std::vector<int>> need;
  size_t work = 0;
  for (int a = 0; a < np; ++a)
    for (int b = a; b < np; ++b) {
      tiles.push_back(make_int2(a, b));
      need.emplace_back();
      for (int c = 0; c < nc; ++c)
        if (live[(size_t)a * nc + c] && live[(size_t)b * nc + c]) need.back().push_back(c);
      work += need.back().size();
    }
  const int ntl = (int)tiles.size();
  // runs of at most per_job chunks, the smallest length that needs no more
  // than kGramJobs blocks (at most one a CU, a single round); a tile's runs
  // are consecutive jobs, summed in order by the reduction
  int per_job = std::max<int>(1, (int)((work + kGramJobs - 1) / kGramJobs));
  for (;; ++per_job) {
    size_t nj = 0;
    for (auto& nd : need) nj += (nd.size() + per_job - 1) / per_job;
    if (nj <= (size_t)kGramJobs) break;
  }
  std::vector<int2> tile_jobs(ntl);
  std::vector<int4> jobs;
  std::vector<int> chunks, run_of;
  for (int t = 0; t < ntl; ++t) {
    const int nu = (int)need[t].size(), nj = (nu + per_job - 1) / per_job;
    tile_jobs[t] = make_int2((int)jobs.size(), nj);
    for (int r = 0; r < nj; ++r) {
      const int u0 = (int)((long)r * nu / nj), u1 = (int)((long)(r + 1) * nu / nj);
      jobs.push_back(make_int4(tiles[t].x, tiles[t].y, (int)chunks.size(), u1 - u0));
      run_of.push_back(r);
      for (int u = u0; u < u1; ++u) chunks.push_back(need[t][u]);
    }
  }
  const int njobs = (int)jobs.size();
  // run-major job order dealt to the 8 XCDs in consecutive runs: block id
  // 8 i + x (XCD x under the round-robin dispatch) takes the (x per + i)-th job
  std::vector<int> byrun(njobs);
  for (int j = 0; j < njobs; ++j) byrun[j] = j;
  std::stable_sort(byrun.begin(), byrun.end(),
                   [&](int x, int y) { return run_of[x] < run_of[y]; });
  constexpr int kXcds = 8;
  const int per = std::max(1, (njobs + kXcds - 1) / kXcds);
  std::vector<int> order((size_t)per * kXcds, -1);
  for (int x = 0; x < kXcds; ++x)
    for (int i = 0; i < per; ++i)
      if (x * per + i < njobs) order[(size_t)i * kXcds + x] = byrun[x * per + i];
  // one packed table: points, reach, tiles, tile jobs, jobs, block order,
  // chunk runs -- uploaded only when it changes
  size_t off = 0;
  auto sect = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 15) / 16 * 16;
    return o;
  };
  const size_t oP = sect(sizeof(SrcPoint) * ns), oL = sect(live.size()),
               oT = sect(sizeof(int2) * ntl), oTJ = sect(sizeof(int2) * ntl),
               oJ = sect(sizeof(int4) * njobs), oO = sect(sizeof(int) * order.size()),
               oC = sect(sizeof(int) * std::max<size_t>(1, chunks.size()));
  std::vector<char> tab(off, 0);
  std::memcpy(tab.data() + oP, pts.data(), sizeof(SrcPoint) * ns);
  std::memcpy(tab.data() + oL, live.data(), live.size());
  std::memcpy(tab.data() + oT, tiles.data(), sizeof(int2) * ntl);
  std::memcpy(tab.data() + oTJ, tile_jobs.data(), sizeof(int2) * ntl);
  if (njobs) std::memcpy(tab.data() + oJ, jobs.data(), sizeof(int4) * njobs);
  std::memcpy(tab.data() + oO, order.data(), sizeof(int) * order.size());
  if (!chunks.empty()) std::memcpy(tab.data() + oC, chunks.data(), sizeof(int) * chunks.size());
  if (tab != ctx->tcc_tab_host || ctx->tcc_tab.ptr == nullptr) {
    KLT_TRY(ctx->tcc_tab.ensure(off));
    KLT_TRY(ctx->upload(ctx->tcc_tab.ptr, tab.data(), off));
    ctx->tcc_tab_host.swap(tab);
  }
  pl->Kpad = Kpad;
  pl->Tpad = Tpad;
  pl->nc = nc;
  pl->ntl = ntl;
  pl->njobs = njobs;
  pl->nblocks = (int)order.size();
  pl->scale = scale;
  pl->oP = oP;
  pl->oL = oL;
  pl->oT = oT;
  pl->oTJ = oTJ;
  pl->oJ = oJ;
  pl->oO = oO;
  pl->oC = oC;
  return KLT_OK;
}

int tcc_build_device(klt_ctx* ctx, const float* src_host, int srcSize, float outerSigma,
                     float lambda, float NA, float defocus, float nmed, int Lx, int Ly,
                     int Nx, int Ny, int precision, double* tcc_dev) {
  const PupilGeom g = make_geom(srcSize, lambda, NA, defocus, nmed, Lx, Ly, Nx, Ny);
  const int T = g.T;
  if (precision != KLT_TCC_FP64 && precision != KLT_TCC_SPLIT_F16)
    return set_error(KLT_ERR_ARG, "klt_tcc_build: unknown precision");
  if (precision == KLT_TCC_FP64) {
    std::vector<SrcPoint> pts;
    int ns = 0;
    KLT_TRY(build_pupil(ctx, src_host, srcSize, outerSigma, g, &ns, pts));
    if (ns == 0) {
      KLT_HIP(hipMemsetAsync(tcc_dev, 0, sizeof(double) * 2 * (size_t)T * T, ctx->stream));
      return KLT_OK;
    }
    const int nt = (T + 31) / 32;
    std::vector<int2> tiles;
    for (int a = 0; a < nt; ++a)
      for (int b = a; b < nt; ++b) tiles.push_back(make_int2(a, b));
    KLT_TRY(ctx->spec.ensure(sizeof(int2) * tiles.size()));
    KLT_TRY(ctx->upload(ctx->spec.ptr, tiles.data(), sizeof(int2) * tiles.size()));
    KLT_LAUNCH(ctx, "k_gram_f64", k_gram_f64, dim3((unsigned)tiles.size()), dim3(256), 0,
               ctx->misc.as<float2>(), ctx->tcc_pts.as<SrcPoint>(), ns, T,
               ctx->spec.as<int2>(), tcc_dev);
    KLT_HIP(hipGetLastError());
    return KLT_OK;
  }
  // the plan (points, chunk reach, jobs) depends on the source and on the
  // geometry of pupil membership, not on defocus: reuse the last one
  auto& pl = ctx->tcc_plan;
  const std::array<float, 4> fp = {outerSigma, lambda, NA, nmed};
  const std::array<int, 5> ip = {srcSize, Lx, Ly, Nx, Ny};
  const size_t nsrc = (size_t)srcSize * srcSize;
  const bool hit = pl.valid && pl.fp == fp && pl.ip == ip && pl.src.size() == nsrc &&
                   std::memcmp(pl.src.data(), src_host, nsrc * sizeof(float)) == 0;
  if (!hit) {
    pl.valid = false;
    KLT_TRY(plan_split16(ctx, g, src_host, srcSize, outerSigma, &pl));
    pl.src.assign(src_host, src_host + nsrc);
    pl.fp = fp;
    pl.ip = ip;
    pl.valid = true;
  }
  if (pl.ns == 0) {
    KLT_HIP(hipMemsetAsync(tcc_dev, 0, sizeof(double) * 2 * (size_t)T * T, ctx->stream));
    return KLT_OK;
  }
  char* dt = ctx->tcc_tab.as<char>();
  const size_t nel = (size_t)pl.Tpad * pl.Kpad;
  KLT_TRY(ctx->qpart.ensure(2 * nel * sizeof(_Float16)));
  _Float16* base = ctx->qpart.as<_Float16>();
  KLT_LAUNCH(ctx, "k_split_pupil", k_split_pupil, dim3(pl.Tpad / 64, pl.Kpad / 64), dim3(256), 0,
             g, reinterpret_cast<const SrcPoint*>(dt + pl.oP), pl.ns, pl.Kpad, pl.scale,
             reinterpret_cast<const unsigned char*>(dt + pl.oL), pl.nc, base, base + nel);
  KLT_TRY(ctx->tcc_part.ensure((size_t)std::max(1, pl.njobs) * kGT * kGT * sizeof(float2)));
  if (pl.njobs)
    KLT_LAUNCH(ctx, "k_gram_split16", k_gram_split16, dim3((unsigned)pl.nblocks), dim3(256), 0,
               base, base + nel, pl.Kpad, reinterpret_cast<const int*>(dt + pl.oO),
               reinterpret_cast<const int4*>(dt + pl.oJ),
               reinterpret_cast<const int*>(dt + pl.oC), ctx->tcc_part.as<float2>());
  KLT_LAUNCH(ctx, "k_gram_reduce", k_gram_reduce, dim3((unsigned)pl.ntl, kGT / 8), dim3(256), 0,
             ctx->tcc_part.as<float2>(), reinterpret_cast<const int2*>(dt + pl.oTJ),
             reinterpret_cast<const int2*>(dt + pl.oT), T, 1.0 / (double)pl.scale, tcc_dev);
  KLT_HIP(hipGetLastError());
  return KLT_OK;
}

}  // namespace klt

extern "C" {

int klt_tcc_build(klt_ctx* ctx, const float* src, int srcSize, float outerSigma, float lambda,
                  float NA, float defocus, float nmed, int Lx, int Ly, int Nx, int Ny,
                  int precision, double* tcc, unsigned flags) {
  if (!ctx || !src || !tcc) return klt::set_error(KLT_ERR_ARG, "klt_tcc_build: NULL argument");
  if (srcSize <= 0 || Lx <= 0 || Ly <= 0 || Nx < 0 || Ny < 0 || lambda <= 0)
    return klt::set_error(KLT_ERR_ARG, "Error: Invalid input parameters. Please check the values.");
  if (NA < 0 || !(nmed >= 1.0f) || NA > nmed)
    return klt::set_error(KLT_ERR_ARG, nmed == 1.0f ? "Error: NA must be in the range [0, 1]."
                                                    : "Error: NA must not exceed the immersion index.");
  KLT_TRY(ctx->activate());
  std::vector<float> hsrc((size_t)srcSize * srcSize);
  if (flags & KLT_HOST_PTRS) {
    std::memcpy(hsrc.data(), src, sizeof(float) * hsrc.size());
  } else {  // through pinned memory (a pageable read is staged and slower)
    void* pinned = nullptr;
    KLT_TRY(ctx->download_area(sizeof(float) * hsrc.size(), &pinned));
    KLT_HIP(hipMemcpyAsync(pinned, src, sizeof(float) * hsrc.size(), hipMemcpyDeviceToHost,
                           ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(hsrc.data(), pinned, sizeof(float) * hsrc.size());
  }
  const int T = (2 * Nx + 1) * (2 * Ny + 1);
  const size_t bytes = sizeof(double) * 2 * (size_t)T * T;
  double* dst = tcc;
  if (flags & KLT_HOST_PTRS) {
    KLT_TRY(ctx->stage_out.ensure(bytes));
    dst = ctx->stage_out.as<double>();
  }
  KLT_TRY(klt::tcc_build_device(ctx, hsrc.data(), srcSize, outerSigma, lambda, NA, defocus, nmed,
                                Lx, Ly, Nx, Ny, precision, dst));
  if (flags & KLT_HOST_PTRS) {
    KLT_HIP(hipMemcpyAsync(tcc, dst, bytes, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return KLT_OK;
}

}  // extern "C"
