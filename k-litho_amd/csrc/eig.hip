// eig.hip -- top-k eigenpairs of a Hermitian matrix in FP64 on MI355X
// (replaces zheevr_ in calcKernels, K-Litho-TCC/klitho_tcc.cpp:301-377).
//
//   1. Householder reduction to tridiagonal form, one column per step:
//        k_hh_symv : every block rebuilds the reflector v, tau of row j
//                    (x = conj(A[j][j+1:])), then computes its rows of
//                    p = tau A22 v and a partial sum of v^H p;
//        k_hh_her2 : A22 -= v w^H + w v^H, w = p - (tau/2)(v^H p) v.
//      For n <= 3400 one launch per step (k_hh_step) does both: it applies
//      step j-1's rank-2 update to row j (for the reflector) and to the
//      block's rows of the trailing matrix while multiplying them by v;
//      once the trailing block has at most kTail rows, k_hh_tail runs the
//      remaining steps in the LDS of one block per matrix.
//      The off-diagonal entries are complex; a diagonal unitary D maps the
//      tridiagonal T to the real symmetric T_r = D^H T D with off-diagonal
//      |e_j|.
//   2. k_bisect: the k largest eigenvalues of T_r by multisection on
//      eigenvalue counts (64 points per pass, each count from a twisted
//      factorization whose two halves run on two waves).
//   3. k_inverse_iteration: eigenvectors of T_r by inverse iteration with a
//      partially pivoted LU of T_r - lambda I (one thread per vector);
//      eigenvalues closer than 1e-4 ||T|| form a cluster whose vectors are
//      re-orthogonalised in eigenvalue order at every iteration (Gram-Schmidt
//      parallel over the vector length, twice in the last of three rounds),
//      one block per cluster.
//   4. k_back_transform: x = H_0 ... H_{n-3} D z, one block per vector (D by a
//      block scan of the unit phases).
// Textbook algorithms (Golub & Van Loan, ch. 8; Wilkinson's inverse
// iteration); all arithmetic FP64.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "klt_internal.h"

namespace {

constexpr int kThreads = 256;
// most blocks of one Householder step (one partial sum of v^H p each)
constexpr int kMaxParts = 1024;

// Element strides between the matrices of a batch: the arrays of matrix mi
// start at base + mi * stride.  mi is the kernel's last grid dimension.
struct EigStride {
  size_t a;     // A, doubles (2 n^2)
  size_t v;     // V, complex (n^2)
  size_t n;     // tau, diag, e, e2 (doubles) and sub (complex): n
  size_t p;     // p, complex (2 n: two steps' sets)
  size_t part;  // block partials, complex (two steps' sets)
  size_t k;     // eigenvalues: k
  size_t z;     // z (doubles) and output vectors (complex): k n; lu work: 5 k n
  size_t g;     // Gram entries of the back-transform's reflector groups, complex: 3 n
};

struct cd {
  double x, y;
};

__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cmulc(cd a, cd b) {  // conj(a) * b
  return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}

// One DPP move of a double across lanes (both 32-bit halves with the same
// pattern); lanes without a source read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, true);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes, valid in every lane: pairs and quads (quad_perm),
// rows of 16 (row_shr 4, 8: lane 15 of each row holds the row's sum), then
// rows 0..3 into lane 63 (row_bcast 15 and 31), read back from lane 63.
// Register-to-register moves, about a third of the latency of a
// __shfl_xor butterfly.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0xb1>(v);       // quad_perm [1, 0, 3, 2]
  v += dpp_move<0x4e>(v);       // quad_perm [2, 3, 0, 1]
  v += dpp_move<0x114>(v);      // row_shr:4
  v += dpp_move<0x118>(v);      // row_shr:8
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Block sum of (a, b) over 256 threads; valid in every thread.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* red) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  a = red[0] + red[2] + red[4] + red[6];
  b = red[1] + red[3] + red[5] + red[7];
}

// Block sums of N values over 256 threads (red: 4 N doubles); valid in every
// thread.
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double* red) {
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) red[w * N + q] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = red[q] + red[N + q] + red[2 * N + q] + red[3 * N + q];
}

// Step j of the reduction.  A: n x n complex (row-major, full Hermitian
// storage).  Each block rebuilds the reflector of x = conj(A[j][j+1:]) in
// LDS, then computes p_i = tau sum_l A[j+1+i][j+1+l] v_l for rows
// i in [blk*rows, ...) and the partial sum of conj(v_i) p_i.
__global__ __launch_bounds__(kThreads) void k_hh_symv(double* __restrict__ A, int n, int j,
                                                      int rows, cd* __restrict__ V,
                                                      double* __restrict__ tau,
                                                      cd* __restrict__ sub,
                                                      double* __restrict__ diag,
                                                      cd* __restrict__ p,
                                                      cd* __restrict__ part, EigStride st) {
  extern __shared__ cd sv[];  // m entries
  __shared__ double red[8];
  const int m = n - j - 1;
  {
    const size_t mi = blockIdx.y;
    A += mi * st.a;
    V += mi * st.v;
    tau += mi * st.n;
    sub += mi * st.n;
    diag += mi * st.n;
    p += mi * st.p;
    part += mi * st.part;
  }
  const cd* Aj = reinterpret_cast<const cd*>(A) + (size_t)j * n + j + 1;
  double s2 = 0.0, dummy = 0.0;
  for (int i = threadIdx.x; i < m; i += kThreads) {
    const cd a = Aj[i];
    const cd xi = {a.x, -a.y};
    sv[i] = xi;
    s2 += xi.x * xi.x + xi.y * xi.y;
  }
  block_sum2(s2, dummy, red);
  const cd x0 = sv[0];
  const double ax0 = hypot(x0.x, x0.y);
  const double alpha = sqrt(s2);
  const bool skip = (s2 - ax0 * ax0) <= 0.0;  // x already a multiple of e_1
  cd ph = {1.0, 0.0};
  if (ax0 > 0.0) ph = {x0.x / ax0, x0.y / ax0};
  const double t = skip ? 0.0 : 1.0 / (alpha * (alpha + ax0));  // 2 / (v^H v)
  __syncthreads();
  if (threadIdx.x == 0 && !skip) {
    sv[0].x += ph.x * alpha;
    sv[0].y += ph.y * alpha;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    cd* Vj = V + (size_t)j * n + j + 1;
    for (int i = threadIdx.x; i < m; i += kThreads) Vj[i] = skip ? cd{0.0, 0.0} : sv[i];
    if (threadIdx.x == 0) {
      tau[j] = t;
      sub[j] = skip ? x0 : cd{-ph.x * alpha, -ph.y * alpha};
      diag[j] = A[2 * ((size_t)j * n + j)];
    }
  }
  // rows of p, one wave per row
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = blockIdx.x * rows, r1 = min(m, r0 + rows);
  double pr = 0.0, pi = 0.0;  // partial conj(v) . p of this wave
  for (int i = r0 + wave; i < r1; i += 4) {
    const cd* Ai = reinterpret_cast<const cd*>(A) + (size_t)(j + 1 + i) * n + j + 1;
    double sr = 0.0, si = 0.0;
    if (!skip) {
      for (int l = lane; l < m; l += 64) {
        const cd a = Ai[l], v = sv[l];
        sr += a.x * v.x - a.y * v.y;
        si += a.x * v.y + a.y * v.x;
      }
    }
    sr = wave_sum(sr) * t;
    si = wave_sum(si) * t;
    if (lane == 0) {
      p[i] = {sr, si};
      const cd vi = sv[i];
      pr += vi.x * sr + vi.y * si;
      pi += vi.x * si - vi.y * sr;
    }
  }
  block_sum2(pr, pi, red);
  if (threadIdx.x == 0) part[blockIdx.x] = {pr, pi};
}

// A22 -= v w^H + w v^H on 64 x 64 tiles; w = p - (tau/2)(v^H p) v.
// Step j with step j-1's update folded in.  A holds the matrix with the
// updates of steps < j-1 applied.  Step j-1's update A[j.., j..] -= u,
// u(i, l) = v'_i conj(w'_l) + w'_i conj(v'_l) with v' = V[j-1][j..],
// w' = p' - K' v', K' = (tau'/2) sum part', is applied
//   - to row j, by every block, before it builds the reflector x = conj(A[j][j+1:]);
//   - to the block's rows of the trailing matrix A[j+1.., j+1..], written back
//     as they are multiplied by v (p_i = tau sum_l A[j+1+i][j+1+l] v_l).
// Rows j, columns j and the entries left of them are finished and not
// updated.  Dynamic LDS: (3m + 2) complex values, m = n - j - 1.
__global__ __launch_bounds__(kThreads) void k_hh_step(double* __restrict__ A, int n, int j,
                                                      int rows, cd* __restrict__ V,
                                                      double* __restrict__ tau,
                                                      cd* __restrict__ sub,
                                                      double* __restrict__ diag,
                                                      const cd* __restrict__ p_prev,
                                                      const cd* __restrict__ part_prev,
                                                      int nparts_prev, cd* __restrict__ p,
                                                      cd* __restrict__ part, EigStride st) {
  extern __shared__ cd sh[];
  __shared__ double red[8];
  const int m = n - j - 1;
  {
    const size_t mi = blockIdx.y;
    A += mi * st.a;
    V += mi * st.v;
    tau += mi * st.n;
    sub += mi * st.n;
    diag += mi * st.n;
    p_prev += mi * st.p;
    part_prev += mi * st.part;
    p += mi * st.p;
    part += mi * st.part;
  }
  cd* sv = sh;              // row j (raw), then this step's reflector; m entries
  cd* svp = sh + m;         // v', m + 1 entries (global index j + i)
  cd* spp = svp + (m + 1);  // p'; w'_i = p'_i - K' v'_i is formed where used
  const double tp = j > 0 ? tau[j - 1] : 0.0;
  const bool upd = tp != 0.0;  // uniform over the grid
  // the wave's first trailing row (entries lane + 64 c, c < kPre) is loaded
  // first: it does not depend on the reflector, so its latency overlaps the
  // prologue
  constexpr int kPre = 24;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = blockIdx.x * rows, r1 = min(m, r0 + rows);
  cd pre[kPre];
  if (r0 + wave < r1) {
    const cd* Af = reinterpret_cast<const cd*>(A) + (size_t)(j + 1 + r0 + wave) * n + j + 1;
#pragma unroll
    for (int c = 0; c < kPre; ++c) {
      const int l = lane + 64 * c;
      pre[c] = l < m ? Af[l] : cd{0.0, 0.0};
    }
  }
  // every load of the prologue at once, one memory round trip for m < 2048:
  // step j-1's partial sums (at most kMaxParts, kPartsPer a thread), then
  // v', p' and row j, kU entries a thread per pass, all issued before any is
  // used (a loop that stores each entry to LDS as it arrives waits for one
  // round trip per 256 entries)
  constexpr int kPartsPer = kMaxParts / kThreads, kU = 8;
  const cd* Aj = reinterpret_cast<const cd*>(A) + (size_t)j * n + j + 1;
  const cd* vp = V + (size_t)(j > 0 ? j - 1 : 0) * n + j;
  // (the partial sums are loaded unconditionally, clamped to entry 0 of the
  // allocated set, and selected where they are added: a conditional load
  // would let the compiler start the sum, and wait for every load in
  // flight, right after it)
  cd pt[kPartsPer];
#pragma unroll
  for (int u = 0; u < kPartsPer; ++u) {
    const int q = threadIdx.x + u * kThreads;
    pt[u] = part_prev[q < nparts_prev ? q : 0];
  }
  for (int i0 = threadIdx.x; i0 <= m; i0 += kU * kThreads) {
    cd ra[kU], rv[kU], rp[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * kThreads;
      ra[u] = i < m ? Aj[i] : cd{0.0, 0.0};
      rv[u] = upd && i <= m ? vp[i] : cd{0.0, 0.0};
      rp[u] = upd && i <= m ? p_prev[i] : cd{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * kThreads;
      if (i < m) sv[i] = ra[u];
      if (upd && i <= m) {
        svp[i] = rv[u];
        spp[i] = rp[u];
      }
    }
  }
  double kr = 0.0, ki = 0.0;
#pragma unroll
  for (int u = 0; u < kPartsPer; ++u) {
    const bool use = upd && threadIdx.x + u * kThreads < nparts_prev;
    kr += use ? pt[u].x : 0.0;
    ki += use ? pt[u].y : 0.0;
  }
  block_sum2(kr, ki, red);  // (its barriers also publish the LDS writes)
  const cd K = {0.5 * tp * kr, 0.5 * tp * ki};
  auto wprev = [&](int i) {
    const cd v = svp[i], kv = cmul(K, v), pp = spp[i];
    return cd{pp.x - kv.x, pp.y - kv.y};
  };
  // row j after the update: entries l + 1 of step j-1's block, l < m
  const cd v0 = upd ? svp[0] : cd{0.0, 0.0}, w0 = upd ? wprev(0) : cd{0.0, 0.0};
  double s2 = 0.0, dummy = 0.0;
  for (int i = threadIdx.x; i < m; i += kThreads) {
    cd a = sv[i];
    if (upd) {
      const cd vl = svp[i + 1], wl = wprev(i + 1);
      a.x -= (v0.x * wl.x + v0.y * wl.y) + (w0.x * vl.x + w0.y * vl.y);
      a.y -= (v0.y * wl.x - v0.x * wl.y) + (w0.y * vl.x - w0.x * vl.y);
    }
    const cd xi = {a.x, -a.y};
    sv[i] = xi;  // (each thread rewrites only the entries it read)
    s2 += xi.x * xi.x + xi.y * xi.y;
  }
  block_sum2(s2, dummy, red);
  const cd x0 = sv[0];
  const double ax0 = hypot(x0.x, x0.y);
  const double alpha = sqrt(s2);
  const bool skip = (s2 - ax0 * ax0) <= 0.0;  // x already a multiple of e_1
  cd ph = {1.0, 0.0};
  if (ax0 > 0.0) ph = {x0.x / ax0, x0.y / ax0};
  const double t = skip ? 0.0 : 1.0 / (alpha * (alpha + ax0));  // 2 / (v^H v)
  __syncthreads();
  if (threadIdx.x == 0 && !skip) {
    sv[0].x += ph.x * alpha;
    sv[0].y += ph.y * alpha;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    cd* Vj = V + (size_t)j * n + j + 1;
    for (int i = threadIdx.x; i < m; i += kThreads) Vj[i] = skip ? cd{0.0, 0.0} : sv[i];
    if (threadIdx.x == 0) {
      tau[j] = t;
      sub[j] = skip ? x0 : cd{-ph.x * alpha, -ph.y * alpha};
      // u(0, 0) = 2 Re(v'_0 conj(w'_0))
      diag[j] = A[2 * ((size_t)j * n + j)] - 2.0 * (v0.x * w0.x + v0.y * w0.y);
    }
  }
  // the block's rows: update, write back, multiply by v (one wave per row)
  double pr = 0.0, pi = 0.0;  // partial conj(v) . p of this wave
  for (int i = r0 + wave; i < r1; i += 4) {
    cd* Ai = reinterpret_cast<cd*>(A) + (size_t)(j + 1 + i) * n + j + 1;
    const cd vi = upd ? svp[i + 1] : cd{0.0, 0.0}, wi = upd ? wprev(i + 1) : cd{0.0, 0.0};
    double sr = 0.0, si = 0.0;
    auto entry = [&](int l, cd a) {
      if (upd) {
        const cd vl = svp[l + 1], wl = wprev(l + 1);
        a.x -= (vi.x * wl.x + vi.y * wl.y) + (wi.x * vl.x + wi.y * vl.y);
        a.y -= (vi.y * wl.x - vi.x * wl.y) + (wi.y * vl.x - wi.x * vl.y);
        Ai[l] = a;
      }
      const cd v = sv[l];
      sr += a.x * v.x - a.y * v.y;
      si += a.x * v.y + a.y * v.x;
    };
    int lstart = lane;
    if (i == r0 + wave) {  // the prefetched row
#pragma unroll
      for (int c = 0; c < kPre; ++c) {
        const int l = lane + 64 * c;
        if (l < m) entry(l, pre[c]);
      }
      lstart = lane + 64 * kPre;
    }
    for (int l = lstart; l < m; l += 64) entry(l, Ai[l]);
    sr = wave_sum(sr) * t;
    si = wave_sum(si) * t;
    if (lane == 0) {
      p[i] = {sr, si};
      const cd v = sv[i];
      pr += v.x * sr + v.y * si;
      pi += v.x * si - v.y * sr;
    }
  }
  block_sum2(pr, pi, red);
  if (threadIdx.x == 0) part[blockIdx.x] = {pr, pi};
}

__global__ __launch_bounds__(kThreads) void k_hh_her2(double* __restrict__ A, int n, int j,
                                                      const cd* __restrict__ V,
                                                      const double* __restrict__ tau,
                                                      const cd* __restrict__ p,
                                                      const cd* __restrict__ part, int nparts,
                                                      EigStride st) {
  __shared__ cd sv_i[64], sw_i[64], sv_l[64], sw_l[64];
  __shared__ double red[8];
  const int m = n - j - 1;
  {
    const size_t mi = blockIdx.z;
    A += mi * st.a;
    V += mi * st.v;
    tau += mi * st.n;
    p += mi * st.p;
    part += mi * st.part;
  }
  const double t = tau[j];
  if (t == 0.0) return;  // no reflection this step (uniform over the grid)
  double kr = 0.0, ki = 0.0;
  for (int q = threadIdx.x; q < nparts; q += kThreads) {
    kr += part[q].x;
    ki += part[q].y;
  }
  block_sum2(kr, ki, red);
  const cd K = {0.5 * t * kr, 0.5 * t * ki};
  const cd* v = V + (size_t)j * n + j + 1;
  const int i0 = blockIdx.y * 64, l0 = blockIdx.x * 64;
  if (threadIdx.x < 64) {
    const int i = i0 + threadIdx.x;
    cd vi = {0.0, 0.0}, wi = {0.0, 0.0};
    if (i < m) {
      vi = v[i];
      const cd kv = cmul(K, vi);
      wi = {p[i].x - kv.x, p[i].y - kv.y};
    }
    sv_i[threadIdx.x] = vi;
    sw_i[threadIdx.x] = wi;
  } else if (threadIdx.x < 128) {
    const int l = l0 + threadIdx.x - 64;
    cd vl = {0.0, 0.0}, wl = {0.0, 0.0};
    if (l < m) {
      vl = v[l];
      const cd kv = cmul(K, vl);
      wl = {p[l].x - kv.x, p[l].y - kv.y};
    }
    sv_l[threadIdx.x - 64] = vl;
    sw_l[threadIdx.x - 64] = wl;
  }
  __syncthreads();
  const int tl = threadIdx.x & 63, tr = threadIdx.x >> 6;
  const int l = l0 + tl;
  if (l >= m) return;
  const cd vl = sv_l[tl], wl = sw_l[tl];
  for (int r = tr; r < 64; r += 4) {
    const int i = i0 + r;
    if (i >= m) break;
    const cd vi = sv_i[r], wi = sw_i[r];
    cd* a = reinterpret_cast<cd*>(A) + (size_t)(j + 1 + i) * n + j + 1 + l;
    // v_i conj(w_l) + w_i conj(v_l)
    const double ur = (vi.x * wl.x + vi.y * wl.y) + (wi.x * vl.x + wi.y * vl.y);
    const double ui = (vi.y * wl.x - vi.x * wl.y) + (wi.y * vl.x - wi.x * vl.y);
    cd av = *a;
    av.x -= ur;
    av.y -= ui;
    *a = av;
  }
}

// The last steps of the reduction in one block per matrix, from step j0 on,
// once the trailing block A[j0.., j0..] has s = n - j0 <= kTail rows: the
// launches of those steps would each cost the fixed latency of a device-wide
// step for a few rows of work.  The trailing block, padded with zeros to
// 64 x 64, is held in registers of kTW waves: wave q, lane c keeps row
// kTW (jb + i) + q, column c, in W[i] during group jb of kTW steps (the window
// moves down one row per wave after each group, so that register indices
// stay static).  The
// block applies step j0-1's pending rank-2 update (as k_hh_step would), then
// runs steps j0 .. n-3 with the same reflector conventions (V rows, tau, sub,
// diag) and records the last 2 x 2 block (as k_hh_finish).  Step t (row t of
// the block): the wave holding row t forms the reflector of x = conj(row t)
// right of the diagonal and publishes v and tau through LDS; every wave sums
// its rows of p = tau S22 v as column sums (S22 is Hermitian, so
// p_c = tau sum_r conj(S[r][c]) v_r), the waves' sums are added in order
// through LDS, each wave forms w = p - (tau / 2)(v^H p) v in its lanes
// and updates its rows, S[r][c] -= v_r conj(w_c) + w_r conj(v_c), taking v_r
// and w_r from lane r.  Outputs are staged in LDS and written at the end.
constexpr int kTail = 64;
constexpr int kTW = 4;           // waves (measured: 16 is slower at n = 64)
constexpr int kTR = 64 / kTW;    // rows a wave

__device__ __forceinline__ double lane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ cd lane_cd(cd v, int l) { return {lane_d(v.x, l), lane_d(v.y, l)}; }

__global__ __launch_bounds__(kTW * 64) void k_hh_tail(
    const double* __restrict__ A, int n, int j0, cd* __restrict__ V,
    double* __restrict__ tau, cd* __restrict__ sub, double* __restrict__ diag,
    const cd* __restrict__ p_prev, const cd* __restrict__ part_prev, int nparts_prev,
    EigStride st) {
  __shared__ cd vb[2][64];     // v of the step (double-buffered)
  __shared__ double tb[2];     // tau of the step
  __shared__ cd pb[2][kTW][64];  // the waves' column sums of S22 v
  __shared__ cd Vs[64][64];    // row t: v of step t
  __shared__ cd so[64];        // sub
  __shared__ double to[64], dg[64];
  __shared__ double red[2 * kTW];
  const int tid = threadIdx.x, c = tid & 63;
  const int q = __builtin_amdgcn_readfirstlane(tid >> 6);  // the wave (uniform)
  {
    const size_t mi = blockIdx.y;
    A += mi * st.a;
    V += mi * st.v;
    tau += mi * st.n;
    sub += mi * st.n;
    diag += mi * st.n;
    p_prev += mi * st.p;
    part_prev += mi * st.part;
  }
  const int s = n - j0;
  const cd* Ag = reinterpret_cast<const cd*>(A);
  cd W[kTR];
#pragma unroll
  for (int i = 0; i < kTR; ++i) {
    const int r = q + kTW * i;
    W[i] = (r < s && c < s) ? Ag[(size_t)(j0 + r) * n + j0 + c] : cd{0.0, 0.0};
  }
  // step j0-1's update S -= v' w'^H + w' v'^H, w' = p' - K' v'
  const double tp = j0 > 0 ? tau[j0 - 1] : 0.0;
  if (tp != 0.0) {  // (uniform)
    double kr = 0.0, ki = 0.0;
    for (int u = tid; u < nparts_prev; u += kTW * 64) {
      kr += part_prev[u].x;
      ki += part_prev[u].y;
    }
    kr = wave_sum(kr);
    ki = wave_sum(ki);
    if (c == 0) {
      red[2 * q] = kr;
      red[2 * q + 1] = ki;
    }
    __syncthreads();
    kr = 0.0;
    ki = 0.0;
    for (int w = 0; w < kTW; ++w) {
      kr += red[2 * w];
      ki += red[2 * w + 1];
    }
    const cd K = {0.5 * tp * kr, 0.5 * tp * ki};
    const cd vc = c < s ? V[(size_t)(j0 - 1) * n + j0 + c] : cd{0.0, 0.0};
    const cd pc = c < s ? p_prev[c] : cd{0.0, 0.0}, kv = cmul(K, vc);
    const cd wc = {pc.x - kv.x, pc.y - kv.y};
#pragma unroll
    for (int i = 0; i < kTR; ++i) {
      const int r = q + kTW * i;
      if (r >= s) continue;  // (uniform)
      const cd vr = lane_cd(vc, r), wr = lane_cd(wc, r);
      W[i].x -= (vr.x * wc.x + vr.y * wc.y) + (wr.x * vc.x + wr.y * vc.y);
      W[i].y -= (vr.y * wc.x - vr.x * wc.y) + (wr.y * vc.x - wr.x * vc.y);
    }
  }
  for (int jb = 0; jb < kTR; ++jb) {
#pragma unroll
    for (int jq = 0; jq < kTW; ++jq) {
      const int t = kTW * jb + jq, par = jq & 1;
      if (t + 2 >= s) break;  // (uniform)
      if (q == jq) {  // row t is this wave's W[0]
        const cd g = W[0];
        const cd x = c > t ? cd{g.x, -g.y} : cd{0.0, 0.0};
        const double s2 = wave_sum(x.x * x.x + x.y * x.y);
        const cd x0 = lane_cd(x, t + 1);
        const double ax0 = hypot(x0.x, x0.y);
        const double alpha = sqrt(s2);
        const bool skip = (s2 - ax0 * ax0) <= 0.0;  // x already a multiple of e_1
        cd ph = {1.0, 0.0};
        if (ax0 > 0.0) ph = {x0.x / ax0, x0.y / ax0};
        const double tt = skip ? 0.0 : 1.0 / (alpha * (alpha + ax0));  // 2 / (v^H v)
        cd v = x;
        if (c == t + 1 && !skip) {
          v.x += ph.x * alpha;
          v.y += ph.y * alpha;
        }
        vb[par][c] = v;
        Vs[t][c] = skip ? cd{0.0, 0.0} : v;
        if (c == 0) {
          tb[par] = tt;
          to[t] = tt;
          so[t] = skip ? x0 : cd{-ph.x * alpha, -ph.y * alpha};
        }
      }
      __syncthreads();
      const double tt = tb[par];
      if (tt == 0.0) continue;  // (uniform) the trailing block stays as it is
      const cd vc = vb[par][c];
      double ar = 0.0, ai = 0.0;  // this wave's rows of sum_r conj(S[r][c]) v_r
#pragma unroll
      for (int i = 0; i < kTR; ++i) {
        const int r = kTW * (jb + i) + q;
        if (r <= t || r >= s) continue;  // (uniform)
        const cd vr = lane_cd(vc, r);
        ar += W[i].x * vr.x + W[i].y * vr.y;
        ai += W[i].x * vr.y - W[i].y * vr.x;
      }
      pb[par][q][c] = {ar, ai};
      __syncthreads();
      cd pc = {0.0, 0.0};
      if (c > t) {
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int w = 0; w < kTW; ++w) {
          sr += pb[par][w][c].x;
          si += pb[par][w][c].y;
        }
        pc = {tt * sr, tt * si};
      }
      const double kr = wave_sum(vc.x * pc.x + vc.y * pc.y);  // conj(v) . p
      const double ki = wave_sum(vc.x * pc.y - vc.y * pc.x);
      const cd K = {0.5 * tt * kr, 0.5 * tt * ki}, kv = cmul(K, vc);
      const cd wc = {pc.x - kv.x, pc.y - kv.y};
#pragma unroll
      for (int i = 0; i < kTR; ++i) {
        const int r = kTW * (jb + i) + q;
        if (r <= t || r >= s) continue;  // (uniform)
        const cd vr = lane_cd(vc, r), wr = lane_cd(wc, r);
        W[i].x -= (vr.x * wc.x + vr.y * wc.y) + (wr.x * vc.x + wr.y * vc.y);
        W[i].y -= (vr.y * wc.x - vr.x * wc.y) + (wr.y * vc.x - wr.x * vc.y);
      }
    }
    // row kTW jb + q is final: its diagonal entry (and the last 2 x 2 block's
    // off-diagonal one, A[n-1][n-2] = conj(A[n-2][n-1])); then shift the window
    const int r = kTW * jb + q;
    if (r < s) {
      const double d = lane_d(W[0].x, r);
      cd e = {0.0, 0.0};
      if (r == s - 2) e = lane_cd(W[0], s - 1);
      if (c == 0) {
        dg[r] = d;
        if (r == s - 2) so[r] = {e.x, -e.y};
      }
    }
#pragma unroll
    for (int i = 0; i + 1 < kTR; ++i) W[i] = W[i + 1];
  }
  __syncthreads();
  for (int o = tid; o < 64 * 64; o += kTW * 64) {
    const int t = o >> 6, cc = o & 63;
    if (t + 2 < s && cc > t && cc < s) V[(size_t)(j0 + t) * n + j0 + cc] = Vs[t][cc];
  }
  for (int t = tid; t < s; t += kTW * 64) {
    if (t + 2 < s) tau[j0 + t] = to[t];
    if (t + 1 < s) sub[j0 + t] = so[t];
    diag[j0 + t] = dg[t];
  }
}

// The last 2 x 2 block after the reduction.
__global__ void k_hh_finish(const double* __restrict__ A, int n, cd* __restrict__ sub,
                            double* __restrict__ diag, EigStride st) {
  if (threadIdx.x != 0) return;
  A += blockIdx.x * st.a;
  sub += blockIdx.x * st.n;
  diag += blockIdx.x * st.n;
  if (n >= 2) {
    const cd a = reinterpret_cast<const cd*>(A)[(size_t)(n - 2) * n + n - 1];
    sub[n - 2] = {a.x, -a.y};  // A[n-1][n-2] = conj(A[n-2][n-1])
    diag[n - 2] = A[2 * ((size_t)(n - 2) * n + n - 2)];
  }
  diag[n - 1] = A[2 * ((size_t)(n - 1) * n + n - 1)];
}

// Number of eigenvalues of T_r below x from a twisted factorization
// T_r - x I = N Delta N^T with twist index r (Sylvester's law of inertia:
// the count is the number of negative entries of Delta).  Delta holds the
// top-down pivots D+_i, i < r, the bottom-up pivots D-_i, i > r, and
//   gamma_r = d_r - x - e2_{r-1} / D+_{r-1} - e2_r / D-_{r+1},
// so the two halves run as independent chains of n / 2 steps (one wave
// each) instead of one chain of n.  Pivots smaller than pivmin count as
// -pivmin.  d and e2 are in LDS; the loads of 8 steps are issued ahead of
// their recurrence steps.
// 1 / q from the hardware reciprocal estimate and two Newton steps
// (r <- r + r (1 - q r)), each doubling its correct bits: within a few ulps,
// in about half the dependent operations of a correctly rounded division.
// The count stays that of a matrix within a few ulps of T_r - x I.
__device__ __forceinline__ double recip(double q) {
  double r = __builtin_amdgcn_rcp(q);
  double e = fma(-q, r, 1.0);
  r = fma(r, e, r);
  e = fma(-q, r, 1.0);
  return fma(r, e, r);
}

__device__ __forceinline__ void pivot_step(double& q, int& c, double di, double ei, double x,
                                           double pivmin) {
  q = di - x - ei * recip(q);
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0) ++c;
}

// D+_0 .. D+_{r-1}: their negative count; qlast = D+_{r-1}.
__device__ int sturm_top(const double* d, const double* e2, int r, double x, double pivmin,
                         double& qlast) {
  int c = 0;
  double q = 1.0;
  if (r > 0) {
    q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0) ++c;
  }
  int i = 1;
  for (; i + 8 <= r; i += 8) {
    double dv[8], ev[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dv[u] = d[i + u];
      ev[u] = e2[i - 1 + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) pivot_step(q, c, dv[u], ev[u], x, pivmin);
  }
  for (; i < r; ++i) pivot_step(q, c, d[i], e2[i - 1], x, pivmin);
  qlast = q;
  return c;
}

// D-_{n-1} .. D-_{r+1} (D-_i = d_i - x - e2_i / D-_{i+1}): their negative
// count; qlast = D-_{r+1}.
__device__ int sturm_bottom(const double* d, const double* e2, int n, int r, double x,
                            double pivmin, double& qlast) {
  int c = 0;
  double q = 1.0;
  if (r + 1 < n) {
    q = d[n - 1] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0) ++c;
  }
  int i = n - 2;
  for (; i - 7 >= r + 1; i -= 8) {
    double dv[8], ev[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dv[u] = d[i - u];
      ev[u] = e2[i - u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) pivot_step(q, c, dv[u], ev[u], x, pivmin);
  }
  for (; i >= r + 1; --i) pivot_step(q, c, d[i], e2[i], x, pivmin);
  qlast = q;
  return c;
}

// The k largest eigenvalues of T_r: block idx finds eigenvalue idx
// (descending) by multisection of [lo, hi] at 64 points a pass.  Wave 0 runs
// the top halves of the 64 twisted factorizations, wave 1 the bottom halves;
// wave 0 combines them, picks the subinterval and shares it through LDS, so
// that both waves leave the loop together.  Matrix mi (grid y) takes its
// Gershgorin interval and pivot threshold from bounds[3 mi ..].
__global__ __launch_bounds__(128) void k_bisect(const double* __restrict__ d,
                                                const double* __restrict__ e2, int n, int k,
                                                const double* __restrict__ bounds,
                                                double* __restrict__ w, EigStride st) {
  extern __shared__ double sde[];  // d, then e2 (n entries each)
  __shared__ double sq[2][64];
  __shared__ int sc[2][64];
  __shared__ double s_lohi[2];
  __shared__ int s_go;
  const int idx = blockIdx.x, tid = threadIdx.x, lane = tid & 63, half = tid >> 6;
  if (idx >= k) return;
  const size_t mi = blockIdx.y;
  d += mi * st.n;
  e2 += mi * st.n;
  w += mi * st.k;
  const double lo0 = bounds[3 * mi], hi0 = bounds[3 * mi + 1], pivmin = bounds[3 * mi + 2];
  for (int i = tid; i < n; i += 128) {
    sde[i] = d[i];
    sde[n + i] = e2[i];
  }
  __syncthreads();
  const double* ld = sde;
  const double* le2 = sde + n;
  const int target = n - 1 - idx;  // ascending index
  const int r = n / 2;
  double lo = lo0, hi = hi0;
  for (int it = 0; it < 64; ++it) {
    if (hi - lo <= 2.2e-16 * fmax(fabs(lo), fabs(hi)) + pivmin) break;  // (uniform)
    const double h = hi - lo;
    const double x = lo + h * (double)(lane + 1) / 65.0;
    double qe;
    const int c = half == 0 ? sturm_top(ld, le2, r, x, pivmin, qe)
                            : sturm_bottom(ld, le2, n, r, x, pivmin, qe);
    sq[half][lane] = qe;
    sc[half][lane] = c;
    __syncthreads();
    if (half == 0) {
      double g = ld[r] - x;
      if (r > 0) g -= le2[r - 1] / sq[0][lane];
      if (r + 1 < n) g -= le2[r] / sq[1][lane];
      if (fabs(g) < pivmin) g = -pivmin;
      const int cnt = sc[0][lane] + sc[1][lane] + (g < 0 ? 1 : 0);
      const unsigned long long mask = __ballot(cnt > target);
      const int f = mask ? __ffsll(mask) - 1 : 64;  // first point past the eigenvalue
      const double nlo = f == 0 ? lo : lo + h * (double)f / 65.0;
      const double nhi = f == 64 ? hi : lo + h * (double)(f + 1) / 65.0;
      const bool go = nlo > lo || nhi < hi;  // progress left in double precision
      if (lane == 0) {
        s_go = go;
        s_lohi[0] = go ? nlo : lo;
        s_lohi[1] = go ? nhi : hi;
      }
    }
    __syncthreads();
    lo = s_lohi[0];
    hi = s_lohi[1];
    if (!s_go) break;
  }
  if (tid == 0) w[idx] = 0.5 * (lo + hi);
}

// LU of T_r - lam I with partial pivoting between rows i and i+1, into the
// 5n doubles at lu: reciprocals of U's diagonal (so that the back substitution
// multiplies: a division on its serial chain costs about ten dependent FP64
// operations; the LU forms them with recip, which its multipliers share), U's first and second superdiagonals, multipliers, and the swap
// flags.  The sweeps below are serial in i; each runs in chunks
// of kSweep steps whose loads are issued together ahead of the chunk's
// recurrence, so that the memory latency is paid once per chunk instead of
// once per step.
constexpr int kSweep = 16;
constexpr int kGs = 8;  // earlier vectors per Gram-Schmidt round (registers)
constexpr int kInvIters = 3;  // inverse iteration rounds
constexpr int kPipe = 8;      // steps a chunk of the pipelined solves

__device__ void tri_lu(const double* __restrict__ d, const double* __restrict__ e, int n,
                       double lam, double tnorm, double* __restrict__ lu) {
  const double eps = 2.220446049250313e-16;
  double* lu_rd = lu;
  double* lu_u1 = lu + n;
  double* lu_u2 = lu + 2 * n;
  double* lu_l = lu + 3 * n;
  double* piv = lu + 4 * n;
  double a = d[0] - lam, b = n > 1 ? e[0] : 0.0;
  // Pivots smaller than eps ||T|| are replaced by +-eps ||T|| (LAPACK's
  // dlagtf / dstein perturb them the same way); a zero tridiagonal (||T|| = 0)
  // takes unit pivots, so that the solves leave the start vector unchanged.
  // The floor also keeps subnormal pivots away from the reciprocal estimate,
  // which is not accurate there.
  const double pfloor = tnorm > 0.0 ? eps * tnorm : 1.0;
  // one elimination step: row i (a, b, 0) against row i + 1 (cdn, dn, en)
  auto step = [&](double cdn, double dn, double en, double& od, double& ou1, double& ou2,
                  double& ol, double& op) {
    if (fabs(a) >= fabs(cdn) || fabs(cdn) < pfloor) {
      if (fabs(a) < pfloor) a = copysign(pfloor, a);
      const double r = recip(a);
      const double l = cdn * r;
      od = r;
      ou1 = b;
      ou2 = 0.0;
      ol = l;
      op = 0.0;
      a = dn - l * b;
      b = en;
    } else {
      const double r = recip(cdn);
      const double l = a * r;
      od = r;
      ou1 = dn;
      ou2 = en;
      ol = l;
      op = 1.0;
      a = b - l * dn;
      b = -l * en;
    }
  };
  int i = 0;
  for (; i + kSweep < n - 1; i += kSweep) {  // steps i .. i + kSweep - 1, all with i + 1 < n - 1
    double cv[kSweep], dv[kSweep], ev[kSweep];
#pragma unroll
    for (int u = 0; u < kSweep; ++u) {
      cv[u] = e[i + u];
      dv[u] = d[i + 1 + u] - lam;
      ev[u] = e[i + 1 + u];
    }
    double od[kSweep], ou1[kSweep], ou2[kSweep], ol[kSweep], op[kSweep];
#pragma unroll
    for (int u = 0; u < kSweep; ++u) step(cv[u], dv[u], ev[u], od[u], ou1[u], ou2[u], ol[u], op[u]);
#pragma unroll
    for (int u = 0; u < kSweep; ++u) {
      lu_rd[i + u] = od[u];
      lu_u1[i + u] = ou1[u];
      lu_u2[i + u] = ou2[u];
      lu_l[i + u] = ol[u];
      piv[i + u] = op[u];
    }
  }
  for (; i < n - 1; ++i) {
    const double cdn = e[i], dn = d[i + 1] - lam, en = i + 1 < n - 1 ? e[i + 1] : 0.0;
    step(cdn, dn, en, lu_rd[i], lu_u1[i], lu_u2[i], lu_l[i], piv[i]);
  }
  if (fabs(a) < pfloor) a = copysign(pfloor, a);
  lu_rd[n - 1] = recip(a);
}

// x <- (T_r - lam I)^{-1} x with the factors of tri_lu.
__device__ void tri_solve(const double* __restrict__ lu, int n, double* __restrict__ x) {
  const double* lu_rd = lu;
  const double* lu_u1 = lu + n;
  const double* lu_u2 = lu + 2 * n;
  const double* lu_l = lu + 3 * n;
  const double* piv = lu + 4 * n;
  // Both sweeps run in chunks of kPipe steps, the loads of chunk c + 1 issued
  // before chunk c's recurrence (two register sets, used alternately); a
  // chunk's loads never touch the entries that the chunk before it writes.
  // Past the last chunk the loads repeat it (in bounds, unused).
  // L (with the row swaps): forward
  double xi = x[0];
  const int nf = (n - 1) / kPipe;  // full chunks: steps c kPipe .. + kPipe - 1
  auto load_f = [&](int c, double(&pv)[kPipe], double(&lv)[kPipe], double(&xn)[kPipe]) {
    const int b = min(c, nf - 1) * kPipe;
#pragma unroll
    for (int u = 0; u < kPipe; ++u) {
      pv[u] = piv[b + u];
      lv[u] = lu_l[b + u];
      xn[u] = x[b + 1 + u];
    }
  };
  auto sweep_f = [&](int c, const double(&pv)[kPipe], const double(&lv)[kPipe],
                     const double(&xn)[kPipe]) {
    double xo[kPipe];
#pragma unroll
    for (int u = 0; u < kPipe; ++u) {
      if (pv[u] != 0.0) {
        xo[u] = xn[u];
        xi = xi - lv[u] * xn[u];
      } else {
        xo[u] = xi;
        xi = xn[u] - lv[u] * xi;
      }
    }
#pragma unroll
    for (int u = 0; u < kPipe; ++u) x[c * kPipe + u] = xo[u];
  };
  {
    double pa[kPipe], la[kPipe], xa[kPipe], pb[kPipe], lb[kPipe], xb[kPipe];
    if (nf > 0) load_f(0, pa, la, xa);
    int c = 0;
    for (; c + 1 < nf; c += 2) {
      load_f(c + 1, pb, lb, xb);
      sweep_f(c, pa, la, xa);
      load_f(c + 2, pa, la, xa);
      sweep_f(c + 1, pb, lb, xb);
    }
    if (c < nf) sweep_f(c, pa, la, xa);
  }
  for (int i = nf * kPipe; i < n - 1; ++i) {
    const double xn = x[i + 1];
    if (piv[i] != 0.0) {
      x[i] = xn;
      xi = xi - lu_l[i] * xn;
    } else {
      x[i] = xi;
      xi = xn - lu_l[i] * xi;
    }
  }
  x[n - 1] = xi;
  // U: backward
  double x1 = x[n - 1] * lu_rd[n - 1];
  x[n - 1] = x1;
  if (n < 2) return;
  double x0 = (x[n - 2] - lu_u1[n - 2] * x1) * lu_rd[n - 2];
  x[n - 2] = x0;
  const int nb = (n - 2) / kPipe;  // full chunks: steps n-3 - c kPipe - u, u < kPipe
  auto load_b = [&](int c, double(&xv)[kPipe], double(&u1v)[kPipe], double(&u2v)[kPipe],
                    double(&dv)[kPipe]) {
    const int b = n - 3 - min(c, nb - 1) * kPipe;
#pragma unroll
    for (int u = 0; u < kPipe; ++u) {
      xv[u] = x[b - u];
      u1v[u] = lu_u1[b - u];
      u2v[u] = lu_u2[b - u];
      dv[u] = lu_rd[b - u];
    }
  };
  auto sweep_b = [&](int c, double(&xv)[kPipe], const double(&u1v)[kPipe],
                     const double(&u2v)[kPipe], const double(&dv)[kPipe]) {
#pragma unroll
    for (int u = 0; u < kPipe; ++u) {
      const double v = (xv[u] - u1v[u] * x0 - u2v[u] * x1) * dv[u];
      xv[u] = v;
      x1 = x0;
      x0 = v;
    }
    const int b = n - 3 - c * kPipe;
#pragma unroll
    for (int u = 0; u < kPipe; ++u) x[b - u] = xv[u];
  };
  {
    double xa[kPipe], u1a[kPipe], u2a[kPipe], da[kPipe];
    double xb[kPipe], u1b[kPipe], u2b[kPipe], db[kPipe];
    if (nb > 0) load_b(0, xa, u1a, u2a, da);
    int c = 0;
    for (; c + 1 < nb; c += 2) {
      load_b(c + 1, xb, u1b, u2b, db);
      sweep_b(c, xa, u1a, u2a, da);
      load_b(c + 2, xa, u1a, u2a, da);
      sweep_b(c + 1, xb, u1b, u2b, db);
    }
    if (c < nb) sweep_b(c, xa, u1a, u2a, da);
  }
  for (int i = n - 3 - nb * kPipe; i >= 0; --i) {
    const double v = (x[i] - lu_u1[i] * x0 - lu_u2[i] * x1) * lu_rd[i];
    x[i] = v;
    x1 = x0;
    x0 = v;
  }
}

// Re-orthogonalises vector q of a cluster against the cluster's earlier
// vectors (rows first .. q-1 of z; Gram-Schmidt `passes` times, kGs rows a
// round) and normalises it, for n <= 256 C: the block keeps x and each
// round's kGs rows in registers (C entries a thread), so that x is read and
// written once and every round's loads are issued together (clamped,
// unconditional loads: a load in a branch would be waited for there).
template <int C>
__device__ void orthonormalise_reg(double* __restrict__ x, const double* __restrict__ z, int n,
                                   int first, int q, int passes, double (*red)[33],
                                   double* dots) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double xr[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int i = tid + 256 * c;
    const double v = x[min(i, n - 1)];
    xr[c] = i < n ? v : 0.0;
  }
  for (int pass = 0; pass < passes; ++pass)
    for (int r0 = first; r0 < q; r0 += kGs) {
      const int nr = min(kGs, q - r0);
      double zr[kGs][C];
#pragma unroll
      for (int j = 0; j < kGs; ++j)
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const int i = tid + 256 * c;
          const double v = z[(size_t)(r0 + min(j, nr - 1)) * n + min(i, n - 1)];
          zr[j][c] = (j < nr && i < n) ? v : 0.0;
        }
      double acc[kGs];
#pragma unroll
      for (int j = 0; j < kGs; ++j) {
        acc[j] = 0.0;
#pragma unroll
        for (int c = 0; c < C; ++c) acc[j] += zr[j][c] * xr[c];
      }
#pragma unroll
      for (int j = 0; j < kGs; ++j) {
        const double v = wave_sum(acc[j]);
        if (lane == 0) red[wv][j] = v;
      }
      __syncthreads();
      if (tid < nr) dots[tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kGs; ++j) {
        const double dj = dots[j];
#pragma unroll
        for (int c = 0; c < C; ++c) xr[c] -= dj * zr[j][c];  // (zr is 0 for j >= nr)
      }
      __syncthreads();
    }
  double s2 = 0.0;
#pragma unroll
  for (int c = 0; c < C; ++c) s2 += xr[c] * xr[c];
  s2 = wave_sum(s2);
  if (lane == 0) red[wv][32] = s2;
  __syncthreads();
  const double inv = 1.0 / sqrt(red[0][32] + red[1][32] + red[2][32] + red[3][32]);
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int i = tid + 256 * c;
    if (i < n) x[i] = xr[c] * inv;
  }
  __syncthreads();
}

// Eigenvectors of T_r by inverse iteration, one block per cluster of close
// eigenvalues (cl_begin[c] .. cl_begin[c+1]).  Thread t of the block factors
// T_r - lam_q I and solves with it for vector q = first + t; after every
// round of solves the block re-orthogonalises the cluster's vectors in
// eigenvalue order (Gram-Schmidt against kGs earlier vectors at a time,
// parallel over the vector length; applied twice in the last round) and
// normalises them.
// Equal eigenvalues of a cluster are separated slightly in the shifts.
// Three rounds: a shift within 10 eps ||T|| of lambda_q amplifies the wanted
// component over any eigenvalue at least 1e-4 ||T|| away (every eigenvalue
// outside the cluster) by more than 1e10 a round, so the random start
// converges in two and the third leaves the error at rounding level; within a
// cluster the re-orthogonalisation keeps the vectors apart.
// Matrix mi has ncl[mi] clusters starting at cl_begin[mi (k + 1) ..] and the
// tridiagonal norm tnorms[mi]; blocks past its cluster count return.
__global__ __launch_bounds__(256) void k_inverse_iteration(
    const double* __restrict__ d, const double* __restrict__ e, int n,
    const double* __restrict__ w, const int* __restrict__ cl_all,
    const int* __restrict__ ncl, const double* __restrict__ tnorms,
    double* __restrict__ z, double* __restrict__ work, EigStride st) {
  __shared__ double red[4][33];
  __shared__ double dots[32];
  __shared__ double zs[64 * 64];  // n <= 64: the cluster's vectors
  const int c = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const size_t mi = blockIdx.y;
  if (c >= ncl[mi]) return;
  d += mi * st.n;
  e += mi * st.n;
  w += mi * st.k;
  z += mi * st.z;
  work += mi * 5 * st.z;
  const int* cl_begin = cl_all + mi * (st.k + 1);
  const double tnorm = tnorms[mi];
  const int first = cl_begin[c], last = cl_begin[c + 1];
  const double eps = 2.220446049250313e-16;
  for (int q = first + tid; q < last; q += 256) {
    const double lam = w[q] + (double)(q - first) * 10.0 * eps * tnorm;
    tri_lu(d, e, n, lam, tnorm, work + (size_t)q * 5 * n);
    // deterministic start vector
    double* x = z + (size_t)q * n;
    unsigned s = 2463534242u ^ (unsigned)(q * 2654435761u);
    for (int i = 0; i < n; ++i) {
      s ^= s << 13;
      s ^= s >> 17;
      s ^= s << 5;
      x[i] = ((double)(s & 0xffffff) / 16777216.0) - 0.5;
    }
  }
  for (int it = 0; it < kInvIters; ++it) {
    for (int q = first + tid; q < last; q += 256)
      tri_solve(work + (size_t)q * 5 * n, n, z + (size_t)q * n);
    __syncthreads();
    // Gram-Schmidt once in the early rounds (the next solve amplifies the
    // wanted directions again), twice in the last
    const int passes = it + 1 == kInvIters ? 2 : 1;
    if (n <= 64) {
      // a vector fits one wave: wave 0 runs the whole cluster from LDS, lane
      // i holding entry i, with no block barrier between vectors (the same
      // kGs-row rounds: classical within a round, rounds in order)
      const int m = last - first;
      for (int u = tid; u < m * n; u += 256) zs[u] = z[(size_t)first * n + u];
      __syncthreads();
      if (wv == 0) {
        for (int q = 0; q < m; ++q) {
          double x = lane < n ? zs[q * n + lane] : 0.0;
          for (int pass = 0; pass < passes; ++pass)
            for (int r0 = 0; r0 < q; r0 += kGs) {
              double zr[kGs], dd[kGs];
#pragma unroll
              for (int j = 0; j < kGs; ++j) {
                const bool ok = r0 + j < q && lane < n;
                const double v = zs[min(r0 + j, q - 1) * n + min(lane, n - 1)];
                zr[j] = ok ? v : 0.0;
              }
#pragma unroll
              for (int j = 0; j < kGs; ++j) dd[j] = wave_sum(zr[j] * x);
#pragma unroll
              for (int j = 0; j < kGs; ++j) x -= dd[j] * zr[j];
            }
          const double inv = 1.0 / sqrt(wave_sum(x * x));
          if (lane < n) zs[q * n + lane] = x * inv;
        }
      }
      __syncthreads();
      for (int u = tid; u < m * n; u += 256) z[(size_t)first * n + u] = zs[u];
      __syncthreads();
      continue;
    }
    for (int q = first; q < last; ++q) {
      double* x = z + (size_t)q * n;
      if (n <= 256 * 5) {
        orthonormalise_reg<5>(x, z, n, first, q, passes, red, dots);
        continue;
      }
      if (n <= 256 * 8) {
        orthonormalise_reg<8>(x, z, n, first, q, passes, red, dots);
        continue;
      }
      for (int pass = 0; pass < passes; ++pass)
        for (int r0 = first; r0 < q; r0 += kGs) {
          const int nr = min(kGs, q - r0);
          double acc[kGs];
#pragma unroll
          for (int j = 0; j < kGs; ++j) acc[j] = 0.0;
          for (int i = tid; i < n; i += 256) {
            const double xi = x[i];
#pragma unroll
            for (int j = 0; j < kGs; ++j)
              if (j < nr) acc[j] += z[(size_t)(r0 + j) * n + i] * xi;
          }
#pragma unroll
          for (int j = 0; j < kGs; ++j) {
            const double v = wave_sum(acc[j]);
            if (lane == 0) red[wv][j] = v;
          }
          __syncthreads();
          if (tid < nr) dots[tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
          __syncthreads();
          for (int i = tid; i < n; i += 256) {
            double v = x[i];
#pragma unroll
            for (int j = 0; j < kGs; ++j)
              if (j < nr) v -= dots[j] * z[(size_t)(r0 + j) * n + i];
            x[i] = v;
          }
          __syncthreads();
        }
      double s2 = 0.0;
      for (int i = tid; i < n; i += 256) s2 += x[i] * x[i];
      s2 = wave_sum(s2);
      if (lane == 0) red[wv][32] = s2;
      __syncthreads();
      const double inv = 1.0 / sqrt(red[0][32] + red[1][32] + red[2][32] + red[3][32]);
      for (int i = tid; i < n; i += 256) x[i] *= inv;
      __syncthreads();
    }
  }
}

// sy[i] = D_i z[i], D_i = prod_{r < i} sub[r] / |sub[r]| (unit phases; 1 for a
// zero sub[r]), for a block of kThreads threads: each thread forms the
// running products over its own run of indices, a block scan combines the
// runs' totals.  scan: kThreads complex values of LDS.
__device__ void phase_scale(const double* __restrict__ z, int n, const cd* __restrict__ sub,
                            cd* __restrict__ sy, cd* __restrict__ scan) {
  const int tid = threadIdx.x;
  const int run = (n + kThreads - 1) / kThreads, i0 = tid * run, i1 = min(n, i0 + run);
  auto phase = [&](int r) {  // factor between D_r and D_{r+1}
    const cd s = sub[r];
    const double a = hypot(s.x, s.y);
    return a > 0.0 ? cd{s.x / a, s.y / a} : cd{1.0, 0.0};
  };
  // this run's product of the factors r in [i0, i1) (D_{i1} / D_{i0})
  cd tot = {1.0, 0.0};
  for (int r = i0; r < i1 && r + 1 < n; ++r) tot = cmul(tot, phase(r));
  scan[tid] = tot;
  __syncthreads();
  // inclusive scan (Hillis-Steele) of the runs' products
  for (int o = 1; o < kThreads; o <<= 1) {
    const cd prev = tid >= o ? scan[tid - o] : cd{1.0, 0.0};
    __syncthreads();
    if (tid >= o) scan[tid] = cmul(prev, scan[tid]);
    __syncthreads();
  }
  cd dl = tid > 0 ? scan[tid - 1] : cd{1.0, 0.0};  // D_{i0}
  for (int i = i0; i < i1; ++i) {
    sy[i] = {dl.x * z[i], dl.y * z[i]};
    if (i + 1 < n) dl = cmul(dl, phase(i));
  }
  __syncthreads();
}

// x = H_0 ... H_{n-3} D z for vector blockIdx.x; D_i = prod_{r<i} sub[r]/|sub[r]|.
__global__ __launch_bounds__(kThreads) void k_back_transform(const double* __restrict__ zr,
                                                             int n, const cd* __restrict__ V,
                                                             const double* __restrict__ tau,
                                                             const cd* __restrict__ sub,
                                                             cd* __restrict__ out,
                                                             EigStride st) {
  extern __shared__ cd sy[];  // n entries
  __shared__ double red[8];
  const int q = blockIdx.x;
  {
    const size_t mi = blockIdx.y;
    zr += mi * st.z;
    V += mi * st.v;
    tau += mi * st.n;
    sub += mi * st.n;
    out += mi * st.z;
  }
  const double* z = zr + (size_t)q * n;
  __shared__ cd scan[kThreads];
  phase_scale(z, n, sub, sy, scan);
  for (int j = n - 3; j >= 0; --j) {
    const double t = tau[j];
    if (t == 0.0) continue;  // uniform
    const cd* v = V + (size_t)j * n;
    double sr = 0.0, si = 0.0;
    for (int i = j + 1 + threadIdx.x; i < n; i += kThreads) {
      const cd c = cmulc(v[i], sy[i]);
      sr += c.x;
      si += c.y;
    }
    block_sum2(sr, si, red);
    const cd s = {t * sr, t * si};
    for (int i = j + 1 + threadIdx.x; i < n; i += kThreads) {
      const cd c = cmul(v[i], s);
      sy[i].x -= c.x;
      sy[i].y -= c.y;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += kThreads) out[(size_t)q * n + i] = sy[i];
}

// The back-transform applies the reflectors in groups of R: group r holds
// u_a = row j0 - a of V (entries i > j0 - a), a < R, j0 = n - 3 - R r, in
// that order.  With G_ab = u_a^H u_b (b < a) the group's product acts as
//   s_a = tau_a (u_a^H y - sum_{b<a} G_ab s_b),   y <- y - sum_a u_a s_a,
// which equals applying the R reflectors one by one but needs one block
// reduction (of the R dot products u_a^H y) instead of R.  k_reflector_gram
// computes G once for all vectors: G[r][a (a - 1) / 2 + b].  Rows past
// j0 - a < 0 are zero (tau 0).
template <int R>
__global__ __launch_bounds__(kThreads) void k_reflector_gram(int n, const cd* __restrict__ V,
                                                             cd* __restrict__ G, EigStride st) {
  constexpr int P = R * (R - 1) / 2;
  __shared__ double red[4 * 2 * P];
  const int r = blockIdx.x;
  {
    const size_t mi = blockIdx.y;
    V += mi * st.v;
    G += mi * st.g;
  }
  const int j0 = n - 3 - R * r;
  double acc[2 * P];
#pragma unroll
  for (int q = 0; q < 2 * P; ++q) acc[q] = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) {
    cd u[R];
#pragma unroll
    for (int a = 0; a < R; ++a) {
      const int j = j0 - a;
      u[a] = (j >= 0 && i > j) ? V[(size_t)j * n + i] : cd{0.0, 0.0};
    }
#pragma unroll
    for (int a = 1; a < R; ++a)
#pragma unroll
      for (int b = 0; b < a; ++b) {
        const cd c = cmulc(u[a], u[b]);
        acc[2 * (a * (a - 1) / 2 + b)] += c.x;
        acc[2 * (a * (a - 1) / 2 + b) + 1] += c.y;
      }
  }
  block_sum_n<2 * P>(acc, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < P; ++q) G[(size_t)r * P + q] = {acc[2 * q], acc[2 * q + 1]};
  }
}

// x = H_0 ... H_{n-3} D z for n <= 256 C: thread t keeps y[t + 256 c]
// (c < C) in registers for the whole sweep, applies the reflectors in
// groups of R (one block reduction a group) and loads group r + 1 while it
// applies group r.
template <int C, int R>
__global__ __launch_bounds__(kThreads) void k_back_transform_reg(const double* __restrict__ zr,
                                                                 int n,
                                                                 const cd* __restrict__ V,
                                                                 const double* __restrict__ tau,
                                                                 const cd* __restrict__ sub,
                                                                 const cd* __restrict__ G,
                                                                 cd* __restrict__ out,
                                                                 EigStride st) {
  constexpr int P = R * (R - 1) / 2;
  extern __shared__ cd sy[];  // n entries
  __shared__ double red[4 * 2 * R];
  __shared__ cd scan[kThreads];
  const int q = blockIdx.x, tid = threadIdx.x;
  {
    const size_t mi = blockIdx.y;
    zr += mi * st.z;
    V += mi * st.v;
    tau += mi * st.n;
    sub += mi * st.n;
    G += mi * st.g;
    out += mi * st.z;
  }
  phase_scale(zr + (size_t)q * n, n, sub, sy, scan);
  cd y[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int i = tid + kThreads * c;
    y[c] = i < n ? sy[i] : cd{0.0, 0.0};  // (a prvalue: an lvalue zero
    // would make the compiler select between the two addresses and load
    // through a flat pointer)
  }
  const int ngroups = (n - 2 + R - 1) / R;  // reflectors n - 3 .. 0
  // tau and G of a group are loaded one value a lane (lane a: tau_{j0-a}
  // and G entry a) and read back with readlane where they are used: as
  // uniform loads they would go through the scalar cache, whose counter the
  // block reduction's LDS wait also drains, so the next group's values would
  // be waited for inside this group
  const int lq = tid & 63;
  // Every load is unconditional, from an address clamped into the arrays,
  // and the unused values are selected away: a load that a branch may skip
  // leaves the compiler unable to count the loads in flight, and it then
  // waits for all of them (the prefetched group's too) at the loop head.
  auto load_group = [&](int r, cd(&u)[R][C], double& t, cd& g) {
    const int j0 = n - 3 - R * r;
    const int jt = j0 - lq;
    const double tl = tau[min(max(jt, 0), n - 1)];
    t = (lq < R && jt >= 0) ? tl : 0.0;
    const cd gl = G[(size_t)min(r, ngroups - 1) * (P > 0 ? P : 1) + min(lq, P > 0 ? P - 1 : 0)];
    g = lq < P ? gl : cd{0.0, 0.0};
#pragma unroll
    for (int a = 0; a < R; ++a) {
      const int j = j0 - a;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int i = tid + kThreads * c;
        const cd v = V[(size_t)max(j, 0) * n + min(i, n - 1)];
        u[a][c] = (j >= 0 && i > j && i < n) ? v : cd{0.0, 0.0};
      }
    }
  };
  auto lane_value = [](double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
  };
  // group r from (u, t, g) while group r + 1 loads into (un, tn, gn); the
  // loop alternates two buffer sets, so that no register copy waits for the
  // loads in flight
  auto apply_group = [&](int r, const cd(&u)[R][C], double t, cd g, cd(&un)[R][C], double& tn,
                         cd& gn) {
    load_group(r + 1, un, tn, gn);  // (past the last group: zeros)
    double d[2 * R];
#pragma unroll
    for (int a = 0; a < R; ++a) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const cd m = cmulc(u[a][c], y[c]);
        sr += m.x;
        si += m.y;
      }
      d[2 * a] = sr;
      d[2 * a + 1] = si;
    }
    block_sum_n<2 * R>(d, red);
    cd sa[R];
#pragma unroll
    for (int a = 0; a < R; ++a) {
      cd x = {d[2 * a], d[2 * a + 1]};
#pragma unroll
      for (int b = 0; b < a; ++b) {
        const int pi = a * (a - 1) / 2 + b;
        const cd gab = {lane_value(g.x, pi), lane_value(g.y, pi)};
        const cd m = cmul(gab, sa[b]);
        x.x -= m.x;
        x.y -= m.y;
      }
      const double ta = lane_value(t, a);
      sa[a] = {ta * x.x, ta * x.y};
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int a = 0; a < R; ++a) {
        const cd m = cmul(u[a][c], sa[a]);
        y[c].x -= m.x;
        y[c].y -= m.y;
      }
  };
  cd u0[R][C], u1[R][C];
  double t0 = 0.0, t1 = 0.0;
  cd g0 = {0.0, 0.0}, g1 = {0.0, 0.0};
  if (ngroups > 0) load_group(0, u0, t0, g0);
  for (int r = 0; r < ngroups; r += 2) {  // (an odd count ends with an empty group)
    apply_group(r, u0, t0, g0, u1, t1, g1);
    apply_group(r + 1, u1, t1, g1, u0, t0, g0);
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int i = tid + kThreads * c;
    if (i < n) out[(size_t)q * n + i] = y[c];
  }
}

// Kernel files layout (klitho_tcc.cpp:344-362): eigenvalues as float, and
// krn_q[j] = z_q[T-1-j] as float real / imaginary parts.
__global__ void k_kernels_from_vectors(const cd* __restrict__ vec, const double* __restrict__ w,
                                       int k, int T, float* __restrict__ kr,
                                       float* __restrict__ ki, float* __restrict__ scales) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) scales[i] = (float)w[i];
  if (i >= k * T) return;
  const int q = i / T, jj = i - q * T;
  const cd v = vec[(size_t)q * T + (T - 1 - jj)];
  kr[i] = (float)v.x;
  ki[i] = (float)v.y;
}

}  // namespace

namespace klt {

// Top-k eigenpairs of the Hermitian n x n matrix a_dev (device, complex
// double, full storage; not modified).  w_dev: k eigenvalues (descending);
// vec_dev: k x n complex eigenvectors (rows, unit norm).
int eig_topk_batch_device(klt_ctx* ctx, const double* a_dev, int count, int n, int k,
                          double* w_dev, double* vec_dev) {
  if (n <= 0 || k <= 0 || k > n) return set_error(KLT_ERR_ARG, "eig: bad n or k");
  if (count <= 0 || count > 65535) return set_error(KLT_ERR_ARG, "eig: bad batch count");
  if (n > 8192) return set_error(KLT_ERR_ARG, "eig: n above 8192 is not supported");
  const size_t nn = (size_t)n * n;
  // workspace per matrix: A copy, V, tau, sub, diag, p, parts, e, e2, z, lu
  // work, clusters; each array holds the batch's matrices back to back
  const int nparts_max = kMaxParts;
  EigStride es;
  es.a = 2 * nn;
  es.v = nn;
  es.n = (size_t)n;
  es.p = 2 * (size_t)n;
  es.part = 2 * (size_t)nparts_max;
  es.k = (size_t)k;
  es.z = (size_t)k * n;
  es.g = 3 * (size_t)n;
  const size_t B = (size_t)count;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t oA = carve(B * nn * 16), oV = carve(B * nn * 16), oTau = carve(B * n * 8),
               oSub = carve(B * n * 16), oDiag = carve(B * n * 8), oP = carve(B * es.p * 16),
               oPart = carve(B * es.part * 16), oE = carve(B * n * 8), oE2 = carve(B * n * 8),
               oZ = carve(B * es.z * 8), oWork = carve(B * 5 * es.z * 8),
               oCl = carve(B * (k + 1) * 4), oNcl = carve(B * 4), oBnd = carve(B * 3 * 8),
               oTn = carve(B * 8), oG = carve(B * es.g * 16);
  KLT_TRY(ctx->eig_ws.ensure(off));
  char* base = ctx->eig_ws.as<char>();
  double* A = reinterpret_cast<double*>(base + oA);
  cd* V = reinterpret_cast<cd*>(base + oV);
  double* tau = reinterpret_cast<double*>(base + oTau);
  cd* sub = reinterpret_cast<cd*>(base + oSub);
  double* diag = reinterpret_cast<double*>(base + oDiag);
  cd* p = reinterpret_cast<cd*>(base + oP);
  cd* part = reinterpret_cast<cd*>(base + oPart);
  double* e = reinterpret_cast<double*>(base + oE);
  double* e2 = reinterpret_cast<double*>(base + oE2);
  double* z = reinterpret_cast<double*>(base + oZ);
  double* work = reinterpret_cast<double*>(base + oWork);
  int* cl = reinterpret_cast<int*>(base + oCl);
  int* ncl_dev = reinterpret_cast<int*>(base + oNcl);
  double* bnd = reinterpret_cast<double*>(base + oBnd);
  double* tn_dev = reinterpret_cast<double*>(base + oTn);
  cd* G = reinterpret_cast<cd*>(base + oG);
  hipStream_t st = ctx->stream;
  KLT_HIP(hipMemcpyAsync(A, a_dev, B * nn * 16, hipMemcpyDeviceToDevice, st));
  KLT_HIP(hipMemsetAsync(tau, 0, B * n * 8, st));
  KLT_HIP(hipMemsetAsync(sub, 0, B * n * 16, st));
  const size_t lds_v = std::max<size_t>(16, (size_t)n * 16);
  if (lds_v > 64 * 1024) {
    KLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_hh_symv),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_v));
    KLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_back_transform),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_v));
  }
  // one launch per step while v, v' and w' fit in LDS (3 n complex values)
  const size_t lds_step = (size_t)(3 * n + 2) * 16;
  const bool fused = lds_step <= 160 * 1024;
  if (fused && lds_step > 64 * 1024)
    KLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_hh_step),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_step));
  int np_prev = 0;
  bool tail = false;
  for (int j = 0; j + 2 < n; ++j) {
    const int m = n - j - 1;
    if (fused && n - j <= kTail) {  // the remaining steps in one block per matrix
      const cd* pp = p + (size_t)((j + 1) & 1) * n;
      const cd* partp = part + (size_t)((j + 1) & 1) * nparts_max;
      KLT_LAUNCH(ctx, "k_hh_tail", k_hh_tail, dim3(1, count), dim3(kTW * 64), 0, A, n, j, V, tau,
                 sub, diag, pp, partp, np_prev, es);
      tail = true;
      break;
    }
    // the fused step: one row per wave; the two-kernel path: 16 rows a block
    const int nblk = std::min(nparts_max, fused ? (m + 3) / 4 : (m + 15) / 16);
    const int rows = (m + nblk - 1) / nblk;
    const int np = (m + rows - 1) / rows;
    // step j writes p / part set j & 1 and reads step j-1's from the other
    cd* pj = p + (size_t)(j & 1) * n;
    cd* partj = part + (size_t)(j & 1) * nparts_max;
    if (fused) {
      const cd* pp = p + (size_t)((j + 1) & 1) * n;
      const cd* partp = part + (size_t)((j + 1) & 1) * nparts_max;
      KLT_LAUNCH(ctx, "k_hh_step", k_hh_step, dim3(np, count), dim3(kThreads),
                 (size_t)(3 * m + 2) * 16, A, n, j, rows, V, tau, sub, diag, pp, partp, np_prev,
                 pj, partj, es);
      if (j + 3 == n) {  // the last step's update, for the final 2 x 2 block
        KLT_LAUNCH(ctx, "k_hh_her2", k_hh_her2, dim3(1, 1, count), dim3(kThreads), 0, A, n, j,
                   V, tau, pj, partj, np, es);
      }
    } else {
      KLT_LAUNCH(ctx, "k_hh_symv", k_hh_symv, dim3(np, count), dim3(kThreads), (size_t)m * 16,
                 A, n, j, rows, V, tau, sub, diag, pj, partj, es);
      const int tiles = (m + 63) / 64;
      KLT_LAUNCH(ctx, "k_hh_her2", k_hh_her2, dim3(tiles, tiles, count), dim3(kThreads), 0, A,
                 n, j, V, tau, pj, partj, np, es);
    }
    np_prev = np;
  }
  if (!tail)
    KLT_LAUNCH(ctx, "k_hh_finish", k_hh_finish, dim3(count), dim3(64), 0, A, n, sub, diag, es);
  KLT_HIP(hipGetLastError());
  // per matrix: real tridiagonal, Gershgorin bounds, pivot threshold (host: O(n))
  std::vector<double> hd(B * n);
  std::vector<cd> hs(B * n);
  KLT_HIP(hipMemcpyAsync(hd.data(), diag, B * n * 8, hipMemcpyDeviceToHost, st));
  KLT_HIP(hipMemcpyAsync(hs.data(), sub, B * n * 16, hipMemcpyDeviceToHost, st));
  KLT_HIP(hipStreamSynchronize(st));
  std::vector<double> he(B * n, 0.0), he2(B * n, 0.0), hb(B * 3), htn(B);
  for (size_t mi = 0; mi < B; ++mi) {
    const double* dd = hd.data() + mi * n;
    const cd* ss = hs.data() + mi * n;
    double* ee = he.data() + mi * n;
    double* ee2 = he2.data() + mi * n;
    double lo = 1e300, hi = -1e300, tnorm = 0.0;
    for (int i = 0; i < n; ++i) {
      if (i + 1 < n) {
        ee[i] = std::hypot(ss[i].x, ss[i].y);
        ee2[i] = ee[i] * ee[i];
      }
      const double r = (i > 0 ? ee[i - 1] : 0.0) + (i + 1 < n ? ee[i] : 0.0);
      lo = std::min(lo, dd[i] - r);
      hi = std::max(hi, dd[i] + r);
      tnorm = std::max(tnorm, std::fabs(dd[i]) + r);
    }
    if (!std::isfinite(lo) || !std::isfinite(hi))
      return set_error(KLT_ERR_EIG, "eig: non-finite tridiagonal (input has NaN/Inf)");
    const double pivmin = std::max(1e-300, 2.2e-16 * 2.2e-16 * std::max(1.0, tnorm * tnorm));
    hb[3 * mi] = lo - (2.2e-16 * tnorm + pivmin);
    hb[3 * mi + 1] = hi + (2.2e-16 * tnorm + pivmin);
    hb[3 * mi + 2] = pivmin;
    htn[mi] = tnorm;
  }
  KLT_HIP(hipMemcpyAsync(e, he.data(), B * n * 8, hipMemcpyHostToDevice, st));
  KLT_HIP(hipMemcpyAsync(e2, he2.data(), B * n * 8, hipMemcpyHostToDevice, st));
  KLT_HIP(hipMemcpyAsync(bnd, hb.data(), B * 3 * 8, hipMemcpyHostToDevice, st));
  KLT_HIP(hipMemcpyAsync(tn_dev, htn.data(), B * 8, hipMemcpyHostToDevice, st));
  const size_t lds_de = (size_t)2 * n * 8;
  if (lds_de > 64 * 1024)
    KLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bisect),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_de));
  KLT_LAUNCH(ctx, "k_bisect", k_bisect, dim3(k, count), dim3(128), lds_de, diag, e2, n, k, bnd,
             w_dev, es);
  std::vector<double> hw(B * k);
  KLT_HIP(hipMemcpyAsync(hw.data(), w_dev, B * k * 8, hipMemcpyDeviceToHost, st));
  KLT_HIP(hipStreamSynchronize(st));
  // clusters of close eigenvalues (descending order).  Inverse iteration
  // leaves an orthogonality error of about eps ||T|| / gap between two
  // eigenvectors; pairs closer than 1e-4 ||T|| are re-orthogonalised, which
  // bounds the error elsewhere by ~2e-12.
  std::vector<int> hcl(B * (k + 1), k), hncl(B);
  int ncl_max = 0;
  for (size_t mi = 0; mi < B; ++mi) {
    const double* ww = hw.data() + mi * k;
    int* c = hcl.data() + mi * (k + 1);
    int nc = 0;
    c[nc++] = 0;
    for (int q = 1; q < k; ++q)
      if (std::fabs(ww[q - 1] - ww[q]) > 1e-4 * htn[mi]) c[nc++] = q;
    c[nc] = k;
    hncl[mi] = nc;
    ncl_max = std::max(ncl_max, nc);
  }
  KLT_HIP(hipMemcpyAsync(cl, hcl.data(), hcl.size() * 4, hipMemcpyHostToDevice, st));
  KLT_HIP(hipMemcpyAsync(ncl_dev, hncl.data(), B * 4, hipMemcpyHostToDevice, st));
  KLT_LAUNCH(ctx, "k_inverse_iteration", k_inverse_iteration, dim3(ncl_max, count), dim3(256),
             0, diag, e, n, w_dev, cl, ncl_dev, tn_dev, z, work, es);
  // reflector groups of 4 (2 where 8 entries a thread leave fewer registers);
  // n <= 2 has no reflectors
  if (n <= kThreads * 5) {
    if (n > 2)
      KLT_LAUNCH(ctx, "k_reflector_gram", k_reflector_gram<4>, dim3((n - 2 + 3) / 4, count),
                 dim3(kThreads), 0, n, V, G, es);
    KLT_LAUNCH(ctx, "k_back_transform", (k_back_transform_reg<5, 4>), dim3(k, count),
               dim3(kThreads), lds_v, z, n, V, tau, sub, G, reinterpret_cast<cd*>(vec_dev), es);
  } else if (n <= kThreads * 8) {
    if (lds_v > 64 * 1024)
      KLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_back_transform_reg<8, 2>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_v));
    KLT_LAUNCH(ctx, "k_reflector_gram", k_reflector_gram<2>, dim3((n - 2 + 1) / 2, count),
               dim3(kThreads), 0, n, V, G, es);  // (n > 1280 here)
    KLT_LAUNCH(ctx, "k_back_transform", (k_back_transform_reg<8, 2>), dim3(k, count),
               dim3(kThreads), lds_v, z, n, V, tau, sub, G, reinterpret_cast<cd*>(vec_dev), es);
  } else {
    KLT_LAUNCH(ctx, "k_back_transform", k_back_transform, dim3(k, count), dim3(kThreads), lds_v,
               z, n, V, tau, sub, reinterpret_cast<cd*>(vec_dev), es);
  }
  KLT_HIP(hipGetLastError());
  KLT_HIP(hipStreamSynchronize(st));  // host vectors above must outlive the copies
  return KLT_OK;
}

int eig_topk_device(klt_ctx* ctx, const double* a_dev, int n, int k, double* w_dev,
                    double* vec_dev) {
  return eig_topk_batch_device(ctx, a_dev, 1, n, k, w_dev, vec_dev);
}

// KLT_EIG_PSD: the subspace solver first; the Householder path when it
// reports that a matrix is outside its reach.
static int eig_dispatch(klt_ctx* ctx, const double* a_dev, int count, int n, int k, double* w,
                        double* v, unsigned flags, double tol = 0.0) {
  if (flags & KLT_EIG_PSD) {
    const int rc = eig_topk_psd_batch_device(ctx, a_dev, count, n, k, w, v,
                                             (flags & KLT_EIG_WARM) != 0, tol);
    if (rc != KLT_ERR_EIG) return rc;
  }
  return eig_topk_batch_device(ctx, a_dev, count, n, k, w, v);
}

// klt_calc_kernels' body (klitho_tcc.cpp:301-377), with the subspace solver's
// residual target tol (0: 5e-14).
int calc_kernels_device(klt_ctx* ctx, const double* tcc, int T, int nk, float* krn_r,
                        float* krn_i, float* scales, unsigned flags, double tol, bool sync) {
  if (!ctx || !tcc || !krn_r || !krn_i || !scales)
    return set_error(KLT_ERR_ARG, "klt_calc_kernels: NULL argument");
  if (T <= 0 || nk <= 0) return set_error(KLT_ERR_ARG, "klt_calc_kernels: bad size");
  if (nk > T) nk = T;  // klitho_tcc.cpp:488-491
  KLT_TRY(ctx->activate());
  const size_t nn = (size_t)T * T;
  const double* da = tcc;
  if (flags & KLT_HOST_PTRS) {
    KLT_TRY(ctx->stage_in.ensure(nn * 16));
    KLT_HIP(hipMemcpyAsync(ctx->stage_in.ptr, tcc, nn * 16, hipMemcpyHostToDevice, ctx->stream));
    da = ctx->stage_in.as<double>();
  }
  const size_t vb = (size_t)nk * T * 16, wb = (size_t)nk * 8, kb = (size_t)nk * T * 4;
  KLT_TRY(ctx->stage_out.ensure(vb + wb + 2 * kb + nk * 4 + 1024));
  char* o = ctx->stage_out.as<char>();
  double* dv = reinterpret_cast<double*>(o);
  double* dw = reinterpret_cast<double*>(o + vb);
  float* fr = reinterpret_cast<float*>(o + vb + wb);
  float* fi = fr + (size_t)nk * T;
  float* fs = fi + (size_t)nk * T;
  KLT_TRY(eig_dispatch(ctx, da, 1, T, nk, dw, dv, flags, tol));
  const int tot = nk * T;
  KLT_LAUNCH(ctx, "k_kernels_from_vectors", k_kernels_from_vectors, dim3((tot + 255) / 256),
             dim3(256), 0, reinterpret_cast<const cd*>(dv), dw, nk, T, fr, fi, fs);
  const hipMemcpyKind kind =
      (flags & KLT_HOST_PTRS) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  KLT_HIP(hipMemcpyAsync(krn_r, fr, kb, kind, ctx->stream));
  KLT_HIP(hipMemcpyAsync(krn_i, fi, kb, kind, ctx->stream));
  KLT_HIP(hipMemcpyAsync(scales, fs, nk * 4, kind, ctx->stream));
  // (sync false, device pointers: the results are ready in stream order)
  if (sync || (flags & KLT_HOST_PTRS)) KLT_HIP(hipStreamSynchronize(ctx->stream));
  return KLT_OK;
}

}  // namespace klt

extern "C" {

int klt_eig_topk(klt_ctx* ctx, const double* a, int n, int k, double* w, double* vecs,
                 unsigned flags) {
  if (!ctx || !a || !w || !vecs) return klt::set_error(KLT_ERR_ARG, "klt_eig_topk: NULL argument");
  if (n <= 0 || k <= 0) return klt::set_error(KLT_ERR_ARG, "klt_eig_topk: bad size");
  if (k > n) k = n;
  KLT_TRY(ctx->activate());
  const size_t nn = (size_t)n * n;
  const double* da = a;
  double* dw = w;
  double* dv = vecs;
  if (flags & KLT_HOST_PTRS) {
    KLT_TRY(ctx->stage_in.ensure(nn * 16));
    KLT_TRY(ctx->stage_out.ensure((size_t)k * n * 16 + (size_t)k * 8 + 256));
    KLT_HIP(hipMemcpyAsync(ctx->stage_in.ptr, a, nn * 16, hipMemcpyHostToDevice, ctx->stream));
    da = ctx->stage_in.as<double>();
    dv = ctx->stage_out.as<double>();
    dw = dv + (size_t)k * n * 2;
  }
  KLT_TRY(klt::eig_dispatch(ctx, da, 1, n, k, dw, dv, flags));
  if (flags & KLT_HOST_PTRS) {
    KLT_HIP(hipMemcpyAsync(vecs, dv, (size_t)k * n * 16, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipMemcpyAsync(w, dw, (size_t)k * 8, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return KLT_OK;
}

int klt_eig_topk_batch(klt_ctx* ctx, const double* a, int count, int n, int k, double* w,
                       double* vecs, unsigned flags) {
  if (!ctx || !a || !w || !vecs)
    return klt::set_error(KLT_ERR_ARG, "klt_eig_topk_batch: NULL argument");
  if (n <= 0 || k <= 0 || count <= 0)
    return klt::set_error(KLT_ERR_ARG, "klt_eig_topk_batch: bad size");
  if (k > n) k = n;
  KLT_TRY(ctx->activate());
  const size_t nn = (size_t)n * n, B = (size_t)count;
  const double* da = a;
  double* dw = w;
  double* dv = vecs;
  if (flags & KLT_HOST_PTRS) {
    KLT_TRY(ctx->stage_in.ensure(B * nn * 16));
    KLT_TRY(ctx->stage_out.ensure(B * ((size_t)k * n * 16 + (size_t)k * 8) + 256));
    KLT_HIP(hipMemcpyAsync(ctx->stage_in.ptr, a, B * nn * 16, hipMemcpyHostToDevice,
                           ctx->stream));
    da = ctx->stage_in.as<double>();
    dv = ctx->stage_out.as<double>();
    dw = dv + B * k * n * 2;
  }
  KLT_TRY(klt::eig_dispatch(ctx, da, count, n, k, dw, dv, flags));
  if (flags & KLT_HOST_PTRS) {
    KLT_HIP(hipMemcpyAsync(vecs, dv, B * k * n * 16, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipMemcpyAsync(w, dw, B * k * 8, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
  }
  return KLT_OK;
}

int klt_calc_kernels(klt_ctx* ctx, const double* tcc, int T, int nk, float* krn_r,
                     float* krn_i, float* scales, unsigned flags) {
  return klt::calc_kernels_device(ctx, tcc, T, nk, krn_r, krn_i, scales, flags, 0.0);
}

}  // extern "C"
