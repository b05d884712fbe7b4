// eig_subspace.hip -- top-k eigenpairs of positive semidefinite Hermitian
// matrices (every TCC of a non-negative source: TCC = P^H diag(S) P) by
// block subspace iteration with Rayleigh-Ritz, FP64.
// Replaces the eigen step of calcKernels (K-Litho-TCC/klitho_tcc.cpp:301-377,
// zheevr_ for the nk largest) when the caller states that the matrix is PSD.
//
// Per matrix A (n x n), block size p = 64 >= k + 16: a deterministic
// pseudo-random start block, then rounds of X <- cholqr(A^2 X) (G = Y^H Y =
// L L^H, Y <- Y L^{-H}; twice, CholQR2, before each Rayleigh-Ritz step, once
// in the other rounds, where the basis need only stay well conditioned and
// span the same subspace).  Each round shrinks the
// components of the wanted eigenvectors outside the block by
// (lambda_p / lambda_q)^2 (about 0.1 for the TCCs of the tests at q = 31);
// the spread of the columns of A^2 X, (lambda_0 / lambda_{p-1})^2 ~ 2e7 for
// them, stays within what CholQR orthonormalises.  After kFirstCheck rounds,
// and every kCheckEvery rounds after that, Rayleigh-Ritz: Z = A X,
// H = X^H Z = U diag(theta) U^H for the k largest (the Householder solver of
// eig.hip on the p x p matrix), eigenvectors X U_k, residuals
// ||Z U_q - theta_q X U_q||; a matrix is done when every residual is at most
// kTol theta_0.  A matrix not done after kMaxRounds rounds, or whose Cholesky
// factor breaks down, is reported to the caller, which then uses the
// Householder path.  Every kernel takes the batch in its last grid dimension
// and a per-matrix `active` flag, so each matrix runs the same arithmetic and
// stops at the same round alone or in a batch; the split-K factor of the
// GEMMs depends on n only.
#include <algorithm>
#include <cmath>
#include <vector>

#include "klt_internal.h"

namespace {

typedef double2 cd;

constexpr int kP = 64;          // block size (columns)
constexpr int kFirstCheck = 9;  // rounds of X <- orth(A^2 X) before the first check
constexpr int kCheckEvery = 3;  // rounds between later checks
constexpr int kMaxRounds = 24;
constexpr double kTol = 5e-14;  // residual, relative to theta_0
constexpr int kT = 64;          // GEMM tile (rows, columns)
constexpr int kK = 16;          // GEMM K step

__device__ __forceinline__ cd cfma(cd a, cd b, cd c) {  // c + a b
  c.x = fma(a.x, b.x, fma(-a.y, b.y, c.x));
  c.y = fma(a.x, b.y, fma(a.y, b.x, c.y));
  return c;
}

// C[b] = A[b] B[b] (complex, row-major; A: n x K with row stride lda, B: K x p,
// or with TB its transpose stored p x K; C: n x p).  grid (ceil(n / 64),
// nsplit, count); block = 256 threads, a 64 x 64 output tile, thread (ty, tx)
// owns rows ty + 16 i, columns tx + 16 j.  nsplit > 1: the block's share of K
// goes to part[b][ks] (n x p) and k_zgemm_reduce sums the shares.
template <bool TB>
__global__ __launch_bounds__(256) void k_zgemm(const cd* __restrict__ A, int lda, size_t sA,
                                               const cd* __restrict__ B, int ldb, size_t sB,
                                               cd* __restrict__ C, size_t sC, int n, int K,
                                               int p, int kchunk, cd* __restrict__ part,
                                               const int* __restrict__ active) {
  const int b = blockIdx.z;
  if (active && !active[b]) return;
  __shared__ cd sA_[kT][kK + 1];
  __shared__ cd sB_[kK][kT + 1];
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  const int r0 = blockIdx.x * kT, ks = blockIdx.y, nsplit = gridDim.y;
  const int k0 = ks * kchunk, k1 = min(K, k0 + kchunk);
  A += (size_t)b * sA;
  B += (size_t)b * sB;
  cd acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = make_double2(0.0, 0.0);
  for (int kb = k0; kb < k1; kb += kK) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      const int row = e >> 4, kk = e & 15;
      const int gr = r0 + row, gk = kb + kk;
      sA_[row][kk] = (gr < n && gk < k1) ? A[(size_t)gr * lda + gk] : make_double2(0.0, 0.0);
      if (TB) {
        const int q = e >> 4, j = e & 15;
        const int gj = kb + j;
        sB_[j][q] = (q < p && gj < k1) ? B[(size_t)q * ldb + gj] : make_double2(0.0, 0.0);
      } else {
        const int kr = e >> 6, col = e & 63;
        const int gk2 = kb + kr;
        sB_[kr][col] = (col < p && gk2 < k1) ? B[(size_t)gk2 * ldb + col] : make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kK; ++kk) {
      cd a[4], bb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sA_[ty + 16 * i][kk];
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[j] = sB_[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = cfma(a[i], bb[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = r0 + ty + 16 * i;
    if (r >= n) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = tx + 16 * j;
      if (c >= p) continue;
      const size_t o = (size_t)r * p + c;
      if (nsplit > 1)
        part[((size_t)b * nsplit + ks) * n * p + o] = acc[i][j];
      else
        C[(size_t)b * sC + o] = acc[i][j];
    }
  }
}

// Fixed-order sum of the split-K partials; grid (ceil(n p / 256), 1, count)
__global__ __launch_bounds__(256) void k_zgemm_reduce(const cd* __restrict__ part, int nsplit,
                                                      int n, int p, cd* __restrict__ C,
                                                      size_t sC,
                                                      const int* __restrict__ active) {
  const int b = blockIdx.z;
  if (active && !active[b]) return;
  const size_t np = (size_t)n * p, o = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= np) return;
  cd s = make_double2(0.0, 0.0);
  for (int ks = 0; ks < nsplit; ++ks) {
    const cd v = part[((size_t)b * nsplit + ks) * np + o];
    s.x += v.x;
    s.y += v.y;
  }
  C[(size_t)b * sC + o] = s;
}

// Partial Gram G_part[b][chunk] = X[chunk rows]^H Y[chunk rows] (p x p, p <= 64);
// grid (ceil(n / 64), 1, count).
__global__ __launch_bounds__(256) void k_gram(const cd* __restrict__ X, const cd* __restrict__ Y,
                                              size_t s, int n, int p, cd* __restrict__ part,
                                              const int* __restrict__ active) {
  const int b = blockIdx.z;
  if (active && !active[b]) return;
  __shared__ cd sX[kK][kT + 1];
  __shared__ cd sY[kK][kT + 1];
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  const int r0 = blockIdx.x * kT, r1 = min(n, r0 + kT);
  X += (size_t)b * s;
  Y += (size_t)b * s;
  cd acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = make_double2(0.0, 0.0);
  for (int rb = r0; rb < r1; rb += kK) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u, rr = e >> 6, col = e & 63, gr = rb + rr;
      const bool ok = gr < r1 && col < p;
      sX[rr][col] = ok ? X[(size_t)gr * p + col] : make_double2(0.0, 0.0);
      sY[rr][col] = ok ? Y[(size_t)gr * p + col] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < kK; ++rr) {
      cd a[4], bb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = sX[rr][ty + 16 * i];
        a[i].y = -a[i].y;  // conj(X)
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[j] = sY[rr][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = cfma(a[i], bb[j], acc[i][j]);
    }
    __syncthreads();
  }
  cd* out = part + ((size_t)b * gridDim.x + blockIdx.x) * p * p;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = ty + 16 * i, c = tx + 16 * j;
      if (r < p && c < p) out[r * p + c] = acc[i][j];
    }
}

// G[b] = sum over chunks (fixed order); herm: G <- (G + G^H) / 2.
// grid (ceil(p p / 256), 1, count)
__global__ __launch_bounds__(256) void k_gram_reduce(const cd* __restrict__ part, int nchunk,
                                                     int p, cd* __restrict__ G, bool herm,
                                                     const int* __restrict__ active) {
  const int b = blockIdx.z;
  if (active && !active[b]) return;
  const int o = blockIdx.x * 256 + threadIdx.x;
  if (o >= p * p) return;
  const int r = o / p, c = o % p, ot = c * p + r;
  cd s = make_double2(0.0, 0.0), t = make_double2(0.0, 0.0);
  for (int q = 0; q < nchunk; ++q) {
    const cd* P = part + ((size_t)b * nchunk + q) * p * p;
    s.x += P[o].x;
    s.y += P[o].y;
    if (herm) {
      t.x += P[ot].x;
      t.y += P[ot].y;
    }
  }
  if (herm) s = make_double2(0.5 * (s.x + t.x), 0.5 * (s.y - t.y));
  G[(size_t)b * p * p + o] = s;
}

// Per matrix (grid (1, 1, count), 256 threads): G = L L^H (left-looking
// Cholesky in LDS: for column j, thread (i, part) sums a quarter of
// sum_{l<j} L[i][l] conj(L[j][l]), then row i's entry is finished and scaled)
// and Binv = L^{-H} (upper triangular, row-major p x p) for Y <- Y Binv, from
// M = L^{-1} row by row (M[i][c] = (delta_ic - sum_{c<=l<i} L[i][l] M[l][c]) /
// L[i][i], split the same way); fail[b] = 1 when a pivot is not positive.
__global__ __launch_bounds__(256) void k_chol_inv(const cd* __restrict__ G, int p,
                                                  cd* __restrict__ Binv, int* __restrict__ fail,
                                                  const int* __restrict__ active) {
  const int b = blockIdx.z;
  if (active && !active[b]) return;
  __shared__ cd L[kP][kP + 1];
  __shared__ cd M[kP][kP + 1];
  __shared__ cd red[4][kP];
  __shared__ int bad;
  const int tid = threadIdx.x, r = tid & 63, part = tid >> 6;
  for (int o = tid; o < p * p; o += 256) L[o / p][o % p] = G[(size_t)b * p * p + o];
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < p; ++j) {
    cd s = make_double2(0.0, 0.0);
    if (r >= j && r < p)
      for (int l = part; l < j; l += 4) {
        const cd x = L[r][l], y = L[j][l];  // x conj(y)
        s.x += x.x * y.x + x.y * y.y;
        s.y += x.y * y.x - x.x * y.y;
      }
    red[part][r] = s;
    __syncthreads();
    // every finishing thread forms the pivot itself (no barrier for it)
    if (part == 0 && r >= j && r < p) {
      const double d = L[j][j].x - red[0][j].x - red[1][j].x - red[2][j].x - red[3][j].x;
      const double ljj = d > 0.0 ? sqrt(d) : 1.0;
      if (r == j) {
        if (!(d > 0.0)) bad = 1;
        L[j][j] = make_double2(ljj, 0.0);
      } else {
        const cd g = L[r][j];
        const double inv = 1.0 / ljj;
        L[r][j] = make_double2((g.x - red[0][r].x - red[1][r].x - red[2][r].x - red[3][r].x) * inv,
                               (g.y - red[0][r].y - red[1][r].y - red[2][r].y - red[3][r].y) * inv);
      }
    }
    __syncthreads();
  }
  for (int i = 0; i < p; ++i) {
    // M[i][c] for c = r <= i
    cd s = make_double2(0.0, 0.0);
    if (r < i)
      for (int l = r + part; l < i; l += 4) s = cfma(L[i][l], M[l][r], s);
    red[part][r] = s;
    __syncthreads();
    if (part == 0 && r <= i) {
      const double inv = 1.0 / L[i][i].x;
      if (r == i) {
        M[i][i] = make_double2(inv, 0.0);
      } else {
        const cd t = make_double2(red[0][r].x + red[1][r].x + red[2][r].x + red[3][r].x,
                                  red[0][r].y + red[1][r].y + red[2][r].y + red[3][r].y);
        M[i][r] = make_double2(-t.x * inv, -t.y * inv);
      }
    }
    if (part == 1 && r > i && r < p) M[i][r] = make_double2(0.0, 0.0);
    __syncthreads();
  }
  for (int o = tid; o < p * p; o += 256) {  // Binv[r][q] = conj(M[q][r])
    const int rr = o / p, q = o % p;
    const cd v = M[q][rr];
    Binv[(size_t)b * p * p + o] = make_double2(v.x, -v.y);
  }
  if (tid == 0 && bad) fail[b] = 1;
}

// res[b][q] = || Z[:, q] - theta[b][q] X[:, q] ||, q < k; grid (k, 1, count)
__global__ __launch_bounds__(256) void k_residuals(const cd* __restrict__ X,
                                                   const cd* __restrict__ Z, size_t s, int n,
                                                   int p, int k, const double* __restrict__ th,
                                                   double* __restrict__ res,
                                                   const int* __restrict__ active) {
  const int b = blockIdx.z, q = blockIdx.x;
  if (active && !active[b]) return;
  __shared__ double red[4];
  const double t = th[(size_t)b * p + q];
  double acc = 0.0;
  for (int r = threadIdx.x; r < n; r += 256) {
    const cd z = Z[(size_t)b * s + (size_t)r * p + q], x = X[(size_t)b * s + (size_t)r * p + q];
    const double dx = z.x - t * x.x, dy = z.y - t * x.y;
    acc += dx * dx + dy * dy;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) res[(size_t)b * k + q] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

// deterministic start block in [-0.5, 0.5) + i [-0.5, 0.5); grid (ceil(n p / 256), 1, count)
__global__ __launch_bounds__(256) void k_start_block(cd* __restrict__ X, size_t s, int n, int p) {
  const size_t o = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= (size_t)n * p) return;
  unsigned h = 2463534242u ^ (unsigned)(o * 2654435761u);
  h ^= h << 13;
  h ^= h >> 17;
  h ^= h << 5;
  unsigned g = h * 1664525u + 1013904223u;
  g ^= g << 13;
  g ^= g >> 17;
  g ^= g << 5;
  X[(size_t)blockIdx.z * s + o] = make_double2((double)(h & 0xffffff) / 16777216.0 - 0.5,
                                               (double)(g & 0xffffff) / 16777216.0 - 0.5);
}

// outputs: w[b][q] = theta[b][q], vecs[b][q][r] = X[b][r][q], q < k;
// grid (ceil(n k / 256), 1, count)
__global__ __launch_bounds__(256) void k_sub_output(const cd* __restrict__ X, size_t s, int n,
                                                    int p, int k, const double* __restrict__ th,
                                                    double* __restrict__ w,
                                                    cd* __restrict__ vecs) {
  const int b = blockIdx.z;
  const size_t o = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (o < (size_t)k) w[(size_t)b * k + o] = th[(size_t)b * p + o];
  if (o >= (size_t)n * k) return;
  const int q = (int)(o / n), r = (int)(o % n);
  vecs[((size_t)b * k + q) * n + r] = X[(size_t)b * s + (size_t)r * p + q];
}

}  // namespace

namespace klt {

// Returns KLT_OK with every matrix solved, or KLT_ERR_EIG when some matrix
// is outside the method's reach (the caller then uses the Householder path).
int eig_topk_psd_batch_device(klt_ctx* ctx, const double* a_dev, int count, int n, int k,
                              double* w_dev, double* vec_dev) {
  const int p = kP;
  if (k + 16 > p || n < 2 * p) return set_error(KLT_ERR_EIG, "psd subspace: shape not covered");
  const size_t np = (size_t)n * p, nk = (size_t)n * k, B = (size_t)count, pp = (size_t)p * p;
  const int rowtiles = (n + kT - 1) / kT;
  // split-K from n alone (identical arithmetic in batched and single calls)
  int nsplit = std::max(1, std::min(16, (512 + rowtiles - 1) / rowtiles));
  const int kchunk = ((n + nsplit - 1) / nsplit + kK - 1) / kK * kK;
  nsplit = (n + kchunk - 1) / kchunk;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t oX = carve(B * np * 16), oY0 = carve(B * np * 16), oY1 = carve(B * np * 16),
               oZ = carve(B * np * 16), oPart = carve(B * nsplit * np * 16),
               oGp = carve(B * rowtiles * pp * 16), oG = carve(B * pp * 16),
               oBi = carve(B * pp * 16), oU = carve(B * k * p * 16), oTh = carve(B * k * 8),
               oRes = carve(B * k * 8), oFail = carve(B * 4), oAct = carve(B * 4);
  KLT_TRY(ctx->sub_ws.ensure(off));
  char* base = ctx->sub_ws.as<char>();
  cd* X = reinterpret_cast<cd*>(base + oX);
  cd* Y0 = reinterpret_cast<cd*>(base + oY0);
  cd* Y1 = reinterpret_cast<cd*>(base + oY1);
  cd* Z = reinterpret_cast<cd*>(base + oZ);
  cd* part = reinterpret_cast<cd*>(base + oPart);
  cd* Gp = reinterpret_cast<cd*>(base + oGp);
  cd* G = reinterpret_cast<cd*>(base + oG);
  cd* Bi = reinterpret_cast<cd*>(base + oBi);
  cd* U = reinterpret_cast<cd*>(base + oU);
  double* th = reinterpret_cast<double*>(base + oTh);
  double* res = reinterpret_cast<double*>(base + oRes);
  int* fail = reinterpret_cast<int*>(base + oFail);
  int* act = reinterpret_cast<int*>(base + oAct);
  const cd* Ad = reinterpret_cast<const cd*>(a_dev);
  std::vector<int> h_act(count, 1), h_fail(count, 0);
  std::vector<double> h_res(B * k), h_th(B * k);
  KLT_TRY(ctx->upload(act, h_act.data(), B * 4));
  KLT_TRY(ctx->upload(fail, h_fail.data(), B * 4));
  const dim3 gblk(256);
  // C = A M (n x p blocks)
  auto gemm_A = [&](const cd* M, cd* C) {
    KLT_LAUNCH(ctx, "k_zgemm", k_zgemm<false>, dim3(rowtiles, nsplit, count), gblk, 0, Ad, n,
               (size_t)n * n, M, p, np, C, np, n, n, p, kchunk, part, act);
    if (nsplit > 1)
      KLT_LAUNCH(ctx, "k_zgemm_reduce", k_zgemm_reduce,
                 dim3((unsigned)((np + 255) / 256), 1, count), gblk, 0, part, nsplit, n, p, C,
                 np, act);
  };
  // C (n x q) = Yin (n x p) Bm, Bm p x q (tb: Bm holds its transpose, q x p)
  auto gemm_small = [&](const cd* Yin, const cd* Bm, bool tb, int q, cd* C) {
    if (tb)
      KLT_LAUNCH(ctx, "k_zgemm", k_zgemm<true>, dim3(rowtiles, 1, count), gblk, 0, Yin, p, np,
                 Bm, p, (size_t)q * p, C, (size_t)n * q, n, p, q, p, part, act);
    else
      KLT_LAUNCH(ctx, "k_zgemm", k_zgemm<false>, dim3(rowtiles, 1, count), gblk, 0, Yin, p, np,
                 Bm, q, (size_t)p * q, C, (size_t)n * q, n, p, q, p, part, act);
  };
  auto gram = [&](const cd* P, const cd* Q) {
    KLT_LAUNCH(ctx, "k_gram", k_gram, dim3(rowtiles, 1, count), gblk, 0, P, Q, np, n, p, Gp, act);
    KLT_LAUNCH(ctx, "k_gram_reduce", k_gram_reduce, dim3((unsigned)((pp + 255) / 256), 1, count),
               gblk, 0, Gp, rowtiles, p, G, true, act);
  };
  // one CholQR pass: Xout = Yin L^{-H}, Yin^H Yin = L L^H (Xout distinct from
  // Yin); two passes (CholQR2) orthonormalise to rounding level
  auto cholqr = [&](const cd* Yin, cd* Xout) {
    gram(Yin, Yin);
    KLT_LAUNCH(ctx, "k_chol_inv", k_chol_inv, dim3(1, 1, count), dim3(256), 0, G, p, Bi, fail,
               act);
    gemm_small(Yin, Bi, false, p, Xout);
  };
  // Rayleigh-Ritz for the k largest: Z = A X, H = X^H Z = U diag(theta) U^H,
  // Xk = X U_k (in Y0, n x k), Zk = Z U_k (in Y1), residuals; theta,
  // residuals and the Cholesky flags to the host
  auto rr = [&]() -> int {
    gemm_A(X, Z);
    gram(X, Z);
    KLT_TRY(eig_topk_batch_device(ctx, reinterpret_cast<const double*>(G), count, p, k, th,
                                  reinterpret_cast<double*>(U)));
    gemm_small(X, U, true, k, Y0);
    gemm_small(Z, U, true, k, Y1);
    KLT_LAUNCH(ctx, "k_residuals", k_residuals, dim3(k, 1, count), gblk, 0, Y0, Y1, nk, n, k, k,
               th, res, act);
    KLT_HIP(hipMemcpyAsync(h_res.data(), res, B * k * 8, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipMemcpyAsync(h_th.data(), th, B * k * 8, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipMemcpyAsync(h_fail.data(), fail, B * 4, hipMemcpyDeviceToHost, ctx->stream));
    KLT_HIP(hipStreamSynchronize(ctx->stream));
    KLT_HIP(hipGetLastError());
    return KLT_OK;
  };
  // rounds of X <- orth(A^2 X); Rayleigh-Ritz after kFirstCheck rounds and
  // every kCheckEvery after that; a matrix stops at its first converged check
  KLT_LAUNCH(ctx, "k_start_block", k_start_block, dim3((unsigned)((np + 255) / 256), 1, count),
             gblk, 0, Y0, np, n, p);
  cholqr(Y0, Y1);
  cholqr(Y1, X);
  for (int round = 1;; ++round) {
    gemm_A(X, Y0);
    gemm_A(Y0, Y1);
    // one pass keeps the span (to rounding relative to the largest column)
    // and a basis close enough to orthonormal for the next products; two
    // before the Rayleigh-Ritz step, which needs an orthonormal X
    const bool check = round >= kFirstCheck && (round - kFirstCheck) % kCheckEvery == 0;
    if (check) {
      cholqr(Y1, Y0);
      cholqr(Y0, X);
    } else {
      cholqr(Y1, X);
      continue;
    }
    KLT_TRY(rr());
    bool all = true;
    for (int b = 0; b < count; ++b) {
      if (h_fail[b]) return set_error(KLT_ERR_EIG, "psd subspace: Cholesky breakdown");
      if (!h_act[b]) continue;
      double mx = 0.0;
      for (int q = 0; q < k; ++q) mx = std::max(mx, h_res[(size_t)b * k + q]);
      if (mx <= kTol * std::fabs(h_th[(size_t)b * k])) {
        h_act[b] = 0;  // Y0 / Y1 of this matrix hold its result from here on
      } else {
        all = false;
      }
    }
    if (all) break;
    if (round >= kMaxRounds) return set_error(KLT_ERR_EIG, "psd subspace: no convergence");
    KLT_TRY(ctx->upload(act, h_act.data(), B * 4));
  }
  KLT_LAUNCH(ctx, "k_sub_output", k_sub_output, dim3((unsigned)((nk + 255) / 256), 1, count),
             gblk, 0, Y0, nk, n, k, k, th, w_dev, reinterpret_cast<cd*>(vec_dev));
  KLT_HIP(hipGetLastError());
  return KLT_OK;
}

}  // namespace klt
