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
// GEMMs depends on n only.  The products (A X, the tall n x 64 times 64 x 64
// ones and the Grams) run on the matrix cores in FP64 (k_zmm), the Cholesky
// factor and its inverse in registers of one block (k_chol_inv).
#include <algorithm>
#include <cmath>
#include <vector>

#include "klt_internal.h"

namespace {

typedef double2 cd;

constexpr int kP = 64;          // block size (columns)
constexpr int kFirstCheck = 11;  // rounds of X <- orth(A^2 X) before the first check
constexpr int kWarmFirstCheck = 5;  // the same from the previous solve's block
constexpr int kCheckEvery = 3;  // rounds between later checks
constexpr int kMaxRounds = 24;
constexpr double kTol = 5e-14;  // residual, relative to theta_0

typedef double d4 __attribute__((ext_vector_type(4)));

// C = op(A) op(B) (complex FP64) on the matrix cores, v_mfma_f64_16x16x4f64:
// lane l supplies a = op(A)[l % 16][l / 16] and b = op(B)[l / 16][l % 16] and
// holds D[l / 16 + 4 v][l % 16], v < 4.  A complex product takes four of
// them: re += ar br, re' += ai bi, im += ar bi + ai br (re - re' at the end).
// op(A) is M x K: A row-major (lda), or with AT the conjugate transpose of a
// K x M row-major A.  op(B) is K x Q, Q <= 64: B row-major (ldb), or with BT
// the transpose of a Q x K row-major B.  Block: 256 threads, a tile of 16 rows
// x 64 columns; wave w takes k-steps (of 4) w, w + 4, ... of the block's share
// [k0, k1) of K, and the four waves' sums are added in wave order through LDS.
// grid (ceil(M / 16), nsplit, count), kchunk a multiple of 16; nsplit > 1:
// block (x, y) writes part[b][y] (M x 64, row stride 64), else C (row stride
// ldc).  Loads are clamped in bounds and zeroed outside.
template <bool AT, bool BT>
__global__ __launch_bounds__(256, 2) void k_zmm(const cd* __restrict__ A, int lda, size_t sA,
                                             const cd* __restrict__ B, int ldb, size_t sB,
                                             cd* __restrict__ C, int ldc, size_t sC, int M,
                                             int K, int Q, int kchunk, cd* __restrict__ part,
                                             const int* __restrict__ active) {
  const int b = blockIdx.z;
  if (active && !active[b]) return;
  __shared__ cd red[4][16][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int r0 = blockIdx.x * 16, ys = blockIdx.y, nsplit = gridDim.y;
  const int k0 = ys * kchunk, k1 = min(K, k0 + kchunk);
  A += (size_t)b * sA;
  B += (size_t)b * sB;
  const int row = r0 + lr, rowc = min(row, M - 1);
  auto load = [&](int ks, cd& a, cd (&bv)[4]) {
    const int k = ks + lk, kc = min(k, K - 1);
    const bool kin = k < k1;
    cd av = AT ? A[(size_t)kc * lda + rowc] : A[(size_t)rowc * lda + kc];
    if (AT) av.y = -av.y;
    a = (kin && row < M) ? av : make_double2(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = 16 * j + lr, colc = min(col, Q - 1);
      const cd bj = BT ? B[(size_t)colc * ldb + kc] : B[(size_t)kc * ldb + colc];
      bv[j] = (kin && col < Q) ? bj : make_double2(0.0, 0.0);
    }
  };
  d4 re1[4], re2[4], im[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) re1[j] = re2[j] = im[j] = d4{0.0, 0.0, 0.0, 0.0};
  auto mma = [&](const cd& a, const cd (&bv)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) re1[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bv[j].x, re1[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) re2[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, bv[j].y, re2[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) im[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bv[j].y, im[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) im[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, bv[j].x, im[j], 0, 0, 0);
  };
  // this wave's steps i = 0 .. nst - 1 at k = k0 + 4 wave + 16 i (zeros past
  // k1); operands of four steps in a register ring, loaded three steps ahead
  const int nst = (k1 - k0 + 15) / 16, kw = k0 + 4 * wave;
  cd a[4], bv[4][4];
#pragma unroll
  for (int u = 0; u < 3; ++u) load(kw + 16 * u, a[u], bv[u]);
  int i = 0;
  for (; i + 4 <= nst; i += 4) {  // sets 0, 1, 2 hold steps i, i + 1, i + 2 here
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      load(kw + 16 * (i + u + 3), a[(u + 3) & 3], bv[(u + 3) & 3]);
      mma(a[u], bv[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < 3; ++u)
    if (i + u < nst) mma(a[u], bv[u]);
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      red[wave][lk + 4 * v][16 * j + lr] = make_double2(re1[j][v] - re2[j][v], im[j][v]);
  __syncthreads();
  cd* out = nsplit > 1 ? part + ((size_t)b * nsplit + ys) * M * 64 : C + (size_t)b * sC;
  const int ldo = nsplit > 1 ? 64 : ldc;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = tid + 256 * u, rr = e >> 6, col = e & 63;
    const cd v0 = red[0][rr][col], v1 = red[1][rr][col], v2 = red[2][rr][col],
             v3 = red[3][rr][col];
    const cd v = make_double2(((v0.x + v1.x) + v2.x) + v3.x, ((v0.y + v1.y) + v2.y) + v3.y);
    if (r0 + rr < M && col < Q) out[(size_t)(r0 + rr) * ldo + col] = v;
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

// Per matrix (grid (1, 1, count), kCW waves): G = L L^H and Binv = L^{-H}
// (upper triangular, row-major p x p, p <= 64) for Y <- Y Binv.  G is padded
// with the identity to 64 x 64 and held in registers: wave q, lane c keeps
// W[i] = row 4q + i, column c.  Step j of the right-looking factorisation
// turns W into M = L^{-1} one row and column at a time, in place: the wave
// holding row j (which is, conjugated, the trailing column j of the Hermitian
// remainder) forms l_jj = sqrt(W[j][j]) and u[c] = W[j][c] / l_jj
// (u[j] = 1 / l_jj), so that u[c] = conj(L[c][j]) for c > j and
// u[c] = M[j][c] for c <= j; row j becomes M[j][.] (zero right of the
// diagonal); then every row r > j takes W[r][c] -= L[r][j] u[c],
// L[r][j] = conj(u[r]) (column j first cleared), which is the trailing update
// G'[r][c] -= L[r][j] conj(L[c][j]) for c > j and the forward substitution
// M[r][c] -= L[r][j] M[j][c] for c <= j.  The steps go in groups of four,
// the rows of one wave: that wave runs the group's pivots and the updates
// among its own rows (u[r] read from lane r), and publishes the four u
// vectors through LDS; after one barrier every later wave applies them to
// its rows in step order (the operations of the step-by-step order, in that
// order), the wave of the next group first among its work.  16 barriers
// (one a step took 28 us at p = 64); fail[b] = 1 when a pivot is not
// positive.
constexpr int kCW = 16;        // waves
constexpr int kCR = kP / kCW;  // rows a wave: one group of steps
static_assert(kCR == 4, "four rows a wave");

// 1 / sqrt(d) for a normal d > 0: the hardware estimate and two Newton steps
// y <- y + y (1 - d y^2) / 2 (each squares the relative error; the estimate's
// is far below 2^-14, so the result is within rounding of the exact value).
// The IEEE square root and division it replaces are long dependent sequences
// on the critical path of every factorisation step.
__device__ __forceinline__ double rsqrt_nr(double d) {
  double y = __builtin_amdgcn_rsq(d);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-d * y, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}
__device__ __forceinline__ double lane_f64(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

__global__ __launch_bounds__(kCW * 64) void k_chol_inv(const cd* __restrict__ G, int p,
                                                       cd* __restrict__ Binv,
                                                       int* __restrict__ fail,
                                                       const int* __restrict__ active) {
  const int b = blockIdx.z;
  if (active && !active[b]) return;
  __shared__ cd ub[2][kCR][kP];  // a group's u vectors (double-buffered)
  __shared__ cd mt[kP][kP + 1];  // mt[c][r] = conj(M[r][c]), written out at the end
  const int tid = threadIdx.x, c = tid & 63;
  const int q = __builtin_amdgcn_readfirstlane(tid >> 6);  // the wave (uniform)
  const cd* Gb = G + (size_t)b * p * p;
  cd* Bo = Binv + (size_t)b * p * p;
  cd W[kCR];
#pragma unroll
  for (int i = 0; i < kCR; ++i) {
    const int r = kCR * q + i;
    W[i] = (r < p && c < p) ? Gb[(size_t)r * p + c] : make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
  bool bad = false;
  auto update = [&](cd& Wr, const cd& ur, const cd& uc, int j) {  // L[r][j] = conj(ur)
    cd w = c == j ? make_double2(0.0, 0.0) : Wr;
    w.x -= ur.x * uc.x + ur.y * uc.y;
    w.y -= ur.x * uc.y - ur.y * uc.x;
    Wr = w;
  };
  for (int jb = 0; jb < kCW; ++jb) {
    if (q == jb) {  // (uniform) the group's pivots
#pragma unroll
      for (int i = 0; i < kCR; ++i) {
        const int j = kCR * jb + i;
        const cd g = W[i];
        const double d = lane_f64(g.x, j);
        if (!(d > 0.0)) bad = true;
        // (d is uniform; pivots below the normal range, which a Gram of this
        // iteration does not reach, keep the exact operations)
        const double inv = d > 1e-300 ? rsqrt_nr(d) : 1.0 / (d > 0.0 ? sqrt(d) : 1.0);
        const cd gi = make_double2(g.x * inv, g.y * inv);
        const cd u = c == j ? make_double2(inv, 0.0) : gi;
        ub[jb & 1][i][c] = u;
        W[i] = c < j ? gi : make_double2(c == j ? inv : 0.0, 0.0);
#pragma unroll
        for (int i2 = i + 1; i2 < kCR; ++i2) {
          const int r = kCR * jb + i2;
          update(W[i2], make_double2(lane_f64(u.x, r), lane_f64(u.y, r)), u, j);
        }
      }
    }
    __syncthreads();  // the group's u vectors are complete
    if (q > jb) {
#pragma unroll
      for (int i = 0; i < kCR; ++i) {
        const cd* u = ub[jb & 1][i];
        const cd uc = u[c];
#pragma unroll
        for (int i2 = 0; i2 < kCR; ++i2) update(W[i2], u[kCR * q + i2], uc, kCR * jb + i);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kCR; ++i) mt[c][kCR * q + i] = make_double2(W[i].x, -W[i].y);
  __syncthreads();
  for (int o = tid; o < p * p; o += kCW * 64) Bo[o] = mt[o / p][o % p];  // Binv[c][r]
  if (bad && c == 0) fail[b] = 1;
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
                              double* w_dev, double* vec_dev, bool warm, double tol) {
  if (!(tol > 0.0)) tol = kTol;
  const int p = kP;
  if (k + 16 > p || n < 2 * p) return set_error(KLT_ERR_EIG, "psd subspace: shape not covered");
  const size_t np = (size_t)n * p, nk = (size_t)n * k, B = (size_t)count, pp = (size_t)p * p;
  // 16-row tiles; split-K of A M from n alone (identical arithmetic in
  // batched and single calls), about 512 blocks a product
  const int rt = (n + 15) / 16;
  // about 768 blocks a product (measured best between 256 and 1024 at
  // n = 1225); chunks of 64 k, four steps of each wave
  int nsplit = std::max(1, std::min(16, 768 / rt));
  const int kchunk = ((n + nsplit - 1) / nsplit + 63) / 64 * 64;
  nsplit = (n + kchunk - 1) / kchunk;
  int nsg = 1, kchg = n;
  nsg = std::min(16, (n + 15) / 16);  // Gram: 16 shares of the n rows
  kchg = ((n + nsg - 1) / nsg + 15) / 16 * 16;
  nsg = (n + kchg - 1) / kchg;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  const size_t oX = carve(B * np * 16), oY0 = carve(B * np * 16), oY1 = carve(B * np * 16),
               oZ = carve(B * np * 16), oPart = carve(B * nsplit * np * 16),
               oGp = carve(B * nsg * pp * 16), oG = carve(B * pp * 16),
               oBi = carve(B * pp * 16), oU = carve(B * k * p * 16), oTh = carve(B * k * 8),
               oRes = carve(B * k * 8), oFail = carve(B * 4), oAct = carve(B * 4);
  // a warm start needs the block of a previous solve of this order (the
  // workspace layout depends on count and n only, so X is still in place)
  warm = warm && count == 1 && ctx->sub_warm_n == n && ctx->sub_ws.bytes >= off;
  ctx->sub_warm_n = 0;  // (set again only by a solve that converges)
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
    KLT_LAUNCH(ctx, "k_zgemm", (k_zmm<false, false>), dim3(rt, nsplit, count), gblk, 0, Ad, n,
               (size_t)n * n, M, p, np, C, p, np, n, n, p, kchunk, part, act);
    if (nsplit > 1)
      KLT_LAUNCH(ctx, "k_zgemm_reduce", k_zgemm_reduce,
                 dim3((unsigned)((np + 255) / 256), 1, count), gblk, 0, part, nsplit, n, p, C,
                 np, act);
  };
  // C (n x q) = Yin (n x p) Bm, Bm p x q (tb: Bm holds its transpose, q x p)
  auto gemm_small = [&](const cd* Yin, const cd* Bm, bool tb, int q, cd* C) {
    if (tb)
      KLT_LAUNCH(ctx, "k_zgemm_tall", (k_zmm<false, true>), dim3(rt, 1, count), gblk, 0, Yin, p,
                 np, Bm, p, (size_t)q * p, C, q, (size_t)n * q, n, p, q, kP, part, act);
    else
      KLT_LAUNCH(ctx, "k_zgemm_tall", (k_zmm<false, false>), dim3(rt, 1, count), gblk, 0, Yin,
                 p, np, Bm, q, (size_t)p * q, C, q, (size_t)n * q, n, p, q, kP, part, act);
  };
  // G = herm(P^H Q) (p x p)
  auto gram = [&](const cd* P, const cd* Q) {
    KLT_LAUNCH(ctx, "k_gram", (k_zmm<true, false>), dim3(p / 16, nsg, count), gblk, 0, P, p, np,
               Q, p, np, Gp, p, pp, p, n, p, kchg, Gp, act);
    KLT_LAUNCH(ctx, "k_gram_reduce", k_gram_reduce, dim3((unsigned)((pp + 255) / 256), 1, count),
               gblk, 0, Gp, nsg, p, G, true, act);
  };
  // one CholQR pass: Xout = Yin L^{-H}, Yin^H Yin = L L^H (Xout distinct from
  // Yin); two passes (CholQR2) orthonormalise to rounding level
  auto cholqr = [&](const cd* Yin, cd* Xout) {
    gram(Yin, Yin);
    KLT_LAUNCH(ctx, "k_chol_inv", k_chol_inv, dim3(1, 1, count), dim3(kCW * 64), 0, G, p, Bi, fail,
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
  // X: the previous solve's orthonormal block (warm) or the orthonormalised
  // start block
  if (!warm) {
    KLT_LAUNCH(ctx, "k_start_block", k_start_block, dim3((unsigned)((np + 255) / 256), 1, count),
               gblk, 0, Y0, np, n, p);
    cholqr(Y0, Y1);
    cholqr(Y1, X);
  }
  const int first = warm ? kWarmFirstCheck : kFirstCheck;
  for (int round = 1;; ++round) {
    gemm_A(X, Y0);
    gemm_A(Y0, Y1);
    // one pass keeps the span (to rounding relative to the largest column)
    // and a basis close enough to orthonormal for the next products; two
    // before the Rayleigh-Ritz step, which needs an orthonormal X
    const bool check = round >= first && (round - first) % kCheckEvery == 0;
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
      if (mx <= tol * std::fabs(h_th[(size_t)b * k])) {
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
  if (count == 1) ctx->sub_warm_n = n;  // X spans the converged block
  return KLT_OK;
}

}  // namespace klt
