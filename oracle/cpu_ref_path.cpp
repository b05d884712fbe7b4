// cpu_ref_path.cpp -- the reference's CPU imaging path, restated in C++ for the
// CPU baseline of bench.py and checked by tests/test_cpu_ref.py.
//
// TEST AND BASELINE INFRASTRUCTURE ONLY: the product (libklt.so, the klitho_*
// tools) never links or calls this file.
//
// The reference (K-Litho-SOCS/klitho_socs.cpp, K-Litho-TCC/klitho_tcc.cpp)
// runs on FFTW3 (double precision, FFTW_ESTIMATE plans) and LAPACK zheevr_.
// This file repeats those calls in the reference's order on the same library
// interfaces, taken at run time from a library that exports them (MKL's
// libmkl_rt.so exports the FFTW3 interface and zheevr_); without one, every
// entry point reports that the library is unavailable rather than
// substituting another implementation.  No FFTW or LAPACK header is used:
// the few prototypes are declared here from the public interfaces.
//
// Per image (klitho_socs.cpp:338-343):
//   FT_r2c  (:71-121)   shift by (L+1)/2, r2c L x L, centre the half spectrum
//                       (R: rows rotated by (Ly+1)/2; L: conjugate mirror),
//                       divide by Lx Ly;
//   calcSOCS (:196-231) per kernel, K_k * M^ placed at (2Ny, 2Nx) of a
//                       (4Ny+1) x (4Nx+1) grid, c2c backward, accumulate
//                       lambda_k |E|^2; shift by floor(size/2);
//   FI      (:159-194)  shift, r2c of the coarse grid / its size, zero-pad the
//                       half spectrum rows into Ly x (Lx/2+1), c2r, shift.
// Eigen (klitho_tcc.cpp:301-377): the transposed TCC into a, zheevr_('V', 'I',
// 'U') for the nk largest with a workspace query first.
#include <dlfcn.h>

#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace {

using cplx = std::complex<double>;
typedef double fftw_cplx[2];
typedef void* fftw_plan_t;
constexpr int kBackward = +1;
constexpr unsigned kEstimate = 1u << 6;

struct Lib {
  void* h = nullptr;
  fftw_plan_t (*plan_r2c_2d)(int, int, double*, fftw_cplx*, unsigned) = nullptr;
  fftw_plan_t (*plan_c2r_2d)(int, int, fftw_cplx*, double*, unsigned) = nullptr;
  fftw_plan_t (*plan_dft_2d)(int, int, fftw_cplx*, fftw_cplx*, int, unsigned) = nullptr;
  void (*execute)(fftw_plan_t) = nullptr;
  void (*destroy)(fftw_plan_t) = nullptr;
  void (*zheevr)(const char*, const char*, const char*, int*, cplx*, int*, double*, double*,
                 int*, int*, double*, int*, double*, cplx*, int*, int*, cplx*, int*, double*,
                 int*, int*, int*, int*) = nullptr;
  int (*set_threads_local)(int) = nullptr;
  void (*set_threads)(int) = nullptr;
  void (*version)(char*, int) = nullptr;
} g;

std::mutex g_plan_mu;  // the FFTW planner is not re-entrant

template <class F>
bool sym(F& f, const char* name) {
  f = reinterpret_cast<F>(dlsym(g.h, name));
  return f != nullptr;
}

fftw_cplx* fc(cplx* p) { return reinterpret_cast<fftw_cplx*>(p); }

// out[(y + yh) % sy][(x + xh) % sx] = in[y][x], with xh = sx / 2 (half) or
// (sx + 1) / 2 (klitho_socs.cpp:55-68)
template <class A, class B>
void shift2(const A* in, B* out, int sx, int sy, bool half) {
  const int xh = half ? sx / 2 : (sx + 1) / 2, yh = half ? sy / 2 : (sy + 1) / 2;
  for (int y = 0; y < sy; ++y) {
    const int oy = (y + yh) % sy;
    for (int x = 0; x < sx; ++x) out[(size_t)oy * sx + (x + xh) % sx] = (B)in[(size_t)y * sx + x];
  }
}

// One worker's buffers and the plans the reference makes once per run.
struct Worker {
  int Lx, Ly, wt;
  std::vector<double> datar, realLxy;
  std::vector<cplx> datac, mskf, R, Lm, cLxy;
  fftw_plan_t fwd = nullptr, c2r = nullptr;
  Worker(int lx, int ly) : Lx(lx), Ly(ly), wt(lx / 2 + 1) {
    datar.assign((size_t)Lx * Ly, 0.0);
    realLxy.assign((size_t)Lx * Ly, 0.0);
    datac.assign((size_t)wt * Ly, 0.0);
    cLxy.assign((size_t)wt * Ly, 0.0);
    mskf.assign((size_t)Lx * Ly, 0.0);
    R.assign((size_t)wt * Ly, 0.0);
    Lm.assign((size_t)wt * Ly, 0.0);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    fwd = g.plan_r2c_2d(Ly, Lx, datar.data(), fc(datac.data()), kEstimate);
    c2r = g.plan_c2r_2d(Ly, Lx, fc(cLxy.data()), realLxy.data(), kEstimate);
  }
  ~Worker() {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    if (fwd) g.destroy(fwd);
    if (c2r) g.destroy(c2r);
  }

  // FT_r2c of one mask into mskf (klitho_socs.cpp:71-121)
  void ft_r2c(const float* m) {
    shift2(m, datar.data(), Lx, Ly, false);
    g.execute(fwd);
    const int AY = (Ly + 1) / 2, BY = Ly - AY, Lxh = Lx / 2;
    for (int y = 0; y < Ly; ++y) {
      const int src = y < BY ? y + AY : y - BY;
      std::copy(&datac[(size_t)src * wt], &datac[(size_t)src * wt] + wt, &R[(size_t)y * wt]);
    }
    for (int y = 0; y < Ly; ++y) {
      const int ym = (Ly % 2 == 0) ? (y == 0 ? 0 : Ly - y) : Ly - 1 - y;
      for (int x = 0; x < wt; ++x) Lm[(size_t)y * wt + x] = std::conj(R[(size_t)ym * wt + wt - 1 - x]);
    }
    const double area = (double)(Lx * Ly);
    for (int y = 0; y < Ly; ++y) {
      cplx* row = &mskf[(size_t)y * Lx];
      for (int x = 0; x < wt; ++x) row[x] = Lm[(size_t)y * wt + x];
      for (int x = Lxh; x < Lx; ++x) row[x] = R[(size_t)y * wt + x - Lxh];
      for (int x = 0; x < Lx; ++x) row[x] /= area;
    }
  }

  // calcSOCS + FI of mskf into image (klitho_socs.cpp:159-231)
  void socs(const std::vector<std::vector<cplx>>& krn, const float* scales, int nk, int Nx,
            int Ny, float* image) {
    const int sx = 4 * Nx + 1, sy = 4 * Ny + 1, difX = 2 * Nx, difY = 2 * Ny;
    const int Lxh = Lx / 2, Lyh = Ly / 2, W = 2 * Nx + 1;
    std::vector<cplx> in((size_t)sx * sy, 0.0), out((size_t)sx * sy, 0.0);
    std::vector<double> rN((size_t)sx * sy), tmp((size_t)sx * sy, 0.0), tmpp((size_t)sx * sy, 0.0);
    const int swt = sx / 2 + 1;
    std::vector<cplx> cN((size_t)swt * sy);
    fftw_plan_t bwd, r2cN;
    {
      std::lock_guard<std::mutex> lk(g_plan_mu);
      bwd = g.plan_dft_2d(sy, sx, fc(in.data()), fc(out.data()), kBackward, kEstimate);
      r2cN = g.plan_r2c_2d(sy, sx, rN.data(), fc(cN.data()), kEstimate);
    }
    for (int k = 0; k < nk; ++k) {
      for (int y = -Ny; y <= Ny; ++y)
        for (int x = -Nx; x <= Nx; ++x)
          in[(size_t)(difY + y + Ny) * sx + difX + x + Nx] =
              krn[k][(size_t)(y + Ny) * W + x + Nx] * mskf[(size_t)(y + Lyh) * Lx + x + Lxh];
      g.execute(bwd);
      for (size_t i = 0; i < tmp.size(); ++i) tmp[i] += scales[k] * std::norm(out[i]);
    }
    shift2(tmp.data(), tmpp.data(), sx, sy, true);
    // FI (:159-194)
    std::fill(rN.begin(), rN.end(), 0.0);
    std::fill(cN.begin(), cN.end(), 0.0);
    shift2(tmpp.data(), rN.data(), sx, sy, false);
    g.execute(r2cN);
    const double areaN = (double)(sx * sy);
    for (auto& v : cN) v /= areaN;
    std::fill(cLxy.begin(), cLxy.end(), 0.0);
    const int dY = Ly - sy;
    for (int y = 0; y < sy; ++y) {
      const int oy = y < sy / 2 + 1 ? y : y + dY;
      for (int x = 0; x < swt; ++x) cLxy[(size_t)oy * wt + x] = cN[(size_t)y * swt + x];
    }
    std::fill(realLxy.begin(), realLxy.end(), 0.0);
    g.execute(c2r);
    shift2(realLxy.data(), image, Lx, Ly, false);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g.destroy(bwd);
    g.destroy(r2cN);
  }
};

}  // namespace

extern "C" {

// Loads the FFTW3 / LAPACK interfaces from `path`; writes the library's
// version string.  0 on success, -1 (with the reason in `msg`) otherwise.
int cr_load(const char* path, char* msg, int len) {
  if (g.h) return 0;
  g.h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!g.h) {
    snprintf(msg, len, "unavailable: %s", dlerror());
    return -1;
  }
  bool ok = sym(g.plan_r2c_2d, "fftw_plan_dft_r2c_2d") && sym(g.plan_c2r_2d, "fftw_plan_dft_c2r_2d") &&
            sym(g.plan_dft_2d, "fftw_plan_dft_2d") && sym(g.execute, "fftw_execute") &&
            sym(g.destroy, "fftw_destroy_plan") && sym(g.zheevr, "zheevr_");
  if (!ok) {
    snprintf(msg, len, "unavailable: %s lacks the FFTW3 / zheevr_ interface", path);
    dlclose(g.h);
    g.h = nullptr;
    return -1;
  }
  sym(g.set_threads_local, "MKL_Set_Num_Threads_Local");
  sym(g.set_threads, "MKL_Set_Num_Threads");
  sym(g.version, "MKL_Get_Version_String");
  if (g.version) g.version(msg, len);
  else snprintf(msg, len, "%s", path);
  return 0;
}

// SOCS images of nmask masks (Ly x Lx each) with `threads` workers, each
// taking a contiguous share of the masks and running the library single
// threaded.  kr, ki: nk x (2Ny+1)(2Nx+1) kernels (krn_i layout).
int cr_socs_images(const float* masks, int nmask, int Lx, int Ly, const float* kr,
                   const float* ki, const float* scales, int nk, int Nx, int Ny, int threads,
                   float* images) {
  if (!g.h) return -1;
  if (4 * Nx + 1 > Lx || 4 * Ny + 1 > Ly) return -2;  // "ERROR in FI()"
  const int T = (2 * Nx + 1) * (2 * Ny + 1);
  std::vector<std::vector<cplx>> krn(nk, std::vector<cplx>(T));
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < T; ++j) krn[k][j] = cplx(kr[(size_t)k * T + j], ki[(size_t)k * T + j]);
  threads = std::max(1, std::min(threads, nmask));
  auto work = [&](int w) {
    if (g.set_threads_local) g.set_threads_local(1);
    Worker wk(Lx, Ly);
    const int per = (nmask + threads - 1) / threads, b0 = w * per, b1 = std::min(nmask, b0 + per);
    for (int b = b0; b < b1; ++b) {
      wk.ft_r2c(masks + (size_t)b * Lx * Ly);
      wk.socs(krn, scales, nk, Nx, Ny, images + (size_t)b * Lx * Ly);
    }
  };
  if (threads == 1) {
    work(0);
  } else {
    std::vector<std::thread> pool;
    for (int w = 0; w < threads; ++w) pool.emplace_back(work, w);
    for (auto& t : pool) t.join();
  }
  return 0;
}

// calcKernels' eigen step (klitho_tcc.cpp:301-377): the k largest eigenpairs
// of the Hermitian T x T matrix tcc (row-major complex) with `threads`
// library threads.  w: k descending; z: k x T row q the eigenvector of w[q].
// Returns zheevr_'s info (0 on success), or -100 when unavailable.
int cr_zheevr_topk(const double* tcc, int n, int k, int threads, double* w, double* z) {
  if (!g.h) return -100;
  if (g.set_threads) g.set_threads(threads);
  if (g.set_threads_local) g.set_threads_local(threads);
  std::vector<cplx> a((size_t)n * n);
  const cplx* t = reinterpret_cast<const cplx*>(tcc);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = t[(size_t)j * n + i];
  int nn = n, lda = n, ldz = n, il = n - (k - 1), iu = n, m = k, info = 0;
  double vl = 0, vu = 0, abstol = -1.0;
  std::vector<double> wv(n);
  std::vector<cplx> zz((size_t)n * k);
  std::vector<int> isuppz(2 * k);
  int lwork = -1, lrwork = -1, liwork = -1, iwk = 0;
  cplx wk;
  double rwk;
  g.zheevr("V", "I", "U", &nn, a.data(), &lda, &vl, &vu, &il, &iu, &abstol, &m, wv.data(),
           zz.data(), &ldz, isuppz.data(), &wk, &lwork, &rwk, &lrwork, &iwk, &liwork, &info);
  if (info != 0) return info;
  lwork = (int)wk.real();
  lrwork = (int)rwk;
  liwork = iwk;
  std::vector<cplx> work(lwork);
  std::vector<double> rwork(lrwork);
  std::vector<int> iwork(liwork);
  g.zheevr("V", "I", "U", &nn, a.data(), &lda, &vl, &vu, &il, &iu, &abstol, &m, wv.data(),
           zz.data(), &ldz, isuppz.data(), work.data(), &lwork, rwork.data(), &lrwork,
           iwork.data(), &liwork, &info);
  if (info != 0) return info;
  for (int q = 0; q < m; ++q) {
    w[q] = wv[m - 1 - q];
    const cplx* col = &zz[(size_t)(m - 1 - q) * n];
    for (int j = 0; j < n; ++j) {
      z[2 * ((size_t)q * n + j)] = col[j].real();
      z[2 * ((size_t)q * n + j) + 1] = col[j].imag();
    }
  }
  return 0;
}

}  // extern "C"
