// socs_noalloc.hip -- klt_socs_forward_noalloc: the fused SOCS forward pass
// (the kernels of socs_fused.hip) without any device allocation, for callers
// that cannot allocate (code under graph capture, serving loops, harnesses
// that forbid hipMalloc).
//
// Everything the kernels need beyond the caller's masks, images and kernel
// bank lives in one file-scope device array, `scratch`, of kScratchBytes per
// GPU:
//   [0, kTableBytes)            the twiddle tables of the call's geometry, in
//                               the layouts of FusedTables / SocsPlan, written
//                               on the device by k_noalloc_tables at the start
//                               of every call (so that a captured call
//                               regenerates them on each replay);
//   [kTableBytes, kScratchBytes) the band partials, intensity spectra and
//                               kernel-group partials of one sub-batch.
// Every region a call reads is fully written earlier in the same call, so no
// region needs clearing.  Host code never passes the array's name to a
// kernel: its device address comes from hipGetSymbolAddress, once per device
// with that device current (at klt_ctx_create), and is kept per device.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "klt_internal.h"

namespace {

constexpr size_t kScratchBytes = KLT_NOALLOC_SCRATCH_BYTES;  // per GPU, until the process ends
constexpr size_t kTableBytes = 4u << 20;
constexpr int kMaxSide = KLT_NOALLOC_MAX_SIDE;  // Lx, Ly of a call
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

__device__ __forceinline__ void split_half_d(double v, _Float16* hi, _Float16* lo) {
  const _Float16 h = (_Float16)(float)v;
  *hi = h;
  *lo = (_Float16)(float)(v - (double)(float)h);
}

// cos and sin of 2 pi num / den, the argument reduced exactly in integers
__device__ __forceinline__ void cs_frac(long long num, int den, double* c, double* s) {
  long long r = num % den;
  if (r < 0) r += den;
  const double a = 2.0 * M_PI * (double)r / (double)den;
  sincos(a, s, c);
}

struct TableArgs {
  int Lx, Ly, Nx, Ny, Kpad, Lypad, Lxp, NP;
  _Float16* Wf;   // [Kpad / 32][2][2][64][8]
  float2* twy;    // [Ly][NP]
  _Float16* TYh;  // [Lypad][64]
  _Float16* TYl;
  _Float16* Ef;   // [Lxp / 16][4][2][64][8]
  float2* tw64;   // [64]
};

// Every entry of every table (padding included), one thread per entry:
//   Wf   row twiddles of the mask band, rows n = 2 nx (cos) / 2 nx + 1 (-sin)
//        of e^{-2 pi i nx (x - floor(Lx/2))/Lx}, nx <= Nx, zero past Lx, as hi /
//        lo fp16 in the band kernel's fragment order (lane l of step s, column
//        block c: row 16 c + l % 16, columns 32 s + 8 (l / 16) + t);
//   twy  e^{-2 pi i ny (y - floor(Ly/2))/Ly}, ny <= Ny (zero up to NP);
//   TY   w cos / -w sin of 2 pi gy (y - (Ly+1)/2)/Ly at j = 2 gy / 2 gy + 1,
//        w = 1 for gy = 0 else 2, gy <= 2 Ny, hi / lo fp16;
//   Ef   cos / sin of 2 pi gx (x - (Lx+1)/2)/Lx at k = 2 gxi / 2 gxi + 1,
//        gx = gxi - 2 Nx, gxi < 4 Nx + 1, in the synthesis fragment order;
//   tw64 e^{+2 pi i j/64}.
__global__ __launch_bounds__(256) void k_noalloc_tables(TableArgs t) {
  const long long nWf = (long long)t.Kpad * 32, nTwy = (long long)t.Ly * t.NP;
  const long long nTY = (long long)t.Lypad * 64, nEf = (long long)t.Lxp * 128;
  const long long total = nWf + nTwy + nTY + nEf + 64;
  const int cix = t.Lx / 2, ciy = t.Ly / 2, cox = (t.Lx + 1) / 2, coy = (t.Ly + 1) / 2;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    long long i = e;
    if (i < nWf) {
      const int tt = (int)(i % 8), l = (int)((i / 8) % 64), c = (int)((i / 512) % 2);
      const int st = (int)(i / 1024);
      const int n = 16 * c + l % 16, x = 32 * st + 8 * (l / 16) + tt, f = n / 2;
      double v = 0.0;
      if (f <= t.Nx && x < t.Lx) {
        double cs, sn;
        cs_frac((long long)f * (x - cix), t.Lx, &cs, &sn);
        v = (n & 1) ? -sn : cs;
      }
      const size_t o = (((size_t)st * 2 + c) * 2 * 64 + l) * 8 + tt;
      split_half_d(v, &t.Wf[o], &t.Wf[o + 64 * 8]);
      continue;
    }
    i -= nWf;
    if (i < nTwy) {
      const int y = (int)(i / t.NP), n = (int)(i % t.NP);
      float2 w = make_float2(0.f, 0.f);
      if (n <= t.Ny) {
        double cs, sn;
        cs_frac((long long)n * (y - ciy), t.Ly, &cs, &sn);
        w = make_float2((float)cs, (float)(-sn));
      }
      t.twy[i] = w;
      continue;
    }
    i -= nTwy;
    if (i < nTY) {
      const int y = (int)(i / 64), j = (int)(i % 64), gy = j / 2;
      double v = 0.0;
      if (y < t.Ly && gy <= 2 * t.Ny) {
        double cs, sn;
        cs_frac((long long)gy * (y - coy), t.Ly, &cs, &sn);
        const double w = gy == 0 ? 1.0 : 2.0;
        v = (j & 1) ? -w * sn : w * cs;
      }
      split_half_d(v, &t.TYh[i], &t.TYl[i]);
      continue;
    }
    i -= nTY;
    if (i < nEf) {
      const int tt = (int)(i % 8), ln = (int)((i / 8) % 64), st = (int)((i / 512) % 4);
      const int xt = (int)(i / 2048);
      const int x = 16 * xt + ln % 16, k = 32 * st + 8 * (ln / 16) + tt, gxi = k / 2;
      double v = 0.0;
      if (x < t.Lx && gxi < 4 * t.Nx + 1) {
        double cs, sn;
        cs_frac((long long)(gxi - 2 * t.Nx) * (x - cox), t.Lx, &cs, &sn);
        v = (k & 1) ? sn : cs;
      }
      const size_t o = ((((size_t)xt * 4 + st) * 2) * 64 + ln) * 8 + tt;
      split_half_d(v, &t.Ef[o], &t.Ef[o + 64 * 8]);
      continue;
    }
    i -= nEf;
    double cs, sn;
    cs_frac(i, 64, &cs, &sn);
    t.tw64[i] = make_float2((float)cs, (float)sn);
  }
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }
int round_up_i(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace

namespace klt {

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(KLT_ERR_ARG, "no-allocation scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return KLT_OK;
  void* p = nullptr;
  KLT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return set_error(KLT_ERR_HIP, "hipGetSymbolAddress(scratch) gave NULL");
  g_scratch[device] = p;
  return KLT_OK;
}

}  // namespace klt

extern "C" {

int klt_socs_forward_noalloc(klt_ctx* ctx, const klt_kernels* kb, const float* masks,
                             int nmask, int Lx, int Ly, float* images) {
  if (!ctx || !kb || !masks || !images)
    return klt::set_error(KLT_ERR_ARG, "klt_socs_forward_noalloc: NULL argument");
  if (nmask <= 0 || Lx <= 0 || Ly <= 0)
    return klt::set_error(KLT_ERR_ARG, "klt_socs_forward_noalloc: bad size");
  if (4 * kb->Nx + 1 > Lx || 4 * kb->Ny + 1 > Ly)  // klitho_socs.cpp:164-169
    return klt::set_error(KLT_ERR_SIZE, "ERROR in FI()");
  if (!klt::socs_fused_supported(kb))
    return klt::set_error(KLT_ERR_ARG, "klt_socs_forward_noalloc: kernel half-size above 15");
  if (Lx > kMaxSide || Ly > kMaxSide)
    return klt::set_error(KLT_ERR_ARG, "klt_socs_forward_noalloc: Lx and Ly are limited to " +
                                           std::to_string(kMaxSide) +
                                           " by the fixed scratch area");
  KLT_TRY(ctx->activate());
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch[ctx->device];
  }
  if (!base)
    return klt::set_error(KLT_ERR_HIP,
                          "klt_socs_forward_noalloc: scratch address of this device not set");
  const int Nx = kb->Nx, Ny = kb->Ny;
  // tables
  TableArgs t{};
  t.Lx = Lx;
  t.Ly = Ly;
  t.Nx = Nx;
  t.Ny = Ny;
  t.Kpad = round_up_i(Lx, 128);
  t.Lypad = round_up_i(Ly, 32);
  t.Lxp = round_up_i(Lx, 128);
  t.NP = ((std::max(Nx, Ny) + 1 + 7) / 8) * 8;
  char* p = static_cast<char*>(base);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    char* q = p + off;
    off += align256(bytes);
    return q;
  };
  t.Wf = reinterpret_cast<_Float16*>(carve((size_t)t.Kpad * 64 * sizeof(_Float16)));
  t.twy = reinterpret_cast<float2*>(carve((size_t)Ly * t.NP * sizeof(float2)));
  t.TYh = reinterpret_cast<_Float16*>(carve((size_t)t.Lypad * 64 * sizeof(_Float16)));
  t.TYl = reinterpret_cast<_Float16*>(carve((size_t)t.Lypad * 64 * sizeof(_Float16)));
  t.Ef = reinterpret_cast<_Float16*>(carve((size_t)t.Lxp * 256 * sizeof(_Float16)));
  t.tw64 = reinterpret_cast<float2*>(carve(64 * sizeof(float2)));
  if (off > kTableBytes)  // (cannot happen for sides <= kMaxSide)
    return klt::set_error(KLT_ERR_ARG, "klt_socs_forward_noalloc: tables exceed their area");
  // intermediates of one sub-batch
  const size_t work = kScratchBytes - kTableBytes;
  const size_t per1 = klt::fused_scratch_bytes(Lx, Ly, Nx, Ny, 1, 1);
  const int SB = (int)std::min<size_t>((size_t)nmask, work / per1);
  int ks = 1, rs = 1, kg = 1;
  klt::fused_fixed_splits(SB, Lx, Ly, kb->nk, &ks, &rs, &kg);
  if ((size_t)SB * klt::fused_scratch_bytes(Lx, Ly, Nx, Ny, ks, kg) > work) ks = kg = 1;
  const int nblk = (Ly + 127) / 128;
  off = kTableBytes;
  klt::FusedScratch w{};
  w.qpart = reinterpret_cast<float2*>(
      carve((size_t)SB * nblk * ks * (2 * Ny + 1) * 16 * sizeof(float2)));
  w.spec = reinterpret_cast<float2*>(carve((size_t)SB * 32 * (4 * Nx + 1) * sizeof(float2)));
  if (kg > 1) w.ipart = reinterpret_cast<float*>(carve((size_t)SB * kg * 64 * 64 * sizeof(float)));
  if (off > kScratchBytes)  // (SB is chosen so that this cannot happen)
    return klt::set_error(KLT_ERR_ARG, "klt_socs_forward_noalloc: intermediates exceed the area");
  KLT_LAUNCH(ctx, "k_noalloc_tables", k_noalloc_tables, dim3(512), dim3(256), 0, t);
  KLT_HIP(hipGetLastError());
  const klt::FusedView v{t.Kpad, t.Lypad, t.NP, reinterpret_cast<const float4*>(t.Wf), t.twy,
                         t.TYh, t.TYl, reinterpret_cast<const float4*>(t.Ef), t.tw64};
  for (int b0 = 0; b0 < nmask; b0 += SB) {
    const int nb = std::min(SB, nmask - b0);
    KLT_TRY(klt::socs_fused_launch(ctx, kb, masks + (size_t)b0 * Lx * Ly, nb, Lx, Ly,
                                   images + (size_t)b0 * Lx * Ly, v, w, ks, rs, kg));
  }
  return KLT_OK;
}

}  // extern "C"
