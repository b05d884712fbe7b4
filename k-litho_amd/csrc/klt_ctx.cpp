// klt_ctx.cpp -- context, error state, buffers and twiddle plans of libklt.so.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "klt_internal.h"

namespace klt {

int DevBuf::ensure(size_t need) {
  if (need <= bytes && ptr) return KLT_OK;
  release();
  size_t want = need < 256 ? 256 : need;
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    bytes = 0;
    return set_error(KLT_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(want) +
                                        "): " + hipGetErrorString(e));
  }
  bytes = want;
  return KLT_OK;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

int to_device(klt_ctx* ctx, const void* src, size_t bytes, unsigned flags, DevBuf& stage,
              const void** dptr) {
  if (!(flags & KLT_HOST_PTRS)) {
    *dptr = src;
    return KLT_OK;
  }
  KLT_TRY(stage.ensure(bytes));
  KLT_HIP(hipMemcpyAsync(stage.ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  *dptr = stage.ptr;
  return KLT_OK;
}

// exp(sign * 2 pi i * num / den) with the argument reduced exactly in integers
static inline void cexp_frac(long long num, long long den, int sign, float* re, float* im) {
  long long r = num % den;
  if (r < 0) r += den;
  double a = 2.0 * M_PI * (double)r / (double)den;
  *re = (float)std::cos(a);
  *im = (float)(sign * std::sin(a));
}

int get_socs_plan(klt_ctx* ctx, int Lx, int Ly, int Nx, int Ny, SocsPlan** out) {
  auto key = std::make_tuple(Lx, Ly, Nx, Ny);
  auto it = ctx->socs_plans.find(key);
  if (it != ctx->socs_plans.end()) {
    *out = it->second.get();
    return KLT_OK;
  }
  auto p = std::make_unique<SocsPlan>();
  p->Lx = Lx;
  p->Ly = Ly;
  p->Nx = Nx;
  p->Ny = Ny;
  int need = 4 * (Nx > Ny ? Nx : Ny) + 1;
  int M = 16;
  while (M < need) M *= 2;
  p->M = M;
  std::vector<float> h;
  auto upload = [&](DevBuf& buf, const std::vector<float>& v) -> int {
    KLT_TRY(buf.ensure(v.size() * sizeof(float)));
    KLT_HIP(hipMemcpy(buf.ptr, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice));
    return KLT_OK;
  };
  const int NP = ((std::max(Nx, Ny) + 1 + 7) / 8) * 8;
  const int G = ((2 * Ny + 1 + 7) / 8) * 8;
  p->NP = NP;
  p->G = G;
  int cix = Lx / 2, ciy = Ly / 2, cox = (Lx + 1) / 2, coy = (Ly + 1) / 2;
  h.assign((size_t)Ly * NP * 2, 0.f);
  for (int y = 0; y < Ly; ++y)
    for (int n = 0; n <= Ny; ++n)
      cexp_frac((long long)n * (y - ciy), Ly, -1, &h[2 * ((size_t)y * NP + n)],
                &h[2 * ((size_t)y * NP + n) + 1]);
  KLT_TRY(upload(p->twy_in, h));
  h.assign((size_t)Lx * NP * 2, 0.f);
  for (int x = 0; x < Lx; ++x)
    for (int n = 0; n <= Nx; ++n)
      cexp_frac((long long)n * (x - cix), Lx, -1, &h[2 * ((size_t)x * NP + n)],
                &h[2 * ((size_t)x * NP + n) + 1]);
  KLT_TRY(upload(p->twx_in, h));
  h.assign((size_t)Lx * (2 * Nx + 1) * 2, 0.f);
  for (int x = 0; x < Lx; ++x)
    for (int g = 0; g <= 2 * Nx; ++g)
      cexp_frac((long long)g * (x - cox), Lx, +1, &h[2 * ((size_t)x * (2 * Nx + 1) + g)],
                &h[2 * ((size_t)x * (2 * Nx + 1) + g) + 1]);
  KLT_TRY(upload(p->twx_out, h));
  h.assign((size_t)Ly * G * 2, 0.f);
  for (int y = 0; y < Ly; ++y)
    for (int g = 0; g <= 2 * Ny; ++g)
      cexp_frac((long long)g * (y - coy), Ly, +1, &h[2 * ((size_t)y * G + g)],
                &h[2 * ((size_t)y * G + g) + 1]);
  KLT_TRY(upload(p->twy_out, h));
  h.assign((size_t)M * 2, 0.f);
  for (int j = 0; j < M; ++j) cexp_frac(j, M, +1, &h[2 * j], &h[2 * j + 1]);
  KLT_TRY(upload(p->twM, h));
  *out = p.get();
  ctx->socs_plans[key] = std::move(p);
  return KLT_OK;
}

}  // namespace klt

int klt_ctx::upload(void* dev, const void* host, size_t bytes) {
  constexpr size_t kCap = 4u << 20;
  const size_t need = (bytes + 255) / 256 * 256;
  if (need > kCap) {  // larger than the area: a synchronous copy
    KLT_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, stream));
    KLT_HIP(hipStreamSynchronize(stream));
    return KLT_OK;
  }
  if (!up_base) {
    KLT_HIP(hipHostMalloc(reinterpret_cast<void**>(&up_base), kCap, hipHostMallocDefault));
    up_cap = kCap;
    up_used = 0;
  }
  if (up_used + need > up_cap) {  // reuse from the start once the queued copies are done
    if (up_last) KLT_HIP(hipEventSynchronize(up_last));
    up_used = 0;
  }
  char* piece = up_base + up_used;
  up_used += need;
  std::memcpy(piece, host, bytes);
  KLT_HIP(hipMemcpyAsync(dev, piece, bytes, hipMemcpyHostToDevice, stream));
  if (!up_last) KLT_HIP(hipEventCreateWithFlags(&up_last, hipEventDisableTiming));
  KLT_HIP(hipEventRecord(up_last, stream));
  return KLT_OK;
}

int klt_ctx::download_area(size_t bytes, void** out) {
  if (bytes > dl_cap) {  // (its previous reads have all been waited for)
    if (dl_base) KLT_HIP(hipHostFree(dl_base));
    dl_base = nullptr;
    dl_cap = 0;
    const size_t cap = std::max<size_t>(bytes, 64 << 10);
    KLT_HIP(hipHostMalloc(&dl_base, cap, hipHostMallocDefault));
    dl_cap = cap;
  }
  *out = dl_base;
  return KLT_OK;
}

hipEvent_t klt_ctx::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

int klt_ctx::activate() const {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return klt::set_error(KLT_ERR_HIP, hipGetErrorString(e));
  return KLT_OK;
}

extern "C" {

int klt_version(void) { return 1; }

int klt_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int klt_ctx_create(int device, klt_ctx** out) {
  if (!out) return klt::set_error(KLT_ERR_ARG, "klt_ctx_create: out is NULL");
  *out = nullptr;
  int n = klt_device_count();
  if (n <= 0) return klt::set_error(KLT_ERR_NODEV, "no HIP device visible");
  if (device < 0 || device >= n)
    return klt::set_error(KLT_ERR_ARG, "klt_ctx_create: device out of range");
  auto* c = new (std::nothrow) klt_ctx();
  if (!c) return klt::set_error(KLT_ERR_NOMEM, "klt_ctx_create: out of host memory");
  c->device = device;
  int rc = c->activate();
  if (rc != KLT_OK) {
    delete c;
    return rc;
  }
  hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return klt::set_error(KLT_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  c->stream = c->own_stream;
  rc = klt::noalloc_setup(device);
  if (rc != KLT_OK) {
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return rc;
  }
  *out = c;
  return KLT_OK;
}

void klt_ctx_destroy(klt_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& t : ctx->timed) {
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->up_last) (void)hipEventDestroy(ctx->up_last);
  if (ctx->up_base) (void)hipHostFree(ctx->up_base);
  if (ctx->dl_base) (void)hipHostFree(ctx->dl_base);
  ctx->socs_plans.clear();
  ctx->fused_tables.clear();
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
}

int klt_ctx_set_stream(klt_ctx* ctx, void* stream) {
  if (!ctx) return klt::set_error(KLT_ERR_ARG, "klt_ctx_set_stream: ctx is NULL");
  const hipStream_t next = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
  if (next == ctx->stream) return KLT_OK;
  KLT_TRY(ctx->activate());
  // The context's scratch buffers, batch graphs and pinned upload area are
  // shared by all its calls: work on the new stream must not start before
  // what is queued on the old one (this also keeps the upload area's reuse
  // rule, which waits for the last copy recorded on the newest stream).
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  KLT_HIP(hipStreamIsCapturing(ctx->stream, &cap));
  if (cap == hipStreamCaptureStatusNone) {
    hipEvent_t e = ctx->take_event();
    const hipError_t r = hipEventRecord(e, ctx->stream);
    const hipError_t w = r == hipSuccess ? hipStreamWaitEvent(next, e, 0) : r;
    ctx->event_pool.push_back(e);  // (a wait holds the event's state at the call)
    KLT_HIP(w);
  }
  ctx->stream = next;
  return KLT_OK;
}

void* klt_ctx_stream(klt_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int klt_ctx_sync(klt_ctx* ctx) {
  if (!ctx) return klt::set_error(KLT_ERR_ARG, "klt_ctx_sync: ctx is NULL");
  KLT_TRY(ctx->activate());
  KLT_HIP(hipStreamSynchronize(ctx->stream));
  return KLT_OK;
}

int klt_ctx_wait_stream(klt_ctx* ctx, void* stream) {
  if (!ctx) return klt::set_error(KLT_ERR_ARG, "klt_ctx_wait_stream: ctx is NULL");
  const hipStream_t other = static_cast<hipStream_t>(stream);
  if (other == ctx->stream) return KLT_OK;
  KLT_TRY(ctx->activate());
  hipEvent_t e = ctx->take_event();
  const hipError_t r = hipEventRecord(e, other);
  const hipError_t w = r == hipSuccess ? hipStreamWaitEvent(ctx->stream, e, 0) : r;
  ctx->event_pool.push_back(e);  // (a wait holds the event's state at the call)
  KLT_HIP(w);
  return KLT_OK;
}

int klt_ctx_join_stream(klt_ctx* ctx, void* stream) {
  if (!ctx) return klt::set_error(KLT_ERR_ARG, "klt_ctx_join_stream: ctx is NULL");
  const hipStream_t other = static_cast<hipStream_t>(stream);
  if (other == ctx->stream) return KLT_OK;
  KLT_TRY(ctx->activate());
  hipEvent_t e = ctx->take_event();
  const hipError_t r = hipEventRecord(e, ctx->stream);
  const hipError_t w = r == hipSuccess ? hipStreamWaitEvent(other, e, 0) : r;
  ctx->event_pool.push_back(e);
  KLT_HIP(w);
  return KLT_OK;
}

int klt_ctx_set_profiling(klt_ctx* ctx, int on) {
  if (!ctx) return klt::set_error(KLT_ERR_ARG, "klt_ctx_set_profiling: ctx is NULL");
  ctx->profiling = on != 0;
  return KLT_OK;
}

int klt_ctx_set_autotune(klt_ctx* ctx, int on) {
  if (!ctx) return klt::set_error(KLT_ERR_ARG, "klt_ctx_set_autotune: ctx is NULL");
  ctx->autotune = on != 0;
  return KLT_OK;
}

int klt_ctx_set_splits(klt_ctx* ctx, int ks, int rs, int kg) {
  if (!ctx) return klt::set_error(KLT_ERR_ARG, "klt_ctx_set_splits: ctx is NULL");
  if (ks < 0 || rs < 0 || kg < 0)
    return klt::set_error(KLT_ERR_ARG, "klt_ctx_set_splits: negative split factor");
  ctx->pin_ks = ks;
  ctx->pin_rs = rs;
  ctx->pin_kg = kg;
  return KLT_OK;
}

int klt_ctx_set_band_rows(klt_ctx* ctx, int rows) {
  if (!ctx) return klt::set_error(KLT_ERR_ARG, "klt_ctx_set_band_rows: ctx is NULL");
  if (rows != 0 && rows != 64 && rows != 128)
    return klt::set_error(KLT_ERR_ARG, "klt_ctx_set_band_rows: rows must be 0, 64 or 128");
  ctx->pin_rb = rows;
  return KLT_OK;
}

int klt_ctx_tune_report(klt_ctx* ctx, char* buf, int len) {
  if (!ctx || !buf || len <= 0)
    return klt::set_error(KLT_ERR_ARG, "klt_ctx_tune_report: bad argument");
  std::string out;
  for (auto& kv : ctx->fused_tune) {
    char line[256];
    const auto& k = kv.first;
    snprintf(line, sizeof line, "%s %d %d %d %d %d %d %d %d %d %d %.6f\n",
             k[0] == 0 ? "socs_forward" : k[0] == 1 ? "socs_backward" : "socs_vjp", k[1], k[2],
             k[3], k[4], k[5], k[6],
             kv.second.ks, kv.second.rs, kv.second.kg, kv.second.rb, kv.second.ms);
    out += line;
  }
  if ((int)out.size() + 1 > len)
    return klt::set_error(KLT_ERR_ARG, "klt_ctx_tune_report: buffer too small");
  std::memcpy(buf, out.c_str(), out.size() + 1);
  return (int)out.size();
}

int klt_ctx_profile_report(klt_ctx* ctx, char* buf, int len) {
  if (!ctx || !buf || len <= 0)
    return klt::set_error(KLT_ERR_ARG, "klt_ctx_profile_report: bad argument");
  KLT_TRY(ctx->activate());
  KLT_HIP(hipStreamSynchronize(ctx->stream));
  struct Acc {
    int calls = 0;
    double ms = 0.0;
  };
  std::map<std::string, Acc> acc;
  std::vector<std::string> order;
  for (auto& t : ctx->timed) {
    float ms = 0.f;
    KLT_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
    if (!acc.count(t.name)) order.push_back(t.name);
    acc[t.name].calls += 1;
    acc[t.name].ms += ms;
    ctx->event_pool.push_back(t.start);
    ctx->event_pool.push_back(t.stop);
  }
  ctx->timed.clear();
  std::string out;
  for (auto& n : order) {
    char line[256];
    snprintf(line, sizeof line, "%s %d %.6f\n", n.c_str(), acc[n].calls, acc[n].ms);
    out += line;
  }
  if ((int)out.size() + 1 > len)
    return klt::set_error(KLT_ERR_ARG, "klt_ctx_profile_report: buffer too small");
  std::memcpy(buf, out.c_str(), out.size() + 1);
  return (int)out.size();
}

}  // extern "C"
