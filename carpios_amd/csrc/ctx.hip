// Context, error plumbing and kernel timers of libcpgpu.
#include <cstdarg>

#include "core.hpp"

namespace cpgpu {

int set_error(cpgpu_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

EvPair timer_events(cpgpu_ctx* ctx, int fam) {
  EvPair p;
  if (!ctx || !ctx->timer_on[fam]) return p;
  auto& v = ctx->ev[fam];
  size_t& u = ctx->ev_used[fam];
  while (v.size() < u + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return p;
    v.push_back(e);
  }
  p.a = v[u];
  p.b = v[u + 1];
  u += 2;
  return p;
}

}  // namespace cpgpu

extern "C" {

int cpgpu_create(int device, cpgpu_ctx** out) {
  if (!out) return CPGPU_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return CPGPU_EHIP;
  if (hipSetDevice(device) != hipSuccess) return CPGPU_EHIP;
  cpgpu_ctx* c = new cpgpu_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) !=
      hipSuccess) {
    delete c;
    return CPGPU_EHIP;
  }
  *out = c;
  return CPGPU_OK;
}

int cpgpu_destroy(cpgpu_ctx* ctx) {
  if (!ctx) return CPGPU_EINVAL;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& v : ctx->ev)
    for (auto e : v) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CPGPU_OK;
}

const char* cpgpu_last_error(const cpgpu_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int cpgpu_synchronize(cpgpu_ctx* ctx) {
  if (!ctx) return CPGPU_EINVAL;
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CPGPU_OK;
}

void* cpgpu_stream(cpgpu_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int cpgpu_timer_enable(cpgpu_ctx* ctx, int family, int on) {
  if (!ctx || family < 0 || family >= 4) return CPGPU_EINVAL;
  ctx->timer_on[family] = on != 0;
  ctx->ev_used[family] = 0;
  return CPGPU_OK;
}

int cpgpu_timer_read(cpgpu_ctx* ctx, int family, int* launches, double* ms) {
  if (!ctx || family < 0 || family >= 4 || !launches || !ms)
    return CPGPU_EINVAL;
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double tot = 0.0;
  size_t u = ctx->ev_used[family];
  for (size_t k = 0; k + 1 < u; k += 2) {
    float t = 0.f;
    CPGPU_HIP(ctx, hipEventElapsedTime(&t, ctx->ev[family][k],
                                       ctx->ev[family][k + 1]));
    tot += t;
  }
  *launches = (int)(u / 2);
  *ms = tot;
  ctx->ev_used[family] = 0;
  return CPGPU_OK;
}

}  // extern "C"
