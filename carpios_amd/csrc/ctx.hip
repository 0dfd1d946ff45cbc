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
  if (cpgpu::static_scratch_setup(device, &c->scratch) != CPGPU_OK) {
    (void)hipStreamDestroy(c->stream);
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

namespace {
// y = x, 16 bytes per lane, grid-stride: the reference point of the HBM
// figures (a plain device copy moves 2 bytes per byte copied)
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ x,
                                                uint4* __restrict__ y, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    y[k] = x[k];
}
}  // namespace

int cpgpu_copy_bandwidth(cpgpu_ctx* ctx, size_t bytes, int reps, double* gbs, double* us) {
  if (!ctx || bytes < 16 || reps < 1 || !gbs) return CPGPU_EINVAL;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  CPGPU_HIP(ctx, hipMalloc(&a, n * 16));
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return cpgpu::set_error(ctx, CPGPU_ENOMEM, "copy probe buffers");
  }
  int rc = CPGPU_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  int ncu = 256;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
          hipSuccess ||
      ncu < 1)
    ncu = 256;
  if (hipMemsetAsync(a, 0, n * 16, st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = cpgpu::set_error(ctx, CPGPU_EHIP, "copy probe setup");
  } else {
    // the fastest of four grid sizes (2 .. 32 blocks per CU): the grid size
    // changes a streaming copy's rate by up to 25% on gfx950
    static const int per_cu[4] = {2, 4, 8, 32};
    ms = 1e30f;
    for (int c = 0; c < 4 && rc == CPGPU_OK; ++c) {
      const int G = ncu * per_cu[c];
      float t = 0.f;
      hipLaunchKernelGGL(k_copy16, dim3(G), dim3(256), 0, st, (const uint4*)a, (uint4*)b, n);
      (void)hipEventRecord(e0, st);
      for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy16, dim3(G), dim3(256), 0, st, (const uint4*)a, (uint4*)b, n);
      (void)hipEventRecord(e1, st);
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)
        rc = cpgpu::set_error(ctx, CPGPU_EHIP, "copy probe timing");
      else if (t < ms)
        ms = t;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (rc) return rc;
  const double t = ms * 1e-3 / reps;
  *gbs = 2.0 * (double)(n * 16) / t / 1e9;
  if (us) *us = t * 1e6;
  return CPGPU_OK;
}

int cpgpu_timer_enable(cpgpu_ctx* ctx, int family, int on) {
  if (!ctx || family < 0 || family >= cpgpu_ctx::kFamilies) return CPGPU_EINVAL;
  ctx->timer_on[family] = on != 0;
  ctx->ev_used[family] = 0;
  return CPGPU_OK;
}

int cpgpu_timer_read(cpgpu_ctx* ctx, int family, int* launches, double* ms) {
  if (!ctx || family < 0 || family >= cpgpu_ctx::kFamilies || !launches || !ms)
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
