// Error norms of a device field against an exact one:
//   structure::Operation_<2>::Error1 / Error2 / Errori (s_operation.hpp:869-908)
//   Error1 = sum |e| v / sum v
//   Error2 = sqrt(sum e^2 v) / sum v      (the reference's normalisation)
//   Errori = max |e|
// over the normal cells, v = (1.0 * sx_i) * sy_j (Grid_::volume,
// s_grid.hpp:97-104).  One pass over both fields (16 B per cell) with
// deterministic block partials; the maximum is order-independent, the sums
// agree with the reference's sequential order to rounding.
#include <cmath>

#include "poisson.hpp"

namespace cpgpu {
namespace {

constexpr int kNormBlocks = 256;

__device__ __forceinline__ double dmax(double a, double b) { return b > a ? b : a; }

// partials: [0..G) sum |e| v, [G..2G) sum e^2 v, [2G..3G) sum v, [3G..4G) max |e|
__global__ __launch_bounds__(256) void k_err_part(int nx, int ny,
                                                  const double* __restrict__ exa, int lde,
                                                  const double* __restrict__ val, int ldv,
                                                  const double* __restrict__ sx,
                                                  const double* __restrict__ sy,
                                                  double* __restrict__ part) {
  __shared__ double lds[16];
  double v[3] = {0.0, 0.0, 0.0};
  double mx = 0.0;
  const size_t n = (size_t)nx * ny;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(k / ny);
    const int j = (int)(k - (size_t)i * ny);
    double vol = 1.0;
    vol *= sx[i];
    vol *= sy[j];
    const double err = val[(size_t)i * ldv + j] - exa[(size_t)i * lde + j];
    const double ae = std::fabs(err);
    v[0] += ae * vol;
    v[1] += err * err * vol;
    v[2] += vol;
    if (ae > mx) mx = ae;  // NaN errors are skipped, as by the reference's '>'
  }
  store_partials<double, 3>(v, part, gridDim.x, blockIdx.x, lds);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = dmax(mx, __shfl_xor(mx, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0)
    part[3 * gridDim.x + blockIdx.x] = dmax(dmax(lds[0], lds[1]), dmax(lds[2], lds[3]));
}

__global__ __launch_bounds__(256) void k_err_fin(const double* __restrict__ part, int G,
                                                 double* __restrict__ out) {
  __shared__ double lds[16];
  double t[3];
  reduce_partials<double, 3>(part, G, t, lds);
  double mx = 0.0;
  for (int b = threadIdx.x; b < G; b += blockDim.x) mx = dmax(mx, part[3 * G + b]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = dmax(mx, __shfl_xor(mx, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = t[0] / t[2];
    out[1] = std::sqrt(t[1]) / t[2];
    out[2] = dmax(dmax(lds[0], lds[1]), dmax(lds[2], lds[3]));
  }
}

}  // namespace

int error_norms_device(cpgpu_ctx* ctx, int nx, int ny, const double* exa, int lde,
                       const double* val, int ldv, const double* sx, const double* sy,
                       double* part, size_t part_cap, double* out) {
  if (nx < 1 || ny < 1 || lde < ny || ldv < ny || !out) return CPGPU_EINVAL;
  const size_t n = (size_t)nx * ny;
  size_t G = (n + 255) / 256;
  if (G > (size_t)kNormBlocks) G = kNormBlocks;
  if (part_cap < 4 * G + 3) G = (part_cap - 3) / 4;
  if (G < 1) return set_error(ctx, CPGPU_EINVAL, "error norms: partial buffer too small");
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_err_part, dim3((unsigned)G), dim3(256), 0, st, nx, ny, exa, lde,
                     val, ldv, sx, sy, part);
  hipLaunchKernelGGL(k_err_fin, dim3(1), dim3(256), 0, st, (const double*)part, (int)G,
                     part + 4 * G);
  CPGPU_HIP(ctx, hipGetLastError());
  CPGPU_HIP(ctx, hipMemcpyAsync(out, part + 4 * G, 3 * sizeof(double),
                                hipMemcpyDeviceToHost, st));
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  return CPGPU_OK;
}

}  // namespace cpgpu

using namespace cpgpu;

extern "C" int cpgpu_poisson_error_norms(cpgpu_poisson* P, const double* exact_field,
                                         double* out) {
  if (!P || !exact_field || !out) return CPGPU_EINVAL;
  cpgpu_ctx* ctx = P->ctx;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  const int NY = P->ny + 2 * P->gl;
  const size_t NT = (size_t)(P->nx + 2 * P->gl) * NY;
  if (P->err_ref.alloc(NT) || P->dsxi.alloc(P->nx) || P->dsyi.alloc(P->ny))
    return set_error(ctx, CPGPU_ENOMEM, "error norm buffers");
  hipStream_t st = ctx->stream;
  CPGPU_HIP(ctx, hipMemcpyAsync(P->err_ref.p, exact_field, sizeof(double) * NT,
                                hipMemcpyHostToDevice, st));
  CPGPU_HIP(ctx, hipMemcpyAsync(P->dsxi.p, P->sxi.data(), sizeof(double) * P->nx,
                                hipMemcpyHostToDevice, st));
  CPGPU_HIP(ctx, hipMemcpyAsync(P->dsyi.p, P->syi.data(), sizeof(double) * P->ny,
                                hipMemcpyHostToDevice, st));
  return error_norms_device(ctx, P->nx, P->ny,
                            P->err_ref.p + (size_t)P->gl * NY + P->gl, NY, P->x.p, P->ny,
                            P->dsxi.p, P->dsyi.p, P->part64.p, P->part64.n, out);
}
