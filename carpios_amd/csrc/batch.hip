// Batched Jacobi-PCG over many independent Poisson problems of uneven sizes
// (no reference counterpart: the reference solves one problem at a time,
// solver_matrix.hpp:264-374; this batches that algorithm).
//
// The batch is zero-copy over existing cpgpu_poisson problems: a device
// table (BProb) holds each problem's operator and vector pointers, and a
// block map sends every block of a launch to one problem.  Problem p owns
// nblk_p blocks that stride over its TI x TJ tiles (tile t = rank + k nblk_p),
// and its reduction partials sit at fixed slots [pofs_p, pofs_p + nblk_p):
// every block of p reduces exactly those in its prologue, so problems
// converge independently inside one launch, and with one tile per block
// (nblk_p = the single-problem block count) a batched solve is bitwise the
// single-problem one.  When problems finish, the host compacts the block
// map so launches cover only live problems.
//
// Iterations run in hipGraph-captured batches of K (direction, update)
// kernel pairs; the host polls the per-problem states between batches.
// FP64 mode: the reference's J-PCG test every iteration on the true norm.
// Mixed mode: FP32 iterations, FP64 residual replacement and stopping test
// every K iterations (the scheme of the single-problem mixed solver).
// Completion times are stamped on the device (wall clock) per problem.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "poisson.hpp"

namespace cpgpu {

struct BProb {
  int nx, ny, ntj, ntiles;
  int nblk;  // blocks of the problem (fixed for a tile setting)
  int pofs;  // first partial slot (fixed for a tile setting)
  int blk0;  // first block in the current launch layout
  int pad;
  double adj[4];
  float adj32[4];
  const double *KX, *KY;
  const float *KX32, *KY32;
  double *x, *b, *r, *p0, *p1, *q;
  float *x32, *r32, *pa32, *pb32, *q32;
};

}  // namespace cpgpu

using namespace cpgpu;

struct cpgpu_batch {
  cpgpu_ctx* ctx = nullptr;
  int n = 0;
  std::vector<cpgpu_poisson*> P;
  std::vector<BProb> h_probs;
  DevBuf<BProb> probs;
  DevBuf<int> bmap;
  int tpb = 1;     // tiles per block
  int Btot = 0;    // partial slots = blocks of the full layout
  int Bcur = 0;    // blocks of the current layout
  int K = 16;      // iterations per graph batch (even)
  bool use_graphs = true;
  DevBuf<double> save;  // the problems' x during autotuning
  DevBuf<KState> st;
  KState* h_st = nullptr;  // pinned, n entries
  DevBuf<double> partB64, partA64, partR64;
  DevBuf<float> partB32, partA32;
  DevBuf<unsigned long long> tstamp;  // [p] completion, [n] start
  unsigned long long* h_ts = nullptr;  // pinned, n + 1
  double clock_khz = 100000.0;
  // instantiated batch graphs: (mode, Bcur) -> exec
  std::vector<std::pair<std::pair<int, int>, hipGraphExec_t>> graphs;
  bool persistent = false;  // one cooperative launch per batched pass
  DevBuf<unsigned> gbar;    // its grid barrier (generation, abort, per-block flags)
  ~cpgpu_batch() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (h_st) (void)hipHostFree(h_st);
    if (h_ts) (void)hipHostFree(h_ts);
  }
};

namespace cpgpu {
namespace {

template <typename T>
struct BVecs {
  T *x, *r, *pa, *pb, *q;
  const T* b;
};

template <typename T>
__device__ __forceinline__ Sten<T> bsten(const BProb& P);
template <>
__device__ __forceinline__ Sten<double> bsten<double>(const BProb& P) {
  Sten<double> s;
  s.nx = P.nx;
  s.ny = P.ny;
  s.KX = P.KX;
  s.KY = P.KY;
#pragma unroll
  for (int k = 0; k < 4; ++k) s.adj[k] = P.adj[k];
  return s;
}
template <>
__device__ __forceinline__ Sten<float> bsten<float>(const BProb& P) {
  Sten<float> s;
  s.nx = P.nx;
  s.ny = P.ny;
  s.KX = P.KX32;
  s.KY = P.KY32;
#pragma unroll
  for (int k = 0; k < 4; ++k) s.adj[k] = P.adj32[k];
  return s;
}

template <typename T>
__device__ __forceinline__ BVecs<T> bvecs(const BProb& P);
template <>
__device__ __forceinline__ BVecs<double> bvecs<double>(const BProb& P) {
  return BVecs<double>{P.x, P.r, P.p0, P.p1, P.q, P.b};
}
template <>
__device__ __forceinline__ BVecs<float> bvecs<float>(const BProb& P) {
  return BVecs<float>{P.x32, P.r32, P.pa32, P.pb32, P.q32, nullptr};
}

// Totals of the n partials part[k*stride + r], r < n, in the association of
// reduce_partials (so one-problem batches reduce bitwise alike).
template <typename T, int NV>
__device__ __forceinline__ void reduce_prob(const T* __restrict__ part, int stride,
                                            int n, T (&tot)[NV], T* lds) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T s = T(0);
    for (int b = tid; b < n; b += kThreads) s += part[k * stride + b];
    tot[k] = s;
  }
  block_allsum<T, NV>(tot, lds);
}

// Start stamp of a solve.
__global__ void k_b_t0(unsigned long long* ts, int n) {
  if (threadIdx.x == 0) ts[n] = wall_clock64();
}

// r = b - A x and the partials of r.z, r.r, b.b (true norms)
// The kernels' bodies are device functions of the logical block index lb
// (the launch index of a plain launch, the loop index of the persistent
// solve), so both run the same code on the same block map.
template <typename T>
__device__ __forceinline__ void b_init(const BProb* __restrict__ probs, T* __restrict__ partB,
                                       int Btot, int pr, int rank) {
  __shared__ T lds[16];
  const BProb& P = probs[pr];
  const Sten<T> s = bsten<T>(P);
  const BVecs<T> v = bvecs<T>(P);
  T acc[3] = {T(0), T(0), T(0)};
  for (int t = rank; t < P.ntiles; t += P.nblk) {
    int i0, j0;
    tile_origin(t, P.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i >= s.nx || j >= s.ny) continue;
      const size_t idx = (size_t)i * s.ny + j;
      const T dg = diag_at(s, i, j);
      const T bv = v.b[idx];
      const T rv = bv - apply_global(s, v.x, i, j, dg);
      v.r[idx] = rv;
      const T z = rv * (T(1) / dg);
      acc[SLOT_RZ] += rv * z;
      acc[SLOT_RR] += rv * rv;
      acc[SLOT_BB] += bv * bv;
    }
  }
  store_partials<T, 3>(acc, partB + P.pofs, Btot, rank, lds);
}

template <typename T>
__global__ __launch_bounds__(256) void k_b_init(const BProb* __restrict__ probs,
                                                const int* __restrict__ bmap,
                                                T* __restrict__ partB, int Btot) {
  const int p = bmap[blockIdx.x];
  b_init<T>(probs, partB, Btot, p, (int)blockIdx.x - probs[p].blk0);
}

// Direction kernel of the live problem's iteration it = st->it_dir:
// prologue (stopping test of iteration it-1, beta), then p = z + beta p_old
// over an LDS tile with halo, q = A p, partial p.q.  The body is the
// single-problem direction kernel's.
template <typename T>
__device__ __forceinline__ void b_dir(const BProb* __restrict__ probs, KState* __restrict__ sts,
                                      const T* __restrict__ partB, T* __restrict__ partA,
                                      const double* __restrict__ partR64, int Btot, int flip,
                                      unsigned long long* __restrict__ ts, int p, int rank) {
  KState* st = sts + p;
  if (st->done) return;
  const BProb& P = probs[p];
  __shared__ T lds[16];
  __shared__ double lds64[4];
  __shared__ TileLds<T> L;
  const int it = st->it_dir;
  T tot[3];
  reduce_prob<T, 3>(partB + P.pofs, Btot, P.nblk, tot, lds);
  const bool lead = rank == 0 && threadIdx.x == 0 && threadIdx.y == 0;
  bool conv = false, tested = false, replaced = false;
  double resid = 0.0;
  if (st->mode == 0) {
    double normb;
    if (it == 1) {
      normb = std::sqrt((double)tot[SLOT_BB]);
      if (normb == 0.0) normb = 1;
      if (lead) st->normb = normb;
    } else {
      normb = st->normb;
    }
    resid = std::sqrt((double)tot[SLOT_RR]) / normb;
    conv = resid <= st->tol;
    tested = true;
  } else if (it > 1 && (it - 1) % st->refresh_every == 0) {
    replaced = true;
    double rr[1];
    reduce_prob<double, 1>(partR64 + P.pofs, Btot, P.nblk, rr, lds64);
    resid = std::sqrt(rr[0]) / st->normb;
    conv = resid <= st->tol;
    tested = true;
  }
  if (conv || it > st->max_iter) {
    if (lead) {
      st->iter = it - 1;
      if (tested) st->resid = resid;
      st->done = conv ? 1 : 2;
      ts[p] = wall_clock64();
    }
    return;
  }
  if (lead && tested) st->resid = resid;
  const T rho = tot[SLOT_RZ];
  // direction restart after a replacement that shows a stagnated FP32
  // recurrence (the single-problem rule)
  const bool first = it == 1 || (replaced && (double)rho > 2.0 * st->rho[(it - 1) & 1]);
  const T beta = first ? T(0) : rho / (T)st->rho[(it - 1) & 1];
  if (lead) {
    st->rho[it & 1] = (double)rho;
    st->it_upd = it;
  }
  const Sten<T> s = bsten<T>(P);
  const BVecs<T> v = bvecs<T>(P);
  const T* __restrict__ pold = flip ? v.pb : v.pa;
  T* __restrict__ pnew = flip ? v.pa : v.pb;
  T acc[1] = {T(0)};
  for (int t = rank; t < P.ntiles; t += P.nblk) {
    int i0, j0;
    tile_origin(t, P.ntj, i0, j0);
    TileStage<T> b;
    stage_load<T>(s, v.r, pold, i0, j0, b);
    __syncthreads();  // previous tile fully consumed
    stage_store<T, true>(s, b, i0, j0, beta, first, pnew, L);
    __syncthreads();
    tile_apply<T>(s, i0, j0, L, v.q, acc[0]);
  }
  store_partials<T, 1>(acc, partA + P.pofs, Btot, rank, lds);
}

template <typename T>
__global__ __launch_bounds__(256) void k_b_dir(const BProb* __restrict__ probs,
                                               const int* __restrict__ bmap,
                                               KState* __restrict__ sts,
                                               const T* __restrict__ partB,
                                               T* __restrict__ partA,
                                               const double* __restrict__ partR64,
                                               int Btot, int flip,
                                               unsigned long long* __restrict__ ts) {
  const int p = bmap[blockIdx.x];
  b_dir<T>(probs, sts, partB, partA, partR64, Btot, flip, ts, p, (int)blockIdx.x - probs[p].blk0);
}

// Update kernel of iteration it = st->it_upd: alpha = rho / p.q; x += alpha p;
// r -= alpha q; partials of r.z, r.r.
template <typename T>
__device__ __forceinline__ void b_upd(const BProb* __restrict__ probs, KState* __restrict__ sts,
                                      const T* __restrict__ partA, T* __restrict__ partB,
                                      int Btot, int flip, int p, int rank) {
  KState* st = sts + p;
  if (st->done) return;
  const BProb& P = probs[p];
  __shared__ T lds[16];
  const int it = st->it_upd;
  T pq[1];
  reduce_prob<T, 1>(partA + P.pofs, Btot, P.nblk, pq, lds);
  const T alpha = (T)st->rho[it & 1] / pq[0];
  const Sten<T> s = bsten<T>(P);
  const BVecs<T> v = bvecs<T>(P);
  const T* __restrict__ pv = flip ? v.pa : v.pb;
  T acc[3] = {T(0), T(0), T(0)};
  constexpr int NR = TI / BY;
  const int nx = s.nx, ny = s.ny;
  for (int t = rank; t < P.ntiles; t += P.nblk) {
    int i0, j0;
    tile_origin(t, P.ntj, i0, j0);
    // every load of the tile first (clamped, valid addresses), as k_cg_upd
    T xv[NR], pw[NR], rw[NR], qv[NR], kx0[NR], kx1[NR];
    const int jc = min(j0 + (int)threadIdx.x, ny - 1);
    const T ky0 = s.KY[jc], ky1 = s.KY[jc + 1];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int ic = min(i0 + (int)threadIdx.y + BY * k, nx - 1);
      const size_t idx = (size_t)ic * ny + jc;
      xv[k] = v.x[idx];
      pw[k] = pv[idx];
      rw[k] = v.r[idx];
      qv[k] = v.q[idx];
      kx0[k] = s.KX[ic];
      kx1[k] = s.KX[ic + 1];
    }
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i >= nx || j >= ny) continue;
      const size_t idx = (size_t)i * ny + j;
      v.x[idx] = xv[k] + alpha * pw[k];
      const T rv = rw[k] - alpha * qv[k];
      v.r[idx] = rv;
      const T z = rv * (T(1) / diag_from(s, i, j, kx0[k], kx1[k], ky0, ky1));
      acc[SLOT_RZ] += rv * z;
      acc[SLOT_RR] += rv * rv;
    }
  }
  store_partials<T, 3>(acc, partB + P.pofs, Btot, rank, lds);
  if (rank == 0 && threadIdx.x == 0 && threadIdx.y == 0) st->it_dir = it + 1;
}

template <typename T>
__global__ __launch_bounds__(256) void k_b_upd(const BProb* __restrict__ probs,
                                               const int* __restrict__ bmap,
                                               KState* __restrict__ sts,
                                               const T* __restrict__ partA,
                                               T* __restrict__ partB, int Btot, int flip) {
  const int p = bmap[blockIdx.x];
  b_upd<T>(probs, sts, partA, partB, Btot, flip, p, (int)blockIdx.x - probs[p].blk0);
}

// x64 += x32; x32 = 0 on live problems (every problem when all != 0)
__device__ __forceinline__ void b_merge(const BProb* __restrict__ probs,
                                        const KState* __restrict__ sts, int all, int p,
                                        int rank) {
  if (!all && sts[p].done) return;
  const BProb& P = probs[p];
  for (int t = rank; t < P.ntiles; t += P.nblk) {
    int i0, j0;
    tile_origin(t, P.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i >= P.nx || j >= P.ny) continue;
      const size_t idx = (size_t)i * P.ny + j;
      P.x[idx] = P.x[idx] + (double)P.x32[idx];
      P.x32[idx] = 0.f;
    }
  }
}

__global__ __launch_bounds__(256) void k_b_merge(const BProb* __restrict__ probs,
                                                 const int* __restrict__ bmap,
                                                 const KState* __restrict__ sts, int all) {
  const int p = bmap[blockIdx.x];
  b_merge(probs, sts, all, p, (int)blockIdx.x - probs[p].blk0);
}

// r32 = (float)(b - A x64) with the FP64 operator on live problems;
// partials of the FP32 r.z, r.r and of the FP64 r.r.
__device__ __forceinline__ void b_refresh(const BProb* __restrict__ probs,
                                          const KState* __restrict__ sts,
                                          float* __restrict__ partB,
                                          double* __restrict__ partR64, int Btot, int p,
                                          int rank) {
  if (sts[p].done) return;
  const BProb& P = probs[p];
  __shared__ float lds[16];
  __shared__ double lds64[4];
  const Sten<double> s64 = bsten<double>(P);
  const Sten<float> s32 = bsten<float>(P);
  float v[3] = {0.f, 0.f, 0.f};
  double w[1] = {0.0};
  for (int t = rank; t < P.ntiles; t += P.nblk) {
    int i0, j0;
    tile_origin(t, P.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
    constexpr int NR = TI / BY;
    // every load of the rows first (clamped addresses)
    Nbr<double> xn[NR], kc[NR];
    Nbr<float> kf[NR];
    double bv[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      load_nbr(s64, (const double*)P.x, i, j, xn[k], kc[k]);
      const int ic = min(i, P.nx - 1), jc = min(j, P.ny - 1);
      bv[k] = P.b[(size_t)ic * P.ny + jc];
      kf[k].kx0 = s32.KX[ic];
      kf[k].kx1 = s32.KX[ic + 1];
      kf[k].ky0 = s32.KY[jc];
      kf[k].ky1 = s32.KY[jc + 1];
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i >= P.nx || j >= P.ny) continue;
      const size_t idx = (size_t)i * P.ny + j;
      const double rv = bv[k] - apply_regs_diff(s64, i, j, xn[k], kc[k]);
      w[0] += rv * rv;
      const float rf = (float)rv;
      P.r32[idx] = rf;
      const float z = rf * (1.f / diag_from(s32, i, j, kf[k].kx0, kf[k].kx1, kf[k].ky0, kf[k].ky1));
      v[SLOT_RZ] += rf * z;
      v[SLOT_RR] += rf * rf;
    }
  }
  store_partials<float, 3>(v, partB + P.pofs, Btot, rank, lds);
  store_partials<double, 1>(w, partR64 + P.pofs, Btot, rank, lds64);
}

__global__ __launch_bounds__(256) void k_b_refresh(const BProb* __restrict__ probs,
                                                   const int* __restrict__ bmap,
                                                   const KState* __restrict__ sts,
                                                   float* __restrict__ partB,
                                                   double* __restrict__ partR64, int Btot) {
  const int p = bmap[blockIdx.x];
  b_refresh(probs, sts, partB, partR64, Btot, p, (int)blockIdx.x - probs[p].blk0);
}

// partials of b.b (FP64) for the mixed mode's ||b||
__device__ __forceinline__ void b_bb(const BProb* __restrict__ probs,
                                     double* __restrict__ partR64, int Btot, int pr, int rank) {
  __shared__ double lds64[4];
  const BProb& P = probs[pr];
  double w[1] = {0.0};
  for (int t = rank; t < P.ntiles; t += P.nblk) {
    int i0, j0;
    tile_origin(t, P.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i >= P.nx || j >= P.ny) continue;
      const double bv = P.b[(size_t)i * P.ny + j];
      w[0] += bv * bv;
    }
  }
  store_partials<double, 1>(w, partR64 + P.pofs, Btot, rank, lds64);
}

__global__ __launch_bounds__(256) void k_b_bb(const BProb* __restrict__ probs,
                                              const int* __restrict__ bmap,
                                              double* __restrict__ partR64, int Btot) {
  const int p = bmap[blockIdx.x];
  b_bb(probs, partR64, Btot, p, (int)blockIdx.x - probs[p].blk0);
}

// st[p].normb = ||b|| (1 when b = 0) of problem p
__device__ __forceinline__ void b_setnorm(const BProb* __restrict__ probs,
                                          const double* __restrict__ partR64, int Btot,
                                          KState* __restrict__ sts, int p) {
  __shared__ double lds64[4];
  const BProb& P = probs[p];
  double bb[1];
  reduce_prob<double, 1>(partR64 + P.pofs, Btot, P.nblk, bb, lds64);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const double nb = std::sqrt(bb[0]);
    sts[p].normb = nb == 0.0 ? 1.0 : nb;
  }
}

// one block per problem
__global__ __launch_bounds__(256) void k_b_setnorm(const BProb* __restrict__ probs,
                                                   const double* __restrict__ partR64,
                                                   int Btot, KState* __restrict__ sts) {
  b_setnorm(probs, partR64, Btot, sts, blockIdx.x);
}

// x32 = 0 on every problem
__device__ __forceinline__ void b_zero32(const BProb* __restrict__ probs, int pr, int rank) {
  const BProb& P = probs[pr];
  for (int t = rank; t < P.ntiles; t += P.nblk) {
    int i0, j0;
    tile_origin(t, P.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < P.nx && j < P.ny) P.x32[(size_t)i * P.ny + j] = 0.f;
    }
  }
}

__global__ __launch_bounds__(256) void k_b_zero32(const BProb* __restrict__ probs,
                                                  const int* __restrict__ bmap) {
  const int p = bmap[blockIdx.x];
  b_zero32(probs, p, (int)blockIdx.x - probs[p].blk0);
}

const dim3 kBlock(TJ, BY);

// ---------------------------------------------------------------------
// Persistent batched solve: a whole batched pass (set-up, every iteration,
// the mixed mode's merges and replacements, the final merge) in one
// cooperative launch of G co-resident blocks.  The phases of the plain
// launches become loops of the same device functions over the logical
// blocks of the same block map (logical block lb on physical block
// lb mod G), separated by a grid barrier; partial slots, tile assignment and
// arithmetic are those of the plain launches, so the results are the same
// bits.  After every K iterations (and the mixed mode's replacement) the
// blocks read the problems' stop flags, written before the last barrier, and
// all leave together once every member has stopped.

// Grid barrier over the G co-resident blocks of a cooperative launch, with
// one arrival flag per block (a gather by block 0, no contended atomics:
// device-scope read-modify-writes on one counter serialise, each a memory
// round trip).  Barrier number k (counted from 1) is complete when every
// flag holds k: thread 0 of every block stores k into its flag with a
// release (its block's writes, ordered before it by the block barrier,
// reach device scope); block 0's threads wait for all the flags, then its
// thread 0 stores k into the generation word with a release; the other
// blocks wait for the generation and acquire.  Flags and generation only
// grow, so nothing is ever reset.  Waits poll relaxed loads with sleeps,
// every 32nd one an acquire load: a relaxed device-scope load may be
// served by the XCD's own L2, which the other XCDs' writes reach only
// through an acquire's invalidation, and invalidating on every poll slows
// the blocks still at work.  A wait longer than `ticks` of the constant wall
// clock raises the abort flag, which every waiting thread polls too, so no
// block can wait for ever: all of them leave and the kernel ends (the host
// then reports an error).
constexpr int kBarStride = 32;                 // unsigned per slot (128 bytes)
constexpr int kBarGen = 0, kBarAbort = kBarStride;
constexpr int kBarFlags = 2 * kBarStride;      // flag of block b at kBarFlags + b * kBarStride
inline size_t bar_words(int G) { return (size_t)kBarFlags + (size_t)G * kBarStride; }

// waits until *w != `until_not` (returns true) or the wait is aborted
__device__ __forceinline__ bool bar_wait(unsigned* w, unsigned until_not, unsigned* abort_w,
                                         unsigned long long t0, unsigned long long ticks) {
  for (unsigned poll = 1;; ++poll) {
    const unsigned v = (poll & 31u) == 0u
                           ? __hip_atomic_load(w, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)
                           : __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v != until_not) return true;
    if (__hip_atomic_load(abort_w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
        wall_clock64() - t0 > ticks) {
      __hip_atomic_store(abort_w, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

__device__ bool grid_sync(unsigned* gb, unsigned G, unsigned& gen,
                          unsigned long long ticks) {
  __shared__ int ok_s;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const unsigned k = gen + 1;
  __syncthreads();
  if (tid == 0) {
    ok_s = 1;
    __hip_atomic_store(&gb[kBarFlags + blockIdx.x * kBarStride], k, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const unsigned long long t0 = wall_clock64();
  if (blockIdx.x == 0) {
    bool ok = true;
    for (unsigned b = tid; b < G && ok; b += kThreads)
      ok = bar_wait(&gb[kBarFlags + b * kBarStride], k - 1u, &gb[kBarAbort], t0, ticks);
    if (!ok) ok_s = 0;
    __syncthreads();
    if (tid == 0 && ok_s) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __hip_atomic_store(&gb[kBarGen], k, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  } else if (tid == 0) {
    if (!bar_wait(&gb[kBarGen], k - 1u, &gb[kBarAbort], t0, ticks)) ok_s = 0;
  }
  if (tid == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  gen = k;
  __syncthreads();
  return ok_s != 0;
}

// The live problems' logical blocks: problem p (stop flag 0) holds its
// nblk_p consecutive indices from pre[p] (pre has n + 1 entries, in LDS),
// and the count is returned.  Every block builds the same table from the
// same stop flags (read after a grid barrier), so a pass covers only the
// live problems, as the host's compaction of the launched batches does; the
// ranks and partial slots are the fixed ones of the layout.
__device__ int live_table(const BProb* __restrict__ probs, const KState* __restrict__ sts,
                          int n, int* pre) {
  __shared__ int scan[kThreads];
  __shared__ int carry;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += kThreads) {
    const int p = base + tid;
    const int v = p < n && sts[p].done == 0 ? probs[p].nblk : 0;
    scan[tid] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {  // inclusive scan
      const int add = tid >= off ? scan[tid - off] : 0;
      __syncthreads();
      scan[tid] += add;
      __syncthreads();
    }
    if (p < n) pre[p] = carry + scan[tid] - v;
    __syncthreads();
    if (tid == 0) carry += scan[kThreads - 1];
    __syncthreads();
  }
  if (tid == 0) pre[n] = carry;
  __syncthreads();
  return pre[n];
}

// the problem of live logical block lb: pre[p] <= lb < pre[p + 1]
__device__ __forceinline__ int live_problem(const int* pre, int n, int lb) {
  int lo = 0, hi = n;  // pre[lo] <= lb < pre[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= lb) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <int MODE>  // 0 FP64 Jacobi-PCG, 1 FP32 iterations with FP64 replacement
__global__ __launch_bounds__(256) void k_b_persist(
    const BProb* __restrict__ probs, KState* __restrict__ sts, int n,
    double* __restrict__ partB64, double* __restrict__ partA64, float* __restrict__ partB32,
    float* __restrict__ partA32, double* __restrict__ partR64, int Btot,
    unsigned long long* __restrict__ ts, unsigned* gb, int K, long long rounds,
    unsigned long long ticks) {
  extern __shared__ int pre[];  // n + 1 live-table entries
  const unsigned G = gridDim.x;
  unsigned gen = 0;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  if (blockIdx.x == 0 && tid == 0) ts[n] = wall_clock64();
  int Blive = live_table(probs, sts, n, pre);
#define FOR_LB                                                  \
  for (int lb = blockIdx.x, p = 0, rank = 0; lb < Blive &&      \
                                             (p = live_problem(pre, n, lb), rank = lb - pre[p], true); \
       lb += G)
#define SYNC \
  if (!grid_sync(gb, G, gen, ticks)) return
  if constexpr (MODE == 0) {
    FOR_LB b_init<double>(probs, partB64, Btot, p, rank);
  } else {
    FOR_LB {
      b_zero32(probs, p, rank);
      b_bb(probs, partR64, Btot, p, rank);
    }
    SYNC;
    for (int q = blockIdx.x; q < n; q += G) b_setnorm(probs, partR64, Btot, sts, q);
    SYNC;
    FOR_LB b_refresh(probs, sts, partB32, partR64, Btot, p, rank);
  }
  SYNC;
  for (long long round = 0; round < rounds; ++round) {
    for (int k = 0; k < K; ++k) {
      const int flip = k & 1;
      if constexpr (MODE == 0) {
        FOR_LB b_dir<double>(probs, sts, partB64, partA64, partR64, Btot, flip, ts, p, rank);
        SYNC;
        FOR_LB b_upd<double>(probs, sts, partA64, partB64, Btot, flip, p, rank);
      } else {
        FOR_LB b_dir<float>(probs, sts, partB32, partA32, partR64, Btot, flip, ts, p, rank);
        SYNC;
        FOR_LB b_upd<float>(probs, sts, partA32, partB32, Btot, flip, p, rank);
      }
      SYNC;
    }
    if constexpr (MODE == 1) {
      FOR_LB b_merge(probs, sts, 0, p, rank);
      SYNC;
      FOR_LB b_refresh(probs, sts, partB32, partR64, Btot, p, rank);
      SYNC;
    }
    // the stop flags were written before the last barrier: every block
    // builds the same table and takes the same decision
    Blive = live_table(probs, sts, n, pre);
    if (Blive == 0) break;
  }
#undef FOR_LB
#undef SYNC
}

// Block layout for `tpb` tiles per block: fixed partial slots per problem
// and, for the launch, the live problems' blocks (all when live is empty).
void build_layout(cpgpu_batch* B, const std::vector<char>* live, std::vector<int>& bmap) {
  bmap.clear();
  int b = 0;
  for (int p = 0; p < B->n; ++p) {
    BProb& q = B->h_probs[p];
    if (live && !(*live)[p]) {
      q.blk0 = -1;
      continue;
    }
    q.blk0 = b;
    for (int k = 0; k < q.nblk; ++k) bmap.push_back(p);
    b += q.nblk;
  }
}

int set_tiles(cpgpu_batch* B, int tpb) {
  cpgpu_ctx* ctx = B->ctx;
  if (tpb < 1 || tpb > 64) return set_error(ctx, CPGPU_EINVAL, "tiles per block out of range");
  B->tpb = tpb;
  int ofs = 0;
  for (int p = 0; p < B->n; ++p) {
    BProb& q = B->h_probs[p];
    const int nb = (q.ntiles + tpb - 1) / tpb;
    q.nblk = std::max(1, std::min(kMaxBlocks, nb));
    q.pofs = ofs;
    ofs += q.nblk;
  }
  B->Btot = ofs;
  if (B->partB64.alloc(3 * (size_t)ofs) || B->partA64.alloc(ofs) || B->partR64.alloc(ofs) ||
      B->partB32.alloc(3 * (size_t)ofs) || B->partA32.alloc(ofs) || B->bmap.alloc(ofs))
    return set_error(ctx, CPGPU_ENOMEM, "batch partials");
  for (auto& g : B->graphs) (void)hipGraphExecDestroy(g.second);
  B->graphs.clear();
  return CPGPU_OK;
}

// Uploads the problem table and the block map of the given layout.
int upload_layout(cpgpu_batch* B, const std::vector<int>& bmap) {
  cpgpu_ctx* ctx = B->ctx;
  hipStream_t st = ctx->stream;
  B->Bcur = (int)bmap.size();
  CPGPU_HIP(ctx, hipMemcpyAsync(B->probs.p, B->h_probs.data(), sizeof(BProb) * B->n,
                                hipMemcpyHostToDevice, st));
  if (B->Bcur > 0)
    CPGPU_HIP(ctx, hipMemcpyAsync(B->bmap.p, bmap.data(), sizeof(int) * bmap.size(),
                                  hipMemcpyHostToDevice, st));
  CPGPU_HIP(ctx, hipStreamSynchronize(st));  // host vectors are reused
  return CPGPU_OK;
}

// K iterations of every live problem (mode 0 FP64, 1 mixed), then in mixed
// mode the merge and the FP64 residual replacement.
void launch_iters(cpgpu_batch* B, hipStream_t s, int mode) {
  const dim3 g(B->Bcur);
  for (int k = 0; k < B->K; ++k) {
    const int flip = k & 1;
    if (mode == 0) {
      hipLaunchKernelGGL(k_b_dir<double>, g, kBlock, 0, s, B->probs.p, B->bmap.p, B->st.p,
                         (const double*)B->partB64.p, B->partA64.p,
                         (const double*)B->partR64.p, B->Btot, flip, B->tstamp.p);
      hipLaunchKernelGGL(k_b_upd<double>, g, kBlock, 0, s, B->probs.p, B->bmap.p, B->st.p,
                         (const double*)B->partA64.p, B->partB64.p, B->Btot, flip);
    } else {
      hipLaunchKernelGGL(k_b_dir<float>, g, kBlock, 0, s, B->probs.p, B->bmap.p, B->st.p,
                         (const float*)B->partB32.p, B->partA32.p,
                         (const double*)B->partR64.p, B->Btot, flip, B->tstamp.p);
      hipLaunchKernelGGL(k_b_upd<float>, g, kBlock, 0, s, B->probs.p, B->bmap.p, B->st.p,
                         (const float*)B->partA32.p, B->partB32.p, B->Btot, flip);
    }
  }
  if (mode == 1) {
    hipLaunchKernelGGL(k_b_merge, g, kBlock, 0, s, B->probs.p, B->bmap.p,
                       (const KState*)B->st.p, 0);
    hipLaunchKernelGGL(k_b_refresh, g, kBlock, 0, s, B->probs.p, B->bmap.p,
                       (const KState*)B->st.p, B->partB32.p, B->partR64.p, B->Btot);
  }
}

hipGraphExec_t iter_graph(cpgpu_batch* B, int mode) {
  const std::pair<int, int> key{mode, B->Bcur};
  for (auto& g : B->graphs)
    if (g.first == key) return g.second;
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    launch_iters(B, cs, mode);
    if (hipStreamEndCapture(cs, &graph) == hipSuccess && graph) {
      if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
        exec = nullptr;
      (void)hipGraphDestroy(graph);
    }
  }
  (void)hipStreamDestroy(cs);
  if (exec) B->graphs.push_back({key, exec});
  return exec;
}

}  // namespace
}  // namespace cpgpu

extern "C" {

int cpgpu_batch_create(cpgpu_ctx* ctx, int n, cpgpu_poisson* const* problems,
                       cpgpu_batch** out) {
  if (!ctx || n < 1 || !problems || !out) return CPGPU_EINVAL;
  *out = nullptr;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<cpgpu_batch> B(new cpgpu_batch());
  B->ctx = ctx;
  B->n = n;
  B->P.assign(problems, problems + n);
  B->h_probs.resize(n);
  for (int p = 0; p < n; ++p) {
    cpgpu_poisson* Q = problems[p];
    if (!Q || Q->ctx != ctx) return set_error(ctx, CPGPU_EINVAL, "problem %d: not on this context", p);
    if (Q->general)
      return set_error(ctx, CPGPU_EUNSUPPORTED, "problem %d: cell sizes vary (single solves only)", p);
    int neumann = 0;
    for (int k = 0; k < 4; ++k) neumann += Q->seg[k].type == CPGPU_BC_NEUMANN;
    if (neumann == 4)
      return set_error(ctx, CPGPU_EINVAL, "problem %d: all-Neumann (singular) problems are not batched", p);
    int rc = poisson_prepare_f32(Q);
    if (rc) return rc;
    BProb& q = B->h_probs[p];
    std::memset(&q, 0, sizeof(q));
    q.nx = Q->nx;
    q.ny = Q->ny;
    q.ntj = Q->tl.ntj;
    q.ntiles = Q->tl.ntiles;
    for (int k = 0; k < 4; ++k) {
      q.adj[k] = Q->adj[k];
      q.adj32[k] = (float)Q->adj[k];
    }
    q.KX = Q->dKX.p;
    q.KY = Q->dKY.p;
    q.KX32 = Q->fKX.p;
    q.KY32 = Q->fKY.p;
    q.x = Q->x.p;
    q.b = Q->b.p;
    q.r = Q->r.p;
    q.p0 = Q->p0.p;
    q.p1 = Q->p1.p;
    q.q = Q->q.p;
    q.x32 = Q->x32.p;
    q.r32 = Q->r32.p;
    q.pa32 = Q->pa32.p;
    q.pb32 = Q->pb32.p;
    q.q32 = Q->q32.p;
  }
  if (B->probs.alloc(n) || B->st.alloc(n) || B->tstamp.alloc(n + 1))
    return set_error(ctx, CPGPU_ENOMEM, "batch tables");
  CPGPU_HIP(ctx, hipHostMalloc((void**)&B->h_st, sizeof(KState) * n, 0));
  CPGPU_HIP(ctx, hipHostMalloc((void**)&B->h_ts, sizeof(unsigned long long) * (n + 1), 0));
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) == hipSuccess &&
      khz > 0)
    B->clock_khz = khz;
  int rc = set_tiles(B.get(), 1);
  if (rc) return rc;
  *out = B.release();
  return CPGPU_OK;
}

int cpgpu_batch_destroy(cpgpu_batch* B) {
  if (!B) return CPGPU_EINVAL;
  (void)hipSetDevice(B->ctx->device);
  (void)hipStreamSynchronize(B->ctx->stream);
  delete B;
  return CPGPU_OK;
}

int cpgpu_batch_set_tiles(cpgpu_batch* B, int tiles_per_block) {
  if (!B) return CPGPU_EINVAL;
  return set_tiles(B, tiles_per_block);
}

int cpgpu_batch_set_graphs(cpgpu_batch* B, int use_graphs) {
  if (!B) return CPGPU_EINVAL;
  B->use_graphs = use_graphs != 0;
  return CPGPU_OK;
}

// The batched iteration over the members (mode 0 FP64, 1 mixed); others
// are marked stopped (done = 3) and left alone.  States and stamps end in
// B->h_st and B->h_ts.
static int run_batched(cpgpu_batch* B, int mode, int max_iter, double tol,
                       const std::vector<char>& member) {
  cpgpu_ctx* ctx = B->ctx;
  hipStream_t st = ctx->stream;
  const int n = B->n;
  for (int p = 0; p < n; ++p) {
    KState& k = B->h_st[p];
    std::memset(&k, 0, sizeof(k));
    k.tol = tol;
    k.max_iter = max_iter;
    k.it_dir = 1;
    k.mode = mode;
    k.refresh_every = B->K;
    k.normb = 1.0;
    if (!member[p]) k.done = 3;
  }
  CPGPU_HIP(ctx, hipMemcpyAsync(B->st.p, B->h_st, sizeof(KState) * n, hipMemcpyHostToDevice, st));
  std::vector<char> live(member);
  std::vector<int> bmap;
  build_layout(B, &live, bmap);
  int rc = upload_layout(B, bmap);
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_t0, dim3(1), dim3(64), 0, st, B->tstamp.p, n);
  if (B->Bcur == 0) {
    CPGPU_HIP(ctx, hipStreamSynchronize(st));
    return CPGPU_OK;
  }
  const dim3 g(B->Bcur);
  if (mode == 0) {
    hipLaunchKernelGGL(k_b_init<double>, g, kBlock, 0, st, B->probs.p, B->bmap.p,
                       B->partB64.p, B->Btot);
  } else {
    hipLaunchKernelGGL(k_b_zero32, g, kBlock, 0, st, B->probs.p, B->bmap.p);
    hipLaunchKernelGGL(k_b_bb, g, kBlock, 0, st, B->probs.p, B->bmap.p, B->partR64.p, B->Btot);
    hipLaunchKernelGGL(k_b_setnorm, dim3(n), kBlock, 0, st, B->probs.p,
                       (const double*)B->partR64.p, B->Btot, B->st.p);
    hipLaunchKernelGGL(k_b_refresh, g, kBlock, 0, st, B->probs.p, B->bmap.p,
                       (const KState*)B->st.p, B->partB32.p, B->partR64.p, B->Btot);
  }
  CPGPU_HIP(ctx, hipGetLastError());
  // batches until every member stopped (the kernels stop a problem at
  // max_iter + 1 by themselves; one batch past that bounds the loop)
  const long long nbatch_max = ((long long)max_iter + 2) / B->K + 2;
  int nlive = 0;
  for (int p = 0; p < n; ++p) nlive += live[p] ? 1 : 0;
  for (long long nb = 0; nb < nbatch_max && nlive > 0; ++nb) {
    hipGraphExec_t ge = B->use_graphs ? iter_graph(B, mode) : nullptr;
    if (ge) {
      CPGPU_HIP(ctx, hipGraphLaunch(ge, st));
    } else {
      launch_iters(B, st, mode);
      CPGPU_HIP(ctx, hipGetLastError());
    }
    CPGPU_HIP(ctx, hipMemcpyAsync(B->h_st, B->st.p, sizeof(KState) * n, hipMemcpyDeviceToHost, st));
    CPGPU_HIP(ctx, hipStreamSynchronize(st));
    int now = 0, blocks_live = 0;
    for (int p = 0; p < n; ++p) {
      if (B->h_st[p].done) live[p] = 0;
      if (live[p]) {
        ++now;
        blocks_live += B->h_probs[p].nblk;
      }
    }
    nlive = now;
    // compact the launch once the live problems hold under half its blocks
    if (nlive > 0 && 2 * blocks_live < B->Bcur) {
      build_layout(B, &live, bmap);
      if ((rc = upload_layout(B, bmap))) return rc;
    }
  }
  if (mode == 1) {  // fold the increments of members stopped at max_iter
    build_layout(B, &member, bmap);
    if ((rc = upload_layout(B, bmap))) return rc;
    hipLaunchKernelGGL(k_b_merge, dim3(B->Bcur), kBlock, 0, st, B->probs.p, B->bmap.p,
                       (const KState*)B->st.p, 1);
    CPGPU_HIP(ctx, hipGetLastError());
  }
  CPGPU_HIP(ctx, hipMemcpyAsync(B->h_st, B->st.p, sizeof(KState) * n, hipMemcpyDeviceToHost, st));
  CPGPU_HIP(ctx, hipMemcpyAsync(B->h_ts, B->tstamp.p, sizeof(unsigned long long) * (n + 1),
                                hipMemcpyDeviceToHost, st));
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  return CPGPU_OK;
}

// The batched pass of run_batched as one persistent cooperative launch
// (k_b_persist); same states, layout and results.
static int run_persistent(cpgpu_batch* B, int mode, int max_iter, double tol,
                          const std::vector<char>& member) {
  cpgpu_ctx* ctx = B->ctx;
  hipStream_t st = ctx->stream;
  const int n = B->n;
  for (int p = 0; p < n; ++p) {
    KState& k = B->h_st[p];
    std::memset(&k, 0, sizeof(k));
    k.tol = tol;
    k.max_iter = max_iter;
    k.it_dir = 1;
    k.mode = mode;
    k.refresh_every = B->K;
    k.normb = 1.0;
    if (!member[p]) k.done = 3;
  }
  CPGPU_HIP(ctx, hipMemcpyAsync(B->st.p, B->h_st, sizeof(KState) * n, hipMemcpyHostToDevice, st));
  std::vector<int> bmap;
  build_layout(B, &member, bmap);
  int rc = upload_layout(B, bmap);
  if (rc) return rc;
  if (mode != 0)
    return set_error(ctx, CPGPU_EUNSUPPORTED, "persistent batch: FP64 Jacobi-PCG only");
  if (B->Bcur > 0) {
    int coop = 0, ncu = 0, per_cu = 0;
    CPGPU_HIP(ctx, hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, ctx->device));
    if (!coop) return set_error(ctx, CPGPU_EUNSUPPORTED, "device without cooperative launches");
    CPGPU_HIP(ctx, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const void* kern = mode == 0 ? (const void*)k_b_persist<0> : (const void*)k_b_persist<1>;
    const size_t dyn = sizeof(int) * ((size_t)n + 1);  // live table
    if (dyn > 32768) return set_error(ctx, CPGPU_EINVAL, "persistent batch: at most 8191 problems");
    CPGPU_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads, dyn));
    if (per_cu < 1) return set_error(ctx, CPGPU_EHIP, "persistent batch kernel does not fit a CU");
    // every block co-resident (the barrier needs it), no more than the work
    const int G = std::max(1, std::min(B->Bcur, per_cu * ncu));
    if (B->gbar.alloc(bar_words(G))) return set_error(ctx, CPGPU_ENOMEM, "grid barrier");
    CPGPU_HIP(ctx, hipMemsetAsync(B->gbar.p, 0, bar_words(G) * sizeof(unsigned), st));
    const BProb* probs = B->probs.p;
    int nn = n, Btot = B->Btot, K = B->K;
    KState* sts = B->st.p;
    double *pB64 = B->partB64.p, *pA64 = B->partA64.p, *pR64 = B->partR64.p;
    float *pB32 = B->partB32.p, *pA32 = B->partA32.p;
    unsigned long long* ts = B->tstamp.p;
    unsigned* gb = B->gbar.p;
    long long rounds = ((long long)max_iter + 2) / B->K + 2;
    // a barrier wait longer than 20 s aborts the launch
    unsigned long long ticks = (unsigned long long)(B->clock_khz * 1e3 * 20.0);
    void* args[] = {(void*)&probs, (void*)&sts,  (void*)&nn,   (void*)&pB64, (void*)&pA64,
                    (void*)&pB32,  (void*)&pA32, (void*)&pR64, (void*)&Btot, (void*)&ts,
                    (void*)&gb,    (void*)&K,    (void*)&rounds, (void*)&ticks};
    CPGPU_HIP(ctx, hipLaunchCooperativeKernel(kern, dim3(G), kBlock, args, (unsigned)dyn, st));
    unsigned abort_flag = 0;
    CPGPU_HIP(ctx, hipMemcpyAsync(&abort_flag, B->gbar.p + kBarAbort, sizeof(unsigned),
                                  hipMemcpyDeviceToHost, st));
    CPGPU_HIP(ctx, hipStreamSynchronize(st));
    if (abort_flag) return set_error(ctx, CPGPU_EHIP, "persistent batch: grid barrier timed out");
  } else {
    hipLaunchKernelGGL(k_b_t0, dim3(1), dim3(64), 0, st, B->tstamp.p, n);
  }
  CPGPU_HIP(ctx, hipMemcpyAsync(B->h_st, B->st.p, sizeof(KState) * n, hipMemcpyDeviceToHost, st));
  CPGPU_HIP(ctx, hipMemcpyAsync(B->h_ts, B->tstamp.p, sizeof(unsigned long long) * (n + 1),
                                hipMemcpyDeviceToHost, st));
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  return CPGPU_OK;
}

int cpgpu_batch_set_persistent(cpgpu_batch* B, int on) {
  if (!B) return CPGPU_EINVAL;
  B->persistent = on != 0;
  return CPGPU_OK;
}

int cpgpu_batch_solve(cpgpu_batch* B, int method, int max_iter, double tol, int* iters,
                      double* resid, int* codes, double* latency_ms) {
  if (!B || max_iter < 0 || !(tol >= 0.0)) return CPGPU_EINVAL;
  if (method != CPGPU_JPCG && method != CPGPU_MIXED && method != CPGPU_FCG_UNET)
    return set_error(B->ctx, CPGPU_EUNSUPPORTED,
                     "batched methods: CPGPU_JPCG, CPGPU_MIXED, CPGPU_FCG_UNET");
  cpgpu_ctx* ctx = B->ctx;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  const int n = B->n;
  const auto t0 = std::chrono::steady_clock::now();
  const auto since = [&]() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  };
  std::vector<int> it_sum(n, 0);
  int all_conv = 1;
  // one batched pass over `member`, results recorded; latencies offset by
  // the host time at the pass's start
  auto batched = [&](const std::vector<char>& member, int mode) -> int {
    const double off = since();
    // the persistent launch serves the FP64 pass; the mixed-precision pass
    // stays on launched batches (its persistent form, k_b_persist<1>, hit a
    // grid barrier time-out on the 256-problem set whose cause is not
    // established, so it is not run)
    int rc = B->persistent && mode == 0 ? run_persistent(B, mode, max_iter, tol, member)
                                        : run_batched(B, mode, max_iter, tol, member);
    if (rc) return rc;
    for (int p = 0; p < n; ++p) {
      if (!member[p]) continue;
      const KState& k = B->h_st[p];
      const int code = k.done == 1 ? CPGPU_OK : CPGPU_MAXITER;
      if (code != CPGPU_OK) all_conv = 0;
      it_sum[p] += k.iter;
      if (iters) iters[p] = it_sum[p];
      if (resid) resid[p] = k.resid;
      if (codes) codes[p] = code;
      if (latency_ms)
        latency_ms[p] = k.done ? off + (double)(B->h_ts[p] - B->h_ts[n]) / B->clock_khz : -1.0;
    }
    return CPGPU_OK;
  };
  if (method != CPGPU_FCG_UNET) {
    const int rc = batched(std::vector<char>(n, 1), method == CPGPU_JPCG ? 0 : 1);
    if (rc) return rc;
    return all_conv ? CPGPU_OK : CPGPU_MAXITER;
  }
  // CPGPU_FCG_UNET: members with a U-Net attached as preconditioner run
  // flexible CG one after another (smallest first, one shape after another:
  // the network's activations are laid out per shape); then one batched
  // mixed pass takes the others together with the learned members whose
  // flexible CG stalled (from their iterates).  One pass: a pass's wall time
  // follows its slowest member, and two passes in turn (the others first)
  // would add the stalled members' thousands of iterations after it.
  std::vector<char> member(n, 1);
  std::vector<int> learned;
  for (int p = 0; p < n; ++p)
    if (B->P[p]->unet && (B->P[p]->unet_flags & CPGPU_UNET_PRECOND)) {
      member[p] = 0;
      learned.push_back(p);
    }
  int rc = 0;
  std::sort(learned.begin(), learned.end(), [&](int a, int b) {
    const cpgpu_poisson *A = B->P[a], *Q = B->P[b];
    return A->N != Q->N ? A->N < Q->N : (A->nx != Q->nx ? A->nx < Q->nx : a < b);
  });
  // flexible-CG iterations a learned member gets before the batched pass
  // takes over from its iterate (an application of the network costs tens
  // of batched iterations, so a slowly converging member is cheaper there)
  constexpr int kFcgBudget = 150;
  for (int p : learned) {
    cpgpu_poisson* P = B->P[p];
    int mi = std::min(max_iter, kFcgBudget), hl = 0;
    double t = tol;
    P->fcg_fallback = false;
    const int code = poisson_solve_internal(P, CPGPU_FCG_UNET, &mi, &t, CPGPU_NORM_TRUE, 0,
                                            nullptr, 0, &hl);
    P->fcg_fallback = true;
    if (code < 0) return code;
    it_sum[p] = mi;
    if (iters) iters[p] = mi;
    if (resid) resid[p] = t;
    if (codes) codes[p] = code;
    if (latency_ms) latency_ms[p] = since();
    if (code != CPGPU_OK) member[p] = 1;  // the batched pass goes on from its iterate
  }
  if ((rc = batched(member, 1))) return rc;
  return all_conv ? CPGPU_OK : CPGPU_MAXITER;
}

// Tile autotuner: for each candidate tiles-per-block count, the time of
// `reps` batches of iterations of `method` from the problems' current
// solutions (restored afterwards); keeps the fastest setting.
int cpgpu_batch_autotune(cpgpu_batch* B, int method, int reps, int* best_tpb,
                         float* ms_per_iter) {
  if (!B || reps < 1) return CPGPU_EINVAL;
  cpgpu_ctx* ctx = B->ctx;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  size_t total = 0;
  for (auto* Q : B->P) total += Q->N;
  if (B->save.alloc(total)) return set_error(ctx, CPGPU_ENOMEM, "autotune save");
  auto copy_x = [&](bool to_save) -> int {
    size_t o = 0;
    for (auto* Q : B->P) {
      double* a = B->save.p + o;
      CPGPU_HIP(ctx, hipMemcpyAsync(to_save ? a : Q->x.p, to_save ? Q->x.p : a,
                                    sizeof(double) * Q->N, hipMemcpyDeviceToDevice, st));
      o += Q->N;
    }
    return CPGPU_OK;
  };
  int rc = copy_x(true);
  if (rc) return rc;
  static const int cand[4] = {1, 2, 4, 8};
  hipEvent_t e0, e1;
  CPGPU_HIP(ctx, hipEventCreate(&e0));
  CPGPU_HIP(ctx, hipEventCreate(&e1));
  float best = 1e30f;
  int best_t = B->tpb;
  const int iters = reps * B->K;
  for (int c = 0; c < 4 && rc == CPGPU_OK; ++c) {
    if ((rc = set_tiles(B, cand[c]))) break;
    // untimed pass builds the graph; then the timed one
    float ms = 0.f;
    for (int pass = 0; pass < 2 && rc == CPGPU_OK; ++pass) {
      if ((rc = copy_x(false))) break;
      if (hipEventRecord(e0, st) != hipSuccess) rc = CPGPU_EHIP;
      const int code = cpgpu_batch_solve(B, method, iters, 0.0, nullptr, nullptr, nullptr, nullptr);
      if (code < 0) rc = code;
      if (rc == CPGPU_OK && (hipEventRecord(e1, st) != hipSuccess ||
                             hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = set_error(ctx, CPGPU_EHIP, "autotune timing");
    }
    if (rc) break;
    const float per = ms / (float)iters;
    if (ms_per_iter) ms_per_iter[c] = per;
    if (per < best) {
      best = per;
      best_t = cand[c];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if ((rc = copy_x(false))) return rc;
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  if ((rc = set_tiles(B, best_t))) return rc;
  if (best_tpb) *best_tpb = best_t;
  return CPGPU_OK;
}

// Random problem set: per problem log2(nx) ~ U[lo, hi] and log2(ny) ~
// U[max(lo, log2 nx - 2), min(hi, log2 nx + 2)], each rounded to the nearest
// integer (aspect ratio <= 4); each wall Dirichlet (1) or Neumann (2) with
// probability 1/2, redrawn while all four are Neumann; a blob seed.  One
// Xoshiro stream from `seed`, drawn problem by problem in that order.
int cpgpu_batch_draw(unsigned long long seed, int n, int lo, int hi, int* nx, int* ny,
                     int* bc_type, unsigned long long* blob_seed) {
  if (n < 1 || lo < 0 || hi < lo || hi > 20 || !nx || !ny || !bc_type || !blob_seed)
    return CPGPU_EINVAL;
  Xoshiro g(seed);
  for (int p = 0; p < n; ++p) {
    const int lx = (int)std::floor(lo + (hi - lo) * g.uniform() + 0.5);
    const int ylo = std::max(lo, lx - 2), yhi = std::min(hi, lx + 2);
    const int ly = (int)std::floor(ylo + (yhi - ylo) * g.uniform() + 0.5);
    nx[p] = 1 << lx;
    ny[p] = 1 << ly;
    unsigned bits;
    do {
      bits = (unsigned)(g.next() >> 60);
    } while (bits == 15u);
    for (int k = 0; k < 4; ++k)
      bc_type[4 * p + k] = (bits >> k) & 1u ? CPGPU_BC_NEUMANN : CPGPU_BC_DIRICHLET;
    blob_seed[p] = g.next();
  }
  return CPGPU_OK;
}

}  // extern "C"
