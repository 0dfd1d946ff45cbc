// Host side of a device-resident Poisson problem, shared by the Poisson API
// and the NS projection step.
#pragma once

#include <vector>

#include "stencil.hpp"

namespace cpgpu {

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  int alloc(size_t count) {
    if (n >= count && p) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, sizeof(T) * (count ? count : 1)) != hipSuccess) return -1;
    n = count;
    return 0;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// b = -(vol*src + ghost constants) (s_poisson.hpp:79-97,
// s_ghost.hpp:161-175, s_solver.hpp:122)
struct RhsSeg {
  int type;       // 1 Dirichlet, 2 Neumann
  double coe;     // coefficient of the ghost term (negative)
  double dl, d1;  // distances of the substitution
};
// The constant a ghost term leaves in its row (s_ghost.hpp:171-176):
// Dirichlet coe*dl*val/d1, Neumann -val*dl*coe, added to cst.
__device__ __forceinline__ double ghost_const(const RhsSeg& g, double cst, double val) {
  if (g.type == CPGPU_BC_DIRICHLET) return cst + (g.coe * g.dl * val / g.d1);
  return cst + (-val * g.dl * g.coe);
}

struct RhsDesc {
  int nx, ny;
  const double* vol;    // nx: (1.0*sx)*sx
  const double* bcval;  // packed interior lines: x- ny | x+ ny | y- nx | y+ nx
  RhsSeg seg[4];
};

// Sums gathered for the refinement loop and the mean projections.
struct RState {
  double bb, rr, rsum, xsum;
};

// Solver_Jacobi_ state (device): done = 1 converged, 2 stopped at max_iter.
struct JState {
  double normb, resid;
  int iter, done;
};

// Dia_BiCGSTAB state (device): the scalars of solver_matrix.hpp:599-672,
// the iterations done, the history entries written (the full iterations and
// a half-step exit), and the exit: stop once the run has ended, half when it
// ended at the half-step test, code the status it returns.
struct BState {
  double normb, tol, resid, rho_1, rho_2, alpha, omega, beta;
  int it, hl, max_iter, hist_cap, stop, half, code, pad;
};

// Flexible-CG state (device, mirrored to pinned host memory).
struct FState {
  double normb;   // ||b|| of the stopping test
  double pq;      // p.q of the last direction (for the next beta)
  double resid;   // ||r|| / ||b|| after the last update
  int iter;       // updates done
  int breakdown;  // p.q <= 0 or not finite
};

// The part of FState the host needs after every flexible-CG iteration,
// written by the device into host memory (mapped, coherent) with the update
// count as a sequence number stored last, so the host polls it instead of a
// copy and a stream synchronisation.
struct FMail {
  double resid;
  int breakdown;
  int seq;  // FState::iter of the values above; stored last, with a release
};

// SplitMix64-seeded xoshiro256** (Blackman & Vigna); doubles from the top
// 53 bits; normals by Box-Muller.  The generator of every seeded input
// (blob sources, batches of random problems).
struct Xoshiro {
  unsigned long long s[4];
  static unsigned long long splitmix(unsigned long long& x) {
    unsigned long long z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  explicit Xoshiro(unsigned long long seed) {
    for (int k = 0; k < 4; ++k) s[k] = splitmix(seed);
  }
  static unsigned long long rotl(unsigned long long x, int k) {
    return (x << k) | (x >> (64 - k));
  }
  unsigned long long next() {
    const unsigned long long r = rotl(s[1] * 5, 7) * 9;
    const unsigned long long t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double normal() {
    const double u1 = 1.0 - uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

// Gaussian-blob parameters from Xoshiro(seed): per blob cx, cy ~ U[0,1),
// w ~ U[wmin, wmax), A ~ N(0,1) (poisson.hip).
void blob_params(unsigned long long seed, int nblobs, double wmin, double wmax,
                 double* out);

// Direction-kernel variant of a new problem (the column march with 16 rows
// per wavefront) until cpgpu_poisson_set_variant or the autotuner changes it.
constexpr int kDefaultDirVariant = 3;

// Identity of a captured batch of Krylov iterations.
struct GraphKey {
  const void* x;
  const void* hist;
  size_t tsize;
  int pc, variant, iters, norm_ref, hist_cap, reliable;
  bool operator==(const GraphKey& o) const {
    return x == o.x && hist == o.hist && tsize == o.tsize && pc == o.pc &&
           variant == o.variant && iters == o.iters &&
           norm_ref == o.norm_ref && hist_cap == o.hist_cap &&
           reliable == o.reliable;
  }
};

}  // namespace cpgpu

struct cpgpu_poisson {
  cpgpu_ctx* ctx = nullptr;
  int nx = 0, ny = 0, gl = 0;
  size_t N = 0;
  double beta = 1.0;
  double shift = 0.0;  // Helmholtz vol*alpha (0: Poisson)
  // cell sizes that vary along an axis: the operator is held as per-face
  // coefficient arrays and per-line ghost data (general.hip), and only the
  // FP64 Jacobi-PCG / CG of the reference run on it
  bool general = false;
  cpgpu::DevBuf<double> gFX, gFY, gD;
  cpgpu::DevBuf<cpgpu::RhsSeg> gseg;
  // where b came from, so that a new alpha can rebuild it: 0 none, 1 the
  // problem's own source field (src), 2 a uniform value (src_uni), 3 a
  // caller's device field (not kept)
  int src_kind = 0;
  double src_uni = 0.0;
  cpgpu::Tiling tl{};
  // host copies of the descriptor
  std::vector<double> KX, KY, vol, bcv;
  std::vector<double> cxi, cyi;  // interior cell centres
  std::vector<double> sxi, syi;  // interior cell sizes
  double adj[4] = {0, 0, 0, 0};
  cpgpu::RhsSeg seg[4];
  // device operator
  cpgpu::DevBuf<double> dKX, dKY, dvol, dbcv;
  cpgpu::DevBuf<float> fKX, fKY;
  // FP64 vectors; xlo: the low part of the mixed solves' solution x + xlo
  cpgpu::DevBuf<double> x, b, r, p0, p1, q, src, xlo;
  // FP32 vectors of the inner solve
  cpgpu::DevBuf<float> b32, x32, r32, pa32, pb32, q32;
  // reductions and states
  cpgpu::DevBuf<double> partA64, partB64, part64, partR64;
  cpgpu::DevBuf<float> partA32, partB32;
  cpgpu::DevBuf<cpgpu::KState> kst;
  cpgpu::DevBuf<cpgpu::RState> rst;
  cpgpu::DevBuf<double> dhist;
  cpgpu::KState* h_kst = nullptr;  // pinned: [0] staging, [1..2] batch ring
  cpgpu::RState* h_rst = nullptr;
  // batched Krylov iterations: replayed graphs and the polling events
  bool use_graphs = true;
  bool fcg_graphs_on = false;  // flexible-CG iterations as replayed graphs too
  int batch = 16;  // iterations per batch (even)
  std::vector<std::pair<cpgpu::GraphKey, hipGraphExec_t>> graphs;
  // flexible-CG iterations (preconditioner cycle + Krylov kernels):
  // configuration key -> replayed graph
  std::vector<std::pair<std::vector<long long>, hipGraphExec_t>> fcg_graphs;
  hipEvent_t ev_batch[2] = {nullptr, nullptr};
  // direction-kernel variant per precision (0 LDS tile, 1..4 column march
  // with H = 4, 8, 16, 32); set by heuristic, cpgpu_poisson_set_variant or
  // the autotuner
  int dir_variant64 = cpgpu::kDefaultDirVariant, dir_variant32 = cpgpu::kDefaultDirVariant;
  double inner_tol = 1e-4;
  int max_outer = 40;
  // attached U-Net (warm start / flexible-CG preconditioner) and the
  // flexible-CG state
  cpgpu_unet* unet = nullptr;
  int unet_flags = 0;
  int fcg_iters = 0;  // flexible-CG iterations of the last CPGPU_FCG_UNET solve
  // CPGPU_FCG_UNET finishes a stalled solve with Jacobi-PCG (off: it returns
  // CPGPU_MAXITER at the stall; the batched solver finishes such members)
  bool fcg_fallback = true;
  // preconditioner cycle of CPGPU_FCG_UNET: weighted-Jacobi sweeps before
  // and after the network (its coarse correction)
  int pc_pre = 2, pc_post = 2;
  double pc_omega = 0.8;
  cpgpu::DevBuf<double> pf[2];
  // the network input's sum-of-squares partials (one per residual block)
  cpgpu::DevBuf<double> npart;
  // Dia_BiCGSTAB (rt, p, ph, v, s, sh, t) and Solver_Jacobi_ vectors
  cpgpu::DevBuf<double> bv[8];
  cpgpu::DevBuf<cpgpu::JState> jst;
  cpgpu::DevBuf<cpgpu::BState> bst;
  cpgpu::DevBuf<cpgpu::FState> fst;
  cpgpu::FState* h_fst = nullptr;  // pinned
  cpgpu::FMail* h_mail = nullptr;  // host-mapped, coherent (null: copy instead)
  cpgpu::FMail* d_mail = nullptr;  // its device address
  // error norms: the exact field and the interior cell sizes on the device
  cpgpu::DevBuf<double> err_ref, dsxi, dsyi;

  cpgpu::Sten<double> sten64() const {
    cpgpu::Sten<double> s;
    s.nx = nx;
    s.ny = ny;
    s.KX = dKX.p;
    s.KY = dKY.p;
    for (int k = 0; k < 4; ++k) s.adj[k] = adj[k];
    s.shift = shift;
    return s;
  }
  cpgpu::Sten<float> sten32() const {
    cpgpu::Sten<float> s;
    s.nx = nx;
    s.ny = ny;
    s.KX = fKX.p;
    s.KY = fKY.p;
    for (int k = 0; k < 4; ++k) s.adj[k] = (float)adj[k];
    s.shift = (float)shift;
    return s;
  }
  ~cpgpu_poisson() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    for (auto& g : fcg_graphs) (void)hipGraphExecDestroy(g.second);
    for (auto e : ev_batch)
      if (e) (void)hipEventDestroy(e);
    if (h_kst) (void)hipHostFree(h_kst);
    if (h_rst) (void)hipHostFree(h_rst);
    if (h_fst) (void)hipHostFree(h_fst);
    if (h_mail) (void)hipHostFree(h_mail);
  }
};

namespace cpgpu {
// Solves with the problem's own vectors (used by the NS step as well).
int poisson_solve_internal(cpgpu_poisson* P, int method, int* max_iter,
                           double* tol, int norm_mode, int opts, double* hist,
                           int hist_cap, int* hist_len);
// b from a device source field (interior-only, j contiguous).
int poisson_rhs_from_device(cpgpu_poisson* P, const double* dsrc,
                            double src_uniform);
// Grids whose cell sizes vary (general.hip): the operator from the
// descriptor, b, y = A x on device vectors, the reference's Jacobi-PCG / CG,
// and the true relative residual.
int general_build(cpgpu_poisson* P, const cpgpu_grid2d* g);
int general_rhs(cpgpu_poisson* P, const double* dsrc, double src_uniform);
int general_apply(cpgpu_poisson* P, const double* dx, double* dy);
int general_solve(cpgpu_poisson* P, int method, int* max_iter, double* tol, int norm_ref,
                  double* hist, int hist_cap, int* hist_len);
int general_residual(cpgpu_poisson* P, double* rel);
// FP32 operator coefficients and FP32 work vectors of the problem.
int poisson_prepare_f32(cpgpu_poisson* P);
// Plain FP64 Jacobi-PCG from the current phi (the reference's test norm
// when norm_ref, else the Euclidean one); *iters in: limit, out: used.
int poisson_jpcg64(cpgpu_poisson* P, int* iters, double tol, int norm_ref,
                   double* resid, int* done);
// carpio::Dia_BiCGSTAB and carpio::Solver_Jacobi_ on the device (krylov.hip);
// max_iter, tol in/out with the reference's meaning, reference status codes.
int run_bicgstab(cpgpu_poisson* P, int* max_iter, double* tol, int norm_ref,
                 double* hist, int hist_cap, int* hist_len);
int run_jacobi(cpgpu_poisson* P, int* max_iter, double* tol, int norm_ref,
               double* hist, int hist_cap, int* hist_len);
// Flexible CG with the attached U-Net as preconditioner (fcg.hip).
int run_fcg_unet(cpgpu_poisson* P, int max_iter, double tol, int norm_ref,
                 int project, double* hist, int hist_cap, int* hist_len,
                 int* iters, double* resid);
// phi = s net(b / s), s = rms(b), with the attached U-Net (fcg.hip).
int unet_warm_start(cpgpu_poisson* P);
// out = s net(in / s), s = rms(in), on H x W device FP32 arrays, with the
// wall types `walls` (pack_walls) as the network's second input (unet.hip).
// sq_part: sq_G partial sums of in^2 formed by the caller's kernel (the
// input's rms is then reduced from them); null: formed here from in.
int unet_apply_device(cpgpu_unet* U, int H, int W, const float* in, float* out, int walls,
                      const double* sq_part = nullptr, int sq_G = 0);
// The wall types of a problem packed for unet_apply_device (unet.hip).
int pack_walls(const cpgpu_poisson* P);
// Error1 / Error2 / Errori of val against exa (interior arrays with row
// strides lde / ldv, j contiguous) on the cells of the sizes sx (nx), sy (ny);
// part holds at least 7 doubles (norms.hip).  out: host, 3 values.
int error_norms_device(cpgpu_ctx* ctx, int nx, int ny, const double* exa, int lde,
                       const double* val, int ldv, const double* sx, const double* sy,
                       double* part, size_t part_cap, double* out);
// The context a U-Net was created on (unet.hip).
cpgpu_ctx* unet_context(const cpgpu_unet* U);
// changes whenever an application captured in a graph would be stale
long long unet_generation(const cpgpu_unet* U);
// drop every captured iteration graph of P (their kernel arguments hold
// the operator's coefficients)
void drop_graphs(cpgpu_poisson* P);
}  // namespace cpgpu
