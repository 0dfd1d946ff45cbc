// A compiled C++ caller of the C-ABI through carpios_amd/binding: the calls
// a structure::Solver_GPU_<2>::solve() and a device-resident NS_explicit_<2>
// make, on Scalar_-layout buffers ((nx+2gl) x (ny+2gl) doubles, j
// contiguous).  Reads its inputs from and writes its results to a directory
// (raw little-endian doubles and a text line per case); tests/
// test_gpu_binding.py compares them with the oracle.
//   caller <dir>
//     poisson_<case>_src.bin  ->  poisson_<case>_phi.bin, poisson_<case>.txt
//     ns_u.bin, ns_v.bin, ns_p.bin, ns.txt (a 48^2 cavity, 5 steps)
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../carpios_amd/binding/structured_gpu.hpp"

namespace {

bool read_bin(const std::string& path, std::vector<double>& v) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  const size_t got = std::fread(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
  return got == v.size();
}

bool write_bin(const std::string& path, const std::vector<double>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const size_t put = std::fwrite(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
  return put == v.size();
}

// One Poisson_ solve as Solver_GPU_::solve() makes it: J-PCG from phi = 0 to
// 1e-10 with the reference's nrm2 test.
int poisson_case(const std::string& dir, const std::string& name, int nx, int ny,
                 double xhi, const int type[4], const double value[4]) {
  const int gl = 1;
  cpgpu_bind::Axis x = cpgpu_bind::uniform_axis(nx, gl, 0.0, xhi);
  cpgpu_bind::Axis y = cpgpu_bind::uniform_axis(ny, gl, 0.0, 1.0);
  cpgpu_bind::Walls w = cpgpu_bind::constant_walls(x, y, type, value);
  std::vector<double> src((size_t)(nx + 2 * gl) * (ny + 2 * gl), 0.0), phi(src.size(), 0.0);
  if (!read_bin(dir + "/poisson_" + name + "_src.bin", src)) return 2;
  int max_iter = 100000;
  double tol = 1e-10;
  std::string err;
  const int code = cpgpu_bind::solve_poisson(x, y, 1.0, w, src.data(), 0.0, phi.data(),
                                             CPGPU_JPCG, max_iter, tol, CPGPU_NORM_REF, &err);
  if (code < 0) {
    std::fprintf(stderr, "poisson %s: %d %s\n", name.c_str(), code, err.c_str());
    return 3;
  }
  if (!write_bin(dir + "/poisson_" + name + "_phi.bin", phi)) return 4;
  FILE* f = std::fopen((dir + "/poisson_" + name + ".txt").c_str(), "w");
  if (!f) return 4;
  std::fprintf(f, "%d %d %.17g\n", code, max_iter, tol);
  std::fclose(f);
  std::printf("poisson %s %dx%d: code %d, %d iterations, residual %.3e\n", name.c_str(), nx,
              ny, code, max_iter, tol);
  return 0;
}

// A 48^2 lid-driven cavity on [-0.5, 0.5]^2 (Re 1000: rho 1000, mu 1; van
// Leer; CFL 0.5; diffusion number 0.25 with coefficient 1/Re; dt0 = h/2),
// five NS_explicit_ steps with J-PCG projections to 1e-10.
int ns_case(const std::string& dir) {
  const int n = 48, gl = 2;
  cpgpu_bind::Axis x = cpgpu_bind::uniform_axis(n, gl, -0.5, 0.5);
  cpgpu_bind::Axis y = cpgpu_bind::uniform_axis(n, gl, -0.5, 0.5);
  const int d1[4] = {CPGPU_BC_DIRICHLET, CPGPU_BC_DIRICHLET, CPGPU_BC_DIRICHLET,
                     CPGPU_BC_DIRICHLET};
  const double lid[4] = {0.0, 0.0, 0.0, 1.0}, still[4] = {0.0, 0.0, 0.0, 0.0};
  cpgpu_bind::ExplicitStep S(x, y, cpgpu_bind::constant_walls(x, y, d1, lid),
                             cpgpu_bind::constant_walls(x, y, d1, still), 1000.0, 1.0, 0.5,
                             0.25, 1e-3, CPGPU_LIM_VANLEER, CPGPU_JPCG, 100000, 1e-10,
                             CPGPU_NORM_REF, 0.5 / n);
  if (S.status() != CPGPU_OK) return 5;
  std::vector<int> its;
  const int code = S.step(5, &its);
  if (code < 0) return 6;
  std::vector<double> f((size_t)(n + 2 * gl) * (n + 2 * gl));
  const int ids[3] = {CPGPU_F_U, CPGPU_F_V, CPGPU_F_P};
  const char* names[3] = {"u", "v", "p"};
  for (int k = 0; k < 3; ++k) {
    if (S.get(ids[k], f.data()) != CPGPU_OK) return 7;
    if (!write_bin(dir + "/ns_" + names[k] + ".bin", f)) return 8;
  }
  double t = 0.0, dt = 0.0;
  S.time(t, dt);
  FILE* o = std::fopen((dir + "/ns.txt").c_str(), "w");
  if (!o) return 8;
  for (int v : its) std::fprintf(o, "%d ", v);
  std::fprintf(o, "%.17g %.17g\n", t, dt);
  std::fclose(o);
  std::printf("ns 48^2: 5 steps, projection iterations");
  for (int v : its) std::printf(" %d", v);
  std::printf(", t %.6g\n", t);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: caller <dir>\n");
    return 1;
  }
  const std::string dir = argv[1];
  const int dir4[4] = {CPGPU_BC_DIRICHLET, CPGPU_BC_DIRICHLET, CPGPU_BC_DIRICHLET,
                       CPGPU_BC_DIRICHLET};
  const double zero[4] = {0.0, 0.0, 0.0, 0.0};
  const int mixed[4] = {CPGPU_BC_DIRICHLET, CPGPU_BC_NEUMANN, CPGPU_BC_NEUMANN,
                        CPGPU_BC_DIRICHLET};
  const double mixed_val[4] = {0.5, -0.25, 1.0, 0.3};
  int rc = poisson_case(dir, "c1", 128, 128, 1.0, dir4, zero);
  if (rc) return rc;
  rc = poisson_case(dir, "mixed", 60, 45, 1.3, mixed, mixed_val);
  if (rc) return rc;
  return ns_case(dir);
}
