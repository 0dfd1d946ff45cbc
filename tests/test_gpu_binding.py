"""The compiled C++ callers of the C-ABI.

tests/cpp/caller (carpios_amd/binding/structured_gpu.hpp) drives
cpgpu_poisson2d_solve and cpgpu_ns_* on Scalar_-layout buffers the way a
structure::Solver_GPU_<2> and a device-resident NS_explicit_<2> would
(s_solver.hpp:261-291, s_ns_explicit.hpp:118-155); its results are compared
with the oracle.  oracle/_ref/algebra_shim (built from the reference tree,
present only where it was) runs the reference's own CG, Dia_BiCGSTAB and
Solver_Jacobi_ and their GPU counterparts (carpios_amd/binding/
algebra_gpu.hpp) on the reference's own MatrixSCR_ / ArrayListV."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "cpp", "_build", "caller")
SHIM = os.path.join(ROOT, "oracle", "_ref", "algebra_shim")
GOLD = os.path.join(ROOT, "tests", "golden", "golden_solvers.npz")


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_cpp_caller_matches_the_oracle(tmp_path):
    if not os.path.exists(CALLER):
        pytest.fail(f"{CALLER} is missing: build with __graft_entry__.build()")
    cases = {
        "c1": (128, 128, 1.0, [1, 1, 1, 1], [0.0, 0.0, 0.0, 0.0]),
        "mixed": (60, 45, 1.3, [1, 2, 2, 1], [0.5, -0.25, 1.0, 0.3]),
    }
    grids = {}
    for name, (nx, ny, xhi, types, vals) in cases.items():
        g = O.StructGrid(nx, ny, 1, 0.0, xhi, 0.0, 1.0)
        src = g.field()
        if name == "c1":
            X, Y = g.centres()
            g.interior(src)[:] = -2 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y)
        else:
            g.interior(src)[:] = np.random.default_rng(8).standard_normal((nx, ny))
        src.tofile(str(tmp_path / f"poisson_{name}_src.bin"))
        grids[name] = (g, src, types, vals)
    out = subprocess.run([CALLER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    for name, (g, src, types, vals) in grids.items():
        code, it, tol = open(tmp_path / f"poisson_{name}.txt").read().split()
        phi = np.fromfile(str(tmp_path / f"poisson_{name}_phi.bin")).reshape(g.shape)
        A = O.Matrix(g, O.make_bc(g, types, vals), 1.0, src)
        co, xo, ito, to, _ = O.solve("JPCG", A, A.b, max_iter=100000, tol=1e-10)
        print(f"{name}: caller {code}/{it}, oracle {co}/{ito}", flush=True)
        assert int(code) == co == 0 and abs(int(it) - ito) <= 1
        assert rel_l2(O.field_to_x(g, phi), xo) <= 1e-12
    # the 48^2 cavity, five steps
    g = O.StructGrid(48, 48, 2, -0.5, 0.5, -0.5, 0.5)
    ora = O.NS(g, O.make_bc(g, [1] * 4, [0, 0, 0, 1.0]), O.make_bc(g, [1] * 4), rho=1000.0,
               mu=1.0, cfl=0.5, dnum=0.25, dcoef=1e-3, limiter="VanLeer", tol=1e-10,
               dt=0.5 / 48)
    its_o = [ora.step()[1] for _ in range(5)]
    line = open(tmp_path / "ns.txt").read().split()
    its_c, t_c = [int(v) for v in line[:5]], float(line[5])
    print(f"ns: caller {its_c}, oracle {its_o}", flush=True)
    assert all(abs(a - b) <= 1 for a, b in zip(its_c, its_o))
    assert t_c == pytest.approx(ora.st.t, rel=1e-14)
    for name in ("u", "v", "p"):
        f = np.fromfile(str(tmp_path / f"ns_{name}.bin")).reshape(g.shape)
        assert rel_l2(g.interior(f), g.interior(ora.f[name])) <= 1e-12, name


def write_csr(path, rp, col, val, b):
    with open(path, "w") as f:
        f.write(f"{len(rp) - 1} {len(col)}\n")
        for arr, fmt in ((rp, "%d"), (col, "%d"), (val, "%.17g"), (b, "%.17g")):
            f.write(" ".join(fmt % v for v in arr) + "\n")


def run_shim(path, max_iter, tol):
    out = subprocess.run([SHIM, str(path), str(max_iter), repr(tol)], capture_output=True,
                         text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    res = {}
    for ln in out.stdout.strip().splitlines():
        t = ln.split()
        res[t[0]] = {"ref": (int(t[2]), int(t[3]), float(t[4]), int(t[5])),
                     "gpu": (int(t[7]), int(t[8]), float(t[9]), int(t[10])),
                     "dx": float(t[12])}
    return res


def test_algebra_shim_against_the_reference_solvers(tmp_path):
    if not os.path.exists(SHIM):
        pytest.skip("oracle/_ref/algebra_shim is built only where the reference tree is")
    # a mixed-boundary Poisson matrix assembled by the oracle (24 x 20), b random
    g = O.StructGrid(24, 20, 1, 0.0, 1.2, 0.0, 1.0)
    A = O.Matrix(g, O.make_bc(g, [1, 2, 1, 2], [0.5, 0.0, 0.0, -1.0]), 1.0)
    b = np.random.default_rng(2).standard_normal(A.n)
    write_csr(tmp_path / "poisson.csr", A.row_ptr, A.col, A.val, b)
    r = run_shim(tmp_path / "poisson.csr", 20000, 1e-10)
    for name in ("CG", "Solver_Jacobi_"):
        ref, gpu = r[name]["ref"], r[name]["gpu"]
        assert ref[0] == gpu[0] == 0
        assert abs(ref[1] - gpu[1]) <= 1 and abs(ref[3] - gpu[3]) <= 1
        assert r[name]["dx"] <= 1e-10
    # Dia_BiCGSTAB's stopping step moves with rounding: the device run is
    # held to the band the oracle (the reference's arithmetic to the bit)
    # shows under 1e-15 perturbations of b (77..92 steps here)
    from test_gpu_krylov import oracle_band
    _, lo, hi, spread, _ = oracle_band("Dia_BiCGSTAB", A, b, runs=40, max_iter=20000, tol=1e-10)
    ref, gpu = r["Dia_BiCGSTAB"]["ref"], r["Dia_BiCGSTAB"]["gpu"]
    assert ref[0] == gpu[0] == 0 and lo - 1 <= gpu[3] <= hi + 1
    assert r["Dia_BiCGSTAB"]["dx"] <= max(1e-12, 10 * spread)
    # the reference test's 685_bus matrix (test/algebra/_test_solver.hpp), b = 1
    z = np.load(GOLD, allow_pickle=False)
    n = len(z["bus_row_ptr"]) - 1
    write_csr(tmp_path / "bus.csr", z["bus_row_ptr"], z["bus_col"], z["bus_val"], np.ones(n))
    r = run_shim(tmp_path / "bus.csr", 5000, 1e-6)
    for name in ("CG", "Dia_BiCGSTAB", "Solver_Jacobi_"):
        ref, gpu = r[name]["ref"], r[name]["gpu"]
        assert ref[0] == gpu[0]
        assert r[name]["dx"] <= 1e-4, name
