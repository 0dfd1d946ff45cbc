"""Writes tests/golden/golden_solvers.npz: outputs of the reference's own
lib/algebra code (compiled from the reference tree into
oracle/_ref/libcarpios_ref.so by oracle/Makefile) on small inputs.

Run in a container that has the reference tree:
    make -C oracle && python tests/golden/make_golden.py

Contents (all plain arrays, loadable with numpy.load(allow_pickle=False)):
  dot_x, dot_y, dot_ref, nrm2_ref ........ BLAS_Level1 dot / nrm2 on a seeded
                                           vector, plus nrm2([3,4]) and
                                           nrm2([-2]) special cases
  c1_* .................................... 32x32 Dirichlet Poisson
                                           (f = -2 pi^2 sin sin) CSR + b as
                                           assembled by the oracle, and per
                                           solver: code, iterations, final
                                           tol, solution, residual history (first and
                                           last 200 entries and its length)
  bus_* ................................... the reference test's 685_bus.mtx
                                           (test/input_files) expanded to CSR
                                           and solved from x=1, b=1 as
                                           test/algebra/_test_solver.hpp:
                                           129-163 does (Solver_Jacobi_ to
                                           1e-6), plus CG / Jacobi-PCG /
                                           Dia_BiCGSTAB on the same system
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "golden_solvers.npz")


def read_mtx_symmetric(path):
    """MatrixMarket coordinate real symmetric -> CSR with both triangles,
    columns sorted per row (the expansion mm_read_mtx_sparse performs,
    io/mmio.h:255)."""
    rows, cols, vals = [], [], []
    with open(path) as f:
        header = f.readline()
        assert "coordinate" in header and "symmetric" in header
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        n, m, nz = (int(t) for t in line.split())
        for _ in range(nz):
            i, j, v = f.readline().split()
            i, j, v = int(i) - 1, int(j) - 1, float(v)
            rows.append(i); cols.append(j); vals.append(v)
            if i != j:
                rows.append(j); cols.append(i); vals.append(v)
    order = np.lexsort((np.array(cols), np.array(rows)))
    r = np.array(rows)[order]
    c = np.array(cols, dtype=np.int32)[order]
    v = np.array(vals)[order]
    rp = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rp, r + 1, 1)
    return n, np.cumsum(rp).astype(np.int32), c, v


class CsrHolder:
    def __init__(self, n, rp, col, val, b):
        self.n, self.row_ptr, self.col, self.val, self.b = n, rp, col, val, b
        self.c = O.CSR(n, rp.ctypes.data_as(C.POINTER(C.c_int)),
                       col.ctypes.data_as(C.POINTER(C.c_int)),
                       val.ctypes.data_as(C.POINTER(C.c_double)))


def run_all(prefix, A, x0, methods, max_iter, tol, out):
    for m in methods:
        code, x, it, t, h = O.ref_solve(m, A, A.b, x0=x0, max_iter=max_iter,
                                        tol=tol)
        key = f"{prefix}_{m}"
        out[key + "_code"] = np.array(code)
        out[key + "_iters"] = np.array(it)
        out[key + "_tol"] = np.array(t)
        out[key + "_x"] = x
        out[key + "_hist_len"] = np.array(len(h))
        out[key + "_hist_head"] = h[:200]
        out[key + "_hist_tail"] = h[-200:]


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref not built: run make -C oracle")
    out = {}
    rng = np.random.default_rng(12345)
    x = rng.standard_normal(1000)
    y = rng.standard_normal(1000)
    R = O.ref()
    dp = C.POINTER(C.c_double)
    out["dot_x"], out["dot_y"] = x, y
    out["dot_ref"] = np.array(R.ref_dot(1000, x.ctypes.data_as(dp),
                                        y.ctypes.data_as(dp)))
    out["nrm2_ref"] = np.array(R.ref_nrm2(1000, x.ctypes.data_as(dp)))
    v34 = np.array([3.0, 4.0])
    out["nrm2_34"] = np.array(R.ref_nrm2(2, v34.ctypes.data_as(dp)))
    vm2 = np.array([-2.0])
    out["nrm2_m2"] = np.array(R.ref_nrm2(1, vm2.ctypes.data_as(dp)))

    # C1 operator at 32^2
    n = 32
    g = O.StructGrid(n, n, 1)
    X, Y = g.centres()
    src = g.field()
    g.interior(src)[:] = -2 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y)
    A = O.Matrix(g, O.make_bc(g, [1, 1, 1, 1]), 1.0, src)
    out["c1_row_ptr"], out["c1_col"], out["c1_val"], out["c1_b"] = \
        A.row_ptr, A.col, A.val, A.b
    run_all("c1", A, None, ["CG", "JPCG", "Dia_BiCGSTAB", "Solver_Jacobi_",
                            "IC_CGS"], 20000, 1e-10, out)
    # SpMV of a seeded vector
    xs = rng.standard_normal(A.n)
    ys = np.zeros(A.n)
    R.ref_spmv(A.n, A.row_ptr.ctypes.data_as(C.POINTER(C.c_int)),
               A.col.ctypes.data_as(C.POINTER(C.c_int)),
               A.val.ctypes.data_as(dp), xs.ctypes.data_as(dp),
               ys.ctypes.data_as(dp))
    out["c1_spmv_x"], out["c1_spmv_y"] = xs, ys

    # 685_bus from the reference's tests
    nb, rp, col, val = read_mtx_symmetric(
        os.path.join(REF, "test/input_files/685_bus.mtx"))
    B = CsrHolder(nb, rp, col, val, np.ones(nb))
    out["bus_row_ptr"], out["bus_col"], out["bus_val"] = rp, col, val
    run_all("bus", B, np.ones(nb), ["Solver_Jacobi_", "CG", "JPCG",
                                    "Dia_BiCGSTAB"], 100000, 1e-6, out)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: out[k] for k in out if k.endswith("_iters")})


if __name__ == "__main__":
    main()
