"""Pins the CPU oracle: golden vectors from the reference's own lib/algebra
(tests/golden/golden_solvers.npz), the compiled reference itself where
oracle/_ref is built, closed forms, invariants and brute force."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "golden_solvers.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def c1_problem(n, gl=1):
    g = O.StructGrid(n, n, gl)
    X, Y = g.centres()
    src = g.field()
    g.interior(src)[:] = -2 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y)
    A = O.Matrix(g, O.make_bc(g, [1, 1, 1, 1]), 1.0, src)
    return g, A, X, Y


# ---------------------------------------------------------------- BLAS-1
def test_nrm2_reference_quirks(gold):
    # BLAS_Level1.hpp:569 drops the last element; n == 1 returns |x0|
    assert O.lib().orc_nrm2_ref(2, np.array([3.0, 4.0]).ctypes.data_as(O._dp)) == 3.0
    assert float(gold["nrm2_34"]) == 3.0
    assert float(gold["nrm2_m2"]) == 2.0
    x = gold["dot_x"]
    assert O.lib().orc_nrm2_ref(len(x), x.ctypes.data_as(O._dp)) == float(gold["nrm2_ref"])


def test_dot_matches_reference_bitwise(gold):
    x, y = gold["dot_x"], gold["dot_y"]
    d = O.lib().orc_dot(len(x), x.ctypes.data_as(O._dp), y.ctypes.data_as(O._dp))
    assert d == float(gold["dot_ref"])
    # a plausible mistake -- pairwise summation -- would differ in the last bits
    assert abs(d - float(np.dot(x, y))) < 1e-12


# ---------------------------------------------------------------- grid
def test_uniform_axis_closed_form():
    a = O.Axis(128, 2, -0.5, 0.5)
    h = 1.0 / 128
    assert np.all(a.s == h)
    assert np.allclose(a.c, -0.5 + (np.arange(132) - 2 + 0.5) * h, atol=1e-15)
    # ghost centres mirror the interior about the wall
    assert abs((a.c[1] + a.c[2]) / 2 - (-0.5)) < 1e-15


# ---------------------------------------------------------------- operator
def test_c1_assembly_matches_golden(gold):
    g, A, X, Y = c1_problem(32)
    assert np.array_equal(A.row_ptr, gold["c1_row_ptr"])
    assert np.array_equal(A.col, gold["c1_col"])
    assert np.array_equal(A.val, gold["c1_val"])
    assert np.array_equal(A.b, gold["c1_b"])
    assert np.array_equal(A.mult(gold["c1_spmv_x"]), gold["c1_spmv_y"])


def test_c1_eigenvector_closed_form():
    # sin(pi x) sin(pi y) is an eigenvector of the Dirichlet operator, so the
    # discrete solution is known exactly.
    for n in (16, 128):
        g, A, X, Y = c1_problem(n)
        h = 1.0 / n
        fac = np.pi ** 2 * h ** 2 / (4 * np.sin(np.pi * h / 2) ** 2)
        exact = fac * np.sin(np.pi * X) * np.sin(np.pi * Y)
        xe = O.field_to_x(g, np.pad(exact, 1))
        assert np.max(np.abs(A.mult(xe) - A.b)) < 1e-12 * np.max(np.abs(A.b)) * 10
        # unpreconditioned CG converges in one iteration on an eigenvector
        code, x, it, tol, h_ = O.solve("CG", A, A.b, tol=1e-10)
        assert code == 0 and it == 1
        assert np.max(np.abs(x - xe)) < 1e-13


def test_operator_symmetry_and_neumann_nullspace():
    g = O.StructGrid(12, 9, 1)
    rng = np.random.default_rng(0)
    for types in ([2, 2, 2, 2], [1, 2, 1, 2], [1, 1, 1, 1]):
        A = O.Matrix(g, O.make_bc(g, types), 1.0)
        D = A.dense()
        assert np.array_equal(D, D.T)
        x, y = rng.standard_normal(A.n), rng.standard_normal(A.n)
        assert abs(y @ A.mult(x) - x @ A.mult(y)) < 1e-11
    An = O.Matrix(g, O.make_bc(g, [2, 2, 2, 2]), 1.0)
    assert np.max(np.abs(An.mult(np.ones(An.n)))) < 1e-12


def test_quadratic_exactness_dirichlet():
    # a quadratic phi with Dirichlet data from the exact solution: the
    # 5-point operator is exact in the interior for quadratics
    n = 20
    g = O.StructGrid(n, n, 1)
    X, Y = g.centres()
    phi = lambda x, y: 2 * x ** 2 + y ** 2
    src = g.field()
    g.interior(src)[:] = 6.0  # laplacian
    NX, NY = g.shape
    # boundary values at the reference's sample points (face coordinate along
    # the normal, cell centre along the wall for interior lines)
    xm = np.array([phi(0.0, g.ay.c[J]) for J in range(NY)])
    xp = np.array([phi(1.0, g.ay.c[J]) for J in range(NY)])
    ym = np.array([phi(g.ax.c[I], 0.0) for I in range(NX)])
    yp = np.array([phi(g.ax.c[I], 1.0) for I in range(NX)])
    A = O.Matrix(g, O.make_bc(g, [1, 1, 1, 1], np.concatenate([xm, xp, ym, yp])),
                 1.0, src)
    xe = O.field_to_x(g, np.pad(phi(X, Y), 1))
    r = A.mult(xe) - A.b
    interior = np.zeros((n, n), bool)
    interior[1:-1, 1:-1] = True
    rr = O.x_to_field(g, r)[1:-1, 1:-1]
    assert np.max(np.abs(rr[interior])) < 1e-10


def test_jpcg_vs_dense_brute_force():
    g = O.StructGrid(10, 7, 1)
    rng = np.random.default_rng(3)
    src = g.field()
    g.interior(src)[:] = rng.standard_normal((10, 7))
    A = O.Matrix(g, O.make_bc(g, [1, 2, 1, 1]), 1.0, src)
    xd = np.linalg.solve(A.dense(), A.b)
    for m in ("CG", "JPCG", "Dia_BiCGSTAB"):
        code, x, it, tol, h = O.solve(m, A, A.b, tol=1e-13, norm=O.NORM_TRUE)
        assert code == 0
        assert np.max(np.abs(x - xd)) < 1e-10 * np.max(np.abs(xd))


# ---------------------------------------------------------------- solvers vs the reference
GOLD_C1 = ["CG", "JPCG", "Dia_BiCGSTAB", "Solver_Jacobi_"]


@pytest.mark.parametrize("method", GOLD_C1)
def test_solvers_match_reference_golden_c1(gold, method):
    g, A, X, Y = c1_problem(32)
    m = "Jacobi" if method == "Solver_Jacobi_" else method
    code, x, it, tol, h = O.solve(m, A, A.b, max_iter=20000, tol=1e-10)
    key = f"c1_{method}"
    assert code == int(gold[key + "_code"])
    assert it == int(gold[key + "_iters"])
    assert tol == float(gold[key + "_tol"])
    assert np.array_equal(x, gold[key + "_x"])
    assert len(h) == int(gold[key + "_hist_len"])
    assert np.array_equal(h[:200], gold[key + "_hist_head"])
    assert np.array_equal(h[-200:], gold[key + "_hist_tail"])


@pytest.mark.parametrize("method", ["Solver_Jacobi_", "CG", "JPCG"])
def test_solvers_match_reference_golden_685_bus(gold, method):
    n = len(gold["bus_row_ptr"]) - 1

    class Holder:
        pass
    A = Holder()
    A.n = n
    A.row_ptr, A.col, A.val = gold["bus_row_ptr"], gold["bus_col"], gold["bus_val"]
    import ctypes as C
    A.c = O.CSR(n, A.row_ptr.ctypes.data_as(C.POINTER(C.c_int)),
                A.col.ctypes.data_as(C.POINTER(C.c_int)),
                A.val.ctypes.data_as(O._dp))
    m = "Jacobi" if method == "Solver_Jacobi_" else method
    code, x, it, tol, h = O.solve(m, A, np.ones(n), x0=np.ones(n),
                                  max_iter=100000, tol=1e-6)
    key = f"bus_{method}"
    assert code == int(gold[key + "_code"])
    assert it == int(gold[key + "_iters"])
    assert tol == float(gold[key + "_tol"])
    assert np.array_equal(x, gold[key + "_x"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("method", ["CG", "JPCG", "Dia_BiCGSTAB"])
def test_solvers_bitwise_vs_compiled_reference(method):
    # Neumann + Dirichlet mixed, random rhs, warm start: every iterate bitwise
    g = O.StructGrid(24, 17, 1)
    rng = np.random.default_rng(7)
    src = g.field()
    g.interior(src)[:] = rng.standard_normal((24, 17))
    bc = O.make_bc(g, [1, 2, 2, 1], [0.3, 0.0, 0.0, -1.0])
    A = O.Matrix(g, bc, 1.0, src)
    x0 = rng.standard_normal(A.n)
    a = O.solve(method, A, A.b, x0=x0, max_iter=5000, tol=1e-11)
    b = O.ref_solve(method, A, A.b, x0=x0, max_iter=5000, tol=1e-11)
    assert a[0] == b[0] and a[2] == b[2] and a[3] == b[3]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[4], b[4])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
def test_jacobi_returns_previous_iterate_like_reference():
    g, A, X, Y = c1_problem(8)
    a = O.solve("Jacobi", A, A.b, max_iter=10000, tol=1e-6)
    b = O.ref_solve("Solver_Jacobi_", A, A.b, max_iter=10000, tol=1e-6)
    assert np.array_equal(a[1], b[1]) and a[2] == b[2]
    # the converged call returns x_{k-1}: the same vector a max_iter=k-1 run
    # (which copies every iterate) ends with
    code, x_prev, *_ = O.solve("Jacobi", A, A.b, max_iter=a[2] - 1, tol=1e-6)
    assert code == 1 and np.array_equal(x_prev, a[1])


def test_mixed_precision_reaches_fp64_tolerance():
    g, A, X, Y = c1_problem(64)
    code, x, outer, inner, res = O.jpcg_mixed(A, A.b, tol=1e-10, inner_tol=1e-4)
    assert code == 0 and res <= 1e-10 and outer >= 2
    xd = O.solve("JPCG", A, A.b, tol=1e-13, norm=O.NORM_TRUE)[1]
    assert np.max(np.abs(x - xd)) < 1e-7 * np.max(np.abs(xd))


def test_threaded_baseline_agrees_and_one_thread_stays_serial():
    # The CPU baseline runs the oracle on many threads: row loops and
    # vector updates split over threads compute every value as before; dots
    # and norms sum chunks in parallel (another order).  One thread must
    # reproduce the serial result bit for bit.
    n = 96
    g = O.StructGrid(n, n, 1)
    A = O.Matrix(g, O.make_bc(g, [1, 2, 1, 2]), 1.0)
    b = np.random.default_rng(4).standard_normal(n * n)
    ser = O.solve("JPCG", A, b, max_iter=5000, tol=1e-10)
    try:
        O.set_threads(4)
        par = O.solve("JPCG", A, b, max_iter=5000, tol=1e-10)
    finally:
        O.set_threads(1)
    again = O.solve("JPCG", A, b, max_iter=5000, tol=1e-10)
    assert ser[0] == par[0] == 0 and abs(ser[2] - par[2]) <= 2
    assert np.linalg.norm(par[1] - ser[1]) <= 1e-8 * np.linalg.norm(ser[1])
    assert np.array_equal(again[1], ser[1]) and again[2] == ser[2]
