"""Pins of the flexible-CG oracle (oracle/fcg.py) on the CPU."""
import numpy as np

import oracle as O
import oracle.fcg as F


def problem(n, types):
    og = O.StructGrid(n, n, 1)
    A = O.Matrix(og, O.make_bc(og, types), 1.0)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(A.n)
    if all(t == 2 for t in types):
        b -= b.mean()
    return A, F.csr(A), b


def test_fcg_with_jacobi_is_the_oracle_jpcg():
    # FCG(1) with a fixed SPD preconditioner is PCG: with M = D^-1 it takes
    # the oracle's Jacobi-PCG iterates
    A, S, b = problem(48, [1, 2, 1, 2])
    d = S.diagonal()
    x, hist = F.fcg(S, b, np.zeros(A.n), lambda r: r / d, tol=1e-10,
                    max_iter=5000, norm_ref=False)
    code, xo, it, tol, ho = O.solve("JPCG", A, b, tol=1e-10, norm=O.NORM_TRUE,
                                    max_iter=5000)
    assert abs(len(hist) - it) <= 1
    m = min(len(hist), it)
    assert np.allclose(hist[:m], ho[:m], rtol=1e-6)
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)


def test_smoothing_cycle_beats_jacobi_and_converges_on_neumann():
    A, S, b = problem(64, [2, 2, 2, 2])
    d = S.diagonal()
    _, hj = F.fcg(S, b, np.zeros(A.n), lambda r: r / d, tol=1e-8, max_iter=5000)
    _, hc = F.fcg(S, b, np.zeros(A.n), lambda r: F.jacobi_cycle(S, d, r, 2, 2, 0.8),
                  tol=1e-8, max_iter=5000)
    assert hc[-1] <= 1e-8 and hj[-1] <= 1e-8
    # four sweeps per application: about half the Jacobi-PCG iterations
    assert len(hc) < 0.7 * len(hj)
