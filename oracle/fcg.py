"""CPU restatement of the flexible-CG solve with the two-level preconditioner
cycle (carpios_amd/csrc/fcg.hip), for checking the device path.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  No reference
counterpart (the reference's Krylov solvers are lib/algebra/solver_matrix.hpp);
parity is pinned by the classical identity that FCG(1) with a fixed
symmetric positive definite preconditioner is preconditioned CG: with the
Jacobi preconditioner its iterates match the oracle's J-PCG
(tests/test_fcg_oracle.py).

Vectors are in the oracle's x-fastest unknown order; ``A`` is an oracle
Matrix.
"""
import numpy as np


def csr(A):
    import scipy.sparse as sp
    return sp.csr_matrix((A.val, A.col, A.row_ptr), shape=(A.n, A.n))


def f32(v):
    return np.asarray(v, dtype=np.float32).astype(np.float64)


def jacobi_cycle(S, d, r, pre=2, post=2, omega=0.8, coarse=None):
    """z = M(r): `pre` weighted-Jacobi sweeps from zero, plus coarse(r - A z)
    (the network; None = 0), then `post` sweeps; FP32 like the device."""
    r = f32(r)
    z = np.zeros_like(r)
    for _ in range(pre):
        z = f32(z + f32(omega * f32(f32(r - S @ z) / d)))
    if coarse is not None:
        z = f32(z + f32(coarse(f32(r - S @ z))))
    for _ in range(post):
        z = f32(z + f32(omega * f32(f32(r - S @ z) / d)))
    return z


def fcg(S, b, x0, precond, tol=1e-8, max_iter=1000, norm_ref=False):
    """FCG(1): p_k = z_k - (z_k.q_{k-1})/(p_{k-1}.q_{k-1}) p_{k-1},
    alpha = (p.r)/(p.q).  Returns (x, residual history ||r_k|| / ||b||)."""
    x = np.array(x0, dtype=np.float64)
    r = b - S @ x
    nrm = (lambda v: np.linalg.norm(v[:-1])) if norm_ref else np.linalg.norm
    nb = nrm(b) or 1.0
    hist = []
    p = q = None
    pq = None
    for _ in range(max_iter):
        if nrm(r) / nb <= tol:
            break
        z = precond(r)
        p = z.copy() if p is None else z - (z @ q) / pq * p
        q = S @ p
        pq = p @ q
        alpha = (p @ r) / pq
        x = x + alpha * p
        r = r - alpha * q
        hist.append(nrm(r) / nb)
    return x, np.array(hist)
