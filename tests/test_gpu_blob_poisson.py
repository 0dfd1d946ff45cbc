"""The precision configuration on the device: an all-Neumann Poisson
problem whose right-hand side is 64 zero-meaned Gaussian blobs (seed 0),
solved by FP32 Jacobi-PCG inside FP64 residual replacement (CPGPU_MIXED)
with the zero-mean projection, against plain FP64 Jacobi-PCG and, at
256^2, against the oracle.  The projection singularity it models is the
reference's all-Neumann pressure Poisson (s_poisson.hpp:349-351,
s_ns.hpp:222)."""
import time

import numpy as np
import pytest

import carpios_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu


def zm_rel(a, b):
    a, b = a - a.mean(), b - b.mean()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def blob_problem(ctx, n):
    P = ca.Poisson(ctx, ca.uniform_grid_desc(n, n, 1, [ca.BC_NEUMANN] * 4))
    P.set_source_blobs(seed=0, nblobs=64, wmin=0.01, wmax=0.1)
    P.fill_phi(0.0)
    return P


def test_blob_256_mixed_matches_oracle(ctx):
    n = 256
    P = blob_problem(ctx, n)
    og = O.StructGrid(n, n, 1)
    b = O.field_to_x(og, P.get_vec(1))
    A = O.Matrix(og, O.make_bc(og, [2] * 4), 1.0)
    code_o, x_o, it_o, res_o, _ = O.solve("JPCG", A, b - b.mean(), tol=1e-13,
                                          max_iter=100000, norm=O.NORM_TRUE)
    assert code_o == 0
    code, it, res, _ = P.solve("MIXED", max_iter=200000, tol=1e-10,
                               norm=ca.NORM_TRUE, opts=ca.OPT_PROJECT_MEAN)
    x = O.field_to_x(og, P.get_phi())
    err = zm_rel(x, x_o)
    print(f"blob 256^2: mixed {it} inner iterations, residual {res:.2e}, "
          f"vs oracle 1e-13 FP64 J-PCG ({it_o} it): rel L2 {err:.2e}", flush=True)
    assert code == 0 and res <= 1e-10 and P.residual() <= 1e-10
    assert abs(x.mean()) <= 1e-12 * np.abs(x).max()
    assert err <= 1e-8
    P.close()


def test_blob_4096_mixed_against_fp64(ctx):
    n = 4096
    P = blob_problem(ctx, n)
    ctx.sync()
    t0 = time.perf_counter()
    code64, it64, res64, _ = P.solve("JPCG", max_iter=10 ** 6, tol=1e-10,
                                     norm=ca.NORM_TRUE, opts=ca.OPT_PROJECT_MEAN)
    ctx.sync()
    t64 = time.perf_counter() - t0
    x64 = P.get_phi()
    print(f"blob 4096^2 FP64 J-PCG: {it64} iterations, {t64:.2f} s, "
          f"residual {res64:.2e}", flush=True)
    assert code64 == 0 and P.residual() <= 1e-10
    P.fill_phi(0.0)
    ctx.sync()
    t0 = time.perf_counter()
    code, it, res, _ = P.solve("MIXED", max_iter=10 ** 6, tol=1e-10,
                               norm=ca.NORM_TRUE, opts=ca.OPT_PROJECT_MEAN)
    ctx.sync()
    t32 = time.perf_counter() - t0
    x = P.get_phi()
    true_res = P.residual()
    err = zm_rel(x[1:-1, 1:-1], x64[1:-1, 1:-1])
    print(f"blob 4096^2 FP32 + FP64 refinement: {it} iterations, {t32:.2f} s, "
          f"true residual {true_res:.2e}, rel L2 vs FP64 {err:.2e}", flush=True)
    assert code == 0 and true_res <= 1e-10
    inner = x[1:-1, 1:-1]
    assert abs(inner.mean()) <= 1e-12 * np.abs(inner).max()
    assert err <= 1e-8
    P.close()
