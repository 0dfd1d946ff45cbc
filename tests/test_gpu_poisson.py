"""Parity of the HIP Poisson path (through the C-ABI) with the CPU oracle."""
import numpy as np
import pytest

import carpios_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu


def both(nx, ny, gl, types, vals=None, src_fn=None, seed=0, lims=((0, 1), (0, 1))):
    og = O.StructGrid(nx, ny, gl, *lims[0], *lims[1])
    bc = O.make_bc(og, types, vals)
    src = og.field()
    if src_fn is None:
        rng = np.random.default_rng(seed)
        og.interior(src)[:] = rng.standard_normal((nx, ny))
    else:
        X, Y = og.centres()
        og.interior(src)[:] = src_fn(X, Y)
    A = O.Matrix(og, bc, 1.0, src)
    desc = ca.uniform_grid_desc(nx, ny, gl, types, bc._keep, 1.0,
                                xlim=lims[0], ylim=lims[1])
    return og, A, src, desc


CASES = [
    (128, 128, 1, [1, 1, 1, 1]),
    (37, 53, 1, [1, 2, 2, 1]),
    (130, 65, 2, [2, 2, 2, 2]),
    (1, 1, 1, [1, 1, 1, 1]),
    (2, 3, 1, [2, 1, 2, 1]),
    (64, 200, 1, [2, 2, 1, 2]),
]


@pytest.mark.parametrize("nx,ny,gl,types", CASES)
def test_operator_and_rhs_bitwise(ctx, nx, ny, gl, types):
    vals = [0.25, -1.5, 0.0, 2.0]
    og, A, src, desc = both(nx, ny, gl, types, vals)
    P = ca.Poisson(ctx, desc)
    P.set_source(src)
    b = og.interior(P.get_vec(1))
    assert np.array_equal(O.field_to_x(og, np.pad(b, gl)), A.b)
    rng = np.random.default_rng(1)
    xf = og.field()
    og.interior(xf)[:] = rng.standard_normal((nx, ny))
    y = P.apply(xf)
    assert np.array_equal(O.field_to_x(og, y), A.mult(O.field_to_x(og, xf)))
    P.close()


def c1(n):
    return both(n, n, 1, [1, 1, 1, 1],
                src_fn=lambda X, Y: -2 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y))


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("norm", [ca.NORM_REF, ca.NORM_TRUE])
def test_c1_jpcg_parity(ctx, norm):
    n = 128
    og, A, src, desc = c1(n)
    code_o, x_o, it_o, tol_o, h_o = O.solve("JPCG", A, A.b, tol=1e-10, norm=norm)
    P = ca.Poisson(ctx, desc)
    P.set_source(src)
    P.set_phi(og.field())
    code, it, tol, hist = P.solve("JPCG", max_iter=10000, tol=1e-10, norm=norm,
                                  hist_cap=10000)
    x = O.field_to_x(og, P.get_phi())
    print(f"C1 128^2 JPCG: gpu {it} it {tol:.3e}  oracle {it_o} it {tol_o:.3e}"
          f"  relL2 {rel_l2(x, x_o):.2e}")
    assert code == 0 == code_o
    assert abs(it - it_o) <= 1
    assert rel_l2(x, x_o) <= 1e-12
    assert len(hist) == it
    m = min(it, it_o)
    assert np.allclose(hist[:m], h_o[:m], rtol=1e-6, atol=0)
    # against the exact discrete solution
    X, Y = og.centres()
    h = 1.0 / n
    exact = np.pi ** 2 * h ** 2 / (4 * np.sin(np.pi * h / 2) ** 2) * np.sin(np.pi * X) * np.sin(np.pi * Y)
    assert np.max(np.abs(og.interior(P.get_phi()) - exact)) < 1e-8
    P.close()


def test_c1_cg_single_iteration(ctx):
    og, A, src, desc = c1(128)
    P = ca.Poisson(ctx, desc)
    P.set_source(src)
    code, it, tol, _ = P.solve("CG", tol=1e-10)
    assert code == 0 and it == 1
    x_o = O.solve("CG", A, A.b, tol=1e-10)[1]
    assert rel_l2(O.field_to_x(og, P.get_phi()), x_o) <= 1e-14
    P.close()


@pytest.mark.parametrize("nx,ny,gl,types", CASES[1:])
def test_jpcg_parity_mixed_boundaries(ctx, nx, ny, gl, types):
    og, A, src, desc = both(nx, ny, gl, types, [0.5, 0.0, 0.0, -1.0], seed=5)
    rng = np.random.default_rng(9)
    x0 = rng.standard_normal(A.n)
    if all(t == 2 for t in types):
        # all-Neumann: shift the source by the constant that makes b sum to
        # zero (b = -(v*src + ghost constants) with the source weight
        # v = sx*sx of StencilPoisson_, s_poisson.hpp:71-77), so that the
        # singular system is consistent and Jacobi-PCG converges as on the
        # others
        vol = og.ax.s[gl] * og.ax.s[gl]
        og.interior(src)[:] += A.b.sum() / (A.n * vol)
        A = O.Matrix(og, O.make_bc(og, types, [0.5, 0.0, 0.0, -1.0]), 1.0, src)
        assert abs(A.b.sum()) < 1e-12 * np.abs(A.b).sum()
    code_o, x_o, it_o, tol_o, _ = O.solve("JPCG", A, A.b, x0=x0, tol=1e-11,
                                          max_iter=20000)
    P = ca.Poisson(ctx, desc)
    P.set_source(src)
    P.set_phi(O.x_to_field(og, x0))
    code, it, tol, _ = P.solve("JPCG", max_iter=20000, tol=1e-11)
    x = O.field_to_x(og, P.get_phi())
    assert code == code_o
    assert abs(it - it_o) <= 1
    assert rel_l2(x, x_o) <= 1e-12
    P.close()


def test_solves_are_bitwise_deterministic(ctx):
    og, A, src, desc = both(200, 150, 1, [1, 2, 1, 2], [0, 0, 1.0, 0], seed=2)
    out = []
    for _ in range(2):
        P = ca.Poisson(ctx, desc)
        P.set_source(src)
        P.solve("JPCG", tol=1e-10)
        out.append(P.get_phi())
        P.close()
    assert np.array_equal(out[0], out[1])


def test_maxiter_code_and_reference_tol_semantics(ctx):
    og, A, src, desc = c1(64)
    P = ca.Poisson(ctx, desc)
    P.set_source(src)
    code, it, tol, hist = P.solve("JPCG", max_iter=5, tol=1e-12, hist_cap=16)
    code_o, x_o, it_o, tol_o, h_o = O.solve("JPCG", A, A.b, max_iter=5, tol=1e-12)
    assert code == 1 == code_o and it == 5 == it_o
    assert abs(tol - tol_o) <= 1e-12 * tol_o
    P.close()


def test_mixed_precision_refinement(ctx):
    og, A, src, desc = c1(256)
    P = ca.Poisson(ctx, desc)
    P.set_source(src)
    P.set_mixed(1e-4, 40)
    code, inner, res, hist = P.solve("MIXED", max_iter=200000, tol=1e-10,
                                     hist_cap=64)
    assert code == 0 and res <= 1e-10
    assert P.residual() <= 1e-10  # true FP64 residual recomputed
    x_ref = O.solve("JPCG", A, A.b, tol=1e-13, norm=O.NORM_TRUE)[1]
    err = rel_l2(O.field_to_x(og, P.get_phi()), x_ref)
    print(f"mixed 256^2: {inner} inner iterations, {len(hist)} sweeps, err {err:.2e}")
    assert err < 1e-7
    P.close()


def test_all_neumann_projected_mixed(ctx):
    n = 96
    og = O.StructGrid(n, n, 1)
    rng = np.random.default_rng(11)
    src = og.field()
    f = rng.standard_normal((n, n))
    og.interior(src)[:] = f - f.mean()
    desc = ca.uniform_grid_desc(n, n, 1, [2, 2, 2, 2])
    P = ca.Poisson(ctx, desc)
    P.set_source(src)
    code, inner, res, _ = P.solve("MIXED", max_iter=100000, tol=1e-10,
                                  opts=ca.OPT_PROJECT_MEAN)
    assert code == 0 and res <= 1e-10
    phi = og.interior(P.get_phi())
    assert abs(phi.mean()) < 1e-12 * np.abs(phi).max()
    P.close()


def test_one_shot_solve_matches_problem_object(ctx):
    og, A, src, desc = c1(64)
    phi = og.field()
    import ctypes as C
    mi, t = C.c_int(10000), C.c_double(1e-10)
    hl = C.c_int(0)
    code = ca.lib().cpgpu_poisson2d_solve(ctx.h, C.byref(desc), ca._p(src),
                                          C.c_double(0), ca._p(phi), ca.JPCG,
                                          C.byref(mi), C.byref(t), ca.NORM_REF,
                                          None, 0, C.byref(hl))
    assert code == 0
    x_o = O.solve("JPCG", A, A.b, tol=1e-10)[1]
    assert rel_l2(O.field_to_x(og, phi), x_o) <= 1e-12


@pytest.mark.parametrize("nx,ny", [(37, 53), (130, 200), (256, 64), (300, 130)])
def test_direction_kernel_variants_bitwise_equal(ctx, nx, ny):
    og, A, src, desc = both(nx, ny, 1, [1, 2, 1, 2], [0.3, 0, 0, -0.2], seed=3)
    sols = []
    for variant in range(17):
        P = ca.Poisson(ctx, desc)
        P.set_variant(variant, variant)
        P.set_source(src)
        code, it, tol, _ = P.solve("JPCG", tol=1e-11)
        sols.append((it, P.get_phi()))
        P.close()
    # the operator is the same arithmetic; the block partition of the dot
    # products differs between variants, so results agree to rounding
    x0 = O.field_to_x(og, sols[0][1])
    for it, phi in sols[1:]:
        assert abs(it - sols[0][0]) <= 1
        assert rel_l2(O.field_to_x(og, phi), x0) <= 1e-12


@pytest.mark.parametrize("nx,ny", [(300, 128), (130, 1030), (70, 2048)])
def test_block_march_variants_in_fp32_and_fp64(ctx, nx, ny):
    # the block-wide march (variants 11..16, 16-byte vectors: 2 FP64 or 4
    # FP32 columns per lane, 512 / 1024 columns per block; 15 and 16 with
    # non-temporal stores in it and in the update kernel) against the
    # streaming march of variant 9, in the FP64 iteration and in the FP32
    # iterations of the mixed solve; ny = 1030 takes the 16-byte path in
    # FP64 and the fallback in FP32, ny = 2048 spans several blocks
    og, A, src, desc = both(nx, ny, 1, [1, 2, 1, 2], [0.3, 0, 0, -0.2], seed=5)
    out = {}
    for variant in (9, 11, 12, 13, 14, 15, 16):
        P = ca.Poisson(ctx, desc)
        P.set_variant(variant, variant)
        P.set_source(src)
        code, it, tol, _ = P.solve("JPCG", tol=1e-11)
        assert code == 0
        x64 = O.field_to_x(og, P.get_phi())
        P.fill_phi(0.0)
        code, itm, tolm, _ = P.solve("MIXED", tol=1e-10, norm=ca.NORM_TRUE)
        assert code == 0 and tolm <= 1e-10
        out[variant] = (it, x64, itm, O.field_to_x(og, P.get_phi()))
        P.close()
    it9, x9, itm9, xm9 = out[9]
    for variant in (11, 12, 13, 14, 15, 16):
        it, x64, itm, xm = out[variant]
        assert abs(it - it9) <= 1
        assert rel_l2(x64, x9) <= 1e-12
        # the FP32 recurrences differ in the order of their dot products
        # only, and their stop is checked every 16 iterations; across the
        # eleven older variants the mixed solve of the 300 x 128 problem
        # takes 1008 or 1152 iterations (a restart after a residual
        # replacement, or none), so counts are held to 20%
        assert abs(itm - itm9) <= max(32, itm9 // 5)
        assert rel_l2(xm, x9) <= 1e-6 and rel_l2(xm9, x9) <= 1e-6


def test_autotune_keeps_phi_and_picks_a_variant(ctx):
    og, A, src, desc = c1(256)
    P = ca.Poisson(ctx, desc)
    P.set_source(src)
    phi0 = og.field()
    og.interior(phi0)[:] = 0.125
    P.set_phi(phi0)
    res = P.autotune(reps=3)
    assert 0 <= res["fp64"][0] <= 16 and 0 <= res["fp32"][0] <= 16
    assert np.array_equal(og.interior(P.get_phi()), og.interior(phi0))
    P.close()


@pytest.mark.parametrize("method", ["MIXED", "MIXED_RESTART"])
@pytest.mark.parametrize("types", [[1, 1, 1, 1], [2, 1, 2, 1]])
def test_mixed_methods_reach_fp64_tolerance(ctx, method, types):
    og, A, src, desc = both(300, 200, 1, types, [0.0, 0.5, 0.0, -0.25], seed=8)
    P = ca.Poisson(ctx, desc)
    P.set_source(src)
    code, it, res, hist = P.solve(method, max_iter=400000, tol=1e-10,
                                  norm=ca.NORM_TRUE, hist_cap=256)
    assert code == 0 and res <= 1e-10
    assert P.residual() <= 1e-10
    x_ref = O.solve("JPCG", A, A.b, tol=1e-13, max_iter=100000, norm=O.NORM_TRUE)[1]
    err = rel_l2(O.field_to_x(og, P.get_phi()), x_ref)
    it64 = O.solve("JPCG", A, A.b, tol=1e-10, max_iter=100000, norm=O.NORM_TRUE)[2]
    print(f"{method} {types}: {it} FP32 iterations (FP64 J-PCG {it64}), "
          f"err {err:.2e}, residual history {len(hist)}")
    assert err < 1e-7
    if method == "MIXED":
        # no restarts: the iteration count stays near the FP64 one
        assert it <= 1.5 * it64 + 32
    P.close()
