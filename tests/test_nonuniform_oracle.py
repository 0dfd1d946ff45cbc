"""Grids whose cell sizes vary along both axes (structure::Grid_ built from
size arrays, s_grid.hpp:397-434): the axis construction, and the oracle's
Poisson rows on them (s_poisson.hpp:98-135 with the distances between the
accumulated centres, s_ghost.hpp:140-179) pinned by symmetry, exactness
for linear fields with Dirichlet and nonzero Neumann walls, and consistency
with the NS step's own Grad and Div (s_operation.hpp:46-120)."""
import numpy as np
import pytest

import carpios_amd as ca
import oracle as O


def stretched(n, ratio, lo=0.0, L=1.0):
    """n sizes growing geometrically by `ratio`, scaled to length L."""
    s = ratio ** np.arange(n)
    return L * s / s.sum()


def grid(nx=14, ny=11, gl=1):
    return O.StructGrid(nx, ny, gl, 0.0, None, 0.0, None,
                        xsizes=stretched(nx, 1.15), ysizes=stretched(ny, 0.9, L=0.7))


def test_axis_from_sizes_matches_the_library_and_the_uniform_axis():
    for n, gl, v0 in ((7, 1, 0.0), (12, 2, -0.5), (3, 3, 2.0)):
        sizes = stretched(n, 1.3, L=1.7)
        a = O.SizedAxis(sizes, gl, v0)
        c, s = ca.axis_from_sizes(sizes, gl, v0)
        assert np.array_equal(c, a.c) and np.array_equal(s, a.s)
        # ghosts repeat the end sizes; centres are the cumulative sums
        assert np.all(a.s[:gl] == sizes[0]) and np.all(a.s[gl + n:] == sizes[-1])
        edges = v0 + np.concatenate([[0.0], np.cumsum(sizes)])
        assert np.allclose(a.c[gl:gl + n], 0.5 * (edges[:-1] + edges[1:]), rtol=0, atol=1e-14)
        # equal sizes give the uniform axis bit for bit
        v1 = v0 + 1.7
        u = O.Axis(n, gl, v0, v1)
        e = O.SizedAxis(np.full(n, (v1 - v0) / n), gl, v0)
        assert np.array_equal(u.c, e.c) and np.array_equal(u.s, e.s)


def test_stretched_rows_are_symmetric_and_exact_for_linear_fields():
    g = grid()
    A = O.Matrix(g, O.make_bc(g, [1, 1, 1, 1]), 1.0)
    M = A.dense()
    assert np.allclose(M, M.T, rtol=1e-13, atol=0)
    # phi = a x + b y + c: Dirichlet on x-, Neumann (val = -dphi/dn) elsewhere
    a, b, c = 0.7, -1.3, 0.4
    NX, NY = g.shape
    X, Y = g.centres()
    xf = g.ax.c[g.gl] - 0.5 * g.ax.s[g.gl]  # the x- wall
    vals = np.concatenate([a * xf + b * g.ay.c + c, np.full(NY, -a), np.full(NX, b),
                           np.full(NX, -b)])
    bc = O.make_bc(g, [O.BC1, O.BC2, O.BC2, O.BC2], vals)
    A = O.Matrix(g, bc, 1.0)
    x = np.linalg.solve(A.dense(), A.b)
    want = O.field_to_x(g, np.pad(a * X + b * Y + c, g.gl))
    assert np.max(np.abs(x - want)) < 1e-12


def test_stretched_rows_are_minus_volume_times_div_grad():
    g = grid(15, 12, 2)
    rng = np.random.default_rng(4)
    phi = g.field()
    g.interior(phi)[:] = rng.standard_normal((15, 12))
    O.apply_bc(g, phi, O.make_bc(g, [2, 2, 2, 2]))
    gx, gy = O.grad(g, phi)
    dv = g.interior(O.div(g, gx, gy))
    G = g.gl
    vol = np.outer(g.ax.s[G:G + 15], g.ay.s[G:G + 12])
    A = O.Matrix(g, O.make_bc(g, [2, 2, 2, 2]), 1.0)
    lhs = O.field_to_x(g, np.pad(dv * vol, G))
    rhs = -A.mult(O.field_to_x(g, phi))
    assert np.allclose(lhs, rhs, rtol=0, atol=1e-12 * np.abs(rhs).max())
