"""Pins of the oracle's NS_kim_ restatement (s_ns_kim.hpp:130-287): the
Helmholtz rows of Poisson_::set_uniform_alpha against the discrete closed
form, the upwind advection against exact one-sided differences, and the
whole step against the Taylor-Green vortex of example/ns_sin/main.cpp:40-51."""
import numpy as np
import pytest

import oracle as O


def test_helmholtz_closed_form_rectangular_cells():
    # -beta lap_h phi + alpha phi = -f on Dirichlet-0 walls: sin(pi x/Lx)
    # sin(pi y/Ly) at the centres is an eigenvector of the discrete operator
    # (ghost = -phi_0), eigenvalue 4 sin^2(th/2)/h^2 per axis.  Rectangular
    # cells (hx != hy) separate v = hx*hy from the Poisson row's hx*hx.
    nx, ny, Lx, Ly = 24, 16, 2.0, 0.5
    g = O.StructGrid(nx, ny, 1, 0.0, Lx, 0.0, Ly)
    X, Y = g.centres()
    hx, hy = Lx / nx, Ly / ny
    mode = np.sin(np.pi * X / Lx) * np.sin(np.pi * Y / Ly)
    bc = O.make_bc(g, [O.BC1] * 4)
    for alpha, beta in ((1.0, 0.05), (-3.0, 0.2), (7.5, 1e-3)):
        f = g.field()
        g.interior(f)[:] = mode
        M = O.Matrix(g, bc, beta=beta, src_field=f, alpha=alpha)
        x = np.linalg.solve(M.dense(), M.b)
        lam = (4 * np.sin(np.pi * hx / Lx / 2) ** 2 / hx ** 2 +
               4 * np.sin(np.pi * hy / Ly / 2) ** 2 / hy ** 2)
        amp = -1.0 / (beta * lam + alpha)
        got = x.reshape(ny, nx).T  # unknowns run i fastest
        assert np.max(np.abs(got - amp * mode)) < 1e-12 * max(1.0, abs(amp))


def test_helmholtz_is_poisson_plus_diagonal_on_square_cells():
    g = O.StructGrid(10, 10, 1, 0.0, 1.0, 0.0, 1.0)
    bc = O.make_bc(g, [O.BC1, O.BC2, O.BC1, O.BC2], [1.0, 0.5, -2.0, 0.25])
    P = O.Matrix(g, bc, beta=0.3)
    H = O.Matrix(g, bc, beta=0.3, alpha=2.0)
    D = H.dense() - P.dense()
    assert np.allclose(D, np.eye(100) * 2.0 * 0.1 * 0.1, rtol=0, atol=1e-14)
    assert np.array_equal(P.b, H.b)


def test_upwind1_one_sided_differences():
    g = O.StructGrid(12, 9, 2, 0.0, 1.2, 0.0, 0.9)
    XF, YF = np.meshgrid(g.ax.c, g.ay.c, indexing="ij")
    hx, hy = 0.1, 0.1
    phi = XF ** 2 + 3.0 * YF ** 2
    X, Y = g.centres()
    for U, V in ((0.7, 0.0), (-0.7, 0.0), (0.0, 1.3), (0.0, -1.3), (0.5, -2.0)):
        uf = np.full(g.shape, U)
        vf = np.full(g.shape, V)
        got = O.upwind1_field(g, uf, vf, phi)
        # backward difference for positive velocity, forward for negative
        dx = 2 * X - hx if U > 0 else 2 * X + hx
        dy = 6 * Y - 3 * hy if V > 0 else 6 * Y + 3 * hy
        want = U * dx * (U != 0) + V * dy * (V != 0)
        assert np.max(np.abs(got - want)) < 1e-12


def tg_exact(X, Y, t):
    u = -np.cos(X) * np.sin(Y) * np.exp(-2.0 * t)
    v = np.sin(X) * np.cos(Y) * np.exp(-2.0 * t)
    p = -0.25 * (np.cos(2 * X) + np.cos(2 * Y)) * np.exp(-4.0 * t)
    return u, v, p


def tg_bc_values(g, t):
    """Packed per-ghost-line Dirichlet values of the exact velocity at the
    wall points (x-, x+ along y; y-, y+ along x)."""
    yc, xc = g.ay.c, g.ax.c
    out = []
    for comp in (0, 1):
        segs = []
        for (x, y) in ((0 * yc, yc), (np.pi + 0 * yc, yc), (xc, 0 * xc),
                       (xc, np.pi + 0 * xc)):
            u, v, _ = tg_exact(x, y, t)
            segs.append(u if comp == 0 else v)
        out.append(np.concatenate(segs))
    return out


def taylor_green(n=32, dt=0.01, steps=20):
    g = O.StructGrid(n, n, 2, 0.0, np.pi, 0.0, np.pi)
    vu, vv = tg_bc_values(g, 0.0)
    bcu = O.make_bc(g, [O.BC1] * 4, vu)
    bcv = O.make_bc(g, [O.BC1] * 4, vv)
    ns = O.NSKim(g, bcu, bcv, rho=1.0, mu=1.0, tol=1e-10, dt=dt)
    X, Y = g.centres()
    u0, v0, _ = tg_exact(X, Y, 0.0)
    g.interior(ns.f["u"])[:] = u0
    g.interior(ns.f["v"])[:] = v0
    XF, YF = np.meshgrid(g.ax.c, g.ay.c, indexing="ij")
    # face velocities of the initial field: x+ face of cell i at x_i + h/2
    h = np.pi / n
    ns.f["uf"][:] = tg_exact(XF + h / 2, YF, 0.0)[0]
    ns.f["vf"][:] = tg_exact(XF, YF + h / 2, 0.0)[1]
    for k in range(steps):
        t1 = (k + 1) * dt
        vu, vv = tg_bc_values(g, t1)
        ns.bcu = O.make_bc(g, [O.BC1] * 4, vu)
        ns.bcv = O.make_bc(g, [O.BC1] * 4, vv)
        ns.step()
    return g, ns, X, Y, steps * dt


def test_taylor_green_decay():
    g, ns, X, Y, t = taylor_green()
    ue, ve, pe = tg_exact(X, Y, t)
    u, v = g.interior(ns.f["u"]), g.interior(ns.f["v"])
    err = max(np.max(np.abs(u - ue)), np.max(np.abs(v - ve)))
    assert err < 0.02 * np.exp(-2 * t), err
    # kinetic energy follows exp(-4t)
    E0 = np.sum(tg_exact(X, Y, 0.0)[0] ** 2 + tg_exact(X, Y, 0.0)[1] ** 2)
    E = np.sum(u ** 2 + v ** 2)
    assert abs(E / E0 / np.exp(-4 * t) - 1) < 0.02
    # pressure up to its constant
    p = g.interior(ns.f["p"])
    # away from the walls: the Neumann projection leaves its pressure error
    # in the wall rows (18% there at this size, 4% two cells in)
    pi, pei = p[2:-2, 2:-2], pe[2:-2, 2:-2]
    pd, ped = pi - pi.mean(), pei - pei.mean()
    assert np.linalg.norm(pd - ped) / np.linalg.norm(ped) < 0.08
    assert ns.st.t == pytest.approx(t)
    assert all(a >= 1 and b >= 1 for a, b in ns.helm_iters)


def test_kim_planted_mistake_is_caught():
    # the same run without viscosity (mu = 0) must miss the decay: the
    # Taylor-Green check above is sensitive to the diffusion terms
    g = O.StructGrid(32, 32, 2, 0.0, np.pi, 0.0, np.pi)
    X, Y = g.centres()
    g2, ns, X, Y, t = taylor_green()
    ns.prm.mu = 0.0
    before = np.sum(g2.interior(ns.f["u"]) ** 2)
    for _ in range(5):
        ns.step()
    after = np.sum(g2.interior(ns.f["u"]) ** 2)
    # with mu = 0 only the implicit half of the diffusion remains
    assert after / before > np.exp(-4 * 5 * 0.01) * 1.02
