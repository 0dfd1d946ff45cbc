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


def tg_errors(n, dt, steps):
    g, ns, X, Y, t = taylor_green(n, dt, steps)
    ue, ve, pe = tg_exact(X, Y, t)
    u, v = g.interior(ns.f["u"]), g.interior(ns.f["v"])
    amp = np.exp(-2 * t)
    emax = max(np.max(np.abs(u - ue)), np.max(np.abs(v - ve))) / amp
    el2 = np.sqrt(np.mean((u - ue) ** 2 + (v - ve) ** 2)) / amp
    E0 = np.sum(tg_exact(X, Y, 0.0)[0] ** 2 + tg_exact(X, Y, 0.0)[1] ** 2)
    E = np.sum(u ** 2 + v ** 2) / E0 / np.exp(-4 * t) - 1
    # pressure up to its constant, away from the walls (the Neumann
    # projection leaves its pressure error in the wall rows)
    p = g.interior(ns.f["p"])
    pi, pei = p[2:-2, 2:-2], pe[2:-2, 2:-2]
    pd, ped = pi - pi.mean(), pei - pei.mean()
    ep = np.linalg.norm(pd - ped) / np.linalg.norm(ped)
    return g, ns, t, emax, el2, E, ep


def test_taylor_green_decay():
    # The predictor's walls are Neumann (its Poisson_ has no boundary
    # conditions of its own) while u's ghosts are the exact Dirichlet values:
    # the wall mismatch keeps the scheme near first order at the walls.
    # Measured at 32^2: max 3.5%, rms 1.4%, energy 1.5%, pressure 14%.
    g, ns, t, emax, el2, E, ep = tg_errors(32, 0.01, 20)
    assert emax < 0.05, emax
    assert el2 < 0.02, el2
    assert abs(E) < 0.025, E
    assert ep < 0.2, ep
    assert ns.st.t == pytest.approx(t)
    assert all(a >= 1 and b >= 1 for a, b in ns.helm_iters)
    # refinement (h and dt halved) shrinks the rms error (1.8x measured)
    _, _, _, emax16, el2_16, _, _ = tg_errors(16, 0.02, 10)
    assert el2_16 / el2 > 1.5, (el2_16, el2)


def test_kim_predictor_rows_are_all_neumann(planted):
    # s_ns_kim.hpp:105 builds the predictor as a Poisson_ of its own, whose
    # empty BoundaryIndex (s_equation.hpp:104) answers every wall with the
    # default BC2, value 0 (s_boundary.hpp:29-36, 195-207).  On square
    # cells of side h its rows are alpha h^2 I + beta L_N, L_N the Neumann
    # graph Laplacian (diagonal = number of interior neighbours), unknowns
    # i fastest; a Dirichlet wall would put 3 (edge) or 6 (corner) there.
    n, m, h = 6, 5, 0.1
    g = O.StructGrid(n, m, 2, 0.0, n * h, 0.0, m * h)
    bc_lid = O.make_bc(g, [O.BC1] * 4, [0, 0, 0, 1.0])
    dt, rho = 0.01, 2.0
    beta = dt * 0.5 / rho

    def lap1(k):
        T = 2 * np.eye(k) - np.eye(k, k=1) - np.eye(k, k=-1)
        T[0, 0] = T[-1, -1] = 1.0
        return T
    want = h * h * np.eye(n * m) + beta * (np.kron(np.eye(m), lap1(n)) +
                                           np.kron(lap1(m), np.eye(n)))
    ns = O.NSKim(g, bc_lid, O.make_bc(g, [O.BC1] * 4), rho=rho, mu=1.0, dt=dt)
    for H in (ns.Hu, ns.Hv):
        assert np.allclose(H.dense(), want, rtol=0, atol=1e-15)
    planted("kim_pred_walls")
    ns = O.NSKim(g, bc_lid, O.make_bc(g, [O.BC1] * 4), rho=rho, mu=1.0, dt=dt)
    assert not np.allclose(ns.Hu.dense(), want, rtol=0, atol=1e-15)


def test_kim_ab2_reads_the_previous_steps_ghosts(planted):
    # Step 1 of NS_kim_: um is u copied before apply_bc(u) refreshes u's
    # ghosts (s_ns_kim.hpp:132, :225-227), so the AB2 term's old level reads
    # the ghost values u came in with.  Cell (0, 5) of a 10^2 grid, h = 0.1:
    # u = 1 inside, ghosts 7, wall value 0 (fresh ghost 2*0 - 1 = -1),
    # uf = 0.5, vf = 0, mu = 0.01, rho = 1, dt = 1e-3:
    #   advn   = 0.5 (1 - (-1)) / 0.1 = 10
    #   advnm1 = 0.5 (1 - 7) / 0.1    = -30
    #   adv    = 1.5 * 10 - 0.5 * (-30) = 30
    #   dif    = 0.01 ((1 - 1)/0.1 - (1 - (-1))/0.1) / 0.1 = -2
    #   src    = (adv - 0.5 dif / rho) dt - u = (30 + 1) 1e-3 - 1 = -0.969
    # (refreshed ghosts for both levels: adv = 10, src = -0.989)
    def source():
        g = O.StructGrid(10, 10, 2, 0.0, 1.0, 0.0, 1.0)
        ns = O.NSKim(g, O.make_bc(g, [O.BC1] * 4), O.make_bc(g, [O.BC1] * 4),
                     rho=1.0, mu=0.01, dt=1e-3)
        ns.f["u"][:] = 7.0
        g.interior(ns.f["u"])[:] = 1.0
        ns.f["uf"][:] = 0.5
        ns.nstep = 1
        ns.step()
        return ns.kim_src[0][2 + 0, 2 + 5]
    assert source() == pytest.approx(-0.969, rel=1e-12)
    planted("kim_ab2_ghosts")
    assert source() == pytest.approx(-0.989, rel=1e-12)


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
