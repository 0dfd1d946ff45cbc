"""Pins the oracle's structured operations and the NS_explicit_ step:
exactness on constant/linear/quadratic fields, limiter special values,
Div(Grad) consistency with the assembled Poisson operator, and the
projection property of one step (s_operation.hpp, s_ns_explicit.hpp)."""
import numpy as np
import pytest

import oracle as O


def grid(n=24, m=20, gl=2, lo=-0.5, hi=0.5):
    return O.StructGrid(n, m, gl, lo, hi, lo, hi)


def full_centres(g):
    return np.meshgrid(g.ax.c, g.ay.c, indexing="ij")


@pytest.mark.parametrize("lim", list(O.LIM))
def test_tvd_of_constant_field_is_constant_times_face_divergence(lim):
    g = grid()
    phi = np.full(g.shape, 3.25)
    rng = np.random.default_rng(0)
    uf = rng.standard_normal(g.shape)
    vf = rng.standard_normal(g.shape)
    adv = O.tvd_field(g, uf, vf, phi, lim)
    # (vf(i)*c - vf(i-1)*c)/h: a divergence term times the constant
    hx, hy = g.ax.s[0], g.ay.s[0]
    expect = np.zeros((g.nx, g.ny))
    G = g.gl
    for i in range(g.nx):
        for j in range(g.ny):
            I, J = i + G, j + G
            expect[i, j] = ((uf[I, J] * 3.25 - uf[I - 1, J] * 3.25) / hx +
                            (vf[I, J] * 3.25 - vf[I, J - 1] * 3.25) / hy)
    assert np.allclose(adv, expect, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("lim", ["VanLeer", "superbee", "WAHYD", "QUICK", "CUI"])
def test_tvd_exact_for_linear_field_constant_velocity(lim):
    # r = 1 everywhere and psi(1) = 1 for these limiters: central faces,
    # exact for a linear phi
    g = grid(16, 16)
    X, Y = full_centres(g)
    phi = 0.7 * X - 1.3 * Y + 0.2
    for (a, b) in ((0.9, -0.4), (-0.3, 0.6)):
        uf = np.full(g.shape, a)
        vf = np.full(g.shape, b)
        adv = O.tvd_field(g, uf, vf, phi, lim)
        assert np.allclose(adv, a * 0.7 + b * (-1.3), atol=1e-9)


def test_upwind2_is_second_order_upwind():
    # limiter r: face = C + r/R (D - C) = C + (C - U)/2 on a uniform grid
    g = grid(12, 12)
    X, Y = full_centres(g)
    phi = X ** 2
    uf = np.full(g.shape, 1.0)
    vf = np.zeros(g.shape)
    adv = O.tvd_field(g, uf, vf, phi, "upwind2")
    h = g.ax.s[0]
    Xi = X[g.gl:-g.gl, g.gl:-g.gl]
    # second-order upwind face C + (C-U)/2, with the reference's SMALL in
    # r = (C-U)/((D-C)+SMALL): where D == C (x^2 at +-h/2) the face is C
    def face(U, Cc, D):
        return Cc + (Cc - U) * (D - Cc) / ((D - Cc) + 1e-10) / 2
    G = g.gl
    P = phi[:, G:G + g.ny]
    fp = face(P[G - 1:G - 1 + g.nx], P[G:G + g.nx], P[G + 1:G + 1 + g.nx])
    fm = face(P[G - 2:G - 2 + g.nx], P[G - 1:G - 1 + g.nx], P[G:G + g.nx])
    assert np.allclose(adv, (fp - fm) / h, atol=1e-9)
    # away from the symmetric centre it is the unregularised upwind formula
    far = np.abs(Xi[:, 0]) > 2 * h
    exact = ((Xi ** 2 + (Xi ** 2 - (Xi - h) ** 2) / 2) -
             ((Xi - h) ** 2 + ((Xi - h) ** 2 - (Xi - 2 * h) ** 2) / 2)) / h
    assert np.allclose(adv[far], exact[far], atol=1e-6)


def test_vanleer_falls_back_to_upwind_at_extremum():
    # at a local maximum r < 0 -> psi = 0 -> face value = upwind cell value
    g = grid(9, 9)
    phi = np.zeros(g.shape)
    G = g.gl
    phi[G + 4, :] = 1.0
    uf = np.full(g.shape, 1.0)
    vf = np.zeros(g.shape)
    adv = O.tvd_field(g, uf, vf, phi, "VanLeer")
    h = g.ax.s[0]
    # cell 4 (max, r<0 on its p face): fp = 1, fm from cell 3 (r = 0 -> 0)
    assert adv[4, 3] == pytest.approx((1.0 - 0.0) / h, rel=1e-12)
    # cell 5: fp = phi5 (0) + ..., fm = phi4 = 1 (upwind, limiter 0)
    assert adv[5, 3] == pytest.approx((0.0 - 1.0) / h, rel=1e-12)


def test_diffusion_exact_for_quadratic():
    g = grid(14, 11)
    X, Y = full_centres(g)
    phi = X ** 2 + 2 * Y ** 2
    d = O.diffusion_field(g, 0.5, phi)
    assert np.allclose(d, 0.5 * (2 + 4), atol=1e-8)


def test_div_grad_equals_neumann_poisson_operator():
    g = O.StructGrid(15, 12, 2)
    rng = np.random.default_rng(4)
    phi = g.field()
    g.interior(phi)[:] = rng.standard_normal((15, 12))
    O.apply_bc(g, phi, O.make_bc(g, [2, 2, 2, 2]))
    gx, gy = O.grad(g, phi)
    dv = g.interior(O.div(g, gx, gy))
    A = O.Matrix(g, O.make_bc(g, [2, 2, 2, 2]), 1.0)
    vol = g.ax.s[0] * g.ay.s[0]
    lhs = O.field_to_x(g, np.pad(dv, 2)) * vol
    rhs = -A.mult(O.field_to_x(g, phi))
    assert np.allclose(lhs, rhs, atol=1e-10 * np.abs(rhs).max())


def test_c2f_f2c_linear_and_ghost_slot():
    g = grid(10, 10)
    X, Y = full_centres(g)
    uc, vc = 2 * X + 1, -Y
    uf, vf = O.c2f(g, uc, vc)
    G = g.gl
    h = g.ax.s[0]
    # x faces at i = -1..nx-1 hold the linear value at the face
    fx = 2 * (X[G - 1:G + g.nx, G:G + g.ny] + h / 2) + 1
    assert np.allclose(uf[G - 1:G + g.nx, G:G + g.ny], fx, atol=1e-12)
    cx, cy = O.f2c(g, uf, vf)
    assert np.allclose(g.interior(cx), g.interior(uc), atol=1e-12)


def test_ghost_fill_closed_forms():
    g = O.StructGrid(8, 8, 2)
    f = g.field()
    g.interior(f)[:] = 5.0
    O.apply_bc(g, f, O.make_bc(g, [1, 2, 1, 1], [2.0, 0.5, 0.0, 1.0]))
    G = g.gl
    h = g.ax.s[0]
    # BC1: g1 = 2v - phi0, g2 = 4v - 3 phi0
    assert np.allclose(f[G - 1, G:-G], 2 * 2.0 - 5.0)
    assert np.allclose(f[G - 2, G:-G], 4 * 2.0 - 3 * 5.0)
    # BC2: g = phi0 - v*dl with dl = h and 2h
    assert np.allclose(f[G + 8, G:-G], 5.0 - 0.5 * h)
    assert np.allclose(f[G + 9, G:-G], 5.0 - 0.5 * 2 * h)


def cavity(n, Re=1000.0, steps=0, lim="VanLeer", tol=1e-10):
    g = O.StructGrid(n, n, 2, -0.5, 0.5, -0.5, 0.5)
    bcu = O.make_bc(g, [1, 1, 1, 1], [0, 0, 0, 1.0])
    bcv = O.make_bc(g, [1, 1, 1, 1])
    h = 1.0 / n
    ns = O.NS(g, bcu, bcv, rho=Re, mu=1.0, cfl=0.5, dnum=0.25, dcoef=1 / Re,
              limiter=lim, tol=tol, dt=0.5 * h)
    for _ in range(steps):
        ns.step()
    return g, ns


def test_projection_makes_faces_divergence_free():
    g, ns = cavity(24, steps=3)
    dv = g.interior(O.div(g, ns.f["uf"], ns.f["vf"]))
    dstar = g.interior(ns.f["div"])
    assert np.abs(dv).max() < 1e-7 * max(np.abs(dstar).max(), 1e-300)
    # lid drives the top row forward
    assert g.interior(ns.f["u"])[:, -1].mean() > 0


def test_cavity_dt_follows_restrict_time():
    g, ns = cavity(32, steps=2)
    h = 1.0 / 32
    # dt only shrinks: min(dt0 = 0.5 h, DN/kappa/sum(1/h^2)), CFL not binding
    # while max|u| < 1 (s_ns.hpp:146-169)
    assert ns.dt == min(0.5 * h, 0.25 / (1 / 1000.0) / (2.0 / h ** 2))
    g2 = O.StructGrid(512, 512, 2, -0.5, 0.5, -0.5, 0.5)
    A = O.make_bc(g2, [1, 1, 1, 1])
    prm = O.NSParams(1000.0, 1.0, 0.5, 0.25, 1 / 1000.0, 0, 1, 1e-10, 0)
    st = O.NSState(*([None] * 11), 0.5 / 512, 0.0)
    u = g2.field()
    st.u, st.v = O._ptr(u), O._ptr(u)
    O.lib().orc_ns_restrict_time(O.C.byref(g2.c), O.C.byref(prm), O.C.byref(st))
    h2 = 1.0 / 512
    assert st.dt == 0.25 / (1 / 1000.0) / (1.0 / (h2 * h2) + 1.0 / (h2 * h2))
    assert st.dt < 0.5 * h2


def _stretched(nx=12, ny=7, gl=2):
    """A grid whose cell sizes vary along both axes (sizes written in
    place into the axis arrays the oracle grid points at)."""
    g = O.StructGrid(nx, ny, gl, 0.0, 2.0, 0.0, 1.0)
    g.ax.s[:] = 0.1 + 0.01 * np.arange(nx + 2 * gl)
    g.ay.s[:] = 0.2 + 0.03 * np.arange(ny + 2 * gl)
    return g


def test_error_norms_closed_forms():
    # s_operation.hpp:869-908: Error1 = sum |e| v / V, Error2 =
    # sqrt(sum e^2 v) / V (the reference's normalisation), Errori = max |e|
    g = _stretched()
    gl = g.gl
    sx = g.ax.s[gl:gl + g.nx]
    sy = g.ay.s[gl:gl + g.ny]
    V = sx.sum() * sy.sum()
    exa = np.random.default_rng(0).standard_normal(g.shape)
    val = exa.copy()
    g.interior(val)[:] -= 0.3
    n = O.error_norms(g, exa, val)
    assert n == pytest.approx([0.3, 0.3 * np.sqrt(V) / V, 0.3], rel=1e-13)
    # one cell: the volume is sx[i] * sy[j] (a transposed index fails)
    val = exa.copy()
    val[gl + 3, gl + 5] += 2.5
    v = sx[3] * sy[5]
    n = O.error_norms(g, exa, val)
    assert n == pytest.approx([2.5 * v / V, 2.5 * np.sqrt(v) / V, 2.5], rel=1e-13)
    # ghosts are not normal cells; Errori takes |e| of either sign
    val[0, :] = 1e9
    val[:, -1] = -1e9
    val[gl, gl] -= 4.0
    n2 = O.error_norms(g, exa, val)
    assert n2[2] == pytest.approx(4.0, rel=1e-14)
    assert n2[0] == pytest.approx((2.5 * v + 4.0 * sx[0] * sy[0]) / V, rel=1e-13)
    # homogeneity: errors scaled by k scale every norm by |k|
    n3 = O.error_norms(g, exa, exa + (-3.0) * (val - exa))
    assert n3 == pytest.approx(3.0 * n2, rel=1e-12)


# ---- NS_explicit_ pins with planted-mistake twins ----

def _tvd_line(uf_m, uf_p, lim="VanLeer"):
    """Advection of one cell (i = 3, j = 1) of a 7 x 3 grid (h = 0.1) whose
    phi varies along x only: phi(i) = -1, 0, 1, 3, 4, 4.5, 5; its m face
    carries uf_m, its p face uf_p, and vf = 0."""
    g = O.StructGrid(7, 3, 2, 0.0, 0.7, 0.0, 0.3)
    phi = g.field()
    G = g.gl
    for i, val in enumerate((-1.0, 0.0, 1.0, 3.0, 4.0, 4.5, 5.0)):
        phi[G + i, :] = val
    uf = np.zeros(g.shape)
    uf[G + 2, :] = uf_m
    uf[G + 3, :] = uf_p
    return O.tvd_field(g, uf, np.zeros(g.shape), phi, lim)[3, 1]


def test_tvd_upwind_side_is_chosen_once_per_cell(planted):
    # s_operation.hpp:556-577: the cell-averaged velocity veo picks U/C/D
    # for both faces.  van Leer on a uniform grid (R = 2):
    # psi = 2r / (1 + r), face = C + psi/2 (D - C), r = (C - U) / (D - C).
    # (a) uf_m = -1, uf_p = 3: veo = 1 > 0, both faces from the left:
    #     p face C=3 D=4 U=1: r = 2, psi = 4/3, f = 3 + 2/3 = 11/3
    #     m face C=1 D=3 U=0: r = 1/2, psi = 2/3, f = 1 + 2/3 = 5/3
    #     adv = (3 * 11/3 + 5/3) / 0.1 = 380/3
    #     (face-centric: m face from the right, C=3 D=1 U=4, f = 7/3,
    #      adv = 400/3)
    # (b) uf_m = 2, uf_p = -5: veo = -1.5 < 0, both faces from the right:
    #     p face C=4 D=3 U=4.5: r = 1/2, f = 4 - 1/3 = 11/3
    #     m face C=3 D=1 U=4: r = 1/2, f = 3 - 2/3 = 7/3
    #     adv = (-5 * 11/3 - 2 * 7/3) / 0.1 = -230
    #     (face-centric: m face from the left, f = 5/3, adv = -650/3)
    assert _tvd_line(-1.0, 3.0) == pytest.approx(380 / 3, rel=1e-8)
    assert _tvd_line(2.0, -5.0) == pytest.approx(-230.0, rel=1e-8)
    planted("face_upwind")
    assert _tvd_line(-1.0, 3.0) == pytest.approx(400 / 3, rel=1e-8)
    assert _tvd_line(2.0, -5.0) == pytest.approx(-650 / 3, rel=1e-8)


def _linear_rows_error():
    # phi = a x + b y + c with a Dirichlet x- wall (the face value) and
    # Neumann walls elsewhere.  The ghost of a Neumann wall is
    # phi_o - val * dl (s_ghost.hpp:171-175, s_equation.hpp:370-374), exact
    # for a linear phi when val = -dphi/dn (outward normal): x+ -a, y- +b,
    # y+ -b.  The 5-point rows are then exact, so A^-1 b is phi at the
    # centres to rounding.
    a, b, c = 0.7, -1.3, 0.4
    g = O.StructGrid(6, 5, 1, 0.0, 0.6, 0.0, 0.5)
    NX, NY = g.shape
    yc = g.ay.c
    vals = np.concatenate([b * yc + c, np.full(NY, -a), np.full(NX, b),
                           np.full(NX, -b)])
    bc = O.make_bc(g, [O.BC1, O.BC2, O.BC2, O.BC2], vals)
    M = O.Matrix(g, bc, beta=1.0)
    x = np.linalg.solve(M.dense(), M.b)
    X, Y = g.centres()
    want = O.field_to_x(g, np.pad(a * X + b * Y + c, 1))
    return np.max(np.abs(x - want))


def test_neumann_rows_exact_for_linear_field(planted):
    assert _linear_rows_error() < 1e-12
    planted("neumann_sign")
    assert _linear_rows_error() > 1e-2


def _tg_explicit(n, dt, steps, rho=2.0, mu=2.0):
    """Taylor-Green vortex (example/ns_sin/main.cpp:40-51, nu = mu/rho = 1)
    stepped by NS_explicit_ with van Leer; exact Dirichlet velocity walls at
    the time each step starts (apply_bc uses Time_::current_time,
    s_equation.hpp:352-357).  Returns the errors at the end relative to
    the amplitudes: max |u - ue|, and the interior pressure up to its
    constant (rel L2)."""
    from test_kim_oracle import tg_bc_values, tg_exact
    g = O.StructGrid(n, n, 2, 0.0, np.pi, 0.0, np.pi)
    vu, vv = tg_bc_values(g, 0.0)
    ns = O.NS(g, O.make_bc(g, [O.BC1] * 4, vu), O.make_bc(g, [O.BC1] * 4, vv),
              rho=rho, mu=mu, cfl=10.0, dnum=1e10, limiter="VanLeer",
              tol=1e-12, dt=dt)
    X, Y = g.centres()
    u0, v0, _ = tg_exact(X, Y, 0.0)
    g.interior(ns.f["u"])[:] = u0
    g.interior(ns.f["v"])[:] = v0
    XF, YF = np.meshgrid(g.ax.c, g.ay.c, indexing="ij")
    h = np.pi / n
    ns.f["uf"][:] = tg_exact(XF + h / 2, YF, 0.0)[0]
    ns.f["vf"][:] = tg_exact(XF, YF + h / 2, 0.0)[1]
    with np.errstate(all="ignore"):
        for k in range(steps):
            vu, vv = tg_bc_values(g, k * dt)
            ns.bcu = O.make_bc(g, [O.BC1] * 4, vu)
            ns.bcv = O.make_bc(g, [O.BC1] * 4, vv)
            ns.step()
            if not np.isfinite(ns.f["u"]).all():
                return np.inf, np.inf
        t = ns.st.t
        ue, ve, pe = tg_exact(X, Y, t)
        u, v = g.interior(ns.f["u"]), g.interior(ns.f["v"])
        eu = max(np.max(np.abs(u - ue)), np.max(np.abs(v - ve))) / np.exp(-2 * t)
        p = g.interior(ns.f["p"])[2:-2, 2:-2]
        pe = pe[2:-2, 2:-2]
        pd, ped = p - p.mean(), pe - pe.mean()
        ep = np.linalg.norm(pd - ped) / np.linalg.norm(ped)
    return eu, ep


def test_explicit_step_taylor_green(planted):
    # u* = (-adv + (dif + src)/rho) dt + u (s_ns_explicit.hpp:196-211),
    # projection, p = phi/dt: second order in u (measured 2.1e-2 at 16^2,
    # 5.6e-3 at 32^2, t = 0.2, dt = h^2/5 scale) and the pressure, whose
    # balance with the advection term carries the sign of adv (the TG
    # nonlinear term is a gradient, so a flipped adv leaves u but flips p)
    eu32, ep32 = _tg_explicit(32, 0.002, 100)
    eu16, ep16 = _tg_explicit(16, 0.008, 25)
    assert eu32 < 8e-3, eu32
    assert ep32 < 0.04, ep32
    assert eu16 / eu32 > 3.5, (eu16, eu32)
    planted("adv_sign")
    eu, ep = _tg_explicit(32, 0.002, 100)
    assert ep > 1.5, ep
    planted("dif_rho")
    eu, ep = _tg_explicit(32, 0.002, 100)
    assert not eu < 8e-3, eu
