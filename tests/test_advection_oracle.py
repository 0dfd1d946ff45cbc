"""Pins of the oracle's Advection_ step (s_advection.hpp:185-212): the
in-place, i-fastest sweep against its closed form, the plain schemes on
fields where they are exact, and the reference example's steady band."""
import numpy as np
import pytest

import oracle as O


def one_d(nx=8, scheme="upwind1", u=1.0, dt=0.05):
    # phi = 0, an inflow value 1 on the x- wall (Dirichlet), v = 0
    g = O.StructGrid(nx, 1, 2, 0.0, 1.0, 0.0, 1.0)
    A = O.Advection(g, O.make_bc(g, [1, 2, 2, 2], [1.0, 0.0, 0.0, 0.0]), scheme, dt=dt)
    A.f["u"][:] = u
    A.initial(O.make_bc(g, [2] * 4), O.make_bc(g, [2] * 4))
    return g, A


def test_upwind_sweep_is_in_place_and_forward(planted):
    # Dirichlet 1 at the x- wall gives the ghost 2 (its centre mirrors the
    # first cell's about the wall), so the first cell becomes 2c (c = u dt /
    # h).  Swept in place in increasing i, cell k reads cell k-1's new value:
    # phi_k = c phi_{k-1}, i.e. 2 c^(k+1).  A two-buffer update leaves every
    # cell but the first at 0; a sweep in decreasing i does the same.
    nx, dt = 8, 0.05
    c = 1.0 * dt * nx
    g, A = one_d(nx, dt=dt)
    A.step(1)
    want = np.array([2 * c ** (k + 1) for k in range(nx)])
    got = g.interior(A.f["phi"]).ravel()
    assert np.allclose(got, want, rtol=1e-14, atol=0)
    planted("adv_two_buffer")
    g2, A2 = one_d(nx, dt=dt)
    A2.step(1)
    assert not np.allclose(g2.interior(A2.f["phi"]).ravel(), want, rtol=1e-6)


@pytest.mark.parametrize("scheme", list(O.ADV))
def test_every_scheme_keeps_a_constant_field(scheme):
    # phi = 3 with Dirichlet 3 walls in a uniform velocity: every scheme's
    # term vanishes (the TVD form is conservative, (v_p f_p - v_m f_m) / h,
    # so it needs div v = 0; its ratio r is 0 / (0 + small))
    g = O.StructGrid(12, 9, 2, 0.0, 1.2, 0.0, 0.9)
    A = O.Advection(g, O.make_bc(g, [1] * 4, [3.0] * 4), scheme, dt=0.01)
    A.f["u"][:] = -0.6
    A.f["v"][:] = 0.9
    A.initial(O.make_bc(g, [2] * 4), O.make_bc(g, [2] * 4))
    g.interior(A.f["phi"])[:] = 3.0
    A.step(3)
    assert np.all(g.interior(A.f["phi"]) == 3.0)


def test_center_schemes_on_a_linear_field():
    # phi = a x + b y on a stretched grid: the centred difference is exact, so
    # VdotNabla_center = u a + v b; VdotNabla_center4 leaves the velocity out
    # (the reference forms it and does not apply it): a + b
    xs = np.linspace(0.5, 1.5, 10) / 10.0
    ys = np.linspace(1.2, 0.8, 8) / 8.0
    g = O.StructGrid(len(xs), len(ys), 2, xsizes=xs, ysizes=ys)
    X, Y = np.meshgrid(g.ax.c, g.ay.c, indexing="ij")
    a, b, u, v = 0.7, -1.3, 2.0, 3.0
    phi = a * X + b * Y
    uf, vf = g.field(), g.field()
    uf[:] = u
    vf[:] = v
    inner = (slice(2, -2), slice(2, -2))
    c = O.adv_field(g, uf, vf, phi, "center")
    assert np.allclose(c, u * a + v * b, rtol=1e-12)
    c4 = O.adv_field(g, uf, vf, phi, "center4")
    assert np.allclose(c4[inner], a + b, rtol=1e-12)


def example_band(scheme, steps=1000):
    # example/advection_2d_TVD/main.cpp: 50 x 50 on [0, 1]^2, u = v = 1,
    # dt = 0.002, phi = 1 entering through x- for 0 <= y <= 0.3, Dirichlet 0
    # elsewhere; after 1000 steps (t = 2) the steady band x in [y - 0.3, y]
    n = 50
    g = O.StructGrid(n, n, 2, 0.0, 1.0, 0.0, 1.0)
    yc = g.ay.c
    NX, NY = g.shape
    vals = np.zeros(2 * NY + 2 * NX)
    vals[:NY] = np.where((yc >= 0.0) & (yc <= 0.3), 1.0, 0.0)  # x- lines (per j)
    A = O.Advection(g, O.make_bc(g, [1] * 4, vals), scheme, dt=0.002)
    A.f["u"][:] = 1.0
    A.f["v"][:] = 1.0
    A.initial(O.make_bc(g, [2] * 4), O.make_bc(g, [2] * 4))
    A.step(steps)
    X, Y = g.centres()
    exact = ((X >= Y - 0.3) & (X <= Y)).astype(float)
    err = np.abs(g.interior(A.f["phi"]) - exact).mean()
    return err, A, g


def test_example_reaches_the_steady_band():
    e1, _, _ = example_band("upwind1")
    ev, _, _ = example_band("VanLeer")
    print(f"mean |phi - band| at t = 2: upwind1 {e1:.4f}, VanLeer {ev:.4f}")
    # a 50 x 50 grid smears the band's two edges over a few cells; the
    # limited scheme keeps them at least twice as sharp as first-order upwind
    assert e1 < 0.2 and ev < 0.08 and ev < 0.5 * e1
