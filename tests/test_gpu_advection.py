"""Parity of the device Advection_ driver (cpgpu_advection, through the
C-ABI) with the oracle's in-place sweep (s_advection.hpp:185-212): every
scheme of set_advection_scheme, mixed boundaries, tile-ragged grids, and the
reference example's steady band."""
import numpy as np
import pytest

import carpios_amd as ca
import oracle as O
from test_advection_oracle import example_band

pytestmark = pytest.mark.gpu


def edges(f, gl):
    # interior plus edge ghosts (corner ghosts are read by no cross stencil)
    m = np.ones(f.shape, dtype=bool)
    m[:gl, :gl] = m[:gl, -gl:] = m[-gl:, :gl] = m[-gl:, -gl:] = False
    return f[m]


def pair(ctx, nx, ny, scheme, seed, dt=0.004):
    rng = np.random.default_rng(seed)
    xlim, ylim = (0.0, 1.3), (-0.2, 0.9)
    g = O.StructGrid(nx, ny, 2, *xlim, *ylim)
    NX, NY = g.shape
    types = [1, 2, 1, 2]
    vals = rng.uniform(-1, 1, 2 * NY + 2 * NX)   # per-line boundary values
    ora = O.Advection(g, O.make_bc(g, types, vals), scheme, dt=dt)
    dev = ca.Advection(ctx, nx, ny, 2, xlim, ylim, bc_phi=(types, vals),
                       bcu=((2, 1, 2, 1), [0.1, -0.2, 0.0, 0.3]),
                       bcv=((1, 2, 1, 2), [0.2, 0.0, -0.1, 0.1]), scheme=scheme, dt=dt)
    X, Y = g.centres()
    u, v = g.field(), g.field()
    g.interior(u)[:] = np.sin(3 * X + 1) + 0.3 * rng.standard_normal(X.shape)
    g.interior(v)[:] = np.cos(2 * Y - X) + 0.3 * rng.standard_normal(X.shape)
    phi = g.field()
    g.interior(phi)[:] = np.exp(-((X - 0.6) ** 2 + (Y - 0.3) ** 2) * 8) + 0.1 * rng.standard_normal(X.shape)
    ora.f["u"][:], ora.f["v"][:], ora.f["phi"][:] = u, v, phi
    ora.initial(O.make_bc(g, [2, 1, 2, 1], [0.1, -0.2, 0.0, 0.3]),
                O.make_bc(g, [1, 2, 1, 2], [0.2, 0.0, -0.1, 0.1]))
    dev.set_field("u", u)
    dev.set_field("v", v)
    dev.set_field("phi", phi)
    dev.initial()
    return g, ora, dev


@pytest.mark.parametrize("scheme", list(ca.ADV_SCHEMES))
@pytest.mark.parametrize("nx,ny", [(37, 53), (64, 48)])
def test_sweep_matches_the_oracle_bitwise(ctx, scheme, nx, ny):
    g, ora, dev = pair(ctx, nx, ny, scheme, seed=nx * 7 + len(scheme))
    for name in ("uf", "vf"):
        assert np.array_equal(edges(dev.get_field(name), 2), edges(ora.f[name], 2)), name
    ora.step(5)
    dev.step(5)
    got, want = dev.get_field("phi"), ora.f["phi"]
    nd = int(np.count_nonzero(edges(got, 2) != edges(want, 2)))
    print(f"advection {scheme} {nx}x{ny}: {nd} values differ, "
          f"max |phi| {np.abs(g.interior(want)).max():.3f}", flush=True)
    assert nd == 0
    t, k = dev.time()
    assert k == 5 and t == pytest.approx(5 * ora.dt)
    dev.close()


def test_large_grid_many_tiles(ctx):
    # 32 x 32 tiles, 63 tile diagonals per step
    g, ora, dev = pair(ctx, 512, 500, "VanLeer", seed=3, dt=5e-4)
    ora.step(2)
    dev.step(2)
    assert np.array_equal(edges(dev.get_field("phi"), 2), edges(ora.f["phi"], 2))
    dev.close()


@pytest.mark.parametrize("scheme", ["upwind1", "VanLeer"])
def test_reference_example_steady_band(ctx, scheme):
    # example/advection_2d_TVD/main.cpp on the device, 1000 steps, against
    # the oracle's run bit for bit and against the band x in [y - 0.3, y]
    err, ora, g = example_band(scheme)
    n = 50
    NX, NY = g.shape
    vals = np.zeros(2 * NY + 2 * NX)
    vals[:NY] = np.where((g.ay.c >= 0.0) & (g.ay.c <= 0.3), 1.0, 0.0)
    dev = ca.Advection(ctx, n, n, 2, (0.0, 1.0), (0.0, 1.0), bc_phi=((1, 1, 1, 1), vals),
                       scheme=scheme, dt=0.002)
    one = g.field()
    one[:] = 1.0
    dev.set_field("u", one)
    dev.set_field("v", one)
    dev.initial()
    dev.step(1000)
    got = dev.get_field("phi")
    print(f"example {scheme}: mean |phi - band| {err:.4f}", flush=True)
    assert np.array_equal(g.interior(got), g.interior(ora.f["phi"]))
    dev.close()
