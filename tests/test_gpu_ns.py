"""Parity of the device-resident NS_explicit_ projection step (through the
C-ABI) with the oracle's restatement of the reference step."""
import numpy as np
import pytest

import carpios_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def pair(ctx, nx, ny, Re=1000.0, lim="VanLeer", lid=1.0, vtypes=(1, 1, 1, 1),
         tol=1e-10, norm=O.NORM_REF):
    g = O.StructGrid(nx, ny, 2, -0.5, 0.5, -0.5, 0.5)
    h = 1.0 / max(nx, ny)
    bcu_v = [0, 0, 0, lid]
    bcu = O.make_bc(g, [1, 1, 1, 1], bcu_v)
    bcv = O.make_bc(g, vtypes)
    kw = dict(rho=Re, mu=1.0, cfl=0.5, dnum=0.25, dcoef=1 / Re, limiter=lim,
              tol=tol, dt=0.5 * h)
    ora = O.NS(g, bcu, bcv, norm=norm, **kw)
    gpu = ca.NS(ctx, nx, ny, 2, (-0.5, 0.5), (-0.5, 0.5),
                bcu=((1, 1, 1, 1), bcu_v), bcv=(vtypes, None), solver="JPCG",
                norm=norm, **kw)
    return g, ora, gpu


def compare(g, ora, gpu, tol=1e-12):
    G, nx, ny = g.gl, g.nx, g.ny
    out = {}
    for name in ("u", "v", "p"):
        a, b = g.interior(gpu.get_field(name)), g.interior(ora.f[name])
        out[name] = rel(a, b)
    # faces: interior faces and the m-face slot of the boundary row/column
    uf_g, uf_o = gpu.get_field("uf"), ora.f["uf"]
    vf_g, vf_o = gpu.get_field("vf"), ora.f["vf"]
    out["uf"] = rel(uf_g[G - 1:G + nx, G:G + ny], uf_o[G - 1:G + nx, G:G + ny])
    out["vf"] = rel(vf_g[G:G + nx, G - 1:G + ny], vf_o[G:G + nx, G - 1:G + ny])
    # edge ghosts of u (corners excluded)
    ug, uo = gpu.get_field("u"), ora.f["u"]
    out["u_ghost"] = rel(np.concatenate([ug[:G, G:G + ny].ravel(), ug[G + nx:, G:G + ny].ravel(),
                                         ug[G:G + nx, :G].ravel(), ug[G:G + nx, G + ny:].ravel()]),
                         np.concatenate([uo[:G, G:G + ny].ravel(), uo[G + nx:, G:G + ny].ravel(),
                                         uo[G:G + nx, :G].ravel(), uo[G:G + nx, G + ny:].ravel()]))
    return out


@pytest.mark.parametrize("nx,ny,lim", [(48, 48, "VanLeer"), (37, 53, "superbee"),
                                       (64, 32, "QUICK"), (40, 40, "upwind2"),
                                       (44, 36, "WAHYD"), (30, 50, "CUI"),
                                       (33, 130, "VanLeer"), (17, 67, "superbee")])
def test_cavity_steps_match_oracle(ctx, nx, ny, lim):
    g, ora, gpu = pair(ctx, nx, ny, lim=lim)
    steps = 12
    it_o = []
    for _ in range(steps):
        code, it, res = ora.step()
        assert code == 0
        it_o.append(it)
    it_g, res_g = gpu.step(steps)
    print(f"{nx}x{ny} {lim}: gpu its {list(it_g)} oracle {it_o}")
    assert np.all(np.abs(np.array(it_g) - np.array(it_o)) <= 1)
    errs = compare(g, ora, gpu)
    print("rel L2:", {k: f"{v:.1e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v <= 1e-12, (k, v)
    # dt = CFL h / max|u| inherits the fields' agreement (the projections'
    # dot products are summed in another order, so u agrees to ~1e-14, and
    # dt may differ from the oracle's in its last bits)
    t, dt = gpu.time()
    assert abs(dt - ora.dt) <= 1e-12 * ora.dt and abs(t - ora.st.t) <= 1e-12 * ora.st.t
    gpu.close()


@pytest.mark.parametrize("nx,ny", [(33, 29), (33, 130), (70, 65), (16, 64), (19, 200)])
def test_first_step_bitwise_until_projection(ctx, nx, ny):
    # before the Krylov solve everything is elementwise in the reference
    # order: the divergence agrees bit for bit (the fused predictor's 16 x 64
    # tiles: ny = 130 and 200 span several tile columns, nx = 33 has a tile
    # row whose stencils reach the ghost rows without owning them)
    g, ora, gpu = pair(ctx, nx, ny)
    ora.step()
    gpu.step(1)
    for name in ("div",):
        assert np.array_equal(g.interior(gpu.get_field(name)), g.interior(ora.f[name]))
    gpu.close()


def test_neumann_v_walls_and_true_norm(ctx):
    g, ora, gpu = pair(ctx, 32, 32, vtypes=(2, 2, 1, 1), norm=O.NORM_TRUE)
    for _ in range(6):
        ora.step()
    gpu.step(6)
    errs = compare(g, ora, gpu)
    for k, v in errs.items():
        assert v <= 1e-12, (k, v)
    gpu.close()


def test_projection_divergence_free_at_solver_tolerance(ctx):
    n = 128
    gpu = ca.NS(ctx, n, n, 2, (-0.5, 0.5), (-0.5, 0.5),
                bcu=((1, 1, 1, 1), [0, 0, 0, 1.0]), rho=1000.0, mu=1.0,
                cfl=0.5, dnum=0.25, dcoef=1e-3, tol=1e-10, dt=0.5 / n)
    gpu.step(5)
    og = O.StructGrid(n, n, 2, -0.5, 0.5, -0.5, 0.5)
    dv = og.interior(O.div(og, gpu.get_field("uf"), gpu.get_field("vf")))
    dstar = og.interior(gpu.get_field("div"))
    assert np.abs(dv).max() < 1e-7 * np.abs(dstar).max()
    gpu.close()


def test_cavity_512_three_steps_match_oracle(ctx):
    # the 512^2 cavity of the plain-method configuration (Re 1000, van Leer,
    # CFL 0.5, dt0 = 0.5 h, diffusion number 0.25, J-PCG to 1e-10) for three
    # steps from rest, device against the oracle
    g, ora, gpu = pair(ctx, 512, 512)
    it_o = []
    for k in range(3):
        code, it, res = ora.step()
        assert code == 0
        it_o.append(it)
        print(f"512^2 cavity: oracle step {k}: {it} iterations", flush=True)
    it_g, res_g = gpu.step(3)
    print(f"512^2 cavity: gpu its {list(it_g)} oracle {it_o}", flush=True)
    assert np.all(np.abs(np.array(it_g) - np.array(it_o)) <= 1)
    errs = compare(g, ora, gpu)
    print("rel L2:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= 1e-12, (k, v)
    t, dt = gpu.time()
    assert dt == ora.dt
    gpu.close()
