"""Operation_::Error1 / Error2 / Errori (s_operation.hpp:869-908) on the
device against the oracle, on solver output and on NS fields."""
import numpy as np
import pytest

import carpios_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu


def check(gpu, ora):
    print(f"gpu {gpu}  oracle {ora}")
    # sums in block order vs the reference's sequential order; the maximum
    # is order-independent
    assert np.allclose(gpu[:2], ora[:2], rtol=1e-13, atol=0)
    assert gpu[2] == ora[2]


@pytest.mark.parametrize("n", [1, 7, 128])
def test_poisson_error_norms_against_exact_solution(ctx, n):
    og = O.StructGrid(n, n, 1)
    X, Y = og.centres()
    src = og.field()
    og.interior(src)[:] = -2 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y)
    bc = O.make_bc(og, [1, 1, 1, 1])
    desc = ca.uniform_grid_desc(n, n, 1, [1, 1, 1, 1], bc._keep, 1.0)
    P = ca.Poisson(ctx, desc)
    P.set_source(src)
    P.solve("JPCG", max_iter=10000, tol=1e-12, norm=ca.NORM_TRUE)
    exact = og.field()
    og.interior(exact)[:] = np.sin(np.pi * X) * np.sin(np.pi * Y)
    gpu = P.error_norms(exact)
    check(gpu, O.error_norms(og, exact, P.get_phi()))
    # the discrete solution is c sin(pi x) sin(pi y) with c = (pi h / 2)^2 /
    # sin(pi h / 2)^2: Errori = (c - 1) max |sin sin|
    h = 1.0 / n
    c = (np.pi * h / 2) ** 2 / np.sin(np.pi * h / 2) ** 2
    s = np.sin(np.pi * X) * np.sin(np.pi * Y)
    assert gpu[2] == pytest.approx((c - 1) * np.abs(s).max(), rel=1e-6)
    P.close()


@pytest.mark.parametrize("nx,ny,xlim,ylim", [(40, 24, (0.0, 2.0), (-0.5, 0.5)),
                                             (3, 130, (0.0, 1.0), (0.0, 3.0))])
def test_ns_error_norms_every_field(ctx, nx, ny, xlim, ylim):
    S = ca.NS(ctx, nx, ny, 2, xlim, ylim)
    og = O.StructGrid(nx, ny, 2, *xlim, *ylim)
    rng = np.random.default_rng(nx)
    for name in ("u", "p", "vf"):
        f = rng.standard_normal(S.shape)
        S.set_field(name, f)
        exact = f + rng.standard_normal(S.shape) * 1e-3
        exact[0, 0] = 1e300  # ghosts are not normal cells
        check(S.error_norms(name, exact), O.error_norms(og, exact, f))
    S.close()


def test_error_norms_full_size_closed_form(ctx):
    # 4096^2: a constant error c gives (|c|, |c| / sqrt(V), |c|); one cell
    # raised by e moves Errori to |e|
    n = 4096
    desc = ca.uniform_grid_desc(n, n, 1, [2, 2, 2, 2], None, 1.0, xlim=(0, 2), ylim=(0, 2))
    P = ca.Poisson(ctx, desc)
    P.fill_phi(0.25)
    exact = np.zeros(P.shape)
    exact[1 + 1000, 1 + 77] = -3.0
    out = P.error_norms(exact)
    V = 4.0
    v = (2.0 / n) ** 2
    assert out[0] == pytest.approx(0.25 + 3.0 * v / V, rel=1e-12)
    assert out[1] == pytest.approx(np.sqrt(0.0625 * (V - v) + 3.25 ** 2 * v) / V, rel=1e-12)
    assert out[2] == 3.25
    P.close()


def test_ns_field_change_is_the_steady_state_test(ctx):
    # EventStopScalarError_ (s_event.hpp:520-566): Error(prev, current), then
    # the snapshot moves to the current field
    nx, ny = 33, 20
    S = ca.NS(ctx, nx, ny, 2, (0.0, 1.0), (0.0, 0.5))
    og = O.StructGrid(nx, ny, 2, 0.0, 1.0, 0.0, 0.5)
    rng = np.random.default_rng(3)
    f1, f2, f3 = (rng.standard_normal(S.shape) for _ in range(3))
    S.set_field("v", f1)
    assert (S.field_change("v") == -1).all()
    S.set_field("v", f2)
    check(S.field_change("v"), O.error_norms(og, f1, f2))
    S.set_field("v", f3)
    check(S.field_change("v"), O.error_norms(og, f2, f3))
    S.close()
