"""General CSR solvers on the device (cpgpu_csr) against the reference's own
golden vectors and the oracle.

The free functions of lib/algebra/solver_matrix.hpp on a MatrixSCR_: CG
(:264-316), Jacobi-PCG (IC_CG's loop with DiaPre, :318-374), Dia_BiCGSTAB
(:599-672) and Solver_Jacobi_ (:820-910), and MatrixSCR_::operator*
(matrix_SparCompRow.hpp:236-249).  The golden vectors come from the compiled
reference (tests/golden/make_golden.py): the C1 Poisson operator at 32^2 and
the reference test's 685_bus.mtx (x = 1, b = 1, tol 1e-6, as
test/algebra/_test_solver.hpp:129-163).  The product sums each row in CSR
order like the reference, so it is compared bit for bit; the solvers' dot
products are associated per block on the device, so their iterates agree
with the reference to rounding, held to the spread the oracle itself shows
under rounding-size perturbations of b where the method's stopping step is
that sensitive.
"""
import ctypes as C
import os

import numpy as np
import pytest

import carpios_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "golden_solvers.npz")


class Holder:
    """CSR arrays in the form the oracle's solvers take."""

    def __init__(self, rp, col, val):
        self.row_ptr = np.ascontiguousarray(rp, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(self.row_ptr) - 1
        self.c = O.CSR(self.n, self.row_ptr.ctypes.data_as(C.POINTER(C.c_int)),
                       self.col.ctypes.data_as(C.POINTER(C.c_int)),
                       self.val.ctypes.data_as(C.POINTER(C.c_double)))

    def mult(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        O.lib().orc_csr_mult(C.byref(self.c), x.ctypes.data_as(C.POINTER(C.c_double)),
                             y.ctypes.data_as(C.POINTER(C.c_double)))
        return y


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def golden():
    return np.load(GOLD, allow_pickle=False)


def case(g, name):
    if name == "c1":
        H = Holder(g["c1_row_ptr"], g["c1_col"], g["c1_val"])
        return H, np.array(g["c1_b"]), None, 20000, 1e-10
    H = Holder(g["bus_row_ptr"], g["bus_col"], g["bus_val"])
    return H, np.ones(H.n), np.ones(H.n), 100000, 1e-6


def band(method, H, b, x0, runs=6, **kw):
    """The oracle's run, the range of its history length and the spread of
    its solution under 1e-15 relative perturbations of b, and the history
    prefix every perturbed run shares with it to 1e-8."""
    base = O.solve(method, H, b, x0=x0, **kw)
    rng = np.random.default_rng(7)
    lens, spread, stable, h0 = [len(base[4])], 0.0, len(base[4]), base[4]
    for _ in range(runs):
        o = O.solve(method, H, b * (1 + 1e-15 * rng.standard_normal(b.size)), x0=x0, **kw)
        lens.append(len(o[4]))
        spread = max(spread, rel_l2(o[1], base[1]))
        m = min(len(o[4]), len(h0))
        bad = (np.abs(o[4][:m] / np.where(h0[:m] > 0, h0[:m], 1) - 1) > 1e-8) | (h0[:m] < 1e-13)
        stable = min(stable, int(np.argmax(bad)) if bad.any() else m)
    return base, min(lens), max(lens), spread, stable


def reassociated_cg(H, b, x0, max_iter, tol, precond, parts):
    """The oracle's CG / Jacobi-PCG step for step (the same product, the
    same vector updates) with its dot products summed in another order:
    numpy's (parts = 0) or as `parts` strided partial sums added pairwise,
    like a grid-stride device reduction.  Returns (iterations, x, residual
    history)."""
    d = None
    if precond:
        d = np.zeros(H.n)
        for i in range(H.n):
            for k in range(H.row_ptr[i], H.row_ptr[i + 1]):
                if H.col[k] == i:
                    d[i] = 1.0 / H.val[k]
                    break

    def dot(a, c):
        if not parts:
            return float(np.dot(a, c))
        return float(np.sum([np.sum(a[t::parts] * c[t::parts]) for t in range(parts)]))

    def nrm(v):  # carpio::nrm2 without the last element
        return np.sqrt(dot(v[:-1], v[:-1]))

    x = np.zeros(H.n) if x0 is None else x0.copy()
    r = b - H.mult(x)
    normb = nrm(b)
    p, rho1, hist = None, 0.0, []
    for i in range(1, max_iter + 1):
        z = r * d if precond else r.copy()
        rho = dot(r, z)
        p = z.copy() if i == 1 else z + (rho / rho1) * p
        q = H.mult(p)
        alpha = rho / dot(p, q)
        x = x + alpha * p
        r = r - alpha * q
        hist.append(nrm(r) / normb)
        if hist[-1] <= tol:
            return i, x, np.array(hist)
        rho1 = rho
    return max_iter, x, np.array(hist)


def reassociated_bicgstab(H, b, x0, max_iter, tol, parts):
    """The oracle's Dia_BiCGSTAB step for step with its dot products summed
    in another order (as reassociated_cg).  Returns (history length, x,
    residual history)."""
    d = np.zeros(H.n)
    for i in range(H.n):
        for k in range(H.row_ptr[i], H.row_ptr[i + 1]):
            if H.col[k] == i:
                d[i] = 1.0 / H.val[k]
                break

    def dot(a, c):
        if not parts:
            return float(np.dot(a, c))
        return float(np.sum([np.sum(a[t::parts] * c[t::parts]) for t in range(parts)]))

    def nrm(v):
        return np.sqrt(dot(v[:-1], v[:-1]))

    x = np.zeros(H.n) if x0 is None else x0.copy()
    r = b - H.mult(x)
    rt = r.copy()
    normb = nrm(b)
    p = np.zeros(H.n)
    v = np.zeros(H.n)
    rho2 = alpha = omega = 0.0
    hist = []
    for i in range(1, max_iter + 1):
        rho1 = dot(rt, r)
        if rho1 == 0:
            break
        p = r.copy() if i == 1 else r + ((rho1 / rho2) * (alpha / omega)) * (p - omega * v)
        ph = p * d
        v = H.mult(ph)
        alpha = rho1 / dot(rt, v)
        sv = r - alpha * v
        if nrm(sv) / normb < tol:
            x = x + alpha * ph
            hist.append(nrm(sv) / normb)
            break
        sh = sv * d
        t = H.mult(sh)
        omega = dot(t, sv) / dot(t, t)
        x = x + alpha * ph + omega * sh
        r = sv - omega * t
        rho2 = rho1
        hist.append(nrm(r) / normb)
        if hist[-1] < tol or omega == 0:
            break
    return len(hist), x, np.array(hist)


def test_product_is_the_reference_product(ctx):
    g = golden()
    H, *_ = case(g, "c1")
    A = ca.CSR(ctx, H.row_ptr, H.col, H.val)
    y = A.mult(g["c1_spmv_x"])
    assert np.array_equal(y, g["c1_spmv_y"])  # the reference's own output
    Hb, *_ = case(g, "bus")
    Ab = ca.CSR(ctx, Hb.row_ptr, Hb.col, Hb.val)
    x = np.random.default_rng(3).standard_normal(Hb.n)
    assert np.array_equal(Ab.mult(x), Hb.mult(x))
    A.close()
    Ab.close()


GOLD_NAMES = {"CG": "CG", "JPCG": "JPCG", "Dia_BiCGSTAB": "Dia_BiCGSTAB",
              "Jacobi": "Solver_Jacobi_"}


@pytest.mark.parametrize("name", ["c1", "bus"])
@pytest.mark.parametrize("method", ["CG", "JPCG", "Dia_BiCGSTAB", "Jacobi"])
def test_solvers_against_reference_golden(ctx, name, method):
    g = golden()
    H, b, x0, max_iter, tol = case(g, name)
    key = f"{name}_{GOLD_NAMES[method]}"
    A = ca.CSR(ctx, H.row_ptr, H.col, H.val)
    x_in = np.zeros(H.n) if x0 is None else x0
    code, x, it, t, hist = A.solve(method, x_in, b, max_iter=max_iter, tol=tol,
                                   hist_cap=max_iter + 2)
    A.close()
    gcode, git = int(g[key + "_code"]), int(g[key + "_iters"])
    ghl, gx = int(g[key + "_hist_len"]), g[key + "_x"]
    print(f"{key}: gpu code {code} it {it} hist {len(hist)} tol {t:.4e}; reference code {gcode} "
          f"it {git} hist {ghl}; relL2 {rel_l2(x, gx):.2e}")
    assert code == gcode
    if method == "Jacobi":
        # the sweeps are the reference's arithmetic exactly (rows in CSR
        # order); only the stopping test's norm is summed differently
        assert abs(it - git) <= 1 and abs(len(hist) - ghl) <= 1
        m = min(len(hist), ghl, 200)
        assert np.allclose(hist[:m], g[key + "_hist_head"][:m], rtol=1e-10, atol=0)
        assert t <= tol
        if it == git:
            assert np.array_equal(x, gx)
        return
    # CG, Jacobi-PCG, Dia_BiCGSTAB: their iterates carry the dot products'
    # rounding, so the device is held to the spread of the oracle (the
    # reference's run to the bit on these inputs) under rounding-size
    # perturbations of b
    o, lo, hi, spread, stable = band(method, H, b, x0, runs=24, max_iter=max_iter, tol=tol)
    assert o[0] == gcode and len(o[4]) == ghl and np.array_equal(o[1], gx)
    # other summation orders of the same dot products move the stopping step
    # further than rounding-size perturbations of b do (CG on 685_bus: the
    # reference's 554 becomes 544..546; its residual swings by orders of
    # magnitude within ten iterations from x = 1, and the histories of
    # different orders part after six)
    # (BiCGSTAB on C1: the stopping step of 60 perturbed runs spreads over
    # 34..41, of 67 summation orders over 35..42)
    h0 = o[4]
    for parts in (0, 2, 3, 5, 7, 16, 64, 256, H.n):
        if method == "Dia_BiCGSTAB":
            ia, xa, ha = reassociated_bicgstab(H, b, x0, max_iter, tol, parts)
        else:
            ia, xa, ha = reassociated_cg(H, b, x0, max_iter, tol, method == "JPCG", parts)
        lo, hi = min(lo, ia), max(hi, ia)
        spread = max(spread, rel_l2(xa, gx))
        m = min(len(ha), len(h0))
        bad = (np.abs(ha[:m] / np.where(h0[:m] > 0, h0[:m], 1) - 1) > 1e-8) | (h0[:m] < 1e-13)
        stable = min(stable, int(np.argmax(bad)) if bad.any() else m)
    assert lo - 1 <= len(hist) <= hi + 1
    assert rel_l2(x, gx) <= max(1e-12, 10 * spread)
    if stable:
        assert np.allclose(hist[:stable], o[4][:stable], rtol=1e-6, atol=0)
    if method == "Dia_BiCGSTAB":
        # a full step's exit reports its count (the history's length), the
        # half-step exit keeps max_iter; which one ends the run is as
        # rounding-sensitive as the stopping step
        assert it in (len(hist), max_iter)
    else:
        assert abs(it - git) <= max(1, hi - lo + 1)


def test_jacobi_returns_the_previous_iterate(ctx):
    # Solver_Jacobi_ stops without copying the converged sweep (:898-903):
    # the device's x is the oracle's to rounding, one sweep behind the
    # sweep that met the tolerance
    g = golden()
    H, b, _, _, _ = case(g, "c1")
    A = ca.CSR(ctx, H.row_ptr, H.col, H.val)
    code, x, it, t, hist = A.solve("Jacobi", np.zeros(H.n), b, max_iter=5000, tol=1e-6)
    o = O.solve("Jacobi", H, b, max_iter=5000, tol=1e-6)
    assert code == 0 and o[0] == 0 and abs(it - o[2]) <= 1
    assert rel_l2(x, o[1]) <= 1e-10
    r = b - H.mult(x)
    r_prev = np.linalg.norm(r[:-1]) / np.linalg.norm(b[:-1])  # carpio::nrm2
    assert r_prev > 1e-6 >= t  # x is the iterate before the one that converged
    A.close()


@pytest.mark.parametrize("norm", [ca.NORM_REF, ca.NORM_TRUE])
def test_random_nonsymmetric_against_the_oracle(ctx, norm):
    # a diagonally dominant nonsymmetric matrix with a random pattern:
    # BiCGSTAB and Jacobi against the oracle in both norm modes
    rng = np.random.default_rng(11)
    n = 500
    rows = []
    for i in range(n):
        cols = sorted(set(rng.choice(n, 6, replace=False).tolist()) | {i})
        vals = [rng.uniform(-1, 1) if c != i else 8.0 + rng.uniform() for c in cols]
        rows.append((cols, vals))
    rp = np.cumsum([0] + [len(c) for c, _ in rows]).astype(np.int32)
    col = np.concatenate([c for c, _ in rows]).astype(np.int32)
    val = np.concatenate([v for _, v in rows])
    H = Holder(rp, col, val)
    b = rng.standard_normal(n)
    A = ca.CSR(ctx, rp, col, val)
    for method in ("Dia_BiCGSTAB", "Jacobi"):
        code, x, it, t, hist = A.solve(method, np.zeros(n), b, max_iter=1000, tol=1e-12,
                                       norm=norm)
        o = O.solve(method, H, b, max_iter=1000, tol=1e-12, norm=norm)
        print(f"{method} norm {norm}: gpu {code} {it} {len(hist)} {t:.3e}; oracle {o[0]} {o[2]} "
              f"{len(o[4])} {o[3]:.3e}; relL2 {rel_l2(x, o[1]):.2e}")
        assert code == o[0] == 0
        assert abs(len(hist) - len(o[4])) <= 1
        assert rel_l2(x, o[1]) <= 1e-10
    A.close()


def test_zero_diagonal_and_bad_input_are_errors(ctx):
    rp = np.array([0, 2, 4], dtype=np.int32)
    col = np.array([0, 1, 0, 1], dtype=np.int32)
    val = np.array([1.0, 2.0, 3.0, 0.0])  # a_11 = 0
    A = ca.CSR(ctx, rp, col, val)
    with pytest.raises(ca.CpgpuError) as e:
        A.solve("JPCG", np.zeros(2), np.ones(2))
    assert e.value.code == ca.EZERODIAG
    with pytest.raises(ca.CpgpuError) as e:
        A.solve("MIXED", np.zeros(2), np.ones(2))
    assert e.value.code == -3
    A.close()
    with pytest.raises(ca.CpgpuError):
        ca.CSR(ctx, rp, np.array([0, 1, 0, 2], dtype=np.int32), val)  # column 2 of 2


@pytest.mark.parametrize("method", ["CG", "JPCG"])
def test_device_scalar_cg_stops_inside_a_batch(ctx, method):
    # CG and Jacobi-PCG keep rho, alpha and the residual on the device and
    # launch iterations in batches of 16, reading the state once per batch;
    # iterations after the stop return at entry.  A run capped at k
    # iterations therefore reports k, its history is the first k entries of
    # the uncapped run's bit for bit, and its iterate is the one the
    # uncapped run held after k steps (the oracle's to rounding)
    # (the C1 operator with a random right-hand side: the golden one takes
    # plain CG to its tolerance in a single step)
    g = golden()
    H, _, _, _, _ = case(g, "c1")
    b = np.random.default_rng(21).standard_normal(H.n)
    A = ca.CSR(ctx, H.row_ptr, H.col, H.val)
    x0 = np.zeros(H.n)
    code, x, it, t, full = A.solve(method, x0, b, max_iter=20000, tol=1e-10)
    assert code == 0 and len(full) == it > 40
    for k in (1, 15, 16, 17, 33):
        ck, xk, ik, tk, hk = A.solve(method, x0, b, max_iter=k, tol=1e-10)
        print(f"{method} capped at {k}: code {ck} it {ik} hist {len(hk)} tol {tk:.3e}", flush=True)
        assert ck == ca.MAXITER and ik == k and len(hk) == k
        assert np.array_equal(hk, full[:k]) and tk == full[k - 1]
        _, xo, ho = reassociated_cg(H, b, None, k, 1e-10, method == "JPCG", 0)
        assert rel_l2(xk, xo) <= 1e-10
        assert np.allclose(hk, ho, rtol=1e-8, atol=0)
    # a history shorter than the run: the first hist_cap entries, the count
    # still the run's
    ck, xk, ik, tk, hk = A.solve(method, x0, b, max_iter=20000, tol=1e-10, hist_cap=5)
    assert ck == 0 and ik == it and np.array_equal(hk, full[:5]) and np.array_equal(xk, x)
    # the initial test: a converged start takes no iteration
    ck, xk, ik, tk, hk = A.solve(method, x, b, max_iter=100, tol=1e-6)
    assert ck == 0 and ik == 0 and len(hk) == 0 and np.array_equal(xk, x)
    A.close()


def test_device_scalar_bicgstab_stops_inside_a_batch(ctx):
    # Dia_BiCGSTAB on the CSR matrix keeps its scalars on the device as CG
    # does: capped runs report the cap and the uncapped run's history prefix
    g = golden()
    H, _, _, _, _ = case(g, "c1")
    b = np.random.default_rng(22).standard_normal(H.n)
    A = ca.CSR(ctx, H.row_ptr, H.col, H.val)
    x0 = np.zeros(H.n)
    code, x, it, t, full = A.solve("Dia_BiCGSTAB", x0, b, max_iter=20000, tol=1e-12)
    assert code == 0 and len(full) > 40
    for k in (1, 15, 16, 17, 33):
        ck, xk, ik, tk, hk = A.solve("Dia_BiCGSTAB", x0, b, max_iter=k, tol=1e-12)
        print(f"BiCGSTAB capped at {k}: code {ck} it {ik} hist {len(hk)} tol {tk:.3e}",
              flush=True)
        assert ck == ca.MAXITER and ik == k and len(hk) == k
        assert np.array_equal(hk, full[:k]) and tk == full[k - 1]
        if k <= 17:
            _, xo, ho = reassociated_bicgstab(H, b, None, k, 1e-12, 0)
            assert rel_l2(xk, xo) <= 1e-9
            assert np.allclose(hk, ho, rtol=1e-7, atol=0)
    A.close()
