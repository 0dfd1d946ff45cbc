"""Batched Jacobi-PCG over problems of uneven sizes (cpgpu_batch) against
the single-problem solver (itself pinned to the oracle) and the oracle."""
import numpy as np
import pytest

import carpios_amd as ca
import oracle as O

pytestmark = pytest.mark.gpu

# uneven sizes, edges that are not tile multiples, mixed walls
SHAPES = [(16, 16, [1, 1, 1, 1]), (37, 53, [1, 2, 2, 1]), (64, 200, [2, 2, 1, 2]),
          (130, 65, [1, 2, 1, 2]), (1, 7, [1, 1, 2, 2]), (96, 24, [2, 1, 1, 1]),
          (256, 128, [1, 1, 1, 2]), (5, 300, [2, 1, 2, 1])]


def make(ctx, shapes=SHAPES, seed=0):
    out = []
    for k, (nx, ny, types) in enumerate(shapes):
        desc = ca.uniform_grid_desc(nx, ny, 1, types, None, 1.0)
        P = ca.Poisson(ctx, desc)
        P.set_source_blobs(seed + k, nblobs=16)
        P.fill_phi(0.0)
        out.append(P)
    return out


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("graphs", [True, False, "persistent"])
def test_batch_is_bitwise_the_single_problem_solve(ctx, graphs):
    single = make(ctx)
    ref = []
    for P in single:
        P.set_variant(0, 0)  # the batch's direction kernel is the LDS tile one
        code, it, tol, _ = P.solve("JPCG", max_iter=20000, tol=1e-10, norm=ca.NORM_TRUE)
        ref.append((code, it, tol, P.get_phi()))
    probs = make(ctx)
    B = ca.Batch(ctx, probs)
    if graphs == "persistent":
        B.set_persistent(True)  # the whole solve in one cooperative launch
    else:
        B.set_graphs(graphs)
    code, it, res, codes, lat = B.solve("JPCG", max_iter=20000, tol=1e-10)
    assert code == 0 and (codes == 0).all()
    for k, P in enumerate(probs):
        c0, it0, tol0, phi0 = ref[k]
        print(f"{SHAPES[k][:2]}: batch {it[k]} it {res[k]:.3e} {lat[k]:.3f} ms; single {it0} it {tol0:.3e}")
        assert it[k] == it0 and res[k] == tol0
        assert np.array_equal(P.get_phi(), phi0)
    # latencies: every problem stamped, the longest solve last among equals
    assert (lat > 0).all()
    assert lat[np.argmax(it)] >= np.median(lat)
    B.close()
    for P in single + probs:
        P.close()


def test_batch_against_oracle(ctx):
    probs = make(ctx, SHAPES[:3], seed=7)
    B = ca.Batch(ctx, probs)
    code, it, res, codes, lat = B.solve("JPCG", max_iter=20000, tol=1e-11)
    assert code == 0
    for k, (nx, ny, types) in enumerate(SHAPES[:3]):
        og = O.StructGrid(nx, ny, 1)
        A = O.Matrix(og, O.make_bc(og, types), 1.0, og.field())
        b = O.field_to_x(og, probs[k].get_vec(1))  # the device rhs (blob source)
        A.b[:] = b
        co, xo, ito, tolo, _ = O.solve("JPCG", A, b, tol=1e-11, max_iter=20000,
                                       norm=O.NORM_TRUE)
        x = O.field_to_x(og, probs[k].get_phi())
        print(f"{nx}x{ny}: batch {it[k]} oracle {ito} relL2 {rel_l2(x, xo):.2e}")
        assert co == 0 and abs(int(it[k]) - ito) <= 1
        assert rel_l2(x, xo) <= 1e-12
    B.close()


@pytest.mark.parametrize("tpb", [2, 4, 8])
def test_batch_tiles_per_block(ctx, tpb):
    ref = make(ctx)
    Br = ca.Batch(ctx, ref)
    _, it0, _, _, _ = Br.solve("JPCG", max_iter=20000, tol=1e-10)
    probs = make(ctx)
    B = ca.Batch(ctx, probs)
    B.set_tiles(tpb)
    code, it, res, codes, lat = B.solve("JPCG", max_iter=20000, tol=1e-10)
    assert code == 0 and (res <= 1e-10).all()
    assert np.abs(it - it0).max() <= 1
    for P, Q in zip(probs, ref):
        assert rel_l2(P.get_phi(), Q.get_phi()) <= 1e-9
    for P in ref + probs:
        P.close()


def test_batch_mixed_precision_reaches_fp64_tolerance(ctx):
    ref = make(ctx)
    Br = ca.Batch(ctx, ref)
    _, it64, _, _, _ = Br.solve("JPCG", max_iter=20000, tol=1e-12)
    probs = make(ctx)
    B = ca.Batch(ctx, probs)
    code, it, res, codes, lat = B.solve("MIXED", max_iter=40000, tol=1e-8)
    for k in range(len(probs)):
        print(f"{SHAPES[k][:2]}: code {codes[k]} {it[k]} it resid {res[k]:.3e}")
    assert code == 0 and (codes == 0).all()
    for k, P in enumerate(probs):
        true = P.residual()  # FP64 b - A phi recomputed
        err = rel_l2(P.get_phi(), ref[k].get_phi())
        print(f"{SHAPES[k][:2]}: mixed {it[k]} it resid {res[k]:.2e} true {true:.2e} "
              f"err {err:.1e} (FP64 {it64[k]} it to 1e-12)")
        assert res[k] <= 1e-8 and true <= 1e-8 * 1.0001
        assert err < 1e-5
    for P in ref + probs:
        P.close()


@pytest.mark.parametrize("method", ["JPCG", "MIXED"])
def test_batch_max_iter(ctx, method):
    probs = make(ctx)
    B = ca.Batch(ctx, probs)
    code, it, res, codes, lat = B.solve(method, max_iter=5, tol=1e-14)
    small = [k for k, (nx, ny, _) in enumerate(SHAPES) if nx * ny <= 16]
    assert code == ca.MAXITER
    for k in range(len(SHAPES)):
        if k in small:
            continue  # a 1 x 7 problem can converge within 5 iterations
        assert codes[k] == ca.MAXITER and it[k] == 5
    for P in probs:
        P.close()


def test_batch_autotune_keeps_phi(ctx):
    probs = make(ctx)
    phi0 = [P.get_phi() for P in probs]
    B = ca.Batch(ctx, probs)
    best, ms = B.autotune("JPCG", reps=1)
    print("tiles per block", best, "ms per iteration", ms)
    assert best in (1, 2, 4, 8) and (ms > 0).all()
    for P, f in zip(probs, phi0):
        assert np.array_equal(P.get_phi(), f)
    code, *_ = B.solve("JPCG", max_iter=20000, tol=1e-10)
    assert code == 0
    for P in probs:
        P.close()


def test_all_neumann_problem_is_rejected(ctx):
    P = ca.Poisson(ctx, ca.uniform_grid_desc(32, 32, 1, [2, 2, 2, 2]))
    with pytest.raises(ca.CpgpuError):
        ca.Batch(ctx, [P])
    P.close()


def test_batch_learned_members_reach_tolerance(ctx):
    # members with a U-Net attached run flexible CG (an untrained network:
    # some stall or pass the budget and are finished by the batched pass);
    # the others run batched; every member ends at the tolerance
    shapes = [(64, 64, [1, 2, 2, 1]), (128, 64, [2, 1, 1, 1]), (64, 192, [1, 1, 2, 2]),
              (37, 53, [1, 2, 2, 1])]
    ref = make(ctx, shapes, seed=11)
    Br = ca.Batch(ctx, ref)
    Br.solve("JPCG", max_iter=20000, tol=1e-12)
    probs = make(ctx, shapes, seed=11)
    net = ca.UNet(ctx, seed=1)
    for P, (nx, ny, _) in zip(probs, shapes):
        P.set_fcg(2, 2, 0.8)
        if nx % 64 == 0 and ny % 64 == 0:
            P.set_unet(net, ca.UNET_PRECOND)
    B = ca.Batch(ctx, probs)
    code, it, res, codes, lat = B.solve("FCG_UNET", max_iter=20000, tol=1e-8)
    print("learned batch:", it, res, codes, lat)
    assert code == 0 and (codes == 0).all() and (lat > 0).all()
    for P, Q in zip(probs, ref):
        assert P.residual() <= 1e-8 * 1.0001
        assert rel_l2(P.get_phi(), Q.get_phi()) < 1e-5
    B.close()
    for P in probs:
        P.set_unet(None, 0)
    net.close()
    for P in ref + probs:
        P.close()


MIXED_PERSISTENT_OFF = ("the persistent form of the mixed-precision pass is switched off: "
                        "one of its grid barrier waits timed out on the 256-problem set "
                        "and the cause is not established")


@pytest.mark.parametrize("method,tpb", [
    ("JPCG", 2), ("JPCG", 8),
    pytest.param("MIXED", 1, marks=pytest.mark.skip(reason=MIXED_PERSISTENT_OFF)),
    pytest.param("MIXED", 4, marks=pytest.mark.skip(reason=MIXED_PERSISTENT_OFF))])
def test_persistent_launch_is_bitwise_the_launched_batches(ctx, method, tpb):
    # one cooperative launch (set-up, every iteration, replacements, stops on
    # the device) against the graph-replayed batches: same block map, same
    # partial slots, so the same bits, iteration counts and stop codes; one
    # member stops at max_iter
    out = []
    for persistent in (False, True):
        probs = make(ctx, SHAPES, seed=11)
        B = ca.Batch(ctx, probs)
        B.set_tiles(tpb)
        B.set_persistent(persistent)
        code, it, res, codes, lat = B.solve(method, max_iter=700, tol=1e-9)
        out.append((code, it.copy(), res.copy(), codes.copy(), [P.get_phi() for P in probs], lat))
        B.close()
        for P in probs:
            P.close()
    (c0, it0, r0, k0, phi0, _), (c1, it1, r1, k1, phi1, lat1) = out
    print(method, tpb, "iterations", list(it1), "codes", list(k1))
    assert c0 == c1 and np.array_equal(it0, it1) and np.array_equal(k0, k1)
    assert np.array_equal(r0, r1)
    for a, b in zip(phi0, phi1):
        assert np.array_equal(a, b)
    assert (k1 == ca.MAXITER).any() and (k1 == ca.OK).any()  # both stop kinds exercised
    assert (lat1 > 0).all()
