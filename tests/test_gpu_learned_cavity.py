"""The learned-preconditioner configuration at its size: the 1024^2
lid-driven cavity (Re 1000), a network trained on the GPU by the library's
own kernels (a shortened schedule), then the next steps' pressure solves
by flexible CG with the network against plain FP64 Jacobi-PCG from the same
saved state.  Asserts the true FP64 residual of every learned solve, the
wall-time gain, and the fields: at 1e-8 within what two 1e-8 solves may
differ by, at 1e-10 within 1e-8."""
import sys
import time

import numpy as np
import pytest

import carpios_amd as ca
from carpios_amd import learned

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def fields_diff(fa, fb, g=2):
    pa, pb = fa["p"][g:-g, g:-g], fb["p"][g:-g, g:-g]
    return (rel_l2(fa["u"][g:-g, g:-g], fb["u"][g:-g, g:-g]),
            rel_l2(pa - pa.mean(), pb - pb.mean()))


def test_learned_1024_cavity_gains_threefold_over_fp64_jpcg(ctx):
    n = 1024
    ns = learned.cavity(ctx, n, tol=1e-8)
    P = ns.projection()
    # a fixed direction kernel, not the autotuner's timing-dependent pick:
    # the variants differ in the order of their dot products, the plain
    # solves' iterates (the training samples) differ in their last bits
    # with it, and the network this short schedule trains depends on them
    # (the same run gave 77-84 flexible-CG iterations per step with
    # variants 5 / 0 and fell back to Jacobi-PCG with another pick)
    P.set_variant(5, 0)
    net = ca.UNet(ctx, seed=1)
    tr = learned.CavityTrainer(ns, net, nu=2, omega=0.8)
    t0 = time.perf_counter()
    # the flexible-CG sample rounds are what make the network useful: with this
    # schedule (two of the benchmark's five rounds) a step takes ~60
    # iterations against ~2500 for Jacobi-PCG; after the first fit alone the
    # solves still hand over to Jacobi-PCG
    tr.run(plain=100, first_fit=2000, rounds=2, round_steps=40, round_fit=1500)
    ctx.sync()
    say(f"  1024^2 training: {time.perf_counter() - t0:.1f} s, {tr.slot} samples")
    state = ns.save()

    def run(solver, flags, tol, steps):
        ns.restore(state)
        P.set_unet(net if flags else None, flags)
        ns.set_solver(solver, max_iter=10 ** 7, tol=tol, norm=ca.NORM_TRUE)
        its, true_res, wall = [], [], 0.0
        for _ in range(steps):
            ctx.sync()
            t = time.perf_counter()
            it, r = ns.step(1)
            ctx.sync()
            wall += time.perf_counter() - t
            its.append(int(it[0]))
            true_res.append(P.residual())
        return its, true_res, wall, {k: ns.get_field(k) for k in ("u", "p")}

    it_l, res_l, t_l, f_l = run("FCG_UNET", ca.UNET_PRECOND, 1e-8, 5)
    it_p, res_p, t_p, f_p = run("JPCG", 0, 1e-8, 5)
    du, dp = fields_diff(f_l, f_p)
    say(f"  1024^2 at 1e-8: learned {it_l} {t_l * 1e3:.1f} ms, true residuals "
        f"{max(res_l):.2e}; FP64 J-PCG {it_p} {t_p * 1e3:.1f} ms; gain {t_p / t_l:.1f}x; "
        f"fields u {du:.1e} p {dp:.1e}")
    assert max(res_l) <= 1e-8
    assert t_p / t_l >= 3.0
    # two solutions that each stop at 1e-8 differ by up to kappa * 1e-8
    assert du <= 1e-6 and dp <= 1e-5
    it_l2, res_l2, t_l2, f_l2 = run("FCG_UNET", ca.UNET_PRECOND, 1e-10, 2)
    it_p2, res_p2, t_p2, f_p2 = run("JPCG", 0, 1e-10, 2)
    du2, dp2 = fields_diff(f_l2, f_p2)
    say(f"  1024^2 at 1e-10: learned {it_l2} (true residuals {max(res_l2):.2e}), FP64 J-PCG "
        f"{it_p2} (true residuals {max(res_p2):.2e}), fields u {du2:.1e} p {dp2:.1e}")
    assert du2 <= 1e-8 and dp2 <= 1e-8
    net.close()
    ns.close()
