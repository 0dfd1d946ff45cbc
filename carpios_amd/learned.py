"""Training workflow of the learned preconditioner (no reference counterpart).

The network is trained on errors of the problems it will precondition.  A
sample is an error e with its residual r = A e (formed on the device by
cpgpu_unet_set_sample); the loss is the relative energy norm of the
network's error on it.  Inside the flexible-CG preconditioner cycle
(cpgpu_poisson_set_fcg) the network sees the residual left after `nu`
weighted-Jacobi sweeps, so every sample is stored after the same sweeps.

Samples come from the plain solver and from the learned solver itself:
  * plain steps: phi_n from FP64 Jacobi-PCG; the error of the warm start
    phi_n - phi_{n-1}, the error of a zero start phi_n, and Jacobi-PCG
    iterate errors;
  * flexible-CG rounds: the errors phi_n - x_k of the flexible-CG iterates
    x_k run with the current network, which are exactly the inputs the
    network meets at iteration k + 1 (the training distribution follows the
    solver's own, round after round).
Everything runs through libcpgpu.so.
"""
import time

import numpy as np

from . import NORM_TRUE, NS, UNET_PRECOND


def interior(P, field):
    g = P.gl
    return np.ascontiguousarray(field[g:g + P.nx, g:g + P.ny])


def add_error_samples(P, net, seeds, checkpoints=(0, 16, 64, 256), nblobs=64,
                      wmin=0.01, wmax=0.1, tol=1e-10, slot0=0, smooth=0,
                      omega=0.8):
    """Samples e_k = phi - x_k of Jacobi-PCG on blob right-hand sides.

    For each seed: b from 'nblobs' Gaussian blobs, phi by FP64 Jacobi-PCG to
    the true relative residual 'tol' from zero, then the iterates x_k after
    k = checkpoints Jacobi-PCG iterations from zero; each stored after
    `smooth` weighted-Jacobi sweeps (the preconditioner cycle's pre-smoothing,
    cpgpu_poisson_set_fcg).  Returns the next free slot.  Leaves the
    problem's right-hand side at the last seed's."""
    P.set_unet(None)
    slot = slot0
    for seed in seeds:
        P.set_source_blobs(seed, nblobs=nblobs, wmin=wmin, wmax=wmax)
        P.fill_phi(0.0)
        P.solve("JPCG", max_iter=10 ** 7, tol=tol, norm=NORM_TRUE)
        phi = interior(P, P.get_vec(0))
        for k in checkpoints:
            if k == 0:
                e = phi
            else:
                P.fill_phi(0.0)
                P.solve("JPCG", max_iter=int(k), tol=1e-300, norm=NORM_TRUE)
                e = phi - interior(P, P.get_vec(0))
            net.set_sample(P, slot, e, smooth=smooth, omega=omega)
            slot += 1
    return slot


def train(net, P, steps, batch=4, lr=1e-3, seed=0, decay_at=0.7):
    """Adam at lr for the first decay_at of the steps, then lr / 10.
    Returns the per-step mean losses."""
    n1 = int(steps * decay_at)
    l1 = net.train(P, n1, batch=batch, lr=lr, seed=seed)
    l2 = net.train(P, steps - n1, batch=batch, lr=lr * 0.1, seed=seed + 1)
    return np.concatenate([l1, l2])


# ------------------------------------------------------------------ cavity
def cavity(ctx, n, re=1000.0, solver="JPCG", tol=1e-8, norm=NORM_TRUE):
    """Lid-driven cavity on [-0.5, 0.5]^2, n x n cells, lid velocity 1 on the
    y+ wall, no-slip elsewhere, Re = re (rho = Re, mu = 1), CFL 0.5,
    diffusion number 0.25, van Leer TVD advection, dt0 = h / 2; the
    projection solved to `tol`."""
    return NS(ctx, n, n, 2, (-0.5, 0.5), (-0.5, 0.5),
              bcu=((1, 1, 1, 1), [0.0, 0.0, 0.0, 1.0]), bcv=((1, 1, 1, 1), None),
              rho=re, mu=1.0, cfl=0.5, dnum=0.25, dcoef=1.0 / re,
              limiter="VanLeer", solver=solver, tol=tol, norm=norm,
              dt=0.5 / n)


class CavityTrainer:
    """Collects samples from a cavity run and trains the network for the
    flexible-CG cycle with `nu` pre- and post-sweeps of weight `omega`."""

    def __init__(self, ns, net, nu=2, omega=0.8, tol=1e-8):
        self.ns, self.net, self.nu, self.omega, self.tol = ns, net, nu, omega, tol
        self.P = ns.projection()
        self.P.set_fcg(nu, nu, omega)
        self.slot = 0
        self.log = []

    def _add(self, e):
        self.net.set_sample(self.P, self.slot, e, smooth=self.nu, omega=self.omega)
        self.slot += 1

    def _plain_step(self):
        """One step with the plain FP64 solver; returns (phi_prev, phi)
        as ghost-padded fields."""
        P = self.P
        prev = P.get_vec(0)
        P.set_unet(None, 0)
        self.ns.set_solver("JPCG", max_iter=10 ** 7, tol=self.tol, norm=NORM_TRUE)
        self.ns.step(1)
        return prev, P.get_vec(0)

    def plain_steps(self, steps, every=5, jpcg_checkpoints=(30, 300)):
        """Plain-solver steps: warm-start errors every step; every `every`
        steps also the zero-start error and Jacobi-PCG iterate errors."""
        P = self.P
        t = time.perf_counter()
        for s in range(steps):
            prev, phi_f = self._plain_step()
            phi = interior(P, phi_f)
            self._add(phi - interior(P, prev))
            if s % every == 0:
                self._add(phi)
                for k in jpcg_checkpoints:
                    P.set_phi(prev)
                    P.solve("JPCG", max_iter=int(k), tol=1e-300, norm=NORM_TRUE)
                    self._add(phi - interior(P, P.get_vec(0)))
                P.set_phi(phi_f)
        self.log.append(("plain", steps, self.slot, time.perf_counter() - t))

    def fcg_steps(self, steps, ks=(1, 2, 4, 8, 16)):
        """Plain-solver steps whose phi scores the flexible-CG iterates of
        the current network: errors phi - x_k after k iterations from the
        warm start."""
        P = self.P
        t = time.perf_counter()
        for _ in range(steps):
            prev, phi_f = self._plain_step()
            phi = interior(P, phi_f)
            self._add(phi - interior(P, prev))
            P.set_unet(self.net, UNET_PRECOND)
            for k in ks:
                P.set_phi(prev)
                P.solve("FCG_UNET", max_iter=int(k), tol=1e-300, norm=NORM_TRUE)
                self._add(phi - interior(P, P.get_vec(0)))
            P.set_unet(None, 0)
            P.set_phi(phi_f)
        self.log.append(("fcg", steps, self.slot, time.perf_counter() - t))

    def fit(self, steps, lr, seed, batch=4):
        t = time.perf_counter()
        losses = train(self.net, self.P, steps, batch=batch, lr=lr, seed=seed)
        self.log.append(("train", steps, float(losses[-min(20, steps):].mean()),
                         time.perf_counter() - t))
        return losses

    def validate(self):
        """Flexible-CG iterations of the next cavity step's solve to the
        trainer's tolerance with the current network (the cavity state is
        restored afterwards); a stalled solve counts as its total with the
        Jacobi-PCG finish."""
        P, ns = self.P, self.ns
        state = ns.save()
        P.set_unet(self.net, UNET_PRECOND)
        ns.set_solver("FCG_UNET", max_iter=10 ** 7, tol=self.tol, norm=NORM_TRUE)
        its, _ = ns.step(1)
        P.set_unet(None, 0)
        ns.restore(state)
        return int(its[0])

    def run(self, plain=100, first_fit=2000, rounds=5, round_steps=40,
            round_fit=1500, progress=None, select=True):
        """The schedule used by the benchmark: plain steps, a first fit, then
        `rounds` of flexible-CG sample collection and refits.  With `select`,
        every fit is scored by validate() and the parameters that scored
        best (the later of equals) are kept at the end: the refits' outcome
        depends on the last bits of their samples, and the last fit is not
        always the best.  `progress`, if given, receives a line after every
        phase."""
        def done():
            if progress:
                progress(f"training: {self.log[-1]}")
        best = None

        def score():
            nonlocal best
            if not select:
                return
            t = time.perf_counter()
            it = self.validate()
            self.log.append(("validate", it, time.perf_counter() - t))
            done()
            if best is None or it <= best[0]:
                best = (it, self.net.get_params())
        self.plain_steps(plain)
        done()
        self.fit(first_fit, 1e-3, 0)
        done()
        score()
        for r in range(rounds):
            self.fcg_steps(round_steps)
            done()
            self.fit(round_fit, 5e-4, r + 1)
            done()
            score()
        if best is not None:
            self.net.set_params(best[1])
            self.log.append(("selected", best[0]))
            done()
        return self.log


# ------------------------------------------------------- batches of problems
def train_problem_net(ctx, net, seed=7, count=16, min_cells=1 << 16, lo=6, hi=10,
                      nu=2, omega=0.8, passes=16, fit=60, dagger_at=(4, 8, 12), lr=1e-3,
                      tol=1e-10, checkpoints=(0, 16, 64, 256, 1024), fcg_ks=(2, 4, 8, 16),
                      log=None):
    """Trains `net` as the flexible-CG preconditioner of random problems of the
    batch's kind: the problem set drawn from `seed` (cpgpu_batch_draw: sizes,
    wall types, blob sources; a different seed from the measured set), members
    with at least `min_cells` cells, up to `count` of them.

    Per member: its FP64 solution phi (J-PCG from zero to `tol`) and the
    errors phi - x_k of its J-PCG iterates (k = checkpoints) are kept on the
    host.  Training interleaves the members: `passes` sweeps over them, each
    member's errors set as the samples (stored after `nu` Jacobi sweeps; the
    samples of one member share its operator) and `fit` Adam steps taken on
    them, the learning rate falling to a tenth over the passes.  Before the
    passes in `dagger_at`, the errors of the current network's flexible-CG
    iterates (k = fcg_ks) are added to every member's set.
    Returns (seconds, samples, mean loss per pass)."""
    from . import Poisson, uniform_grid_desc, draw_problems
    nx, ny, bc, bs = draw_problems(seed, 8 * count, lo, hi)
    pick = [p for p in range(len(nx)) if int(nx[p]) * int(ny[p]) >= min_cells][:count]
    t0 = time.perf_counter()
    members = []
    for p in pick:
        P = Poisson(ctx, uniform_grid_desc(int(nx[p]), int(ny[p]), 1,
                                           [int(t) for t in bc[p]]))
        P.set_source_blobs(int(bs[p]))
        P.set_fcg(nu, nu, omega)
        P.fill_phi(0.0)
        P.solve("JPCG", max_iter=10 ** 7, tol=tol, norm=NORM_TRUE)
        phi = interior(P, P.get_vec(0))
        errs = []
        for c in checkpoints:
            if c == 0:
                errs.append(phi)
                continue
            P.fill_phi(0.0)
            P.solve("JPCG", max_iter=int(c), tol=1e-300, norm=NORM_TRUE)
            errs.append(phi - interior(P, P.get_vec(0)))
        members.append((P, phi, errs))
    curve = []
    for ps in range(passes):
        if ps in dagger_at:
            for P, phi, errs in members:
                P.set_unet(net, UNET_PRECOND)
                for c in fcg_ks:
                    P.fill_phi(0.0)
                    P.solve("FCG_UNET", max_iter=int(c), tol=1e-300, norm=NORM_TRUE)
                    errs.append(phi - interior(P, P.get_vec(0)))
                P.set_unet(None, 0)
        rate = lr * 0.1 ** (ps / max(passes - 1, 1))
        losses = []
        for k, (P, phi, errs) in enumerate(members):
            net.clear_samples()
            for slot, e in enumerate(errs):
                net.set_sample(P, slot, e, smooth=nu, omega=omega)
            losses.append(float(np.mean(net.train(P, fit, batch=4, lr=rate,
                                                  seed=1000 * ps + k)[-5:])))
        curve.append(float(np.mean(losses)))
        if log:
            log(f"train_problem_net: pass {ps} lr {rate:.1e} mean loss {curve[-1]:.3f}")
    nsamp = sum(len(errs) for _, _, errs in members)
    net.clear_samples()
    for P, _, _ in members:
        P.close()
    return time.perf_counter() - t0, nsamp, curve
