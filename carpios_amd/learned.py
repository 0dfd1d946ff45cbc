"""Training workflow of the learned preconditioner (no reference counterpart).

The network is trained on errors of the problem family it will precondition:
for a right-hand side b, FP64 Jacobi-PCG gives phi and its own iterates
x_k; every e_k = phi - x_k is a sample whose residual A e_k = b - A x_k the
library forms on the device.  The initial error (x_0 = 0) is smooth; later
errors carry the slowly converging components Jacobi-PCG leaves behind --
the spectrum a flexible-CG preconditioner has to cover.  Everything runs
through libcpgpu.so.
"""
import numpy as np

from . import NORM_TRUE


def interior(P, field):
    g = P.gl
    return np.ascontiguousarray(field[g:g + P.nx, g:g + P.ny])


def add_error_samples(P, net, seeds, checkpoints=(0, 16, 64, 256), nblobs=64,
                      wmin=0.01, wmax=0.1, tol=1e-10, slot0=0):
    """Samples e_k = phi - x_k of Jacobi-PCG on blob right-hand sides.

    For each seed: b from 'nblobs' Gaussian blobs, phi by FP64 Jacobi-PCG to
    the true relative residual 'tol' from zero, then the iterates x_k after
    k = checkpoints Jacobi-PCG iterations from zero.  Returns the next free
    slot.  Leaves the problem's right-hand side at the last seed's."""
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
            net.set_sample(P, slot, e)
            slot += 1
    return slot


def train(net, P, steps, batch=4, lr=1e-3, seed=0, decay_at=0.7):
    """Adam at lr for the first decay_at of the steps, then lr / 10.
    Returns the per-step mean losses."""
    n1 = int(steps * decay_at)
    l1 = net.train(P, n1, batch=batch, lr=lr, seed=seed)
    l2 = net.train(P, steps - n1, batch=batch, lr=lr * 0.1, seed=seed + 1)
    return np.concatenate([l1, l2])
