"""Kernel probe of the Krylov iteration kernels, for profilers.

Sets up the 1024^2 all-Neumann blob problem (64 Gaussian blobs, seed 0) and
either
  --mode times: replays hipGraphs of back-to-back launches of the direction
                and update kernels (cpgpu_poisson_kernel_times), or
  --mode solve: runs one solve (default MIXED) with directly launched
                batches (no hipGraph), the launch pattern of a real solve, or
  --mode unet:  runs --reps forward passes of the U-Net (seed 1) on a
                --size^2 input (the network's kernels alone);
At --size 4096 with a loose --tol, the solve's k_merge_x launches
(x64 += x32, x32 = 0: 12 bytes read and 12 written per cell, over vectors
far larger than the caches) calibrate hardware byte counters.

    python -m carpios_amd.kernel_probe --mode solve --variant 0 --tol 1e-4
"""
import argparse
import json
import sys


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--mode", choices=("times", "solve", "unet"), default="times")
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--method", default="MIXED")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--variant", type=int, default=-1,
                    help="direction-kernel variant (default: autotune)")
    a = ap.parse_args(argv)

    import carpios_amd as ca

    n = a.size
    ctx = ca.Context(0)
    if a.mode == "unet":
        import numpy as np
        net = ca.UNet(ctx, seed=1)
        x = np.random.default_rng(0).standard_normal((n, n)).astype(np.float32)
        ctx.timer_enable(2)
        for _ in range(a.reps):
            net.forward(x)
        k, ms = ctx.timer_read(2)
        net.close()
        ctx.close()
        # net.forward keeps every activation: both finest-level 16 -> 16
        # convolutions store their output (one also its 4 x 4 pooling)
        print(json.dumps({"size": n, "mode": "unet", "reps": a.reps,
                          "conv16_launches": k, "conv16_us": round(ms * 1e3 / max(k, 1), 3)}),
              flush=True)
        return 0
    P = ca.Poisson(ctx, ca.uniform_grid_desc(n, n, 1, [ca.BC_NEUMANN] * 4))
    P.set_source_blobs(seed=0, nblobs=64, wmin=0.01, wmax=0.1)
    tune = None
    if a.variant >= 0:
        P.set_variant(a.variant, a.variant)
    else:
        tune = P.autotune(reps=10)
    out = {"size": n, "mode": a.mode, "autotune": tune, "variant": a.variant}
    if a.mode == "times":
        d, u = P.kernel_times(a.precision, reps=a.reps)
        out.update(precision=a.precision, reps=a.reps, dir_us=round(d, 3),
                   upd_us=round(u, 3))
    else:
        P.set_batch(16, use_graphs=False)
        P.fill_phi(0.0)
        code, it, res, _ = P.solve(a.method, max_iter=10 ** 7, tol=a.tol,
                                   norm=ca.NORM_TRUE, opts=ca.OPT_PROJECT_MEAN)
        out.update(method=a.method, tol=a.tol, code=code, iterations=it,
                   residual=res)
    P.close()
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
