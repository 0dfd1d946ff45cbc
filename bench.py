"""Headline benchmark: Poisson solves/s to true relative residual 1e-8 on a
1024x1024 grid, one MI355X.

A step is one complete pressure-Poisson solve from a zero initial guess on
an all-Neumann 1024^2 grid whose right-hand side is resident in HBM.  The
timed method is the fastest path that reaches the FP64 tolerance (FP32
Jacobi-PCG inside FP64 iterative refinement); plain FP64 Jacobi-PCG on the
same GPU and the CPU oracle are reported beside it.

Prints one JSON line (contract in the repository docs).  --gpus N>1 exits:
this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hbm", action="store_true",
                    help="skip the 4096^2 HBM roofline probe")
    a = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        log("bench.py: only --gpus 1 is supported (single-GPU build)")
        sys.exit(2)

    import numpy as np

    import carpios_amd as ca

    n = a.size
    N = n * n
    ctx = ca.Context(0)
    desc = ca.uniform_grid_desc(n, n, 1, [ca.BC_NEUMANN] * 4)
    P = ca.Poisson(ctx, desc)
    P.set_source_blobs(seed=0, nblobs=64, wmin=0.01, wmax=0.1)
    P.set_mixed(inner_tol=1e-4, max_outer=40)
    log(f"bench: {n}^2 all-Neumann blob problem ready")

    def solve(method):
        P.fill_phi(0.0)
        code, it, res, _ = P.solve(method, max_iter=10 ** 7, tol=a.tol,
                                   norm=ca.NORM_TRUE,
                                   opts=ca.OPT_PROJECT_MEAN if method == "MIXED" else 0)
        if code != 0:
            raise RuntimeError(f"{method} did not converge: code {code} res {res}")
        return it, res

    # ---- headline: FP32 inner + FP64 refinement ----
    for _ in range(a.warmup):
        solve("MIXED")
    ctx.sync()
    ctx.timer_enable(0, True)
    ctx.timer_enable(1, True)
    t0 = time.perf_counter()
    iters = []
    for s in range(a.steps):
        it, res = solve("MIXED")
        iters.append(it)
    ctx.sync()
    t1 = time.perf_counter()
    n0, ms0 = ctx.timer_read(0)
    n1, ms1 = ctx.timer_read(1)
    ctx.timer_enable(0, False)
    ctx.timer_enable(1, False)
    wall = t1 - t0
    value = a.steps / wall
    true_res = P.residual()
    log(f"bench: mixed {value:.2f} solves/s, {iters[-1]} inner it, true res {true_res:.2e}")

    # ---- GPU baseline: plain FP64 Jacobi-PCG ----
    nb = max(1, min(a.steps, 3))
    solve("JPCG")
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(nb):
        it64, res64 = solve("JPCG")
    ctx.sync()
    fp64_rate = nb / (time.perf_counter() - t0)
    log(f"bench: fp64 J-PCG {fp64_rate:.2f} solves/s, {it64} it")

    # dominant kernel family of the headline solve
    fam = [(ms0, n0, 4, "k_cg_dir<float> (p update + A p + p.q)"),
           (ms1, n1, 6, "k_cg_upd<float> (x,r update + r.z, r.r)")]
    ms, launches, nvec, kname = max(fam)
    avg_s = ms / 1e3 / max(launches, 1)
    alg_bytes = nvec * N * 4
    achieved = alg_bytes / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": kname, "bytes_per_launch": alg_bytes,
                "avg_launch_us": round(avg_s * 1e6, 3), "launches": launches,
                "note": "1024^2 FP32 vectors (4 MiB each) stay in the 256 MiB "
                        "Infinity Cache, so this rate can exceed HBM"}

    # ---- HBM-resident probe: FP64 stencil kernel at 4096^2 ----
    hbm = None
    if not a.no_hbm:
        m = 4096
        Pm = ca.Poisson(ctx, ca.uniform_grid_desc(m, m, 1, [ca.BC_NEUMANN] * 4))
        Pm.set_source_blobs(seed=0)
        Pm.fill_phi(0.0)
        Pm.solve("JPCG", max_iter=4, tol=1e-30, norm=ca.NORM_TRUE)
        ctx.sync()
        ctx.timer_enable(0, True)
        ctx.timer_enable(1, True)
        Pm.fill_phi(0.0)
        Pm.solve("JPCG", max_iter=60, tol=1e-30, norm=ca.NORM_TRUE)
        k0, t0ms = ctx.timer_read(0)
        k1, t1ms = ctx.timer_read(1)
        ctx.timer_enable(0, False)
        ctx.timer_enable(1, False)
        Nm = m * m
        d_gbs = 4 * Nm * 8 / (t0ms / 1e3 / k0) / 1e9
        u_gbs = 6 * Nm * 8 / (t1ms / 1e3 / k1) / 1e9
        hbm = {"grid": "4096^2 FP64", "k_cg_dir_GBs": round(d_gbs, 1),
               "k_cg_dir_frac": round(d_gbs / HBM_PEAK_GBS, 4),
               "k_cg_upd_GBs": round(u_gbs, 1),
               "k_cg_upd_frac": round(u_gbs / HBM_PEAK_GBS, 4),
               "iteration_us": round((t0ms / k0 + t1ms / k1) * 1e3, 2)}
        Pm.close()
        log(f"bench: 4096^2 fp64 dir {d_gbs:.0f} GB/s, upd {u_gbs:.0f} GB/s")

    # ---- CPU baseline: the oracle's FP64 J-PCG, bounded sample ----
    cpu = None
    if not a.no_cpu:
        import oracle as O
        og = O.StructGrid(n, n, 1)
        A = O.Matrix(og, O.make_bc(og, [2, 2, 2, 2]), 1.0)
        b = O.field_to_x(og, P.get_vec(1))
        k = 150
        t0 = time.perf_counter()
        O.solve("JPCG", A, b, max_iter=k, tol=1e-300, norm=O.NORM_TRUE)
        t_it = (time.perf_counter() - t0) / k
        cpu = {"value": round(1.0 / (t_it * it64), 5), "unit": "solves/s",
               "cores": 1, "kind": "port",
               "sample": f"{k} FP64 J-PCG iterations of the same {n}^2 problem "
                         f"(oracle C, -O2, 1 thread), {t_it * 1e3:.2f} ms/iteration, "
                         f"extrapolated to the {it64} iterations FP64 J-PCG needs"}
        log(f"bench: cpu {cpu['value']} solves/s")

    out = {
        "metric": "Poisson solves/s to rel-res 1e-8 @1024^2: learned+FP32/FP64-refine vs plain FP64",
        "value": round(value, 4), "unit": "solves/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 inner / fp64 refinement",
        "data": "synthetic",
        "config": {"workload": f"{n}x{n} all-Neumann pressure Poisson, 64 Gaussian "
                               f"blobs (seed 0) zero-mean RHS, zero initial guess, "
                               f"true rel-res {a.tol:g}; FP32 Jacobi-PCG inner "
                               f"(rel 1e-4) in FP64 iterative refinement",
                   "method": "MIXED", "inner_iterations": iters[-1],
                   "true_rel_residual": true_res},
        "gpu_fp64_jpcg": {"solves_per_s": round(fp64_rate, 4), "iterations": it64,
                          "speedup_of_value": round(value / fp64_rate, 3)},
        "roofline": roofline,
        "roofline_hbm_probe": hbm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    P.close()
    ctx.close()


if __name__ == "__main__":
    main()
