"""Headline benchmark: Poisson solves/s to true relative residual 1e-8 on a
1024x1024 grid, one MI355X.

A step is one complete pressure-Poisson solve from a zero initial guess on
an all-Neumann 1024^2 grid whose right-hand side is resident in HBM.  The
timed method is the fastest path that reaches the FP64 tolerance (FP32
Jacobi-PCG inside FP64 iterative refinement); plain FP64 Jacobi-PCG on the
same GPU and the CPU oracle are reported beside it.

The roofline figure is the GPU time per launch of the dominant kernel of
the timed method, measured in this process with HIP events around a
replayed hipGraph of back-to-back launches of that kernel
(cpgpu_poisson_kernel_times), against the kernel's algorithmic bytes.

Prints one JSON line.  --gpus N>1 exits: this build runs on one GPU.
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
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the plain FP64 J-PCG comparison")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the Krylov batches directly instead of as "
                         "replayed hipGraphs (same kernels; for profilers)")
    a = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        log("bench.py: only --gpus 1 is supported (single-GPU build)")
        sys.exit(2)

    import carpios_amd as ca

    n = a.size
    N = n * n
    ctx = ca.Context(0)
    desc = ca.uniform_grid_desc(n, n, 1, [ca.BC_NEUMANN] * 4)
    P = ca.Poisson(ctx, desc)
    P.set_source_blobs(seed=0, nblobs=64, wmin=0.01, wmax=0.1)
    P.set_mixed(inner_tol=1e-4, max_outer=40)
    if a.no_graphs:
        P.set_batch(16, use_graphs=False)
    tune = P.autotune(reps=20)
    log(f"bench: {n}^2 all-Neumann blob problem ready; autotune {tune}")

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
    t0 = time.perf_counter()
    iters = []
    for _ in range(a.steps):
        it, res = solve("MIXED")
        iters.append(it)
    ctx.sync()
    wall = time.perf_counter() - t0
    value = a.steps / wall
    true_res = P.residual()
    log(f"bench: mixed {value:.2f} solves/s, {iters[-1]} inner it, true res {true_res:.2e}")

    # ---- GPU baseline: plain FP64 Jacobi-PCG ----
    fp64 = None
    it64 = None
    if not a.no_fp64:
        nb = max(1, min(a.steps, 3))
        solve("JPCG")
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(nb):
            it64, res64 = solve("JPCG")
        ctx.sync()
        rate = nb / (time.perf_counter() - t0)
        fp64 = {"solves_per_s": round(rate, 4), "iterations": it64,
                "speedup_of_value": round(value / rate, 3)}
        log(f"bench: fp64 J-PCG {rate:.2f} solves/s, {it64} it")

    # ---- dominant kernel of the headline solve: GPU time per launch ----
    # One more solve of the same problem with HIP events around every
    # launch of the two Krylov kernel families (direct launches on the
    # context stream; the timed solves replay the same kernels as graphs).
    ctx.timer_enable(0)
    ctx.timer_enable(1)
    solve("MIXED")
    n_d, ms_d = ctx.timer_read(0)
    n_u, ms_u = ctx.timer_read(1)
    ctx.timer_enable(0, False)
    ctx.timer_enable(1, False)
    d32, u32 = ms_d * 1e3 / max(n_d, 1), ms_u * 1e3 / max(n_u, 1)
    fam = [(ms_d, d32, n_d, 4, f"k_cg_dir (variant {tune['fp32'][0]}) <float>: "
                                "p update + A p + p.q"),
           (ms_u, u32, n_u, 6, "k_cg_upd<float>: x, r update + r.z, r.r")]
    _, us, launches, nvec, kname = max(fam)
    alg_bytes = nvec * N * 4
    achieved = alg_bytes / (us * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": kname, "bytes_per_launch": alg_bytes,
                "avg_launch_us": round(us, 3), "launches": launches,
                "other_kernel_us": round(min(d32, u32), 3),
                "iteration_us": round(d32 + u32, 3),
                "note": f"{n}^2 FP32 vectors ({N * 4 / 2**20:.0f} MiB each) are "
                        "Infinity-Cache resident; see roofline_hbm_probe for HBM"}

    # ---- HBM-resident probe: the same kernels at 4096^2 in FP64 ----
    hbm = None
    if not a.no_hbm:
        m = 4096
        Pm = ca.Poisson(ctx, ca.uniform_grid_desc(m, m, 1, [ca.BC_NEUMANN] * 4))
        Pm.set_source_blobs(seed=0)
        tune_m = Pm.autotune(reps=5)
        d64, u64 = Pm.kernel_times(64, reps=20)
        Nm = m * m
        d_gbs = 4 * Nm * 8 / (d64 * 1e-6) / 1e9
        u_gbs = 6 * Nm * 8 / (u64 * 1e-6) / 1e9
        hbm = {"grid": "4096^2 FP64", "dir_variant": tune_m["fp64"][0],
               "k_cg_dir_us": round(d64, 2), "k_cg_dir_GBs": round(d_gbs, 1),
               "k_cg_dir_frac": round(d_gbs / HBM_PEAK_GBS, 4),
               "k_cg_upd_us": round(u64, 2), "k_cg_upd_GBs": round(u_gbs, 1),
               "k_cg_upd_frac": round(u_gbs / HBM_PEAK_GBS, 4)}
        Pm.close()
        log(f"bench: 4096^2 fp64 dir {d_gbs:.0f} GB/s, upd {u_gbs:.0f} GB/s")

    # ---- CPU baseline: the oracle's FP64 J-PCG, bounded sample ----
    cpu = None
    if not a.no_cpu and it64:
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
                   "autotune": tune, "true_rel_residual": true_res},
        "gpu_fp64_jpcg": fp64,
        "roofline": roofline,
        "roofline_hbm_probe": hbm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    P.close()
    ctx.close()


if __name__ == "__main__":
    main()
