"""Headline benchmark: pressure-Poisson solves/s of a 1024^2 lid-driven cavity
(Re 1000) to the true relative residual 1e-8, on one MI355X.

A step is one projection step of the cavity (NS_explicit_: TVD advection,
explicit diffusion, the pressure-correction Poisson solve warm-started from
the previous step, the velocity correction), all resident in HBM.  The
Poisson solve dominates it.  The timed method is flexible CG preconditioned
by the BF16 U-Net inside a Jacobi-smoothing cycle (FP64 Krylov vectors,
FP32 smoothing, BF16/FP32 network), with FP64 Jacobi-PCG as its fallback.
The network is trained first, on the GPU with the library's own kernels,
from the plain solver's steps of the same run and from its own flexible-CG
iterates (untimed; its cost is reported).  Then the same K steps are timed
with the learned solver, with plain FP64 Jacobi-PCG and with FP32 Jacobi-PCG
in FP64 refinement, each from the same saved state; their fields are
compared.

The roofline figure is the average duration of the dominant U-Net kernel
(the 16 -> 16 3x3 convolution of the finest level that stores its output and
its 4 x 4 pooling) over one more learned step, from timestamps the dispatch
itself carries (hipExtLaunchKernelGGL events), against its algorithmic
bytes.  The same kernel's other launch in an applied network (its epilogue
applies the output layer and stores no activation) is reported beside it.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# HBM bytes per launch of k_dec0_apply at 1024^2, from rocprofv3 PMC passes
# (FETCH_SIZE x 2.0 and WRITE_SIZE x 1.0, the factors a copy kernel of known
# size reads): profiles/k_dec0_apply_1024_traffic_pmc.txt
DEC0_TRAFFIC_1024 = (2 * 19210.3 + 4096.0) * 1024


def cpu_threads():
    """Host threads of the CPU baseline: the GPU's share of the host (16 per
    GPU), or fewer where the machine has fewer."""
    env = os.environ.get("OMP_NUM_THREADS")
    n = int(env) if env and env.isdigit() else 16
    return max(1, min(n, os.cpu_count() or 1))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def batched(ctx, ca, tol, no_cpu, learned_net=True):
    """256 independent Poisson problems (sides 2^6..2^10 drawn from seed 2,
    aspect <= 4, random Dirichlet/Neumann walls, Gaussian-blob sources)
    solved to the true relative residual `tol` from zero by the batched
    solver: plain FP64 Jacobi-PCG and FP32 Jacobi-PCG with FP64 residual
    replacement, each after its tile autotune, and (learned_net) the trained
    U-Net preconditioner on the largest members."""
    t0 = time.perf_counter()
    probs = ca.random_problems(ctx, seed=2, n=256)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    cells = np.array([P.nx * P.ny for P in probs], dtype=np.int64)
    B = ca.Batch(ctx, probs)
    res = {"problems": len(probs), "cells": int(cells.sum()),
           "sides": {"min": int(min(min(P.nx, P.ny) for P in probs)),
                     "max": int(max(max(P.nx, P.ny) for P in probs))},
           "setup_s": round(t_setup, 2)}
    sols, iters = {}, {}
    for method, s in (("JPCG", 8), ("MIXED", 4)):
        tpb, ms = B.autotune(method, reps=2)
        runs = {}
        # graph-replayed batches polled by the host; the FP64 pass also as
        # one persistent cooperative launch (bitwise the same results; the
        # mixed pass has no persistent form in use)
        for mode in (("launched", "persistent") if method == "JPCG" else ("launched",)):
            B.set_persistent(mode == "persistent")
            for P in probs:
                P.fill_phi(0.0)
            ctx.sync()
            t = time.perf_counter()
            code, it, r, codes, lat = B.solve(method, max_iter=10 ** 6, tol=tol)
            wall = time.perf_counter() - t
            if code != ca.OK or r.max() > tol:
                raise RuntimeError(f"batched {method} ({mode}): code {code}, "
                                   f"max residual {r.max():.3e}")
            byts = float((it.astype(np.float64) * 11 * cells * s).sum())
            runs[mode] = ({"solves_per_s": round(len(probs) / wall, 2), "wall_s": round(wall, 4),
                           "cells_per_s": round(float(cells.sum()) / wall, 1),
                           "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
                           "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
                           "iterations_total": int(it.sum()), "iterations_max": int(it.max()),
                           "tiles_per_block": tpb,
                           "algorithmic_GBs": round(byts / wall / 1e9, 1),
                           "hbm_frac": round(byts / wall / 1e9 / HBM_PEAK_GBS, 4)},
                          [P.get_phi() for P in probs], it.copy())
        res[method] = dict(runs["launched"][0],
                           launch="graph-replayed batches of 16 iterations, host-polled",
                           autotune_ms_per_iter=[round(float(v), 4) for v in ms])
        if "persistent" in runs:
            res[method]["persistent_launch"] = dict(
                runs["persistent"][0], launch="one persistent cooperative launch",
                bitwise_equal_to_launched=all(
                    np.array_equal(a_, b_) for a_, b_ in
                    zip(runs["launched"][1], runs["persistent"][1])))
        sols[method] = runs["launched"][1]
        iters[method] = runs["launched"][2]
        log(f"bench: batched {method}: {res[method]}")
    errs = [rel_l2(a_, b_) for a_, b_ in zip(sols["MIXED"], sols["JPCG"])]
    res["MIXED"]["rel_l2_vs_fp64_max"] = float(max(errs))
    res["mixed_speedup"] = round(res["MIXED"]["solves_per_s"] / res["JPCG"]["solves_per_s"], 3)
    if learned_net:
        # the U-Net preconditioner on the largest members (>= 2^19 cells),
        # trained on members of another draw (seed 7); the rest, and learned
        # members that stall or pass 150 flexible-CG iterations, batched
        from carpios_amd import learned
        net = ca.UNet(ctx, seed=1)
        t_tr, n_s, curve = learned.train_problem_net(ctx, net, log=log)
        thr = 1 << 19
        for k, P in enumerate(probs):
            P.set_fcg(2, 2, 0.8)
            P.set_unet(net if cells[k] >= thr else None, ca.UNET_PRECOND if cells[k] >= thr else 0)
            P.fill_phi(0.0)
        ctx.sync()
        t = time.perf_counter()
        code, it, r, codes, lat = B.solve("FCG_UNET", max_iter=10 ** 6, tol=tol)
        wall = time.perf_counter() - t
        if code != ca.OK or r.max() > tol:
            raise RuntimeError(f"batched FCG_UNET: code {code}, max residual {r.max():.3e}")
        m = cells >= thr
        errs = [rel_l2(P.get_phi(), s_) for P, s_ in zip(probs, sols["JPCG"])]
        res["FCG_UNET"] = {"solves_per_s": round(len(probs) / wall, 2), "wall_s": round(wall, 4),
                           "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
                           "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
                           "learned_members": int(m.sum()), "min_cells": thr,
                           "learned_iterations": sorted(int(v) for v in it[m]),
                           "speedup_vs_fp64": round(res["JPCG"]["wall_s"] / wall, 3),
                           "rel_l2_vs_fp64_max": float(max(errs)),
                           "training": {"seconds": round(t_tr, 1), "samples": n_s,
                                        "loss_per_pass": [round(v, 3) for v in curve]}}
        log(f"bench: batched FCG_UNET: {res['FCG_UNET']}")
        for P in probs:
            P.set_unet(None, 0)
        net.close()
    if not no_cpu:
        # the oracle's FP64 J-PCG time per cell-iteration on the member
        # closest to 512^2 cells, times the batch's own FP64 iteration counts
        import oracle as O
        k = int(np.argmin(np.abs(cells - 512 * 512)))
        Pk = probs[k]
        og = O.StructGrid(Pk.nx, Pk.ny, 1)
        A = O.Matrix(og, O.make_bc(og, [int(t) for t in Pk.desc.bc_type]), 1.0)
        b = O.field_to_x(og, Pk.get_vec(1))
        m = 40
        t = time.perf_counter()
        O.solve("JPCG", A, b, max_iter=m, tol=1e-300, norm=O.NORM_TRUE)
        per = (time.perf_counter() - t) / (m * Pk.nx * Pk.ny)
        total = float((iters["JPCG"].astype(np.float64) * cells).sum()) * per
        res["cpu_baseline"] = {"value": round(len(probs) / total, 5), "unit": "solves/s",
                               "kind": "port", "cores": 1,
                               "sample": f"{m} oracle FP64 J-PCG iterations on member {k} "
                                         f"({Pk.nx}x{Pk.ny}), {per * 1e9:.2f} ns per "
                                         f"cell-iteration, times the batch's FP64 "
                                         f"iteration counts ({total:.0f} s)"}
    B.close()
    for P in probs:
        P.close()
    return res


def kim_probe(ctx, ca, no_cpu, n=512, steps=10):
    """The NS_kim_ step (cpgpu_ns_set_kim) on the n^2 lid-driven cavity,
    Re 1000 (rho 1, mu 1e-3), fixed dt = 0.5 h: upwind predictor source, two
    Helmholtz solves with the reference's default Solver_Jacobi_ (1000
    iterations, 1e-4) and the J-PCG projection to 1e-10.  The first step
    from rest is timed alone, then `steps` more; the oracle's first step at
    the same size is the CPU baseline."""
    import time as _t
    h = 1.0 / n
    kw = dict(bcu=((1, 1, 1, 1), [0, 0, 0, 1.0]), bcv=((1, 1, 1, 1), None),
              rho=1.0, mu=1e-3, solver="JPCG", tol=1e-10, dt=0.5 * h)
    warm = ca.NS(ctx, 64, 64, 2, **kw)
    warm.set_kim(True)
    warm.step(2)
    ns = ca.NS(ctx, n, n, 2, **kw)
    ns.set_kim(True)
    ctx.sync()
    t0 = _t.perf_counter()
    it0, _ = ns.step(1)
    ctx.sync()
    first_ms = (_t.perf_counter() - t0) * 1e3
    helm0 = ns.kim_iters()
    t0 = _t.perf_counter()
    its = []
    helm = []
    for _ in range(steps):
        it, _ = ns.step(1)
        its.append(int(it[0]))
        helm.append(ns.kim_iters())
    ctx.sync()
    ms = (_t.perf_counter() - t0) * 1e3 / steps
    out = {"workload": f"{n}x{n} cavity Re=1000, NS_kim_ step, dt=0.5h fixed",
           "first_step_ms": round(first_ms, 3), "ms_per_step": round(ms, 3),
           "steps_per_s": round(1e3 / ms, 2),
           "projection_iterations": [int(it0[0])] + its,
           "helmholtz_iterations": [list(helm0)] + [list(x) for x in helm]}
    if not no_cpu:
        import oracle as O
        g = O.StructGrid(n, n, 2, 0.0, 1.0, 0.0, 1.0)
        ora = O.NSKim(g, O.make_bc(g, [O.BC1] * 4, [0, 0, 0, 1.0]),
                      O.make_bc(g, [O.BC1] * 4), rho=1.0, mu=1e-3, tol=1e-10,
                      dt=0.5 * h)
        t0 = _t.perf_counter()
        ora.step()
        cpu_ms = (_t.perf_counter() - t0) * 1e3
        out["cpu_baseline"] = {"first_step_ms": round(cpu_ms, 1), "cores": 1,
                               "kind": "port",
                               "sample": "the oracle's first NS_kim_ step from rest at the same size"}
        out["first_step_speedup_vs_cpu"] = round(cpu_ms / first_ms, 1)
        out["first_step_iterations_match"] = (ora.iters[0] == int(it0[0]) and
                                              tuple(ora.helm_iters[0]) == tuple(helm0))
    log(f"bench: kim {out}")
    return out


def ns_kernel_times(ctx, ns, N):
    """One more step of `ns` with the NS kernels' own dispatch timestamps
    against their algorithmic bytes per cell: the fused predictor reads u, v,
    u f, v f and writes u*, v*, u*f, v*f, div and the projection's source
    (80 B); the correction reads phi, u*f, v*f, u*, v* and writes u*f, v*f,
    u*, v*, p, phi (88 B)."""
    fam = {4: ("k_ns_predict", 80), 6: ("k_grad_correct", 88)}
    for f in fam:
        ctx.timer_enable(f)
    ns.step(1)
    kern = {}
    for f, (name, bpc) in fam.items():
        k, ms = ctx.timer_read(f)
        ctx.timer_enable(f, False)
        us = ms * 1e3 / max(k, 1)
        gbs = bpc * N / (us * 1e-6) / 1e9
        kern[name] = {"us": round(us, 3), "bytes_per_cell": bpc, "GBs": round(gbs, 1),
                      "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    return kern


def ns_explicit_probe(ctx, ca, no_cpu, n=512, steps=100, warm=2, graphs=True):
    """The plain method on the GPU: the n^2 lid-driven cavity (Re 1000,
    rho = Re, mu = 1, van Leer TVD, CFL 0.5, dt0 = 0.5 h, diffusion number
    0.25) stepped by NS_explicit_ with FP64 Jacobi-PCG projections to 1e-10
    (the reference's nrm2 test), from rest.  `warm` steps, then `steps`
    timed; then one more step with the NS kernels' own dispatch timestamps
    against their algorithmic bytes; the oracle's first step at the same
    size is the CPU baseline."""
    h = 1.0 / n
    kw = dict(bcu=((1, 1, 1, 1), [0, 0, 0, 1.0]), bcv=((1, 1, 1, 1), None),
              rho=1000.0, mu=1.0, cfl=0.5, dnum=0.25, dcoef=1e-3,
              limiter="VanLeer", solver="JPCG", tol=1e-10, norm=ca.NORM_REF,
              dt=0.5 * h)
    ns = ca.NS(ctx, n, n, 2, (-0.5, 0.5), (-0.5, 0.5), **kw)
    P = ns.projection()
    if graphs:
        tune = P.autotune(reps=20)
    else:
        # profiled runs: direct launches, the default direction kernel
        P.set_batch(16, use_graphs=False)
        tune = None
    it0, _ = ns.step(warm)
    ctx.sync()
    t0 = time.perf_counter()
    its, res = ns.step(steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    N = n * n
    kern = ns_kernel_times(ctx, ns, N)
    t, dt = ns.time()
    it_sum = int(np.sum(its))
    out = {"workload": f"{n}x{n} lid-driven cavity Re=1000, NS_explicit_ (van Leer, CFL 0.5, "
                       f"dt0=0.5h), FP64 J-PCG projection to 1e-10 (reference nrm2)",
           "steps": steps, "warmup": warm, "ms_per_step": round(wall / steps * 1e3, 3),
           "steps_per_s": round(steps / wall, 2), "dt": dt, "t_end": t,
           "projection_iterations": {"first": [int(v) for v in it0], "mean": round(it_sum / steps, 1),
                                     "last": int(its[-1])},
           "jpcg_algorithmic_GBs": round(it_sum * 10 * N * 8 / wall / 1e9, 1),
           "autotune": tune, "kernels": kern,
           "note": "kernel figures: L2/MALL-resident at this size (16 fields of 2.1 MB)"}
    # the same kernels where the fields far exceed the caches: one step of
    # the 4096^2 cavity with the projection cut at 16 iterations (the NS
    # kernels' work does not depend on the projection's iteration count)
    m = 4096
    big = ca.NS(ctx, m, m, 2, (-0.5, 0.5), (-0.5, 0.5),
                **dict(kw, dt=0.5 / m, max_iter=16))
    big.step(1)
    out["kernels_4096"] = ns_kernel_times(ctx, big, m * m)
    big.close()
    if not no_cpu:
        import oracle as O
        g = O.StructGrid(n, n, 2, -0.5, 0.5, -0.5, 0.5)
        ora = O.NS(g, O.make_bc(g, [1] * 4, [0, 0, 0, 1.0]), O.make_bc(g, [1] * 4),
                   rho=1000.0, mu=1.0, cfl=0.5, dnum=0.25, dcoef=1e-3, limiter="VanLeer",
                   tol=1e-10, dt=0.5 * h)
        c0 = time.perf_counter()
        _, it_cpu, _ = ora.step()
        cpu_s = time.perf_counter() - c0
        # the same step on every thread of the GPU's host share (another
        # summation order in the dots, so its iteration count may differ by one)
        threads = cpu_threads()
        ora2 = O.NS(g, O.make_bc(g, [1] * 4, [0, 0, 0, 1.0]), O.make_bc(g, [1] * 4),
                    rho=1000.0, mu=1.0, cfl=0.5, dnum=0.25, dcoef=1e-3, limiter="VanLeer",
                    tol=1e-10, dt=0.5 * h)
        O.set_threads(threads)
        try:
            c0 = time.perf_counter()
            _, it_cpu2, _ = ora2.step()
            cpu_s2 = time.perf_counter() - c0
        finally:
            O.set_threads(1)
        out["cpu_baseline"] = {"first_step_ms": round(cpu_s2 * 1e3, 1), "cores": threads,
                               "kind": "port", "iterations": it_cpu2,
                               "single_thread_first_step_ms": round(cpu_s * 1e3, 1),
                               "sample": "the oracle's first NS_explicit_ step from rest at the "
                                         f"same size (C, -O2, OpenMP on {threads} threads; "
                                         "one thread beside it)"}
        out["first_step_iterations_match"] = abs(int(it0[0]) - it_cpu) <= 1
    ns.close()
    log(f"bench: NS_explicit_ {n}^2: {out}")
    return out


def blob_probe(ctx, ca, no_cpu, n=4096, tol=1e-10):
    """The precision configuration: n^2 all-Neumann Poisson, 64 zero-meaned
    Gaussian blobs (seed 0), solved from zero to the true relative residual
    `tol` by plain FP64 Jacobi-PCG and by FP32 Jacobi-PCG inside FP64
    residual replacement with the zero-mean projection; time to solution,
    iterations, algorithmic GB/s (11 N s bytes per iteration, the scope
    table's figure), and the zero-mean L2 difference of the two."""
    P = ca.Poisson(ctx, ca.uniform_grid_desc(n, n, 1, [ca.BC_NEUMANN] * 4))
    P.set_source_blobs(seed=0, nblobs=64, wmin=0.01, wmax=0.1)
    tune = P.autotune(reps=5)
    N = n * n
    res = {"grid": f"{n}x{n}", "tol": tol, "autotune": tune}
    sols = {}
    for method, s in (("JPCG", 8), ("MIXED", 4)):
        P.fill_phi(0.0)
        ctx.sync()
        t0 = time.perf_counter()
        code, it, r, _ = P.solve(method, max_iter=10 ** 7, tol=tol, norm=ca.NORM_TRUE,
                                 opts=ca.OPT_PROJECT_MEAN)
        ctx.sync()
        wall = time.perf_counter() - t0
        true_res = P.residual()
        # the plain method stops on its recurrence residual (the reference's
        # test); the refinement on the true FP64 residual of its solution
        # pair (whose FP64 rounding adds the rounding floor below)
        if code != ca.OK or r > tol:
            raise RuntimeError(f"blob {method}: code {code}, residual {r:.3e}, "
                               f"true residual {true_res:.3e}")
        x = P.get_phi()[1:-1, 1:-1]
        sols[method] = x - x.mean()
        # relative residual that rounding this solution to FP64 alone leaves
        # (a random half-ulp perturbation u: ||A u|| / ||b||)
        u = (np.random.default_rng(0).random(x.shape) - 0.5) * np.spacing(np.abs(x))
        up = np.pad(u, 1, mode="edge")
        Au = 4 * u - up[:-2, 1:-1] - up[2:, 1:-1] - up[1:-1, :-2] - up[1:-1, 2:]
        floor = float(np.linalg.norm(Au) / np.linalg.norm(P.get_vec(1)[1:-1, 1:-1]))
        gbs = it * 11 * N * s / wall / 1e9
        res[method] = {"seconds": round(wall, 4), "iterations": it, "residual": r,
                       "true_residual": true_res, "fp64_rounding_floor": floor,
                       "algorithmic_GBs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"bench: blob {n}^2 {method}: {res[method]}")
    res["MIXED"]["speedup_vs_fp64"] = round(res["JPCG"]["seconds"] / res["MIXED"]["seconds"], 3)
    res["MIXED"]["rel_l2_vs_fp64"] = rel_l2(sols["MIXED"], sols["JPCG"])
    if not no_cpu:
        import oracle as O
        og = O.StructGrid(n, n, 1)
        A = O.Matrix(og, O.make_bc(og, [2] * 4), 1.0)
        b = O.field_to_x(og, P.get_vec(1))
        k = 10
        c0 = time.perf_counter()
        O.solve("JPCG", A, b, max_iter=k, tol=1e-300, norm=O.NORM_TRUE)
        t_it = (time.perf_counter() - c0) / k
        threads = cpu_threads()
        k2 = 100
        O.set_threads(threads)
        try:
            c0 = time.perf_counter()
            O.solve("JPCG", A, b, max_iter=k2, tol=1e-300, norm=O.NORM_TRUE)
            t_it2 = (time.perf_counter() - c0) / k2
        finally:
            O.set_threads(1)
        res["cpu_baseline"] = {"seconds_to_solution": round(t_it2 * res["JPCG"]["iterations"], 1),
                               "ms_per_iteration": round(t_it2 * 1e3, 2), "cores": threads,
                               "kind": "port",
                               "single_thread_ms_per_iteration": round(t_it * 1e3, 2),
                               "sample": f"{k2} oracle FP64 J-PCG iterations at {n}^2 (C, -O2, "
                                         f"OpenMP on {threads} threads; {k} on one thread), "
                                         f"times the GPU FP64 iteration count"}
    P.close()
    return res


def csr_probe(ctx, ca, no_cpu, m=2048, iters=200):
    """The general CSR path (cpgpu_csr: the algebra solvers' MatrixSCR_ on
    the device) on the m^2 five-point Dirichlet Laplacian as an explicit CSR
    matrix (sorted columns, 5 m^2 entries): Jacobi-PCG iterations timed
    without convergence, against the algorithmic bytes of one iteration
    (the product: 12 B per entry + row pointers; the direction formed with
    the product: z and the old direction read, the new direction and the
    product written, 32 B per row; the update: x, r, p, q, d read and x,
    r, z written, 64 B per row -- the kernels here move 16 B per row more,
    a separate direction pass that keeps the product's gather to one
    vector) and the oracle's CPU iterations on the same matrix, on the host
    threads of the CPU baseline."""
    import scipy.sparse as sp
    n = m * m
    T = sp.diags([-np.ones(m - 1), 4.0 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    E = sp.diags([-np.ones(m - 1), -np.ones(m - 1)], [-1, 1])
    A = (sp.kron(sp.identity(m), T) + sp.kron(E, sp.identity(m))).tocsr()
    A.sort_indices()
    rp, col, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    nnz = int(rp[-1])
    b = np.random.default_rng(5).standard_normal(n)
    M = ca.CSR(ctx, rp, col, val)
    def per_iteration(method):
        M.solve(method, np.zeros(n), b, max_iter=5, tol=1e-300, hist_cap=8)  # warm
        walls = {}
        for k in (iters, 2 * iters):
            t0 = time.perf_counter()
            M.solve(method, np.zeros(n), b, max_iter=k, tol=1e-300, hist_cap=k + 2)
            walls[k] = time.perf_counter() - t0
        # the difference of two run lengths: the iterations alone, without
        # the copies of x and b and the initial test
        return (walls[2 * iters] - walls[iters]) / iters, walls[iters]

    t_it, wall = per_iteration("JPCG")
    t_bi, _ = per_iteration("Dia_BiCGSTAB")
    M.close()
    # Dia_BiCGSTAB: two products (12 B per entry + row pointers each), the
    # direction with its product: r, rt, p, v, D^-1 read and p, v written
    # (56 B per row); s with its product: r, v, D^-1 read and s, t written
    # (40 B); the update: D^-1, x, rt, p, s, t read and x, r written (64 B).
    # The kernels here move 40 B per row more (ph and sh written by
    # elementwise passes and gathered by the products).
    alg_bi = 2 * (12 * nnz + 4 * (n + 1)) + 160 * n
    alg = 12 * nnz + 4 * (n + 1) + 32 * n + 64 * n
    out = {"matrix": f"{m}^2 five-point Laplacian, CSR, {n} rows, {nnz} entries",
           "method": "JPCG (IC_CG loop with DiaPre), scalars on the device, state read once "
                     "per 16 iterations",
           "iterations": [iters, 2 * iters], "ms_per_iteration": round(t_it * 1e3, 4),
           "ms_per_iteration_with_copies": round(wall / iters * 1e3, 4),
           "algorithmic_bytes_per_iteration": alg,
           "algorithmic_GBs": round(alg / t_it / 1e9, 1),
           "hbm_frac": round(alg / t_it / 1e9 / HBM_PEAK_GBS, 4),
           "note": "ms_per_iteration from the difference of the two runs' wall times; the "
                   "other figure divides the shorter run's whole call (copies of x and b, the "
                   "initial residual test) by its iterations",
           "dia_bicgstab": {"ms_per_iteration": round(t_bi * 1e3, 4),
                            "algorithmic_bytes_per_iteration": alg_bi,
                            "hbm_frac": round(alg_bi / t_bi / 1e9 / HBM_PEAK_GBS, 4)}}
    if not no_cpu:
        import ctypes as C
        import oracle as O
        H = O.CSR(n, rp.ctypes.data_as(C.POINTER(C.c_int)), col.ctypes.data_as(C.POINTER(C.c_int)),
                  val.ctypes.data_as(C.POINTER(C.c_double)))

        class Wrap:
            pass
        w = Wrap()
        w.c, w.n = H, n
        k = 10
        threads = cpu_threads()
        O.set_threads(threads)
        try:
            c0 = time.perf_counter()
            O.solve("JPCG", w, b, max_iter=k, tol=1e-300)
            t_cpu = (time.perf_counter() - c0) / k
        finally:
            O.set_threads(1)
        out["cpu_baseline"] = {"ms_per_iteration": round(t_cpu * 1e3, 3), "cores": threads,
                               "kind": "port", "sample": f"{k} oracle JPCG iterations on the same CSR"}
        out["speedup_vs_cpu"] = round(t_cpu / t_it, 1)
    log(f"bench: CSR {out}")
    return out


def bicgstab_probe(ctx, ca, m=2048, iters=200):
    """Dia_BiCGSTAB on the structured operator (matrix-free, the m^2
    Dirichlet Laplacian with a random source): time per iteration from the
    difference of two run lengths, against the algorithmic bytes of one
    iteration (the direction kernel's r, rt, p, v, D^-1 read and p, v
    written: 56 B per cell; the s/t kernel's r, v, D^-1 read and s, t
    written: 40 B; the update's D^-1, x, rt, p, s, t read and x, r written:
    64 B)."""
    n = m * m
    desc = ca.uniform_grid_desc(m, m, 1, [1, 1, 1, 1], [0.0, 0.0, 0.0, 0.0], 1.0)
    P = ca.Poisson(ctx, desc)
    src = np.zeros((m + 2, m + 2))
    src[1:-1, 1:-1] = np.random.default_rng(6).standard_normal((m, m))
    P.set_source(src)
    P.fill_phi(0.0)
    P.solve("Dia_BiCGSTAB", max_iter=5, tol=1e-300)  # warm
    walls = {}
    for k in (iters, 2 * iters):
        P.fill_phi(0.0)
        t0 = time.perf_counter()
        code, it, _, _ = P.solve("Dia_BiCGSTAB", max_iter=k, tol=1e-300, hist_cap=k + 2)
        walls[k] = time.perf_counter() - t0
    P.close()
    t_it = (walls[2 * iters] - walls[iters]) / iters
    alg = 160 * n
    out = {"grid": f"{m}^2 Dirichlet Poisson, matrix-free", "iterations": [iters, 2 * iters],
           "ms_per_iteration": round(t_it * 1e3, 4), "algorithmic_bytes_per_iteration": alg,
           "hbm_frac": round(alg / t_it / 1e9 / HBM_PEAK_GBS, 4),
           "method": "Dia_BiCGSTAB, scalars on the device, state read once per 16 iterations"}
    log(f"bench: BiCGSTAB {out}")
    return out


def kernel_probe(ca, mode, size, precision=32, reps=50, method="MIXED", tol=1e-8,
                 variant=-1):
    """Kernels alone, for profilers, on the size^2 all-Neumann blob problem
    (64 Gaussian blobs, seed 0):
      times: replayed graphs of back-to-back launches of the Krylov
             direction and update kernels (cpgpu_poisson_kernel_times);
      solve: one solve (default MIXED) with directly launched batches (the
             launch pattern of a real solve); at 4096^2 with a loose tol its
             k_merge_x launches (x64 += x32 as an exact pair, x32 = 0) stream
             vectors far larger than the caches;
      unet:  `reps` forward passes of the U-Net (seed 1) on a size^2 input;
      apply: `reps` applications of the same network as a preconditioner
             applies it (the fused level-0 kernels), as warm starts of a
             size^2 Poisson problem;
      fcg:   `reps` flexible-CG solves of 20 iterations each with that
             network in the 2 + 2 sweep cycle, on the same problem (the
             wall time per iteration: cycle, network and Krylov kernels)."""
    n = size
    ctx = ca.Context(0)
    if mode == "fcg":
        net = ca.UNet(ctx, seed=1)
        types = [ca.BC_DIRICHLET] * 3 + [ca.BC_NEUMANN]
        P = ca.Poisson(ctx, ca.uniform_grid_desc(n, n, 1, types))
        P.set_source_blobs(seed=0, nblobs=64, wmin=0.01, wmax=0.1)
        P.set_unet(net, ca.UNET_PRECOND)
        P.set_fcg(2, 2, 0.8)
        its, wall = 0, 0.0
        for r in range(reps + 1):
            P.fill_phi(0.0)
            ctx.sync()
            t = time.perf_counter()
            code, it, res, _ = P.solve("FCG_UNET", max_iter=20, tol=1e-300, norm=ca.NORM_TRUE)
            ctx.sync()
            if r:  # the first solve allocates and warms up
                wall += time.perf_counter() - t
                its += it
        P.set_unet(None, 0)
        P.close()
        net.close()
        ctx.close()
        return {"size": n, "mode": "fcg", "reps": reps, "iterations": its,
                "us_per_iteration": round(wall * 1e6 / max(its, 1), 2)}
    if mode == "apply":
        net = ca.UNet(ctx, seed=1)
        types = [ca.BC_DIRICHLET] * 3 + [ca.BC_NEUMANN]
        P = ca.Poisson(ctx, ca.uniform_grid_desc(n, n, 1, types))
        P.set_source_blobs(seed=0, nblobs=64, wmin=0.01, wmax=0.1)
        P.set_unet(net, ca.UNET_WARM)
        for f in (8, 9):
            ctx.timer_enable(f)
        for _ in range(reps):
            P.fill_phi(0.0)
            P.solve("JPCG", max_iter=0, tol=1e-300)
        (ke, mse), (kd, msd) = ctx.timer_read(8), ctx.timer_read(9)
        P.set_unet(None, 0)
        P.close()
        net.close()
        ctx.close()
        return {"size": n, "mode": "apply", "reps": reps, "enc0_launches": ke,
                "enc0_us": round(mse * 1e3 / max(ke, 1), 3), "dec0_launches": kd,
                "dec0_us": round(msd * 1e3 / max(kd, 1), 3)}
    if mode == "unet":
        net = ca.UNet(ctx, seed=1)
        x = np.random.default_rng(0).standard_normal((n, n)).astype(np.float32)
        ctx.timer_enable(2)
        for _ in range(reps):
            net.forward(x)
        k, ms = ctx.timer_read(2)
        net.close()
        ctx.close()
        # net.forward keeps every activation: both finest-level 16 -> 16
        # convolutions store their output (one also its 4 x 4 pooling)
        return {"size": n, "mode": "unet", "reps": reps, "conv16_launches": k,
                "conv16_us": round(ms * 1e3 / max(k, 1), 3)}
    P = ca.Poisson(ctx, ca.uniform_grid_desc(n, n, 1, [ca.BC_NEUMANN] * 4))
    P.set_source_blobs(seed=0, nblobs=64, wmin=0.01, wmax=0.1)
    tune = None
    if variant >= 0:
        P.set_variant(variant, variant)
    else:
        tune = P.autotune(reps=10)
    out = {"size": n, "mode": mode, "autotune": tune, "variant": variant}
    if mode == "times":
        d, u = P.kernel_times(precision, reps=reps)
        out.update(precision=precision, reps=reps, dir_us=round(d, 3), upd_us=round(u, 3))
    else:
        P.set_batch(16, use_graphs=False)
        P.fill_phi(0.0)
        code, it, res, _ = P.solve(method, max_iter=10 ** 7, tol=tol, norm=ca.NORM_TRUE,
                                   opts=ca.OPT_PROJECT_MEAN)
        out.update(method=method, tol=tol, code=code, iterations=it, residual=res)
    P.close()
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--train-plain", type=int, default=100,
                    help="plain-solver cavity steps whose errors seed training")
    ap.add_argument("--rounds", type=int, default=6,
                    help="flexible-CG sample rounds after the first fit")
    ap.add_argument("--first-fit", type=int, default=2000)
    ap.add_argument("--round-fit", type=int, default=2500)
    ap.add_argument("--nu", type=int, default=2,
                    help="weighted-Jacobi pre- and post-sweeps of the learned cycle")
    ap.add_argument("--omega", type=float, default=0.8,
                    help="weight of the learned cycle's Jacobi sweeps")
    ap.add_argument("--round-steps", type=int, default=67,
                    help="cavity steps per flexible-CG sample round (6 samples each: with "
                         "the plain steps' 160, 2572 training samples by default)")
    ap.add_argument("--save-net", default="",
                    help="save the trained network and the cavity state (npz)")
    ap.add_argument("--load-net", default="",
                    help="skip training: load a network and cavity state saved by --save-net")
    ap.add_argument("--learned-only", action="store_true",
                    help="time the learned steps alone and stop (for profilers)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 256-problem batched solve")
    ap.add_argument("--no-batch-learned", action="store_true",
                    help="skip the U-Net variant of the batched solve (its training)")
    ap.add_argument("--no-kim", action="store_true",
                    help="skip the NS_kim_ step probe")
    ap.add_argument("--no-csr", action="store_true",
                    help="skip the general-CSR Jacobi-PCG / Dia_BiCGSTAB and the structured "
                         "Dia_BiCGSTAB probes")
    ap.add_argument("--no-hbm", action="store_true",
                    help="skip the 4096^2 Krylov-kernel HBM probe")
    ap.add_argument("--probe", choices=("times", "solve", "unet", "apply", "fcg"), default="",
                    help="kernels alone on the --size^2 blob problem, for profilers")
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--method", default="MIXED")
    ap.add_argument("--variant", type=int, default=-1,
                    help="direction-kernel variant of --probe (default: autotune)")
    ap.add_argument("--only", default="",
                    help="comma-separated legs to run alone (ns, blob, kim, csr, bicgstab, "
                         "batch)")
    ap.add_argument("--ns-steps", type=int, default=100,
                    help="timed steps of the 512^2 NS_explicit_ cavity leg")
    ap.add_argument("--no-ns", action="store_true",
                    help="skip the 512^2 NS_explicit_ cavity run")
    ap.add_argument("--no-blob", action="store_true",
                    help="skip the 4096^2 blob solve (FP64 vs FP32 + FP64 refinement)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch Jacobi-PCG batches directly instead of as "
                         "replayed hipGraphs (same kernels; for profilers)")
    a = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        log("bench.py: only --gpus 1 is supported (single-GPU build)")
        sys.exit(2)

    import carpios_amd as ca
    from carpios_amd import learned

    if a.probe:
        print(json.dumps(kernel_probe(ca, a.probe, a.size, a.precision, a.reps, a.method,
                                      a.tol, a.variant)), flush=True)
        return

    if a.only:
        # the named legs alone, one JSON line of their results (no headline)
        ctx = ca.Context(0)
        legs = {"ns": lambda: ns_explicit_probe(ctx, ca, a.no_cpu, steps=a.ns_steps,
                                                graphs=not a.no_graphs),
                "blob": lambda: blob_probe(ctx, ca, a.no_cpu),
                "kim": lambda: kim_probe(ctx, ca, a.no_cpu),
                "csr": lambda: csr_probe(ctx, ca, a.no_cpu),
                "bicgstab": lambda: bicgstab_probe(ctx, ca),
                "batch": lambda: batched(ctx, ca, a.tol, a.no_cpu, not a.no_batch_learned)}
        out = {}
        for name in a.only.split(","):
            out[name] = legs[name]()
        print(json.dumps(out), flush=True)
        ctx.close()
        return

    n, K, tol = a.size, a.steps, a.tol
    N = n * n
    ctx = ca.Context(0)
    ns = learned.cavity(ctx, n, tol=tol)
    P = ns.projection()
    if a.no_graphs:
        # the autotuner times its variants with replayed graphs: profiled runs
        # keep the default direction kernel
        P.set_batch(16, use_graphs=False)
    # the training's plain-solver steps run one fixed pair of direction
    # kernels: the autotuner's pick depends on timing, and a network trained
    # on the steps' iterates depends on their last bits
    P.set_variant(5, 0)
    log(f"bench: {n}^2 cavity ready")

    # ---- training (untimed) ----
    net = ca.UNet(ctx, seed=1)
    tr = learned.CavityTrainer(ns, net, nu=a.nu, omega=a.omega, tol=tol)
    t0 = time.perf_counter()
    if a.load_net:
        # a network and cavity state saved by an earlier --save-net run (a
        # profiler run then holds the learned solves without the training)
        z = np.load(a.load_net, allow_pickle=False)
        net.set_params(z["params"])
        ns.restore({"fields": {k: z["f_" + k] for k in ns.STATE},
                    "time": (float(z["t"]), float(z["dt"]))})
        tr.slot = int(z["samples"])
        tr.log = [("loaded", a.load_net)]
    else:
        tr.run(plain=a.train_plain, first_fit=a.first_fit, rounds=a.rounds,
               round_steps=a.round_steps, round_fit=a.round_fit, progress=log)
    ctx.sync()
    t_train = time.perf_counter() - t0
    log(f"bench: trained in {t_train:.1f} s on {tr.slot} samples: {tr.log}")
    tune = None if a.no_graphs else P.autotune(reps=20)
    log(f"bench: Jacobi-PCG kernel autotune {tune}")
    if a.save_net:
        st = ns.save()
        np.savez(a.save_net, params=net.get_params(), t=st["time"][0], dt=st["time"][1],
                 samples=tr.slot, **{"f_" + k: v for k, v in st["fields"].items()})

    # ---- warm-up steps of the learned solver, then the common state ----
    P.set_unet(net, ca.UNET_PRECOND)
    ns.set_solver("FCG_UNET", max_iter=10 ** 7, tol=tol, norm=ca.NORM_TRUE)
    for _ in range(a.warmup):
        ns.step(1)
    state = ns.save()
    t_state = ns.time()[0]

    def timed(solver, flags, tol_=tol):
        ns.restore(state)
        P.set_unet(net if flags else None, flags)
        ns.set_solver(solver, max_iter=10 ** 7, tol=tol_, norm=ca.NORM_TRUE)
        ctx.sync()
        t = time.perf_counter()
        it, r = ns.step(K)  # one call for the K steps (per-step counts returned)
        ctx.sync()
        wall = time.perf_counter() - t
        its, res = [int(v) for v in it], [float(v) for v in r]
        if max(res) > tol_:
            raise RuntimeError(f"{solver}: a step ended at residual {max(res):.3e}")
        fields = {k: ns.get_field(k) for k in ("u", "v", "p")}
        return wall, its, fields

    def field_diff(fa, fb):
        g = 2
        inner = (slice(g, g + n), slice(g, g + n))
        pa_, pb_ = fa["p"][inner], fb["p"][inner]
        return {"u": rel_l2(fa["u"][inner], fb["u"][inner]),
                "v": rel_l2(fa["v"][inner], fb["v"][inner]),
                "p_zero_mean": rel_l2(pa_ - pa_.mean(), pb_ - pb_.mean())}

    wl, itl, fl = timed("FCG_UNET", ca.UNET_PRECOND)
    value = K / wl
    log(f"bench: learned {value:.2f} solves/s, iterations {itl}")
    if a.learned_only:
        print(json.dumps({"metric": "learned solves/s (timed steps only)", "value": round(value, 4),
                          "ms_per_step": round(wl / K * 1e3, 3), "steps": K,
                          "iterations_learned": itl}), flush=True)
        return
    wp, itp, fp = timed("JPCG", 0)
    log(f"bench: plain FP64 J-PCG {K / wp:.2f} solves/s, iterations {itp}")
    wm, itm, fm = timed("MIXED", 0)
    log(f"bench: FP32 J-PCG in FP64 refinement {K / wm:.2f} solves/s, iterations {itm}")
    field_err = field_diff(fl, fp)
    log(f"bench: learned vs plain FP64 fields after {K} steps: {field_err}")
    # the same steps with both solvers at 1e-10: two solutions that each
    # stop at a relative residual tau differ by up to the condition number
    # times tau (~4e5 at 1024^2), so fields that agree to 1e-8 need the
    # tighter stop
    tight = 1e-10
    wl2, itl2, fl2 = timed("FCG_UNET", ca.UNET_PRECOND, tight)
    wp2, itp2, fp2 = timed("JPCG", 0, tight)
    field_tight = {"tol": tight, "rel_l2": field_diff(fl2, fp2),
                   "learned_solves_per_s": round(K / wl2, 3),
                   "plain_fp64_solves_per_s": round(K / wp2, 3),
                   "speedup": round(wp2 / wl2, 3),
                   "iterations_learned": itl2, "iterations_plain_fp64": itp2}
    log(f"bench: at {tight:g}: {field_tight}")

    # ---- dominant kernels of the learned step: timestamps of their own
    # dispatches (the applied network's fused level-0 decoder and encoder)
    ns.restore(state)
    P.set_unet(net, ca.UNET_PRECOND)
    ns.set_solver("FCG_UNET", max_iter=10 ** 7, tol=tol, norm=ca.NORM_TRUE)
    for f in (8, 9):
        ctx.timer_enable(f)
    ns.step(1)
    n_e, ms_e = ctx.timer_read(8)
    n_d, ms_d = ctx.timer_read(9)
    for f in (8, 9):
        ctx.timer_enable(f, False)
    us_d = ms_d * 1e3 / max(n_d, 1)
    us_e = ms_e * 1e3 / max(n_e, 1)
    # decoder: read the level-1 output (N/16 pixels x 32 BF16) and the skip
    # channels (N x 16 BF16), write the FP32 output: 40 B per cell; MFMA
    # work per cell: up-conv 16x32, 3x3 32 -> 16, 3x3 16 -> 16 (2 flop/MAC)
    dec_bytes = 40 * N
    dec_flops = 2 * (16 * 32 + 9 * 16 * 32 + 9 * 16 * 16) * N
    dec_tf = dec_flops / (us_d * 1e-6) / 1e12 if n_d else 0.0
    dec_gbs = dec_bytes / (us_d * 1e-6) / 1e9 if n_d else 0.0
    # encoder: read the FP32 input, write the skip channels and the pooled
    # level-1 input: 4 + 32 + 2 B per cell; first layer 2 -> 16 (FP32 FMA)
    # and 3x3 16 -> 16 on MFMA
    enc_bytes = 38 * N
    enc_flops = 2 * (9 * 2 * 16 + 9 * 16 * 16) * N
    enc_gbs = enc_bytes / (us_e * 1e-6) / 1e9 if n_e else 0.0
    roofline = {"bound": "mfma", "achieved": round(dec_tf, 2), "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(dec_tf / BF16_PEAK_TFLOPS, 4),
                "traffic": int(DEC0_TRAFFIC_1024) if (n == 1024 and DEC0_TRAFFIC_1024) else None,
                "kernel": "k_dec0_apply (fused finest-level decoder of the applied U-Net: "
                          "4x4 up-convolution 32 -> 16, 3x3 32 -> 16, 3x3 16 -> 16, 1x1 "
                          "output layer; BF16 MFMA, persistent 16 x 16 tiles)",
                "flops_per_launch": dec_flops, "bytes_per_launch": dec_bytes,
                "flop_per_byte": round(dec_flops / dec_bytes, 1),
                "avg_launch_us": round(us_d, 3), "launches": n_d,
                "achieved_GBs": round(dec_gbs, 1),
                "hbm_frac": round(dec_gbs / HBM_PEAK_GBS, 4),
                "encoder": {"kernel": "k_enc0_apply (first layer 2 -> 16 + 3x3 16 -> 16 + "
                                      "4x4 pooling)", "avg_launch_us": round(us_e, 3),
                            "launches": n_e, "bytes_per_launch": enc_bytes,
                            "achieved_GBs": round(enc_gbs, 1),
                            "hbm_frac": round(enc_gbs / HBM_PEAK_GBS, 4),
                            "mfma_plus_fma_tflops": round(enc_flops / (us_e * 1e-6) / 1e12, 2)
                            if n_e else None}}
    log(f"bench: fused decoder {us_d:.2f} us/launch ({dec_tf:.0f} TFLOP/s), "
        f"encoder {us_e:.2f} us ({enc_gbs:.0f} GB/s)")

    # ---- Krylov kernels of the plain solve (same step, timed dispatches) ----
    ns.restore(state)
    P.set_unet(None, 0)
    ns.set_solver("JPCG", max_iter=10 ** 7, tol=tol, norm=ca.NORM_TRUE)
    ctx.timer_enable(0)
    ctx.timer_enable(1)
    ns.step(1)
    nd, msd = ctx.timer_read(0)
    nu_, msu = ctx.timer_read(1)
    ctx.timer_enable(0, False)
    ctx.timer_enable(1, False)
    d64, u64 = msd * 1e3 / max(nd, 1), msu * 1e3 / max(nu_, 1)
    krylov = {"k_cg_dir<double>_us": round(d64, 3),
              "k_cg_dir<double>_GBs": round(4 * N * 8 / (d64 * 1e-6) / 1e9, 1),
              "k_cg_upd<double>_us": round(u64, 3),
              "k_cg_upd<double>_GBs": round(6 * N * 8 / (u64 * 1e-6) / 1e9, 1)}

    # ---- HBM-resident probe of the Krylov kernels at 4096^2 FP64 ----
    hbm = None
    if not a.no_hbm:
        m = 4096
        Pm = ca.Poisson(ctx, ca.uniform_grid_desc(m, m, 1, [ca.BC_NEUMANN] * 4))
        Pm.set_source_blobs(seed=0)
        tune_m = Pm.autotune(reps=5)
        dd, uu = Pm.kernel_times(64, reps=20)
        Nm = m * m
        d_gbs = 4 * Nm * 8 / (dd * 1e-6) / 1e9
        u_gbs = 6 * Nm * 8 / (uu * 1e-6) / 1e9
        dd32, uu32 = Pm.kernel_times(32, reps=20)
        copy_gbs, copy_us = ctx.copy_bandwidth(1 << 30, reps=20)
        hbm = {"grid": "4096^2 FP64", "dir_variant": tune_m["fp64"][0],
               "k_cg_dir_us": round(dd, 2), "k_cg_dir_GBs": round(d_gbs, 1),
               "k_cg_dir_frac": round(d_gbs / HBM_PEAK_GBS, 4),
               "k_cg_upd_us": round(uu, 2), "k_cg_upd_GBs": round(u_gbs, 1),
               "k_cg_upd_frac": round(u_gbs / HBM_PEAK_GBS, 4),
               "fp32": {"dir_variant": tune_m["fp32"][0],
                        "k_cg_dir_GBs": round(4 * Nm * 4 / (dd32 * 1e-6) / 1e9, 1),
                        "k_cg_upd_GBs": round(6 * Nm * 4 / (uu32 * 1e-6) / 1e9, 1)},
               "device_copy_1GiB_GBs": round(copy_gbs, 1),
               "k_cg_dir_frac_of_copy": round(d_gbs / copy_gbs, 4),
               "k_cg_upd_frac_of_copy": round(u_gbs / copy_gbs, 4)}
        Pm.close()
        log(f"bench: 4096^2 fp64 dir {d_gbs:.0f} GB/s, upd {u_gbs:.0f} GB/s, copy {copy_gbs:.0f} GB/s")

    # ---- CPU baseline: whole NS_explicit_ steps of the oracle (the
    # reference's plain FP64 Jacobi-PCG) from the timed region's start state,
    # on every host thread of the GPU's share ----
    cpu = None
    if not a.no_cpu:
        import oracle as O
        threads = cpu_threads()
        og = O.StructGrid(n, n, 2, -0.5, 0.5, -0.5, 0.5)
        ora = O.NS(og, O.make_bc(og, [1, 1, 1, 1], [0.0, 0.0, 0.0, 1.0]),
                   O.make_bc(og, [1, 1, 1, 1]), rho=1000.0, mu=1.0, cfl=0.5, dnum=0.25,
                   dcoef=1e-3, limiter="VanLeer", tol=tol, norm=O.NORM_TRUE,
                   max_iter=10 ** 7)
        ns.restore(state)
        for name in ("u", "v", "uf", "vf", "p", "phi"):
            ora.f[name][:] = ns.get_field(name)
        ora.st.t, ora.st.dt = ns.time()
        O.set_threads(threads)
        try:
            t0 = time.perf_counter()
            code_c, it_c, res_c = ora.step()
            t_step = time.perf_counter() - t0
        finally:
            O.set_threads(1)
        # one thread: a bounded sample of the same solve's iterations
        k1 = 40
        t0 = time.perf_counter()
        O.solve("JPCG", ora.A, np.random.default_rng(0).standard_normal(n * n), max_iter=k1,
                tol=1e-300, norm=O.NORM_TRUE)
        t_it1 = (time.perf_counter() - t0) / k1
        cpu = {"value": round(1.0 / t_step, 5), "unit": "solves/s", "cores": threads,
               "kind": "port",
               "sample": f"one whole NS_explicit_ step of the {n}^2 cavity from the timed "
                         f"region's start state: TVD predictor, FP64 Jacobi-PCG projection "
                         f"to true rel-res {tol:g} ({it_c} iterations, the device's plain "
                         f"FP64 solve took {itp[0]}), correction; oracle C (-O2, OpenMP) "
                         f"on {threads} threads, {t_step:.2f} s",
               "ms_per_iteration": round(t_step / max(it_c, 1) * 1e3, 3),
               "single_thread_ms_per_iteration": round(t_it1 * 1e3, 3),
               "single_thread_solves_per_s_est": round(1.0 / (t_it1 * max(it_c, 1)), 5)}
        log(f"bench: cpu {cpu['value']} solves/s on {threads} threads ({it_c} iterations)")

    batch = None
    if not a.no_batch:
        batch = batched(ctx, ca, tol, a.no_cpu, not a.no_batch_learned)
    csr = None if a.no_csr else csr_probe(ctx, ca, a.no_cpu)
    bicg = None if a.no_csr else bicgstab_probe(ctx, ca)
    kim = None if a.no_kim else kim_probe(ctx, ca, a.no_cpu)
    nsx = None if a.no_ns else ns_explicit_probe(ctx, ca, a.no_cpu, steps=a.ns_steps)
    blob = None if a.no_blob else blob_probe(ctx, ca, a.no_cpu)
    out = {
        "metric": "Poisson solves/s to rel-res 1e-8 @1024^2: learned+FP32/FP64-refine vs plain FP64",
        "value": round(value, 4), "unit": "solves/s", "n_gpus": 1,
        "steps": K, "warmup": a.warmup,
        "ms_per_step": round(wl / K * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64 krylov / fp32 smoothing / bf16 u-net",
        "data": "synthetic (cavity from rest, random-init U-Net trained in-run)",
        "config": {"workload": f"{n}x{n} lid-driven cavity Re=1000 (CFL 0.5, van Leer), "
                               f"pressure-correction Poisson solves of consecutive steps "
                               f"warm-started from the previous phi, true rel-res {tol:g}",
                   "method": "FCG_UNET: flexible CG, preconditioner = 2 Jacobi sweeps + "
                             "U-Net + 2 Jacobi sweeps, FP64 J-PCG fallback",
                   "cells": N, "start_time": t_state,
                   "unet": "4 levels 16-32-64-128, 4x4 average pooling and 4x4 stride-4 "
                           "up-convolutions (1024^2 -> 16^2 bottleneck), 610913 parameters, "
                           "about 25 GFLOP per 1024^2 application",
                   "training_samples": tr.slot,
                   "training_set": "error samples (warm-start, zero-start and Jacobi-PCG/"
                                   "flexible-CG iterate errors with their residuals) from "
                                   "the plain solver's cavity steps",
                   "iterations_learned": itl, "iterations_plain_fp64": itp,
                   "iterations_mixed": itm, "autotune": tune,
                   "training": {"seconds": round(t_train, 1), "samples": tr.slot,
                                "log": tr.log}},
        "cells_per_s": round(value * N, 1),
        "gpu_fp64_jpcg": {"solves_per_s": round(K / wp, 4),
                          "speedup_of_value": round(wp / wl, 3)},
        "gpu_mixed_jpcg": {"solves_per_s": round(K / wm, 4),
                           "speedup_of_value": round(wm / wl, 3)},
        "field_rel_l2_vs_fp64": field_err,
        "field_rel_l2_vs_fp64_at_1e-10": field_tight,
        "roofline": roofline,
        "krylov_kernels_1024": krylov,
        "roofline_hbm_probe": hbm,
        "cpu_baseline": cpu,
        "batched_256": batch,
        "csr_jpcg": csr,
        "dia_bicgstab_2048": bicg,
        "ns_kim_512": kim,
        "ns_explicit_512": nsx,
        "blob_4096": blob,
    }
    print(json.dumps(out), flush=True)
    net.close()
    ns.close()
    ctx.close()


if __name__ == "__main__":
    main()
