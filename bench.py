#!/usr/bin/env python3
"""Benchmark: TeaLeaf's implicit heat-conduction solve on one MI355X.

Workload: the reference's layered-rectangle deck (2d/tea.in geometry) at
4096 x 4096 cells, FP64, eps 1e-15, solved three ways from identical
initial states -- CG; Chebyshev after CG presteps (presteps 30); PPCG with
10 inner steps (presteps 30).  A "step" is one implicit timestep solved by
all three; the three chunks stay resident in HBM throughout.

value = solver iterations per second over the timed steps, counting
iterations as the reference host prints them (CG + Chebyshev + PPCG lines).
The JSON line also carries the per-solver rates, the roofline of the
dominant kernel (of calc_pw, timed in every leg, and the temporally
blocked PPCG and Chebyshev loops, timed in theirs -- live, with HIP events
on the suite's stream -- the one with the most total time; the fused PPCG
loop's own roofline beside it), and the reference omp3 CPU path timed on
this host on a bounded sample of the same deck.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells 4096]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0          # MI355X published HBM3E bandwidth
# Algorithmic bytes per interior cell per call, FP64 (each distinct array
# read once + each array written once).
BYTES_PER_CELL = {
    "tlhip2d_ppcg_inner_iteration": 64,   # read sd, kx, ky, r, u; write r, u, sd'
    "tlhip2d_ppcg_inner_iterations": 64,  # the same arrays once for S fused inner steps
    "tlhip2d_cg_calc_pw": 48,             # read p, r, kx, ky; write p', w
    "tlhip2d_cg_calc_pw_cond": 40,        # read p, r, cond; write p', w
    "tlhip2d_cg_calc_pw_cond.no_w": 32,   # read p, r, cond; write p' (w kept in registers)
    "tlhip2d_cg_calc_ur_cond": 48,        # read u, p, r, cond; write u, r
    "tlhip2d_cg_calc_w": 32,              # read p, kx, ky; write w
    "tlhip2d_cg_calc_ur": 48,             # read u, p, r, w; write u, r
    "tlhip2d_cg_calc_p": 24,              # read p, r; write p
    "tlhip2d_cheby_iterate": 64,          # read u, u0, p, kx, ky; write u', p, r
    "tlhip2d_cheby_iterations": 64,       # the same, once per launch of up to 10 iterations
    "tlhip3d_ppcg_inner_iteration": 72,   # read sd, kx, ky, kz, r, u; write r, u, sd'
    "tlhip3d_ppcg_inner_iterations": 72,  # the same arrays once for two inner steps
}


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=4096)
    ap.add_argument("--solvers", default="cg,cheby,ppcg")
    ap.add_argument("--ppcg-fusion", type=int, default=10,
                    help="PPCG inner steps per fused launch (1 = one launch per step)")
    ap.add_argument("--cpu-cg-iters", type=int, default=1020,
                    help="CG phase of the bounded CPU reference samples (must reach the switch "
                         "to Chebyshev/PPCG: 1018 iterations at 4096^2, presteps 30)")
    ap.add_argument("--cpu-cheby-iters", type=int, default=230,
                    help="Chebyshev iterations of the CPU sample after the switch")
    ap.add_argument("--cpu-ppcg-outer", type=int, default=20,
                    help="PPCG outer iterations of the CPU sample after the switch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--refine-cells", type=int, default=8192,
                    help="grid of the FP32-refined vs FP64 time-to-1e-10 comparison (0 = skip)")
    ap.add_argument("--refine-steps", type=int, default=5)
    ap.add_argument("--refine-cheby-steps", type=int, default=2,
                    help="steps of the FP32-refined vs FP64 Chebyshev comparison on the "
                         "--cells deck (0 = skip)")
    ap.add_argument("--refine-only", action="store_true",
                    help="run only the FP32-refined vs FP64 comparison and print its JSON")
    ap.add_argument("--blocks3d", action="store_true",
                    help="run only the 3D 256^3 seeded block-density PPCG workload and print its JSON")
    ap.add_argument("--blocks3d-steps", type=int, default=2)
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint (forward+backward) timing")
    ap.add_argument("--write-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="build the HBM-traffic summary under profiles/ from the two rocprofv3 "
                         "--pmc passes over --profile-run, and exit (no GPU)")
    ap.add_argument("--traffic-commit", default="",
                    help="commit the counter passes ran at, stamped into the summary")
    ap.add_argument("--profile-run", action="store_true",
                    help="counter-collection run: a few launches of each hot kernel at --cells "
                         "and of a copy kernel of known size, then exit")
    return ap.parse_args()


def gpu_run(args):
    import torch
    from tealeaf_amd import deck as tdeck
    from tealeaf_amd.abi import lib
    from tealeaf_amd.problem import Problem

    L = lib()
    L.tlhip_init(0)
    L.tlhip_set_ppcg_fusion(args.ppcg_fusion)          # raises on a bad value
    names = {"cg": tdeck.CG, "cheby": tdeck.CHEBY, "ppcg": tdeck.PPCG}
    solvers = [s for s in args.solvers.split(",") if s]
    total_steps = args.warmup + args.steps
    probs = {}
    for s in solvers:
        text = tdeck.layered_2d(args.cells, total_steps, solver=names[s], eps=1e-15,
                                max_iters=10000, inner=10, presteps=30)
        probs[s] = Problem.from_deck(text, 2)
    n_int = args.cells * args.cells

    for tt in range(args.warmup):
        for s in solvers:
            probs[s].step(tt)
            print(f"[bench] warmup step {tt} {s} done", file=sys.stderr, flush=True)

    by = {}
    total_time = 0.0
    total_iters = 0
    kernel = {}
    # Each timed launch adds two event records to the stream (about 4% of a
    # CG iteration), so: the fused PPCG launch (the kernel with the most time)
    # and PPCG's calc_pw in the PPCG leg, the fused Chebyshev launch and the
    # CG pair in the Chebyshev leg (its 1559 CG iterations; the same kernels
    # at the same size run every CG iteration of all three legs), nothing
    # in the CG leg; the small rest (2-norms, finishers, halo and ring
    # copies) untimed.
    own = {"cg": [],
           "cheby": ["tlhip2d_cheby_iterations", "tlhip2d_cg_calc_pw_cond.no_w",
                     "tlhip2d_cg_calc_ur_cond"],
           "ppcg": [dominant(args), "tlhip2d_cg_calc_pw_cond"]}
    for s in solvers:
        timed = own[s]
        L.tlhip_synchronize()
        torch.cuda.synchronize()
        if timed:
            L.tlhip_kernel_timer_start(",".join(timed).encode())
        t0 = time.perf_counter()
        reps = [probs[s].step(tt) for tt in range(args.warmup, total_steps)]
        L.tlhip_synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        for name in timed:
            ms, nl = ctypes.c_double(), ctypes.c_long()
            L.tlhip_kernel_timer_read_entry(name.encode(), ctypes.addressof(ms),
                                            ctypes.addressof(nl))
            k = kernel.setdefault(name, {"launches": 0, "total_ms": 0.0, "by_solver": {}})
            k["launches"] += nl.value
            k["total_ms"] += ms.value
            if nl.value:
                k["by_solver"][s] = {"launches": nl.value, "ms": ms.value}
        if timed:
            ms, nl = ctypes.c_double(), ctypes.c_long()
            L.tlhip_kernel_timer_read(ctypes.addressof(ms), ctypes.addressof(nl))   # stop
        printed = sum(r["printed_cg_iters"] + r["printed_inner_iters"] for r in reps)
        sweeps = {"cg_steps": sum(r["n_cg_steps"] for r in reps),
                  "cheby_steps": sum(r["n_cheby_steps"] for r in reps),
                  "ppcg_outer": sum(r["n_ppcg_outer"] for r in reps),
                  "ppcg_inner": sum(r["n_ppcg_inner"] for r in reps)}
        by[s] = {"iterations": printed, "seconds": dt, "iters_per_s": printed / dt,
                 "per_step": [{"CG": r["printed_cg_iters"], "inner": r["printed_inner_iters"]}
                              for r in reps], "sweeps": sweeps,
                 "temp": probs[s].field_summary()[3]}
        total_time += dt
        total_iters += printed
        print(f"[bench] {s}: {printed} iterations in {dt:.3f} s", file=sys.stderr, flush=True)
    for p in probs.values():
        p.close()
    for name, k in kernel.items():
        k["avg_ms"] = k["total_ms"] / k["launches"] if k["launches"] else None
        k["gbs"] = (BYTES_PER_CELL[name] * n_int / (k["avg_ms"] * 1e-3) / 1e9
                    if k["launches"] else None)
        k["share_of_timed_time"] = k["total_ms"] * 1e-3 / total_time
    return {"by_solver": by, "kernels": kernel, "time": total_time, "iters": total_iters}


def refine_run(args):
    """FP32-refined vs FP64 PPCG, each timestep solved to ||u0 - A u|| <=
    1e-10 ||u0||: the seeded block-density workload (64x64-cell blocks,
    log-uniform density over a 1e3 or 1e6 contrast, energy U(0.1, 25)),
    dt 0.004 on [0,10]^2, PPCG with 10 inner steps entered through the
    reference's error switch (20 CG iterations for the eigenvalue
    estimate)."""
    import numpy as np
    import torch
    from tealeaf_amd import blocks
    from tealeaf_amd import deck as tdeck
    from tealeaf_amd.abi import lib
    from tealeaf_amd.problem import Problem

    L = lib()
    n = args.refine_cells
    text = tdeck.layered_2d(n, args.refine_steps, solver=tdeck.PPCG, eps=1e-15, max_iters=20000,
                            inner=10, presteps=30, epslim=1e30)
    s, states = tdeck.parse_deck(text, 2)
    out = {}
    for contrast, seed in ((1e3, 1001), (1e6, 1002)):
        dens, en = blocks.block_fields((n, n), 64, seed=seed, rho=(1.0, contrast),
                                       energy=(0.1, 25.0), halo_depth=s.halo_depth)
        res = {}
        u_final = {}
        for prec in (0, 1):
            prob = Problem(s, states)
            prob.inject(dens, en)
            prob.set_solve(prec, 1e-10)
            times, reps = [], []
            # the last step also times its main kernels (HIP events)
            names = (["tlhip2d_mix_calc_pw32", "tlhip2d_ppcg_inner_iterations_f32",
                      "tlhip2d_mix_norm32"] if prec else
                     ["tlhip2d_cg_calc_pw_cond", "tlhip2d_ppcg_inner_iterations",
                      "tlhip2d_calculate_2norm"])
            kern = {}
            for tt in range(args.refine_steps):
                last = tt == args.refine_steps - 1
                L.tlhip_synchronize()
                torch.cuda.synchronize()
                if last:
                    L.tlhip_kernel_timer_start(",".join(names).encode())
                t0 = time.perf_counter()
                reps.append(prob.step(tt))
                L.tlhip_synchronize()
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t0)
                if last:
                    for nm in names:
                        ms, nl = ctypes.c_double(), ctypes.c_long()
                        L.tlhip_kernel_timer_read_entry(nm.encode(), ctypes.addressof(ms),
                                                        ctypes.addressof(nl))
                        kern[nm] = {"launches": nl.value, "total_ms": ms.value,
                                    "avg_ms": ms.value / nl.value if nl.value else None}
                    L.tlhip_kernel_timer_read(ctypes.addressof(ms), ctypes.addressof(nl))
                print(f"[bench] refine contrast {contrast:g} {'fp32' if prec else 'fp64'} step {tt}: "
                      f"{times[-1]:.3f} s, {reps[-1]['n_cg_steps']} CG + {reps[-1]['n_ppcg_outer']} PPCG, "
                      f"rel residual {reps[-1]['rel_residual']:.2e}", file=sys.stderr, flush=True)
            u_final[prec] = prob.get("u")
            prob.close()
            res["fp32_refined" if prec else "fp64"] = {
                "seconds_per_step": times, "mean_s": sum(times) / len(times),
                "cg_iters": [r["n_cg_steps"] for r in reps],
                "ppcg_outer": [r["n_ppcg_outer"] for r in reps],
                "true_residual_evals": [r["n_true_residuals"] for r in reps],
                "rel_residual": [r["rel_residual"] for r in reps],
                "last_step_kernels": kern}
        hd = s.halo_depth
        a = u_final[1].reshape(s.shape())[hd:-hd, hd:-hd]
        b = u_final[0].reshape(s.shape())[hd:-hd, hd:-hd]
        res["rel_error_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        res["time_ratio"] = res["fp32_refined"]["mean_s"] / res["fp64"]["mean_s"]
        out[f"contrast_{contrast:g}"] = res
    out["cheby"] = refine_cheby(args) if args.refine_cheby_steps > 0 else None
    return {"workload": f"2D {n}x{n}, 64x64-cell blocks, log-uniform density [1, C], energy "
                        f"U(0.1,25), dt 0.004, PPCG 10 inner via the error switch, "
                        f"{args.refine_steps} steps each solved to relative residual 1e-10",
            "target_ratio": 0.6, **out}


def refine_cheby(args):
    """FP32-refined vs FP64 Chebyshev on the benchmark's own deck (2D
    4096^2 layered, presteps 30), each timestep solved to the deck's eps
    (1e-15 on the squared residual norm; the FP32-refined solve tests it on
    FP64 true residuals)."""
    import numpy as np
    import torch
    from tealeaf_amd import deck as tdeck
    from tealeaf_amd.abi import lib
    from tealeaf_amd.problem import Problem

    L = lib()
    n, steps = args.cells, args.refine_cheby_steps
    text = tdeck.layered_2d(n, steps, solver=tdeck.CHEBY, eps=1e-15, max_iters=10000,
                            presteps=30)
    s, states = tdeck.parse_deck(text, 2)
    res, u_final = {}, {}
    for prec in (0, 1):
        prob = Problem(s, states)
        prob.set_solve(prec, 0.0)
        times, reps = [], []
        for tt in range(steps):
            L.tlhip_synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps.append(prob.step(tt))
            L.tlhip_synchronize()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            print(f"[bench] refine cheby {'fp32' if prec else 'fp64'} step {tt}: {times[-1]:.3f} s, "
                  f"{reps[-1]['printed_cg_iters']} CG + {reps[-1]['printed_inner_iters']} Chebyshev",
                  file=sys.stderr, flush=True)
        u_final[prec] = prob.get("u")
        prob.close()
        res["fp32_refined" if prec else "fp64"] = {
            "seconds_per_step": times, "mean_s": sum(times) / len(times),
            "printed_cg": [r["printed_cg_iters"] for r in reps],
            "printed_cheby": [r["printed_inner_iters"] for r in reps],
            "true_residual_evals": [r["n_true_residuals"] for r in reps]}
    hd = s.halo_depth
    a = u_final[1].reshape(s.shape())[hd:-hd, hd:-hd]
    b = u_final[0].reshape(s.shape())[hd:-hd, hd:-hd]
    res["rel_error_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    res["time_ratio"] = res["fp32_refined"]["mean_s"] / res["fp64"]["mean_s"]
    res["workload"] = (f"2D {n}x{n} layered deck, Chebyshev (presteps 30), eps 1e-15, "
                       f"{steps} steps")
    return res


def profile_run(args):
    """Counter-collection run (rocprofv3 --pmc): three launches each of every
    hot kernel on synthetic fields -- 2D at --cells^2 (the fused PPCG and
    Chebyshev loops, calc_w, calc_pw, calc_ur, calc_p, the 2-norm,
    field_summary, cg_init and the single-step PPCG sweep; the CG pair from
    the conductivity and PPCG's calc_pw from it; the FP32 fused
    PPCG and Chebyshev loops), 3D at 256^3 (calc_w, calc_ur, calc_p, the CG pair from the
    conductivity, the PPCG inner sweep, the two-step PPCG sweep, the 2-norm, field_summary,
    cg_init) -- then one timestep of the FP32-refined
    2D PPCG solve at 2048^2 for the FP32 CG kernels, and a 1 GiB device copy
    (8 B read + 8 B written per element) whose FETCH_SIZE / WRITE_SIZE
    calibrate the counters."""
    import ctypes as C
    import numpy as np
    from tealeaf_amd import blocks
    from tealeaf_amd import deck as tdeck
    from tealeaf_amd.abi import lib
    from tealeaf_amd.device import DeviceArray
    from tealeaf_amd.problem import Problem

    L = lib()
    L.tlhip_init(0)
    n, hd = args.cells, 2
    x = y = n + 2 * hd
    N, P = x * y, (x + 3) // 4 * 4
    rng = np.random.default_rng(0)
    f64 = {k: DeviceArray.from_host(rng.uniform(-1, 1, N))
           for k in ("u", "u2", "r", "r2", "sd", "sd2", "p", "w")}
    for k in ("kx", "ky"):
        f64[k] = DeviceArray.from_host(rng.uniform(0.05, 0.2, N))
    f64["rho"] = DeviceArray.from_host(rng.uniform(0.5, 2.0, N))
    f64["e"] = DeviceArray.from_host(rng.uniform(0.1, 1.0, N))
    f64["cond"] = DeviceArray(N, np.float64)
    L.tlhip2d_cg_cond(x, y, hd, 2, f64["rho"].ptr, f64["cond"].ptr)
    f64["vol"] = DeviceArray.from_host(np.full(N, 1e-4))
    f32 = {k: DeviceArray.from_host(rng.uniform(-1, 1, P * y).astype(np.float32))
           for k in ("u", "u2", "r", "r2", "sd", "sd2", "p", "w")}
    for k in ("kx", "ky"):
        f32[k] = DeviceArray.from_host(rng.uniform(0.05, 0.2, P * y).astype(np.float32))
    ncopy = (1 << 30) // 8
    src, dst = DeviceArray(ncopy, np.float64), DeviceArray(ncopy, np.float64)
    nb = np.full(4, -1, dtype=np.int32)
    al, be = np.full(10, 0.7), np.full(10, 0.2)
    pw = C.c_double()
    s4 = [C.c_double() for _ in range(4)]
    a4 = [C.addressof(v) for v in s4]
    for _ in range(3):
        L.tlhip_stream_copy(dst.ptr, src.ptr, C.c_size_t(ncopy))
        # the driver's launch: cg_calc_ur + ppcg_init + 10 inner steps (MODE 2)
        L.tlhip2d_ppcg_inner_iterations(x, y, hd, 10, al.ctypes.data, be.ctypes.data, 1.3, 1,
                                        nb.ctypes.data, None, f64["sd2"].ptr, f64["r"].ptr,
                                        f64["r2"].ptr, f64["u"].ptr, f64["u2"].ptr, f64["kx"].ptr,
                                        f64["ky"].ptr, 0.3, f64["p"].ptr, f64["w"].ptr)
        # 10 inner steps continuing from sd (MODE 0)
        L.tlhip2d_ppcg_inner_iterations(x, y, hd, 10, al.ctypes.data, be.ctypes.data, 1.3, 0,
                                        nb.ctypes.data, f64["sd"].ptr, f64["sd2"].ptr, f64["r"].ptr,
                                        f64["r2"].ptr, f64["u"].ptr, f64["u2"].ptr, f64["kx"].ptr,
                                        f64["ky"].ptr, 0.0, None, None)
        L.tlhip2d_cg_calc_pw(x, y, hd, 0.5, C.addressof(pw), f64["p"].ptr, f64["sd"].ptr,
                             f64["r"].ptr, f64["w"].ptr, f64["kx"].ptr, f64["ky"].ptr)
        # the CG pair from the conductivity, then PPCG's calc_pw from it
        L.tlhip2d_cg_calc_pw_cond(x, y, hd, 1, 0.5, 0.01, 0.01, C.addressof(pw), f64["p"].ptr,
                                  f64["sd"].ptr, f64["r"].ptr, None, f64["cond"].ptr)
        L.tlhip2d_cg_calc_ur_cond(x, y, hd, 1e-3, 0.01, 0.01, C.addressof(pw), f64["u"].ptr,
                                  f64["sd"].ptr, f64["r"].ptr, f64["cond"].ptr)
        L.tlhip2d_cg_calc_pw_cond(x, y, hd, 1, 0.5, 0.01, 0.01, C.addressof(pw), f64["p"].ptr,
                                  f64["sd"].ptr, f64["r"].ptr, f64["w"].ptr, f64["cond"].ptr)
        # 10 Chebyshev iterations (u -> u2, p -> sd2; u0 = sd)
        L.tlhip2d_cheby_iterations(x, y, hd, 10, al.ctypes.data, be.ctypes.data, nb.ctypes.data,
                                   f64["u"].ptr, f64["u2"].ptr, f64["sd"].ptr, f64["p"].ptr,
                                   f64["sd2"].ptr, f64["r2"].ptr, f64["kx"].ptr, f64["ky"].ptr)
        L.tlhip2d_ppcg_inner_iterations_f32(x, y, hd, P, 10, al.ctypes.data, be.ctypes.data, 1.3, 1,
                                            nb.ctypes.data, None, f32["sd2"].ptr,
                                            f32["r"].ptr, f32["r2"].ptr, f32["u"].ptr,
                                            f32["u2"].ptr, f32["kx"].ptr, f32["ky"].ptr, 0.3,
                                            f32["p"].ptr, f32["w"].ptr)
        # 10 FP32 Chebyshev iterations (u -> u2, p -> sd2; u0 = sd)
        L.tlhip2d_cheby_iterations_f32(x, y, hd, P, 10, al.ctypes.data, be.ctypes.data,
                                       nb.ctypes.data, f32["u"].ptr, f32["u2"].ptr, f32["sd"].ptr,
                                       f32["p"].ptr, f32["sd2"].ptr, f32["r2"].ptr, f32["kx"].ptr,
                                       f32["ky"].ptr, None)
        L.tlhip2d_ppcg_inner_iteration(x, y, hd, 0.7, 0.2, f64["u"].ptr, f64["r"].ptr, f64["kx"].ptr,
                                       f64["ky"].ptr, f64["sd"].ptr, f64["sd2"].ptr)
        L.tlhip2d_cg_calc_w(x, y, hd, C.addressof(pw), f64["p"].ptr, f64["w"].ptr, f64["kx"].ptr,
                            f64["ky"].ptr)
        L.tlhip2d_cg_calc_ur(x, y, hd, 1e-3, C.addressof(pw), f64["u"].ptr, f64["p"].ptr,
                             f64["r"].ptr, f64["w"].ptr)
        L.tlhip2d_cg_calc_p(x, y, hd, 0.5, f64["p"].ptr, f64["r"].ptr)
        L.tlhip2d_calculate_2norm(x, y, hd, f64["r"].ptr, C.addressof(pw))
        L.tlhip2d_field_summary(x, y, hd, f64["vol"].ptr, f64["rho"].ptr, f64["e"].ptr,
                                f64["u"].ptr, *a4)
        L.tlhip2d_cg_init(x, y, hd, 1, 0.01, 0.01, C.addressof(pw), f64["rho"].ptr, f64["e"].ptr,
                          f64["u2"].ptr, f64["p"].ptr, f64["r2"].ptr, f64["w"].ptr,
                          f64["kx"].ptr, f64["ky"].ptr)
    L.tlhip_synchronize()
    for v in list(f64.values()) + list(f32.values()):
        v.free()
    # 3D at 256^3
    m = 256
    X3 = m + 2 * hd
    N3 = X3 ** 3
    g = {k: DeviceArray.from_host(rng.uniform(-1, 1, N3))
         for k in ("u", "u2", "r", "r2", "sd", "sd2", "p", "w")}
    for k in ("kx", "ky", "kz"):
        g[k] = DeviceArray.from_host(rng.uniform(0.02, 0.1, N3))
    g["rho"] = DeviceArray.from_host(rng.uniform(0.5, 2.0, N3))
    g["e"] = DeviceArray.from_host(rng.uniform(0.1, 1.0, N3))
    g["vol"] = DeviceArray.from_host(np.full(N3, 1e-6))
    g["cond"] = DeviceArray(N3, np.float64)
    L.tlhip3d_cg_cond(X3, X3, X3, hd, 2, g["rho"].ptr, g["cond"].ptr)
    nb6 = np.full(6, -1, dtype=np.int32)
    for _ in range(3):
        L.tlhip3d_cg_solver_calc_w(X3, X3, X3, hd, C.addressof(pw), g["p"].ptr, g["w"].ptr,
                                   g["kx"].ptr, g["ky"].ptr, g["kz"].ptr)
        L.tlhip3d_cg_solver_calc_ur(X3, X3, X3, hd, 1e-3, C.addressof(pw), g["u"].ptr, g["p"].ptr,
                                    g["r"].ptr, g["w"].ptr)
        L.tlhip3d_cg_solver_calc_p(X3, X3, X3, hd, 0.5, g["p"].ptr, g["r"].ptr)
        L.tlhip3d_cg_calc_pw_cond(X3, X3, X3, hd, 1, 0.5, 0.01, 0.01, 0.01, C.addressof(pw),
                                  g["p"].ptr, g["sd2"].ptr, g["r"].ptr, None, g["cond"].ptr)
        L.tlhip3d_cg_calc_ur_cond(X3, X3, X3, hd, 1e-3, 0.01, 0.01, 0.01, C.addressof(pw),
                                  g["u"].ptr, g["sd2"].ptr, g["r"].ptr, g["cond"].ptr)
        L.tlhip3d_ppcg_inner_iteration(X3, X3, X3, hd, 0.7, 0.2, g["u"].ptr, g["r"].ptr,
                                       g["kx"].ptr, g["ky"].ptr, g["kz"].ptr, g["sd"].ptr,
                                       g["sd2"].ptr)
        L.tlhip3d_ppcg_inner_iterations(X3, X3, X3, hd, 2, al.ctypes.data, be.ctypes.data,
                                        nb6.ctypes.data, g["sd"].ptr, g["sd2"].ptr, g["r"].ptr,
                                        g["r2"].ptr, g["u"].ptr, g["u2"].ptr, g["kx"].ptr,
                                        g["ky"].ptr, g["kz"].ptr)
        L.tlhip3d_calculate_2norm(X3, X3, X3, hd, g["r"].ptr, C.addressof(pw))
        L.tlhip3d_field_summary(X3, X3, X3, hd, g["vol"].ptr, g["rho"].ptr, g["e"].ptr,
                                g["u"].ptr, *a4)
        L.tlhip3d_cg_solver_init(X3, X3, X3, hd, 1, 0.01, 0.01, 0.01, C.addressof(pw),
                                 g["rho"].ptr, g["e"].ptr, g["u"].ptr, g["p"].ptr, g["r"].ptr,
                                 g["w"].ptr, g["kx"].ptr, g["ky"].ptr, g["kz"].ptr)
    L.tlhip_synchronize()
    for v in g.values():
        v.free()
    # one FP32-refined PPCG timestep at 2048^2: the FP32 CG kernels
    text = tdeck.layered_2d(2048, 1, solver=tdeck.PPCG, eps=1e-15, max_iters=20000, inner=10,
                            presteps=30, epslim=1e30)
    st, states = tdeck.parse_deck(text, 2)
    dens, en = blocks.block_fields((2048, 2048), 64, seed=1001, rho=(1.0, 1e3),
                                   energy=(0.1, 25.0), halo_depth=st.halo_depth)
    with Problem(st, states) as prob:
        prob.inject(dens, en)
        prob.set_solve(1, 1e-10)
        prob.step(0)
    L.tlhip_synchronize()
    print(json.dumps({"profile_run": True, "cells": n, "cells3d": m,
                      "copy_elements": ncopy}), flush=True)


def dominant(args) -> str:
    return "tlhip2d_ppcg_inner_iterations" if args.ppcg_fusion > 1 else "tlhip2d_ppcg_inner_iteration"


def cpu_info() -> dict:
    """CPU model, the physical cores in this process's affinity mask, and
    the threads the CPU reference may use: those cores, capped by the CPU
    share the job is given (OMP_NUM_THREADS when set, as on a one-GPU
    machine shared with other jobs)."""
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        topo = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
        try:
            cores.add((topo.joinpath("physical_package_id").read_text().strip(),
                       topo.joinpath("core_id").read_text().strip()))
        except OSError:
            cores.add(("?", str(c)))
    physical = len(cores) or len(cpus)
    share = os.environ.get("OMP_NUM_THREADS")
    threads = min(physical, int(share)) if share and share.isdigit() and int(share) > 0 else physical
    return {"cpu_model": model, "logical_cpus_in_mask": len(cpus),
            "physical_cores_in_mask": physical, "omp_num_threads_env": share, "threads": threads}


_PROF = re.compile(r"^(run_\w+|eigenvalue_driver_initialise)\s+(\d+)\s+([0-9.]+)\s*$", re.M)


def _ref_profile(exe: Path, text: str, threads: int) -> dict:
    """Run the reference host + omp3 binary on a deck; return its printed
    iteration lines, Wallclock and -DENABLE_PROFILING kernel table."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as d:
        Path(d, "tea.in").write_text(text)
        Path(d, "tea.problems").write_text((ROOT / "tests" / "golden" / "tea_problems_2d.txt").read_text())
        p = subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, env=env, timeout=900)
    if p.returncode != 0:
        raise RuntimeError(f"{exe.name} exited {p.returncode}: {p.stdout[-1000:]}")
    iters = {k: int(v) for k, v in re.findall(r"^(CG|Cheby|PPCG):\s+(\d+) iterations", p.stdout, re.M)}
    prof = {name: (int(calls), float(sec)) for name, calls, sec in _PROF.findall(p.stdout)}
    m_w = re.search(r"Wallclock:\s+([0-9.]+)s", p.stdout)
    return {"iters": iters, "profile": prof, "wallclock": float(m_w.group(1)) if m_w else None}


def cpu_baseline(cells: int, cg_iters: int, cheby_iters: int, ppcg_outer: int):
    """The reference omp3 CPU path (reference host + omp3 suite, built from
    source into oracle/_ref) on bounded samples of the benchmark deck, per
    solver, timed by the reference's own -DENABLE_PROFILING kernel table:

      * a Chebyshev-deck timestep (presteps 30) cut at max_iters =
        cg_iters + cheby_iters: its CG phase (cg_calc_w + cg_calc_ur +
        cg_calc_p + their halo updates) gives CG iterations/s, its
        Chebyshev phase (cheby_iterate + 2-norms + halo updates) Chebyshev
        iterations/s;
      * a PPCG-deck timestep (presteps 30, 10 inner) cut at
        cg_iters + ppcg_outer: its PPCG phase (outer steps' calc_w, calc_ur,
        calc_p, 2-norm, ppcg_init, the inner sweeps and halo updates) gives
        PPCG outer iterations/s, counted as the reference prints them.

    cg_iters must reach the switch to Chebyshev/PPCG (1018 CG iterations at
    4096^2 with presteps 30); the halo updates are shared out by call count."""
    exe = ROOT / "oracle" / "_ref" / "tealeaf2d_omp3"
    if not exe.exists():
        return None
    from tealeaf_amd import deck as tdeck
    info = cpu_info()
    th = info["threads"]
    t0 = time.perf_counter()
    ch = _ref_profile(exe, tdeck.layered_2d(cells, 1, solver=tdeck.CHEBY, eps=1e-15,
                                            max_iters=cg_iters + cheby_iters, presteps=30), th)
    pp = _ref_profile(exe, tdeck.layered_2d(cells, 1, solver=tdeck.PPCG, eps=1e-15,
                                            max_iters=cg_iters + ppcg_outer, presteps=30,
                                            inner=10), th)
    wall = time.perf_counter() - t0

    def t(prof, *names):
        return sum(prof.get(n, (0, 0.0))[1] for n in names)

    def calls(prof, n):
        return prof.get(n, (0, 0.0))[0]

    out = {}
    P = ch["profile"]
    n_cg = calls(P, "run_cg_calc_w")            # CG steps (the printed line counts them too)
    n_ch = calls(P, "run_cheby_iterate")
    halo = t(P, "run_local_halos") / max(calls(P, "run_local_halos"), 1)
    if n_cg:
        t_cg = t(P, "run_cg_calc_w", "run_cg_calc_ur", "run_cg_calc_p") + n_cg * halo
        out["cg"] = {"iterations": n_cg, "seconds": t_cg, "iters_per_s": n_cg / t_cg}
    if n_ch:
        t_ch = t(P, "run_cheby_iterate", "run_calculate_2norm") + n_ch * halo
        out["cheby"] = {"iterations": n_ch, "seconds": t_ch, "iters_per_s": n_ch / t_ch}
    Q = pp["profile"]
    n_out = pp["iters"].get("PPCG", 0)
    if n_out:
        n_cgq = pp["iters"].get("CG", 0)
        # the outer steps are cg_calc_w/ur/p calls beyond the CG phase
        frac = n_out / max(calls(Q, "run_cg_calc_w"), 1)
        halo_q = t(Q, "run_local_halos") / max(calls(Q, "run_local_halos"), 1)
        n_halo_ppcg = calls(Q, "run_local_halos") - n_cgq - 1
        t_pp = (frac * t(Q, "run_cg_calc_w", "run_cg_calc_ur", "run_cg_calc_p")
                + t(Q, "run_ppcg_inner_iteration", "run_ppcg_init", "run_calculate_2norm",
                    "run_calculate_residual") + max(n_halo_ppcg, 0) * halo_q)
        out["ppcg"] = {"iterations": n_out, "inner_sweeps": calls(Q, "run_ppcg_inner_iteration"),
                       "seconds": t_pp, "iters_per_s": n_out / t_pp}
    if not out:
        return None
    return {"unit": "iters/s", "cores": th, "kind": "reference", **info,
            "by_solver": out, "sample_wall_s": wall,
            "sample": f"reference host + omp3 suite built from source (oracle/_ref/tealeaf2d_omp3), "
                      f"{cells}^2 layered deck, one timestep as Chebyshev (presteps 30) cut at "
                      f"{cg_iters + cheby_iters} iterations and one as PPCG (presteps 30, 10 inner) "
                      f"cut at {cg_iters + ppcg_outer}; per-solver time from the reference's own "
                      f"ENABLE_PROFILING kernel table, OMP_NUM_THREADS={th}, OMP_PROC_BIND=close, "
                      f"OMP_PLACES=cores"}


def load_traffic(kernel: str, init_share: float = 0.0):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size),
    and where they come from (file, commit the counters were taken at).
    For the fused PPCG loop, launches that start from r (ppcg_init folded
    in) and launches that continue from sd are weighted by their share."""
    path = ROOT / TRAFFIC_JSON
    if not path.exists():
        return None, None
    try:
        doc = json.loads(path.read_text())
        kernels = doc["kernels"]
        src = {"file": str(path.relative_to(ROOT)), "commit": doc.get("commit"),
               "date": doc.get("date")}
        if kernel == "tlhip2d_ppcg_inner_iterations":
            a = kernels["tlhip2d_ppcg_inner_iterations (from r)"]["bytes_per_launch"]
            b = kernels["tlhip2d_ppcg_inner_iterations"]["bytes_per_launch"]
            return init_share * a + (1.0 - init_share) * b, src
        return kernels[kernel]["bytes_per_launch"], src
    except Exception:
        return None, None


TRAFFIC_JSON = "profiles/hbm_traffic_2d4096_3d256.json"

# The main kernel(s) of each entry point --profile-run launches, in launch
# order: (entry point, HIP kernel name fragments, compulsory bytes per
# interior cell = each array the call reads once + each it writes once).
_TRAFFIC_2D = [
    ("tlhip2d_ppcg_inner_iterations (from r)", ["k_ppcg_ring2d<double, 10, 2,"], 72,
     "cg_calc_ur + ppcg_init + 10 inner steps: reads r, w, u, p, kx, ky; writes sd, r, u"),
    ("tlhip2d_ppcg_inner_iterations", ["k_ppcg_ring2d<double, 10, 0,"], 64,
     "10 inner steps from sd: reads sd, kx, ky, r, u; writes sd, r, u"),
    ("tlhip2d_cg_calc_pw", ["k_calc_pw<double, true>"], 48,
     "calc_p fused into calc_w: reads p, r, kx, ky; writes p', w"),
    ("tlhip2d_cg_calc_pw_cond.no_w", ["k_calc_pw_cond<true, false>"], 32,
     "calc_p fused into calc_w from the conductivity, w kept in registers: reads p, r, cond; "
     "writes p'"),
    ("tlhip2d_cg_calc_ur_cond", ["k_calc_ur_cond"], 48,
     "cg_calc_ur forming w = A p from p and cond: reads u, p, r, cond; writes u, r"),
    ("tlhip2d_cg_calc_pw_cond", ["k_calc_pw_cond<true, true>"], 40,
     "calc_p fused into calc_w from the conductivity (PPCG): reads p, r, cond; writes p', w"),
    ("tlhip2d_cheby_iterations", ["k_ppcg_ring2d<double, 10, 3,"], 64,
     "10 Chebyshev iterations: reads u, u0, p, kx, ky; writes u', p', r"),
    ("tlhip2d_ppcg_inner_iterations_f32 (from r)", ["k_ppcg_ring2d<float, 10, 2, false>"], 36,
     "FP32 fields, cg_calc_ur + ppcg_init + 10 inner steps"),
    ("tlhip2d_cheby_iterations_f32", ["k_ppcg_ring2d<float, 10, 3,"], 32,
     "FP32 fields, 10 Chebyshev iterations: reads u, u0, p, kx, ky; writes u', p', r"),
    ("tlhip2d_ppcg_inner_iteration", ["k_stencil2d<2, tl::OpPpcgInner>"], 64,
     "one unfused inner step: reads sd, kx, ky, r, u; writes sd', r, u"),
    ("tlhip2d_cg_calc_w", ["k_stencil2d<2, tl::OpCalcW>"], 32, "reads p, kx, ky; writes w"),
    ("tlhip2d_cg_calc_ur", ["k_interior<2, tl::OpCalcUr>"], 48, "reads u, p, r, w; writes u, r"),
    ("tlhip2d_cg_calc_p", ["k_interior<2, tl::OpCalcP>"], 24, "reads p, r; writes p"),
    ("tlhip2d_calculate_2norm", ["k_interior<2, tl::OpNorm>"], 8, "reads r"),
    ("tlhip2d_field_summary", ["k_interior<2, tl::OpFieldSummary>"], 32,
     "reads volume, density, energy, u"),
    ("tlhip2d_cg_init", ["k_cg_init_fields(", "k_stencil2d<2, tl::OpCgInit>"], 64,
     "reads density, energy; writes u, p, r, w, kx, ky (two passes: the coefficients, then "
     "the stencil)"),
]
_TRAFFIC_3D = [
    ("tlhip3d_cg_solver_calc_w", ["k_stencil3d<2, tl::OpCalcW>"], 40,
     "reads p, kx, ky, kz; writes w"),
    ("tlhip3d_cg_solver_calc_ur", ["k_interior<2, tl::OpCalcUr>"], 48,
     "reads u, p, r, w; writes u, r"),
    ("tlhip3d_cg_solver_calc_p", ["k_interior<2, tl::OpCalcP>"], 24, "reads p, r; writes p"),
    ("tlhip3d_cg_calc_pw_cond.no_w", ["k_cg3_cond<false, true>"], 32,
     "calc_p fused into calc_w from the conductivity, w kept in registers: reads p, r, cond; "
     "writes p'"),
    ("tlhip3d_cg_calc_ur_cond", ["k_cg3_cond<true, false>"], 48,
     "cg_calc_ur forming w = A p from p and cond: reads u, p, r, cond; writes u, r"),
    ("tlhip3d_ppcg_inner_iteration", ["k_stencil3d<2, tl::OpPpcgInner>"], 72,
     "reads sd, kx, ky, kz, r, u; writes sd', r, u"),
    ("tlhip3d_ppcg_inner_iterations", ["k_copy_ring3d", "k_copy_ring3d", "k_copy_ring3d",
                                       "k_ppcg3d_pair_lds<8>"], 72,
     "two inner steps (halo rings of sd, r, u copied, then one sweep): reads sd, kx, ky, kz, r, "
     "u; writes sd'', r, u"),
    ("tlhip3d_calculate_2norm", ["k_interior<2, tl::OpNorm>"], 8, "reads r"),
    ("tlhip3d_field_summary", ["k_interior<2, tl::OpFieldSummary>"], 32,
     "reads volume, density, energy, u"),
    ("tlhip3d_cg_solver_init", ["k_cg_init_fields3(", "k_stencil3d<2, tl::OpCgInit>"], 72,
     "reads density, energy; writes u, p, r, w, kx, ky, kz (two passes)"),
]
# FP32 kernels of the FP32-refined PPCG timestep at 2048^2 (many launches
# each): name fragment -> (call, bytes per cell)
_TRAFFIC_MIXED = [
    ("tlhip2d_ppcg_inner_iterations_f32_norm (from r), 2048^2", "k_ppcg_ring2d<float, 10, 2, true>",
     36, "FP32 cg_calc_ur + ppcg_init + 10 inner steps + the 2-norm of r"),
    ("FP32 calc_pw", "k_calc_pw_rows<float, true>", 24, "reads p, r, kx, ky; writes p', w (FP32)"),
    ("FP32 calc_w", "k_calc_pw_rows<float, false>", 16, "reads p, kx, ky; writes w (FP32)"),
    ("FP32 calc_ur", "k_groups<(anonymous namespace)::OpUr32>", 24,
     "reads d, p, r, w; writes d, r (FP32)"),
    ("FP32 2-norm", "k_groups<(anonymous namespace)::OpNorm32>", 4, "reads r (FP32)"),
    ("FP64 residual into FP32", "k_groups<(anonymous namespace)::OpResidual64>", 36,
     "reads u, u0, kx, ky (FP64); writes r (FP32)"),
    ("u += d", "k_groups<(anonymous namespace)::OpPromote>", 24,
     "reads u (FP64), d (FP32); writes u, d"),
]


def write_traffic(fetch_csv: str, write_csv: str, cells: int = 4096, cells3d: int = 256,
                  cells_mixed: int = 2048) -> dict:
    """Builds the HBM-traffic summary from the FETCH_SIZE and WRITE_SIZE
    passes (rocprofv3 --pmc ... --output-format csv) over `bench.py
    --profile-run`: bytes per launch of every kernel it launches, scaled by
    the 1 GiB device copy (1 GiB read + 1 GiB written) that opens each 2D
    round, against the compulsory bytes of the call."""
    import csv

    def rows(path, counter):
        out = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    out[int(r["Dispatch_Id"])] = (r["Kernel_Name"], float(r["Counter_Value"]))
        return out

    fe, wr = rows(fetch_csv, "FETCH_SIZE"), rows(write_csv, "WRITE_SIZE")
    ids = sorted(set(fe) & set(wr))
    disp = [(fe[i][0], fe[i][1], wr[i][1]) for i in ids]
    gib_kib = float(1 << 20)
    calib = [k for k, d in enumerate(disp) if "k_copy_all" in d[0] and d[2] > 0.9 * gib_kib]
    if len(calib) < 3:
        raise SystemExit("write_traffic: the three 1 GiB calibration copies are missing")
    f_kib = sum(disp[k][1] for k in calib) / len(calib)
    w_kib = sum(disp[k][2] for k in calib) / len(calib)
    fs, ws = gib_kib / f_kib, gib_kib / w_kib

    def bytes_of(d):
        return 1024.0 * (fs * d[1] + ws * d[2])

    def walk(start, table, rounds):
        got = {name: [] for name, *_ in table}
        k = start
        for _ in range(rounds):
            for name, frags, *_ in table:
                total, fk, wk = 0.0, 0.0, 0.0
                for frag in frags:
                    while k < len(disp) and frag not in disp[k][0]:
                        k += 1
                    if k == len(disp):
                        raise SystemExit(f"write_traffic: {frag} not found")
                    total += bytes_of(disp[k])
                    fk += disp[k][1]
                    wk += disp[k][2]
                    k += 1
                got[name].append((total, fk, wk))
        return got, k

    g2, k = walk(calib[0], _TRAFFIC_2D, 3)
    g3, k = walk(k, _TRAFFIC_3D, 3)
    kernels = {}

    def add(table, got, n_cells, size):
        for name, frags, bpc, note in table:
            v = got[name]
            b = sum(t for t, _, _ in v) / len(v)
            kernels[name] = {
                "hip_kernels": frags, "cells": n_cells, "launches": len(v), "size": size,
                "FETCH_SIZE_KiB": sum(f for _, f, _ in v) / len(v),
                "WRITE_SIZE_KiB": sum(w for _, _, w in v) / len(v),
                "bytes_per_launch": b, "min_bytes_per_launch": bpc * n_cells,
                "ratio_to_min": b / (bpc * n_cells), "note": note}

    add(_TRAFFIC_2D, g2, cells * cells, f"{cells}^2")
    add(_TRAFFIC_3D, g3, cells3d ** 3, f"{cells3d}^3")
    rest = disp[k:]
    nm = cells_mixed * cells_mixed
    for name, frag, bpc, note in _TRAFFIC_MIXED:
        v = [d for d in rest if frag in d[0]]
        if not v:
            continue
        b = sum(bytes_of(d) for d in v) / len(v)
        kernels[name] = {
            "hip_kernels": [frag], "cells": nm, "launches": len(v), "size": f"{cells_mixed}^2",
            "FETCH_SIZE_KiB": sum(d[1] for d in v) / len(v),
            "WRITE_SIZE_KiB": sum(d[2] for d in v) / len(v),
            "bytes_per_launch": b, "min_bytes_per_launch": bpc * nm,
            "ratio_to_min": b / (bpc * nm), "note": note}
    return {"calibration": {"kernel": "tl::k_copy_all (tlhip_stream_copy), 1 GiB read + 1 GiB "
                                      "written", "launches": len(calib),
                            "FETCH_SIZE_KiB": f_kib, "WRITE_SIZE_KiB": w_kib,
                            "fetch_scale": fs, "write_scale": ws},
            "kernels": kernels}


def traffic_main(fetch_csv: str, write_csv: str, commit: str) -> None:
    """--write-traffic: the summary as JSON (read by load_traffic) and as a
    markdown table beside it."""
    doc = write_traffic(fetch_csv, write_csv)
    cmds = ["rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmc_fetch -o run -- "
            "python3 bench.py --profile-run",
            "rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmc_write -o run -- "
            "python3 bench.py --profile-run",
            f"python3 bench.py --write-traffic {fetch_csv} {write_csv} --traffic-commit {commit}"]
    out = {"workload": "bench.py --profile-run: 2D 4096^2 interior (halo 2), 3D 256^3 interior "
                       "(halo 2), and one FP32-refined PPCG timestep at 2048^2",
           "commit": commit, "date": time.strftime("%Y-%m-%d"), "commands": cmds, **doc}
    path = ROOT / TRAFFIC_JSON
    path.write_text(json.dumps(out, indent=1) + "\n")
    c = doc["calibration"]
    lines = [
        "# HBM traffic per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE)", "",
        f"Kernels as of commit {commit}; MI355X, ROCm 7.2. Commands, from the repository root "
        "(rocprofv3 run from /tmp with TMPDIR=/tmp, one counter per run):", "", "```",
        *cmds, "```", "",
        "`bench.py --profile-run` launches each entry point three times on synthetic fields "
        "(2D at 4096², 3D at 256³), then one FP32-refined PPCG timestep at 2048² (the FP32 "
        "kernels, many launches each), with a 1 GiB device copy opening each 2D round. "
        f"The copy reads and writes 1 GiB: FETCH_SIZE reads {c['FETCH_SIZE_KiB']:.0f} KiB "
        f"(scale {c['fetch_scale']:.3f}), WRITE_SIZE {c['WRITE_SIZE_KiB']:.0f} KiB "
        f"(scale {c['write_scale']:.3f}). Bytes = 1024 × (fetch scale × FETCH_SIZE + write "
        "scale × WRITE_SIZE), averaged over the launches. The minimum is the call's "
        "compulsory traffic: each array it reads once plus each array it writes once, over "
        "the interior cells. Auxiliary launches (one-block reduction finishers, the ring "
        "copies of the fused loops' halo rows) are under 0.3% of any row and are not "
        "included. `" + TRAFFIC_JSON.split("/")[-1] + "` holds the numbers.", "",
        "| call | size | HIP kernel(s) | launches | measured MB | minimum MB | ratio | what it moves |",
        "|---|---|---|---|---|---|---|---|"]
    for name, v in doc["kernels"].items():
        lines.append(f"| `{name}` | {v['size']} | " + ", ".join(f"`{h.rstrip('(')}`" for h in v["hip_kernels"])
                     + f" | {v['launches']} | {v['bytes_per_launch'] / 1e6:.1f} | "
                     f"{v['min_bytes_per_launch'] / 1e6:.1f} | {v['ratio_to_min']:.3f} | {v['note']} |")
    path.with_suffix(".md").write_text("\n".join(lines) + "\n")
    print(f"wrote {path.relative_to(ROOT)} and {path.with_suffix('.md').relative_to(ROOT)}")


def blocks3d_run(args):
    """The 3D workload: 256^3 cells on [0,10]^3, recip conductivity
    (coefficient_inverse_density), 16^3-cell blocks with density
    exp(U(ln 0.1, ln 100)) and energy U(0.1, 25) from the seeded block
    generator, dt 0.004, PPCG with 10 inner steps (presteps 20), each step
    solved to ||u0 - A u|| <= 1e-10 ||u0||.  Reports the wall time per step,
    the iterations the reference would print, and the per-launch time of
    the PPCG inner sweeps: the two-step launch (72 B/cell algorithmic, 144
    B/cell for the two single steps it replaces) and the single step (72
    B/cell)."""
    import numpy as np
    import torch
    from tealeaf_amd import blocks
    from tealeaf_amd import deck as tdeck
    from tealeaf_amd.abi import lib
    from tealeaf_amd.problem import Problem

    L = lib()
    L.tlhip_init(0)
    n = 256
    text = tdeck.cuboid_3d(n, args.blocks3d_steps, solver=tdeck.PPCG, inner=10, coefficient=2,
                           dt=0.004, max_iters=20000)
    s, states = tdeck.parse_deck(text, 3)
    dens, en = blocks.block_fields((n, n, n), 16, seed=256, rho=(0.1, 100.0), energy=(0.1, 25.0),
                                   halo_depth=s.halo_depth)
    timed = ("tlhip3d_ppcg_inner_iterations", "tlhip3d_ppcg_inner_iteration")
    steps = []
    with Problem(s, states) as prob:
        prob.inject(dens, en)
        prob.set_solve(0, 1e-10)
        L.tlhip_kernel_timer_start(",".join(timed).encode())
        for tt in range(args.blocks3d_steps):
            L.tlhip_synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rep = prob.step(tt)
            L.tlhip_synchronize()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            steps.append({"seconds": dt, "printed_cg": rep["printed_cg_iters"],
                          "printed_ppcg": rep["printed_inner_iters"],
                          "ppcg_inner": rep["n_ppcg_inner"], "rel_residual": rep["rel_residual"]})
            print(f"[bench] 3d blocks step {tt}: {dt:.3f} s, {rep['printed_cg_iters']} CG + "
                  f"{rep['printed_inner_iters']} PPCG, rel residual {rep['rel_residual']:.2e}",
                  file=sys.stderr, flush=True)
        kernel = {}
        for name in timed:
            ms, nl = ctypes.c_double(), ctypes.c_long()
            L.tlhip_kernel_timer_read_entry(name.encode(), ctypes.addressof(ms),
                                            ctypes.addressof(nl))
            avg_ms = ms.value / max(nl.value, 1)
            gbs = BYTES_PER_CELL[name] * n ** 3 / (avg_ms * 1e-3) / 1e9 if nl.value else None
            kernel[name] = {"launches": nl.value, "avg_ms": avg_ms, "gbs": gbs,
                            "frac": gbs / HBM_PEAK_GBS if gbs else None}
            if name == timed[0] and gbs:
                kernel[name]["unfused_equiv_gbs"] = 2 * gbs
                kernel[name]["unfused_equiv_frac"] = 2 * gbs / HBM_PEAK_GBS
        L.tlhip_kernel_timer_read(ctypes.addressof(ms), ctypes.addressof(nl))   # stop
        u64 = prob.get("u")
    total = sum(st["seconds"] for st in steps)
    iters = sum(st["printed_cg"] + st["printed_ppcg"] for st in steps)
    # the same steps FP32-refined (FP32 sweeps, FP64 residual replacement)
    steps32 = []
    with Problem(s, states) as prob:
        prob.inject(dens, en)
        prob.set_solve(1, 1e-10)
        for tt in range(args.blocks3d_steps):
            L.tlhip_synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rep = prob.step(tt)
            L.tlhip_synchronize()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            steps32.append({"seconds": dt, "printed_cg": rep["printed_cg_iters"],
                            "printed_ppcg": rep["printed_inner_iters"],
                            "true_residual_evals": rep["n_true_residuals"],
                            "rel_residual": rep["rel_residual"]})
            print(f"[bench] 3d blocks fp32 step {tt}: {dt:.3f} s, {rep['printed_cg_iters']} CG + "
                  f"{rep['printed_inner_iters']} PPCG, rel residual {rep['rel_residual']:.2e}",
                  file=sys.stderr, flush=True)
        u32 = prob.get("u")
    total32 = sum(st["seconds"] for st in steps32)
    hd = s.halo_depth
    a = u32.reshape(s.shape())[hd:-hd, hd:-hd, hd:-hd]
    b = u64.reshape(s.shape())[hd:-hd, hd:-hd, hd:-hd]
    return {"workload": "3D 256^3 [0,10]^3, recip conductivity, 16^3-cell blocks (density "
                        "log-uniform 0.1..100, energy U(0.1,25), seed 256), dt 0.004, PPCG 10 inner "
                        "(presteps 20), relative residual 1e-10, FP64",
            "steps": steps, "seconds_per_step": total / len(steps), "iters_per_s": iters / total,
            "kernel": kernel,
            "fp32_refined": {"steps": steps32, "seconds_per_step": total32 / len(steps32),
                             "time_ratio": total32 / total,
                             "rel_error_vs_fp64": float(np.linalg.norm(a - b) / np.linalg.norm(b))}}


def adjoint_run(args):
    """The adjoint workload: loss = sum over the interior of
    (u_N - u_target)^2 after N = 5 implicit steps (CG, eps 1e-12), u_target
    the uniform-density (1.0) solve perturbed by seeded N(0, 0.01) noise;
    the gradient with respect to every cell's log-density, with per-step u
    checkpoints (tlhip_problem_adjoint: 5 forward solves, then per step back
    one adjoint CG solve and the face-gradient kernel).  2D: the layered
    deck at 2048^2; 3D: the cuboid deck at 128^3.  Times forward+backward,
    and the forward solves alone for comparison."""
    import numpy as np
    import torch
    from tealeaf_amd import deck as tdeck
    from tealeaf_amd.abi import lib
    from tealeaf_amd.problem import Problem

    L = lib()
    L.tlhip_init(0)
    steps = 5
    out = {}
    for dims, n in ((2, 2048), (3, 128)):
        text = (tdeck.layered_2d(n, steps, eps=1e-12, max_iters=20000) if dims == 2 else
                tdeck.cuboid_3d(n, steps, eps=1e-12, max_iters=20000))
        s, states = tdeck.parse_deck(text, dims)
        with Problem(s, states) as p:
            dens, e0 = p.get("density"), p.get("energy0")
        with Problem(s, states) as p:
            p.inject(np.ones_like(dens), e0)
            p.run(steps)
            u_uni = p.get("u")
        hd = s.halo_depth
        inner = tuple(slice(hd, m - hd) for m in s.shape())
        rng = np.random.default_rng(5)
        target = np.zeros(s.shape())
        target[inner] = u_uni[inner] + 0.01 * rng.standard_normal(u_uni[inner].shape)
        with Problem(s, states) as p:
            L.tlhip_synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = p.run(steps)
            L.tlhip_synchronize()
            t_fwd = time.perf_counter() - t0
        with Problem(s, states) as p:
            L.tlhip_synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            loss, grad, areps = p.adjoint(steps, target)
            L.tlhip_synchronize()
            t_adj = time.perf_counter() - t0
        key = f"{dims}d_{n}"
        out[key] = {"forward_backward_s": t_adj, "forward_only_s": t_fwd,
                    "backward_over_forward": (t_adj - t_fwd) / t_fwd,
                    "forward_cg_iters": [r["printed_cg_iters"] for r in reps],
                    "loss": loss, "grad_norm": float(np.linalg.norm(grad))}
        print(f"[bench] adjoint {key}: forward+backward {t_adj:.3f} s (forward {t_fwd:.3f} s)",
              file=sys.stderr, flush=True)
    return {"workload": "loss sum (u_5 - u_target)^2 after 5 CG steps (eps 1e-12), gradient "
                        "w.r.t. per-cell log-density; 2D layered deck 2048^2, 3D cuboid deck 128^3",
            **out}


def kernel_roofline(name: str, k: dict, res: dict, n_int: int) -> dict:
    """Roofline of one timed entry point: achieved = its compulsory HBM
    bytes per launch (each array read once + each written once, over the
    interior cells) / its average launch time from HIP events on the suite's
    stream; traffic = the counter-measured bytes per launch from the
    committed rocprofv3 summary."""
    t = k["avg_ms"] * 1e-3
    out = {}
    init_share = 0.0
    if name == "tlhip2d_ppcg_inner_iterations":
        sw = res["by_solver"]["ppcg"]["sweeps"]
        # A launch that starts an outer iteration (cg_calc_ur + ppcg_init +
        # the inner steps) reads r, w, u, p, kx, ky and writes sd', r', u'
        # (72 B/cell); one that continues from sd reads sd, kx, ky, r, u and
        # writes the same (64 B/cell).
        init_share = min(1.0, sw["ppcg_outer"] / k["launches"])
        min_cell = 64.0 + 8.0 * init_share
        # the bytes the same steps would move unfused: 64 B/cell per inner
        # step plus cg_calc_ur (48) and ppcg_init (16) per outer iteration
        unfused_cell = (64.0 * sw["ppcg_inner"] + (48.0 + 16.0) * sw["ppcg_outer"]) / k["launches"]
        out["inner_steps_per_launch"] = sw["ppcg_inner"] / k["launches"]
        # the rate an unfused sequence of the same steps would need (not a
        # fraction of any peak: one launch does several steps' work)
        out["unfused_equivalent_gbs"] = unfused_cell * n_int / t / 1e9
    else:
        min_cell = float(BYTES_PER_CELL[name])
    achieved = min_cell * n_int / t / 1e9
    traffic, tsrc = load_traffic(name, init_share)
    return {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": tsrc, "bytes_per_cell": min_cell,
            "algorithmic_bytes": min_cell * n_int, "avg_kernel_ms": k["avg_ms"],
            "launches": k["launches"], "share_of_timed_time": k["share_of_timed_time"],
            # what the launch moves through HBM by the counters, and at what rate
            "hbm_gbs": traffic / t / 1e9 if traffic else None,
            "hbm_frac": traffic / t / 1e9 / HBM_PEAK_GBS if traffic else None, **out}


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path "
              "(one chunk per GPU, halo faces and two scalar all-reduces per CG iteration) "
              "is a design, not a build.", file=sys.stderr)
        sys.exit(2)
    if args.write_traffic:
        traffic_main(*args.write_traffic, args.traffic_commit)
        return
    if args.profile_run:
        profile_run(args)
        return
    if args.refine_only:
        print(json.dumps({"refine": refine_run(args)}), flush=True)
        return
    if args.blocks3d:
        print(json.dumps({"blocks3d": blocks3d_run(args)}), flush=True)
        return
    res = gpu_run(args)
    refine = refine_run(args) if args.refine_cells > 0 else None
    adjoint = None if args.no_adjoint else adjoint_run(args)
    value = res["iters"] / res["time"]
    n_int = args.cells * args.cells
    # roofline: the timed kernel with the most total time over the three legs
    timed = {n: k for n, k in res["kernels"].items() if k["launches"]}
    dom = max(timed, key=lambda n: timed[n]["total_ms"]) if timed else None
    roofline = kernel_roofline(dom, timed[dom], res, n_int) if dom else None
    fused = dominant(args)
    roofline_ppcg = (kernel_roofline(fused, timed[fused], res, n_int)
                     if fused in timed and fused != dom else None)
    cpu = None if args.no_cpu else cpu_baseline(args.cells, args.cpu_cg_iters,
                                                 args.cpu_cheby_iters, args.cpu_ppcg_outer)
    if cpu:
        # the CPU time for the same iteration mix as the timed GPU steps
        rate = {kind: v["iters_per_s"] for kind, v in cpu["by_solver"].items()}
        kind_of = {"cg": "cg", "cheby": "cheby", "ppcg": "ppcg"}
        cpu_total, ratio = 0.0, {}
        ok = True
        for sname, b in res["by_solver"].items():
            n_cg = sum(st["CG"] for st in b["per_step"])
            n_in = sum(st["inner"] for st in b["per_step"])
            if "cg" not in rate or (n_in and kind_of[sname] not in rate):
                ok = False
                continue
            t_cpu = n_cg / rate["cg"] + (n_in / rate[kind_of[sname]] if n_in else 0.0)
            cpu_total += t_cpu
            ratio[sname] = t_cpu / b["seconds"]
        cpu["gpu_speedup_by_solver"] = ratio
        cpu["value"] = res["iters"] / cpu_total if ok and cpu_total > 0 else None
        cpu["value_note"] = ("the timed GPU steps' iteration mix (CG, Chebyshev and PPCG "
                             "iterations as printed) at the per-solver CPU rates")
    line = {
        "metric": "solver iters/s + HBM GB/s, 2D 4096²; FP32-refined vs FP64 time-to-1e-10",
        "value": value,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * res["time"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: the reference layered-rectangle deck (2d/tea.in states) generated in-process",
        "config": {"workload": f"2D {args.cells}x{args.cells} layered-rectangle implicit solve, "
                               "FP64, eps 1e-15, solved by CG, Chebyshev (presteps 30) and PPCG "
                               "(10 inner, presteps 30); one step = one timestep of each",
                   "cells": n_int, "ppcg_fusion": args.ppcg_fusion, "solvers": args.solvers, "eps": 1e-15, "max_iters": 10000},
        "roofline": roofline,
        "roofline_ppcg_fused": roofline_ppcg,
        "cpu_baseline": cpu,
        "refine": refine,
        "adjoint": adjoint,
        "by_solver": res["by_solver"],
        "kernels": res["kernels"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
