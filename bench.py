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
dominant kernel (the temporally blocked PPCG inner loop, timed live with HIP events on
the suite's stream), and the reference omp3 CPU path timed on this host on a
bounded sample of the same deck.

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
    "tlhip2d_cg_calc_w": 32,              # read p, kx, ky; write w
    "tlhip2d_cg_calc_ur": 48,             # read u, p, r, w; write u, r
    "tlhip2d_cg_calc_p": 24,              # read p, r; write p
    "tlhip2d_cheby_iterate": 64,          # read u, u0, p, kx, ky; write u', p, r
    "tlhip2d_cheby_iterations": 64,       # the same, once per launch of up to 10 iterations
    "tlhip3d_ppcg_inner_iteration": 72,   # read sd, kx, ky, kz, r, u; write r, u, sd'
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
    ap.add_argument("--refine-only", action="store_true",
                    help="run only the FP32-refined vs FP64 comparison and print its JSON")
    ap.add_argument("--blocks3d", action="store_true",
                    help="run only the 3D 256^3 seeded block-density PPCG workload and print its JSON")
    ap.add_argument("--blocks3d-steps", type=int, default=2)
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint (forward+backward) timing")
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
    for s in solvers:
        timed = dominant(args) if s == "ppcg" else ("tlhip2d_cg_calc_pw" if s == "cg" else "tlhip2d_cheby_iterations")
        L.tlhip_synchronize()
        torch.cuda.synchronize()
        L.tlhip_kernel_timer_start(timed.encode())
        t0 = time.perf_counter()
        reps = [probs[s].step(tt) for tt in range(args.warmup, total_steps)]
        L.tlhip_synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ms, nl = ctypes.c_double(), ctypes.c_long()
        L.tlhip_kernel_timer_read(ctypes.addressof(ms), ctypes.addressof(nl))
        printed = sum(r["printed_cg_iters"] + r["printed_inner_iters"] for r in reps)
        sweeps = {"cg_steps": sum(r["n_cg_steps"] for r in reps),
                  "cheby_steps": sum(r["n_cheby_steps"] for r in reps),
                  "ppcg_outer": sum(r["n_ppcg_outer"] for r in reps),
                  "ppcg_inner": sum(r["n_ppcg_inner"] for r in reps)}
        avg_ms = ms.value / max(nl.value, 1)
        kernel[timed] = {"launches": nl.value, "avg_ms": avg_ms,
                         "gbs": BYTES_PER_CELL[timed] * n_int / (avg_ms * 1e-3) / 1e9 if nl.value else None}
        by[s] = {"iterations": printed, "seconds": dt, "iters_per_s": printed / dt,
                 "per_step": [{"CG": r["printed_cg_iters"], "inner": r["printed_inner_iters"]}
                              for r in reps], "sweeps": sweeps,
                 "temp": probs[s].field_summary()[3]}
        total_time += dt
        total_iters += printed
        print(f"[bench] {s}: {printed} iterations in {dt:.3f} s", file=sys.stderr, flush=True)
    for p in probs.values():
        p.close()
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
            for tt in range(args.refine_steps):
                L.tlhip_synchronize()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                reps.append(prob.step(tt))
                L.tlhip_synchronize()
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t0)
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
                "rel_residual": [r["rel_residual"] for r in reps]}
        hd = s.halo_depth
        a = u_final[1].reshape(s.shape())[hd:-hd, hd:-hd]
        b = u_final[0].reshape(s.shape())[hd:-hd, hd:-hd]
        res["rel_error_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        res["time_ratio"] = res["fp32_refined"]["mean_s"] / res["fp64"]["mean_s"]
        out[f"contrast_{contrast:g}"] = res
    return {"workload": f"2D {n}x{n}, 64x64-cell blocks, log-uniform density [1, C], energy "
                        f"U(0.1,25), dt 0.004, PPCG 10 inner via the error switch, "
                        f"{args.refine_steps} steps each solved to relative residual 1e-10",
            "target_ratio": 0.6, **out}


def profile_run(args):
    """Three launches each of the hot kernels on --cells^2 synthetic fields,
    and of a 1 GiB device copy (8 B read + 8 B written per element) whose
    FETCH_SIZE / WRITE_SIZE calibrate the counters; for rocprofv3 --pmc."""
    import ctypes as C
    import numpy as np
    from tealeaf_amd.abi import lib
    from tealeaf_amd.device import DeviceArray

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
    f32 = {k: DeviceArray.from_host(rng.uniform(-1, 1, P * y).astype(np.float32))
           for k in ("u", "u2", "r", "r2", "sd", "sd2", "p", "w")}
    for k in ("kx", "ky"):
        f32[k] = DeviceArray.from_host(rng.uniform(0.05, 0.2, P * y).astype(np.float32))
    ncopy = (1 << 30) // 8
    src, dst = DeviceArray(ncopy, np.float64), DeviceArray(ncopy, np.float64)
    nb = np.full(4, -1, dtype=np.int32)
    al, be = np.full(10, 0.7), np.full(10, 0.2)
    pw = C.c_double()
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
        # 10 Chebyshev iterations (u -> u2, p -> sd2; u0 = sd)
        L.tlhip2d_cheby_iterations(x, y, hd, 10, al.ctypes.data, be.ctypes.data, nb.ctypes.data,
                                   f64["u"].ptr, f64["u2"].ptr, f64["sd"].ptr, f64["p"].ptr,
                                   f64["sd2"].ptr, f64["r2"].ptr, f64["kx"].ptr, f64["ky"].ptr)
        L.tlhip2d_ppcg_inner_iterations_f32(x, y, hd, P, 10, al.ctypes.data, be.ctypes.data, 1.3, 1,
                                            nb.ctypes.data, None, f32["sd2"].ptr,
                                            f32["r"].ptr, f32["r2"].ptr, f32["u"].ptr,
                                            f32["u2"].ptr, f32["kx"].ptr, f32["ky"].ptr, 0.3,
                                            f32["p"].ptr, f32["w"].ptr)
        L.tlhip2d_ppcg_inner_iteration(x, y, hd, 0.7, 0.2, f64["u"].ptr, f64["r"].ptr, f64["kx"].ptr,
                                       f64["ky"].ptr, f64["sd"].ptr, f64["sd2"].ptr)
        L.tlhip2d_cg_calc_w(x, y, hd, C.addressof(pw), f64["p"].ptr, f64["w"].ptr, f64["kx"].ptr,
                            f64["ky"].ptr)
    L.tlhip_synchronize()
    print(json.dumps({"profile_run": True, "cells": n, "copy_elements": ncopy}), flush=True)


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
    path = ROOT / "profiles" / "ppcg_4096_traffic.json"
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


def blocks3d_run(args):
    """The 3D workload: 256^3 cells on [0,10]^3, recip conductivity
    (coefficient_inverse_density), 16^3-cell blocks with density
    exp(U(ln 0.1, ln 100)) and energy U(0.1, 25) from the seeded block
    generator, dt 0.004, PPCG with 10 inner steps (presteps 20), each step
    solved to ||u0 - A u|| <= 1e-10 ||u0||.  Reports the wall time per step,
    the iterations the reference would print, and the per-launch time of
    the PPCG inner sweep (72 B/cell)."""
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
    timed = "tlhip3d_ppcg_inner_iteration"
    steps = []
    with Problem(s, states) as prob:
        prob.inject(dens, en)
        prob.set_solve(0, 1e-10)
        L.tlhip_kernel_timer_start(timed.encode())
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
        ms, nl = ctypes.c_double(), ctypes.c_long()
        L.tlhip_kernel_timer_read(ctypes.addressof(ms), ctypes.addressof(nl))
    total = sum(st["seconds"] for st in steps)
    iters = sum(st["printed_cg"] + st["printed_ppcg"] for st in steps)
    avg_ms = ms.value / max(nl.value, 1)
    gbs = BYTES_PER_CELL[timed] * n ** 3 / (avg_ms * 1e-3) / 1e9 if nl.value else None
    return {"workload": "3D 256^3 [0,10]^3, recip conductivity, 16^3-cell blocks (density "
                        "log-uniform 0.1..100, energy U(0.1,25), seed 256), dt 0.004, PPCG 10 inner "
                        "(presteps 20), relative residual 1e-10, FP64",
            "steps": steps, "seconds_per_step": total / len(steps), "iters_per_s": iters / total,
            "kernel": {timed: {"launches": nl.value, "avg_ms": avg_ms, "gbs": gbs,
                               "frac": gbs / HBM_PEAK_GBS if gbs else None}}}


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


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path "
              "(one chunk per GPU, halo faces and two scalar all-reduces per CG iteration) "
              "is a design, not a build.", file=sys.stderr)
        sys.exit(2)
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
    dom = dominant(args)
    k = res["kernels"].get(dom)
    roofline = None
    if k and k["launches"]:
        sw = res["by_solver"]["ppcg"]["sweeps"]
        t = k["avg_ms"] * 1e-3
        if dom == "tlhip2d_ppcg_inner_iterations":
            # Compulsory HBM bytes of one launch: every array it reads once and
            # every array it writes once.  A launch that starts an outer
            # iteration (cg_calc_ur + ppcg_init + the inner steps) reads r, w,
            # u, p, kx, ky and writes sd', r', u' (72 B/cell); one that
            # continues from sd reads sd, kx, ky, r, u and writes the same
            # (64 B/cell).
            init_share = min(1.0, sw["ppcg_outer"] / k["launches"])
            min_cell = 64.0 + 8.0 * init_share
            # the bytes the same steps would move unfused: 64 B/cell per inner
            # step plus cg_calc_ur (48) and ppcg_init (16) per outer iteration
            unfused_cell = (64.0 * sw["ppcg_inner"] + (48.0 + 16.0) * sw["ppcg_outer"]) / k["launches"]
        else:
            init_share, min_cell, unfused_cell = 0.0, 64.0, 64.0
        achieved = min_cell * n_int / t / 1e9
        traffic, tsrc = load_traffic(dom, init_share)
        roofline = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                    "traffic_source": tsrc,
                    "bytes_per_cell": min_cell, "algorithmic_bytes": min_cell * n_int,
                    "avg_kernel_ms": k["avg_ms"], "launches": k["launches"],
                    "inner_steps_per_launch": sw["ppcg_inner"] / k["launches"],
                    # what the launch moves through HBM by the counters, and at what rate
                    "hbm_gbs": traffic / t / 1e9 if traffic else None,
                    "hbm_frac": traffic / t / 1e9 / HBM_PEAK_GBS if traffic else None,
                    # the rate an unfused sequence of the same steps would need (not a
                    # fraction of any peak: one launch does several steps' work)
                    "unfused_equivalent_gbs": unfused_cell * n_int / t / 1e9}
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
        "cpu_baseline": cpu,
        "refine": refine,
        "adjoint": adjoint,
        "by_solver": res["by_solver"],
        "kernels": res["kernels"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
