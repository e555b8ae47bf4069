"""Whole-problem runners shared by the parity tests.

  oracle_run(deck_text)         -- the CPU oracle's driver (oracle/oracle_driver.c)
  ref_run(binary, deck_text)    -- a reference host binary built under oracle/_ref/
                                   (omp3 suite, or the hip_cdna4 suite), parsed
  known_answer(settings)        -- the reference's stored temp (tests/golden/tea_problems_*)

Each returns a dict with per-step reports and the final four-field summary.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess
import tempfile
from pathlib import Path

from tealeaf_amd import deck as tdeck

GOLDEN = Path(__file__).resolve().parent / "golden"
REF_DIR = Path(__file__).resolve().parents[1] / "oracle" / "_ref"


def known_answer(s: tdeck.Settings):
    path = GOLDEN / f"tea_problems_{s.dims}d.txt"
    for line in path.read_text().splitlines():
        parts = line.split()
        if not parts:
            continue
        nums = [int(v) for v in parts[:-1]]
        if s.dims == 2 and nums == [s.x_cells, s.y_cells, s.end_step]:
            return float(parts[-1])
        if s.dims == 3 and nums == [s.x_cells, s.y_cells, s.z_cells, s.end_step]:
            return float(parts[-1])
    return None


def reference_check_passes(temp: float, expected: float) -> bool:
    """The reference's own known-answer test (2d/drivers/field_summary_driver.c:43-52)."""
    return abs(100.0 * (temp / expected) - 100.0) < 0.001


def oracle_run(text: str, dims: int | None = None, inject=None) -> dict:
    from oracle.pyoracle import run_problem
    s, states = tdeck.parse_deck(text, dims or (3 if "z_cells" in text else 2))
    return run_problem(s, states, inject=inject)


_ITER = re.compile(r"^(CG|Cheby|PPCG|Jacobi):\s+(\d+) iterations", re.M)


def ref_run(binary: str, text: str, timeout: float = 600.0, threads: int | None = None) -> dict:
    """Run a reference host binary on a deck in a scratch directory."""
    exe = REF_DIR / binary
    if not exe.exists():
        raise FileNotFoundError(exe)
    dims = 3 if "3d" in binary else 2
    with tempfile.TemporaryDirectory() as d:
        Path(d, "tea.in").write_text(text)
        Path(d, "tea.problems").write_text((GOLDEN / f"tea_problems_{dims}d.txt").read_text())
        env = dict(os.environ)
        if threads:
            env["OMP_NUM_THREADS"] = str(threads)
        p = subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, timeout=timeout,
                           env=env)
    if p.returncode != 0:
        raise RuntimeError(f"{binary} exited {p.returncode}: {p.stdout[-2000:]}{p.stderr[-2000:]}")
    steps = []
    for block in p.stdout.split("Timestep ")[1:]:
        it = {k: int(v) for k, v in _ITER.findall(block)}
        m = re.search(r"Wallclock:\s+([0-9.]+)s", block)
        e = re.search(r"Error:\s+(\S+)", block)
        steps.append({"iters": it, "wallclock": float(m.group(1)) if m else None,
                      "error": float(e.group(1)) if e else None})
    m = re.search(r"TL_FIELD_SUMMARY (\S+) (\S+) (\S+) (\S+)", p.stdout)
    summary = [float(v) for v in m.groups()] if m else None
    m = re.search(r"TL_REFORDER_SUMMARY (\S+) (\S+) (\S+) (\S+)", p.stdout)
    refsum = [float(v) for v in m.groups()] if m else None
    return {"steps": steps, "summary": summary, "refsum": refsum, "stdout": p.stdout}


def _cell_terms(s: tdeck.Settings, density, energy0, u):
    """The per-cell terms of 2d/c_kernels/omp3/field_summary.c:31-38 (3D
    3d/c_kernels/omp3/field_summary.c:33-40) over the interior, in the
    reference's loop order: cell volume (dx*dy[*dz], set_chunk_data.c:49),
    mass = volume*density, mass*energy0, mass*u -- each an IEEE product."""
    import numpy as np
    hd = s.halo_depth
    inner = tuple(slice(hd, n - hd) for n in s.shape())
    d = np.asarray(density).reshape(s.shape())[inner].ravel()
    e0 = np.asarray(energy0).reshape(s.shape())[inner].ravel()
    uu = np.asarray(u).reshape(s.shape())[inner].ravel()
    vol = s.dx * s.dy * s.dz if s.dims == 3 else s.dx * s.dy
    v = np.full(d.size, vol)
    mass = v * d
    return v, mass, mass * e0, mass * uu


def reference_order_summary(s: tdeck.Settings, density, energy0, u) -> list[float]:
    """vol, mass, ie, temp summed as the reference's field_summary sums
    them: one cell after another into a running double, in its loop order
    (numpy's add.accumulate is that sequential loop)."""
    import numpy as np
    return [float(np.add.accumulate(t)[-1]) for t in _cell_terms(s, density, energy0, u)]


def exact_summary(s: tdeck.Settings, density, energy0, u) -> list[float]:
    """The same four sums correctly rounded (math.fsum)."""
    import math
    return [math.fsum(t.tolist()) for t in _cell_terms(s, density, energy0, u)]


def printed_iters(rep: dict, solver: int) -> dict:
    """The iteration lines the reference prints for a step, from a report."""
    if solver == tdeck.JACOBI:
        return {"Jacobi": rep["printed_cg_iters"]}
    out = {"CG": rep["printed_cg_iters"]}
    if solver == tdeck.CHEBY:
        out["Cheby"] = rep["printed_inner_iters"]
    if solver == tdeck.PPCG:
        out["PPCG"] = rep["printed_inner_iters"]
    return out


def rel(a: float, b: float) -> float:
    return abs(a - b) / max(abs(b), 1e-300)
