"""Writes tests/golden/omp3_fullsize.json: the reference's omp3 CPU path at
the sizes of its own known answers and of the benchmark decks.

Each case runs the reference host + omp3 suite, built from the reference
sources by oracle/Makefile with the report hook that prints all four
field_summary values, on a deck made by tealeaf_amd.deck, and records per
timestep the iteration lines the host prints, plus the final
vol/mass/ie/temp and the temperature line it checks against tea.problems:

  top level   oracle/_ref/tealeaf{2d,3d}_omp3: the reference's own flags
              (gcc -O3, which contracts a*b+c into fused multiply-adds);
  "nofma"     oracle/_ref/tealeaf{2d,3d}_omp3_nofma: the same with
              -ffp-contract=off, every operation of the C source rounded on
              its own -- the arithmetic the hip_cdna4 suite follows.

The GPU parity tests compare the suite with these records, so the
1-20 minute CPU runs are not repeated on every test run.

Usage: python tests/golden/make_fullsize.py [--nofma] [case ...]   (default: all)
"""
from __future__ import annotations

import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from tealeaf_amd import deck as tdeck  # noqa: E402
from tests.problem_helpers import ref_run  # noqa: E402

OUT = Path(__file__).resolve().parent / "omp3_fullsize.json"

CASES = {
    # the reference's stored answers (2d/tea.problems, 3d/tea.problems), its deck
    "2d_512x20_cg": (2, tdeck.layered_2d(512, 20, max_iters=5000)),
    "2d_1024x20_cg": (2, tdeck.layered_2d(1024, 20, max_iters=5000)),
    "2d_4096x1_cg": (2, tdeck.layered_2d(4096, 1, max_iters=5000)),
    "3d_32x20_cg": (3, tdeck.cuboid_3d(32, 20)),
    "3d_64x20_cg": (3, tdeck.cuboid_3d(64, 20)),
    "3d_256x5_cg": (3, tdeck.cuboid_3d(256, 5)),
    # the benchmark decks: 512^2 x 10 CG (eps 1e-15, max_iters 10000), and
    # 4096^2 one timestep as Chebyshev and PPCG (presteps 30, 10 inner)
    "2d_512x10_cg_maxit10000": (2, tdeck.layered_2d(512, 10, max_iters=10000)),
    "2d_4096x1_cheby": (2, tdeck.layered_2d(4096, 1, solver=tdeck.CHEBY, presteps=30,
                                            max_iters=10000)),
    "2d_4096x1_ppcg": (2, tdeck.layered_2d(4096, 1, solver=tdeck.PPCG, presteps=30, inner=10,
                                           max_iters=10000)),
}

_ACTUAL = re.compile(r"^Actual\s+(\S+)", re.M)


def main(args: list[str]) -> None:
    nofma = "--nofma" in args
    names = [a for a in args if not a.startswith("--")]
    threads = int(os.environ.get("OMP_NUM_THREADS", str(os.cpu_count() or 1)))
    for name in names or list(CASES):
        dims, text = CASES[name]
        t0 = time.time()
        binary = f"tealeaf{dims}d_omp3" + ("_nofma" if nofma else "")
        r = ref_run(binary, text, timeout=7200, threads=threads)
        m = _ACTUAL.search(r["stdout"])
        rec = {
            "threads": threads,
            "steps": [s["iters"] for s in r["steps"]],
            "summary": r["summary"],
            "actual_temp": float(m.group(1)) if m else None,
            "passed": "PASSED" in r["stdout"],
            "wallclock_s": r["steps"][-1]["wallclock"] if r["steps"] else None,
        }
        data = json.loads(OUT.read_text()) if OUT.exists() else {}
        entry = data.setdefault(name, {"dims": dims, "deck": text})
        if nofma:
            entry["nofma"] = rec
        else:
            entry.update(rec)
        OUT.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
        print(f"{name} ({binary}): {time.time() - t0:.1f} s, steps {rec['steps'][-1]}, "
              f"summary {r['summary']}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
