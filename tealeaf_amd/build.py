"""Builds libtealeaf_hip.so for gfx950 in-tree with hipcc.

Flags: -ffp-contract=off keeps every kernel's arithmetic to the same IEEE
operations, in the same order, as the reference's C expressions; the
packed-FP32 feature is switched off for all device code; --no-undefined
makes a missing kernel definition a link error rather than a load error.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_lib" / "libtealeaf_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TLHIP_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"   # the host pass does not know the device feature
    err = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if noise not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{err}")
    if err.strip():
        print(err, file=sys.stderr)


def build_lib(verbose: bool = False) -> Path:
    """Compile every source to an object in parallel, then link the library."""
    OUT.parent.mkdir(parents=True, exist_ok=True)
    objdir = PKG / "_lib" / "obj"
    objdir.mkdir(parents=True, exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = objdir / (src.name + ".o")
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append(f"{' '.join(cmd)}\n{out}")
        elif out.strip():
            print(out, file=sys.stderr)
    if failed:
        raise RuntimeError("compile failed:\n" + "\n\n".join(failed))
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined",
          *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_lib(verbose="-v" in sys.argv))
