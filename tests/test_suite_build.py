"""The suite plugs into the reference's own build: a copy of the reference's
2d/ and 3d/ trees (never the original) with tealeaf_amd/hip_cdna4/{2d,3d}
linked in as c_kernels/hip_cdna4 builds with

    make KERNELS=hip_cdna4 CC=gcc COMPILER=GNU OPTIONS="-DNO_MPI -DENABLE_PROFILING"

(2d/Makefile:36-37 and 3d/Makefile:29-30 include the suite's Makefile and
make.deps), with and without -DDIFFUSE_OVERLOAD.  The binary must export
every run_* entry point of the reference's kernel_interface.h from the
suite's glue and resolve libtealeaf_hip.so.  CPU only (nothing runs on a
GPU); skipped where the reference tree or the built library is absent.
"""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REFERENCE = Path("/root/reference")
LIB = ROOT / "tealeaf_amd" / "_lib" / "libtealeaf_hip.so"

NEEDS = pytest.mark.skipif(
    not (REFERENCE / "2d" / "Makefile").exists() or not LIB.exists()
    or shutil.which("make") is None or shutil.which("gcc") is None,
    reason="needs the reference tree, make, gcc and the built libtealeaf_hip.so")


def _declared(dim: str) -> set[str]:
    text = (REFERENCE / dim / "kernel_interface.h").read_text()
    return set(re.findall(r"\b(run_\w+)\s*\(", text))


@pytest.fixture(scope="module")
def tree(tmp_path_factory):
    dst = tmp_path_factory.mktemp("reftree")
    for dim in ("2d", "3d"):
        shutil.copytree(REFERENCE / dim, dst / dim, symlinks=True,
                        ignore=shutil.ignore_patterns("kokkos*", "raja", "openacc", "omp4*",
                                                      "cuda", "mkl", ".git"))
        (dst / dim / "c_kernels" / "hip_cdna4").symlink_to(ROOT / "tealeaf_amd" / "hip_cdna4" / dim)
    return dst


@NEEDS
@pytest.mark.parametrize("dim", ["2d", "3d"])
@pytest.mark.parametrize("overload", [False, True], ids=["diffuse", "diffuse_overload"])
def test_reference_make_builds_suite(tree, dim, overload):
    d = tree / dim
    subprocess.run(["make", "clean"], cwd=d, capture_output=True)
    opts = "-DNO_MPI -DENABLE_PROFILING" + (" -DDIFFUSE_OVERLOAD" if overload else "")
    cmd = ["make", "-j4", "KERNELS=hip_cdna4", "CC=gcc", "CPP=g++", "COMPILER=GNU",
           f"OPTIONS={opts}"]
    p = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    exe = d / "tealeaf"
    assert exe.exists()
    objdir = d / ("obj/hip_cdna4" if dim == "2d" else "build/obj")
    nm = subprocess.run(["nm", "--defined-only", str(objdir / "kernel_interface.o")],
                        capture_output=True, text=True, check=True).stdout
    defined = set(re.findall(r"\bT (run_\w+)$", nm, re.M))
    missing = _declared(dim) - defined
    assert not missing, missing
    ldd = subprocess.run(["ldd", str(exe)], capture_output=True, text=True).stdout
    assert re.search(r"libtealeaf_hip\.so => \S*/tealeaf_amd/_lib/libtealeaf_hip\.so", ldd), ldd
    main_syms = subprocess.run(["nm", str(exe)], capture_output=True, text=True).stdout
    assert re.search(r"\bT diffuse_overload$", main_syms, re.M)
