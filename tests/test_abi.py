"""The C ABI library loads and exports every entry point include/*.h declares
(CPU only: no compute calls)."""
import ctypes
import re
import subprocess
from pathlib import Path

from tealeaf_amd import abi

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(tlhip\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    assert abi.LIB_PATH.exists(), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    dll = ctypes.CDLL(str(abi.LIB_PATH))
    declared = _declared()
    assert len(declared) >= 50
    missing = [n for n in sorted(declared) if not hasattr(dll, n)]
    assert not missing, missing


def test_every_export_is_declared():
    out = subprocess.run(["nm", "-D", "--defined-only", str(abi.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (tlhip\w+)", out))
    assert exported == _declared()


def test_header_parses_to_ctypes():
    decls = abi.parse_header()
    assert decls["tlhip2d_cg_calc_w"][1][:3] == [ctypes.c_int] * 3
    assert decls["tlhip3d_ppcg_inner_iteration"][1][:4] == [ctypes.c_int] * 4
    assert decls["tlhip_last_error"][0] == "const char*"


def test_library_loads_without_gpu():
    """Loading the library in a fresh process needs no GPU and records no
    error (a subprocess, so that calls made by earlier tests in this one do
    not count)."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, sys.argv[1]); from tealeaf_amd import abi; "
            "print(repr(abi.lib().last_error()))")
    out = subprocess.run([sys.executable, "-c", code, str(abi.ROOT)], capture_output=True,
                         text=True, timeout=120, env={**__import__("os").environ,
                                                      "HIP_VISIBLE_DEVICES": ""})
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "''", out.stdout
