"""ctypes loader for the CPU oracle (oracle/liboracle.so) and, when built,
the reference omp3 kernels compiled from source (oracle/_ref/libomp3_*.so).

Test infrastructure only: the oracle is the checker, never the product.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
REF_DIR = ORACLE_DIR / "_ref"

_SCALAR = {"int": ctypes.c_int, "double": ctypes.c_double, "bool": ctypes.c_bool,
           "long": ctypes.c_long}
_DECL = re.compile(r"^\s*(void|int|long|or_problem\*)\s+(or\w+)\s*\(([^;]*?)\)\s*;", re.S | re.M)


def _atype(decl: str):
    d = " ".join(decl.replace("const ", "").split())
    if d in ("void", ""):
        return None
    if "*" in d or "[" in d:
        return ctypes.c_void_p
    return _SCALAR[d.rsplit(" ", 1)[0]]


class Oracle:
    def __init__(self, path: Path):
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.dll = ctypes.CDLL(str(path))
        text = re.sub(r"/\*.*?\*/", "", (ORACLE_DIR / "oracle.h").read_text(), flags=re.S)
        for ret, name, args in _DECL.findall(text):
            fn = getattr(self.dll, name)
            fn.argtypes = [t for t in (_atype(a) for a in args.split(",")) if t is not None]
            fn.restype = {"void": None, "int": ctypes.c_int, "long": ctypes.c_long}.get(
                ret, ctypes.c_void_p)

    def __getattr__(self, name):
        return getattr(self.dll, name)


_OR = None


def load_oracle() -> Oracle:
    global _OR
    if _OR is None:
        _OR = Oracle(ORACLE_DIR / "liboracle.so")
    return _OR


def load_ref(dims: int):
    """The reference omp3 kernels (serial build), or None if not built here."""
    p = REF_DIR / f"libomp3_{dims}d.so"
    if not p.exists():
        return None
    dll = ctypes.CDLL(str(p))
    return dll


def ptr(a) -> int:
    """Address of a numpy array's data (or 0 for None)."""
    return 0 if a is None else a.ctypes.data
