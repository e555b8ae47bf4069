"""ctypes binding of libtealeaf_hip.so, generated from include/tealeaf_hip.h.

The header is the single source of truth: every ``int name(...);``
declaration in it becomes a checked Python callable here.  Pointer
arguments are passed as integer addresses (device pointers for fields, host
addresses for scalar outputs and small host arrays).

There is no CPU fallback: if the shared library is missing or does not load,
``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent
HEADER = ROOT / "include" / "tealeaf_hip.h"
LIB_PATH = PKG_DIR / "_lib" / "libtealeaf_hip.so"


class TealeafError(RuntimeError):
    """A suite entry point failed (the reference would have called die())."""


_SCALAR = {
    "int": ctypes.c_int,
    "double": ctypes.c_double,
    "bool": ctypes.c_bool,
    "size_t": ctypes.c_size_t,
}

_DECL = re.compile(r"^\s*(int|const char\*|void\*)\s+(tlhip\w+)\s*\(([^;]*?)\)\s*;", re.S | re.M)


def _arg_type(decl: str):
    decl = " ".join(decl.replace("const ", "").split())
    if decl == "void":
        return None
    if "*" in decl:
        return ctypes.c_void_p
    base = decl.rsplit(" ", 1)[0]
    if base not in _SCALAR:
        raise ValueError(f"unsupported argument type in header: {decl!r}")
    return _SCALAR[base]


def parse_header(path: Path = HEADER) -> dict[str, tuple[str, list]]:
    """Return {name: (return_type_text, [ctypes arg types])} for every entry point."""
    text = re.sub(r"/\*.*?\*/", "", path.read_text(), flags=re.S)
    out: dict[str, tuple[str, list]] = {}
    for ret, name, args in _DECL.findall(text):
        types = []
        for a in args.split(","):
            t = _arg_type(a.strip())
            if t is not None:
                types.append(t)
        out[name] = (ret.strip(), types)
    return out


class _Lib:
    def __init__(self, path: Path):
        if not path.exists():
            raise TealeafError(
                f"{path} is missing: build it with __graft_entry__.build() "
                "(the HIP suite has no CPU fallback)")
        self.path = path
        self.dll = ctypes.CDLL(str(path))
        self.decls = parse_header()
        for name, (ret, types) in self.decls.items():
            fn = getattr(self.dll, name)
            fn.argtypes = types
            if ret == "int":
                fn.restype = ctypes.c_int
            elif ret == "const char*":
                fn.restype = ctypes.c_char_p
            else:
                fn.restype = ctypes.c_void_p

    def last_error(self) -> str:
        msg = self.dll.tlhip_last_error()
        return msg.decode() if msg else ""

    def call(self, name: str, *args) -> None:
        """Call an int-returning entry point and raise on a non-zero code."""
        rc = getattr(self.dll, name)(*args)
        if rc != 0:
            raise TealeafError(f"{name} failed ({rc}): {self.last_error()}")

    def __getattr__(self, name: str):
        """int entry points raise TealeafError on a non-zero code; the rest
        (tlhip_last_error) are returned as they are."""
        if name.startswith("tlhip") and name in self.decls:
            if self.decls[name][0] != "int":
                return getattr(self.dll, name)
            def checked(*args, _n=name):
                return self.call(_n, *args)
            return checked
        raise AttributeError(name)


_LIB: _Lib | None = None


def lib() -> _Lib:
    """Load (once) and return the suite library."""
    global _LIB
    if _LIB is None:
        _LIB = _Lib(Path(os.environ.get("TLHIP_LIB", LIB_PATH)))
    return _LIB


def dbl() -> ctypes.c_double:
    return ctypes.c_double(0.0)


def addr(obj) -> int:
    return ctypes.addressof(obj)
