"""ctypes loader for the CPU oracle (oracle/liboracle.so) and, when built,
the reference omp3 kernels compiled from source (oracle/_ref/libomp3_*.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use it, and only as the checker, never as the thing measured.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

ORACLE_DIR = Path(__file__).resolve().parent
ROOT = ORACLE_DIR.parent
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


def run_problem(settings, states, steps=None, inject=None, rel_tol=0.0, fields=()):
    """The oracle driver on a parsed deck: per-step reports + final summary
    (+ the named fields, by oracle field id, after the last step)."""
    from tealeaf_amd import deck as tdeck
    O = load_oracle()
    cs, cst = tdeck.c_settings(settings), tdeck.c_states(states)
    P = O.or_problem_create(ctypes.addressof(cs), ctypes.addressof(cst), len(states))
    if not P:
        raise RuntimeError("oracle problem creation failed")
    try:
        if rel_tol:
            O.or_problem_set_rel_tol(P, float(rel_tol))
        if inject is not None:
            import numpy as np
            dens, en = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in inject)
            if not dens.size == en.size == O.or_problem_field_len(P):
                raise ValueError("injected fields have the wrong size")
            O.or_problem_inject(P, dens.ctypes.data, en.ctypes.data)
        reps = []
        for tt in range(settings.end_step if steps is None else steps):
            rep = tdeck.CStepReport()
            rc = O.or_problem_step(P, tt, ctypes.addressof(rep))
            if rc:
                raise RuntimeError(f"oracle step {tt} failed ({rc})")
            reps.append(rep.as_dict())
        out = (ctypes.c_double * 4)()
        O.or_problem_field_summary(P, ctypes.addressof(out))
        got = {}
        if fields:
            import numpy as np
            n = O.or_problem_field_len(P)
            for name, fid in fields:
                a = np.empty(n)
                O.or_problem_get_field(P, fid, a.ctypes.data)
                got[name] = a
        return {"steps": reps, "summary": list(out), "settings": settings, "fields": got}
    finally:
        O.or_problem_destroy(P)
