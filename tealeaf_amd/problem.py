"""The native timestep driver (tlhip_problem_*) from Python."""
from __future__ import annotations

import ctypes

import numpy as np

from . import deck as tdeck
from .abi import lib

FIELD_IDS = {"density": 0, "energy0": 1, "energy": 2, "u": 3, "u0": 4, "p": 5, "r": 6,
             "w": 7, "kx": 8, "ky": 9, "kz": 10, "sd": 11}


class Problem:
    """One chunk on the GPU, initialised from a deck, stepped natively.

    step(tt) returns the reference's per-timestep report (iteration counts as
    printed by 2d/drivers/*_driver.c, error, eigenvalue estimate, and the
    field summary when tt % summary_frequency == 0).
    """

    def __init__(self, settings: tdeck.Settings, states: list, device: int = 0):
        L = lib()
        L.tlhip_init(device)
        self.settings = settings
        self._cs = tdeck.c_settings(settings)
        self._cst = tdeck.c_states(states)
        h = ctypes.c_void_p()
        L.tlhip_problem_create(ctypes.addressof(self._cs), ctypes.addressof(self._cst),
                               len(states), ctypes.addressof(h))
        self.handle = h.value
        n = ctypes.c_long()
        L.tlhip_problem_field_len(self.handle, ctypes.addressof(n))
        self.n = n.value

    @classmethod
    def from_deck(cls, text: str, dims: int | None = None, device: int = 0) -> "Problem":
        s, states = tdeck.parse_deck(text, dims or (3 if "z_cells" in text else 2))
        return cls(s, states, device)

    def inject(self, density: np.ndarray, energy0: np.ndarray) -> None:
        d = np.ascontiguousarray(density, dtype=np.float64).ravel()
        e = np.ascontiguousarray(energy0, dtype=np.float64).ravel()
        if d.size != self.n or e.size != self.n:
            raise ValueError(f"expected {self.n} cells, got {d.size} / {e.size}")
        lib().tlhip_problem_inject(self.handle, d.ctypes.data, e.ctypes.data)

    def set_solve(self, precision: int = 0, rel_residual: float = 0.0) -> None:
        """precision 0: FP64 (the reference's arithmetic); 1: FP32 CG/PPCG
        sweeps inside an FP64 residual-replacement loop (2D).  rel_residual
        > 0 stops on ||u0 - A u|| <= rel_residual ||u0|| (FP64-confirmed)
        instead of the deck's eps."""
        lib().tlhip_problem_set_solve(self.handle, int(precision), float(rel_residual))

    def adjoint(self, steps: int, target: np.ndarray):
        """Run `steps` timesteps and return (loss, dloss/dln(density), reports)
        for loss = sum over the interior of (u_N - target)^2."""
        t = np.ascontiguousarray(target, dtype=np.float64).ravel()
        if t.size != self.n:
            raise ValueError(f"expected {self.n} cells, got {t.size}")
        grad = np.zeros(self.n)
        loss = ctypes.c_double()
        reps = (tdeck.CStepReport * steps)()
        lib().tlhip_problem_adjoint(self.handle, int(steps), t.ctypes.data, ctypes.addressof(loss),
                                    grad.ctypes.data, ctypes.addressof(reps))
        return loss.value, grad, [r.as_dict() for r in reps]

    def step(self, tt: int) -> dict:
        rep = tdeck.CStepReport()
        lib().tlhip_problem_step(self.handle, int(tt), ctypes.addressof(rep))
        return rep.as_dict()

    def run(self, steps: int | None = None) -> list[dict]:
        return [self.step(tt) for tt in range(steps if steps is not None else self.settings.end_step)]

    def field_summary(self) -> list[float]:
        out = (ctypes.c_double * 4)()
        lib().tlhip_problem_field_summary(self.handle, ctypes.addressof(out))
        return list(out)

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.n)
        lib().tlhip_problem_get_field(self.handle, FIELD_IDS[name], out.ctypes.data)
        return out.reshape(self.settings.shape())

    def field_ptr(self, name: str) -> int:
        p = ctypes.c_void_p()
        lib().tlhip_problem_field_ptr(self.handle, FIELD_IDS[name], ctypes.addressof(p))
        return p.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().tlhip_problem_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
