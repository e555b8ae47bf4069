"""The tea.in problem deck: settings and material states.

Mirrors the reference parser (2d/parse_config.c:77-367, 3d/parse_config.c)
so the same deck file drives this suite, the CPU oracle and the reference
binaries identically, including its quirks:

* a key matches when the line starts with it after leading blanks, first
  match wins in the reference's key order (``epslim`` before ``eps``);
* ``x_cells``/``y_cells``/``z_cells`` are read only while still at their
  default of 10 (2d/parse_config.c:102-107);
* a value is the token starting at the first letter or digit after the key
  (so a leading sign or dot is dropped), converted like atof/atoi;
* state extents are pulled in by a hundredth of a cell
  (2d/parse_config.c:222-260).
"""
from __future__ import annotations

import ctypes
import re
from dataclasses import dataclass, field, fields
from pathlib import Path

JACOBI, CG, CHEBY, PPCG = 0, 1, 2, 3
CONDUCTIVITY, RECIP_CONDUCTIVITY = 1, 2
SOLVER_NAMES = {JACOBI: "Jacobi", CG: "CG", CHEBY: "Chebyshev", PPCG: "PPCG"}


class DeckError(ValueError):
    pass


@dataclass
class Settings:
    """Defaults from 2d/settings.h:8-42."""
    dims: int = 2
    x_cells: int = 10
    y_cells: int = 10
    z_cells: int = 10
    xmin: float = 0.0
    ymin: float = 0.0
    zmin: float = 0.0
    xmax: float = 100.0
    ymax: float = 100.0
    zmax: float = 100.0
    dt_init: float = 0.1
    end_step: int = 2147483647
    max_iters: int = 1000
    eps: float = 1.0e-15
    solver: int = CG
    coefficient: int = CONDUCTIVITY
    presteps: int = 30
    ppcg_inner_steps: int = 10
    eps_lim: float = 1e-5
    error_switch: int = 0
    check_result: int = 1
    summary_frequency: int = 10
    halo_depth: int = 2
    end_time: float = 10.0
    use_c_kernels: bool = False

    @property
    def dx(self) -> float:
        return (self.xmax - self.xmin) / float(self.x_cells)

    @property
    def dy(self) -> float:
        return (self.ymax - self.ymin) / float(self.y_cells)

    @property
    def dz(self) -> float:
        return (self.zmax - self.zmin) / float(self.z_cells)

    def shape(self) -> tuple[int, ...]:
        """Array shape including the halo, slowest axis first."""
        h = 2 * self.halo_depth
        if self.dims == 2:
            return (self.y_cells + h, self.x_cells + h)
        return (self.z_cells + h, self.y_cells + h, self.x_cells + h)


@dataclass
class State:
    density: float = 0.0
    energy: float = 0.0
    x_min: float = 0.0
    y_min: float = 0.0
    z_min: float = 0.0
    x_max: float = 0.0
    y_max: float = 0.0
    z_max: float = 0.0
    radius: float = 0.0
    geometry: int = 0
    defined: int = 0


# ---------------------------------------------------------------- C layouts

class CSettings(ctypes.Structure):
    """Layout shared by oracle/oracle.h (tl_settings) and include/tealeaf_hip.h."""
    _fields_ = [
        ("dims", ctypes.c_int), ("x_cells", ctypes.c_int), ("y_cells", ctypes.c_int),
        ("z_cells", ctypes.c_int),
        ("xmin", ctypes.c_double), ("ymin", ctypes.c_double), ("zmin", ctypes.c_double),
        ("xmax", ctypes.c_double), ("ymax", ctypes.c_double), ("zmax", ctypes.c_double),
        ("dt_init", ctypes.c_double), ("end_step", ctypes.c_int), ("max_iters", ctypes.c_int),
        ("eps", ctypes.c_double), ("solver", ctypes.c_int), ("coefficient", ctypes.c_int),
        ("presteps", ctypes.c_int), ("ppcg_inner_steps", ctypes.c_int),
        ("eps_lim", ctypes.c_double), ("error_switch", ctypes.c_int),
        ("check_result", ctypes.c_int), ("summary_frequency", ctypes.c_int),
        ("halo_depth", ctypes.c_int),
    ]


class CState(ctypes.Structure):
    _fields_ = [
        ("density", ctypes.c_double), ("energy", ctypes.c_double),
        ("x_min", ctypes.c_double), ("y_min", ctypes.c_double), ("z_min", ctypes.c_double),
        ("x_max", ctypes.c_double), ("y_max", ctypes.c_double), ("z_max", ctypes.c_double),
        ("radius", ctypes.c_double), ("geometry", ctypes.c_int), ("defined", ctypes.c_int),
    ]


class CStepReport(ctypes.Structure):
    _fields_ = [
        ("printed_cg_iters", ctypes.c_int), ("printed_inner_iters", ctypes.c_int),
        ("est_iters", ctypes.c_int), ("n_cg_steps", ctypes.c_int),
        ("n_cheby_steps", ctypes.c_int), ("n_ppcg_outer", ctypes.c_int),
        ("n_ppcg_inner", ctypes.c_int), ("n_jacobi", ctypes.c_int),
        ("error", ctypes.c_double), ("exact_error", ctypes.c_double),
        ("eigmin", ctypes.c_double), ("eigmax", ctypes.c_double),
        ("summary_done", ctypes.c_int),
        ("vol", ctypes.c_double), ("mass", ctypes.c_double), ("ie", ctypes.c_double),
        ("temp", ctypes.c_double),
        ("rel_residual", ctypes.c_double), ("n_true_residuals", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def c_settings(s: Settings) -> CSettings:
    c = CSettings()
    for name, _ in CSettings._fields_:
        setattr(c, name, getattr(s, name))
    return c


def c_states(states: list[State]):
    arr = (CState * len(states))()
    for i, st in enumerate(states):
        for f in fields(State):
            setattr(arr[i], f.name, getattr(st, f.name))
    return arr


# ---------------------------------------------------------------- parsing

def _starts_with(word: str, line: str) -> bool:
    return line.lstrip().startswith(word)


def _read_value(line: str, word: str) -> str:
    """Token after `word` starting at the first letter/digit (2d/parse_config.c:317-341)."""
    matched = 0
    n = len(word)
    for i, ch in enumerate(line):
        if matched == n:
            if ch.isalnum():
                return line[i:].split()[0]
        else:
            matched = matched + 1 if ch == word[matched] else 0
    raise DeckError(f"Failed to find a value for key '{word}'")


_NUM = re.compile(r"[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?)")


def _atof(tok: str) -> float:
    m = _NUM.match(tok)
    return float(m.group(0)) if m else 0.0


def _atoi(tok: str) -> int:
    m = re.match(r"[+-]?\d+", tok)
    return int(m.group(0)) if m else 0


def parse_deck(text: str, dims: int = 2) -> tuple[Settings, list[State]]:
    s = Settings(dims=dims)
    lines = text.splitlines()
    for line in lines:
        def get(key, attr, conv):
            if _starts_with(key, line):
                setattr(s, attr, conv(_read_value(line, key)))
                return True
            return False

        if get("initial_timestep", "dt_init", _atof): continue
        if get("end_time", "end_time", _atof): continue
        if get("end_step", "end_step", _atoi): continue
        if get("xmin", "xmin", _atof): continue
        if get("ymin", "ymin", _atof): continue
        if dims == 3 and get("zmin", "zmin", _atof): continue
        if get("xmax", "xmax", _atof): continue
        if get("ymax", "ymax", _atof): continue
        if dims == 3 and get("zmax", "zmax", _atof): continue
        if s.x_cells == 10 and get("x_cells", "x_cells", _atoi): continue
        if s.y_cells == 10 and get("y_cells", "y_cells", _atoi): continue
        if dims == 3 and s.z_cells == 10 and get("z_cells", "z_cells", _atoi): continue
        if get("summary_frequency", "summary_frequency", _atoi): continue
        if get("presteps", "presteps", _atoi): continue
        if get("ppcg_inner_steps", "ppcg_inner_steps", _atoi): continue
        if get("epslim", "eps_lim", _atof): continue
        if get("max_iters", "max_iters", _atoi): continue
        if get("eps", "eps", _atof): continue
        if _starts_with("num_chunks_per_rank", line): continue
        if get("halo_depth", "halo_depth", _atoi): continue
        if _starts_with("check_result", line):
            s.check_result = 1; continue
        if _starts_with("errswitch", line):
            s.error_switch = 1; continue
        if _starts_with("preconditioner_on", line): continue
        if _starts_with("use_fortran_kernels", line):
            s.use_c_kernels = False; continue
        if _starts_with("use_c_kernels", line):
            s.use_c_kernels = True; continue
        if _starts_with("use_jacobi", line):
            s.solver = JACOBI; continue
        if _starts_with("use_cg", line):
            s.solver = CG; continue
        if _starts_with("use_chebyshev", line):
            s.solver = CHEBY; continue
        if _starts_with("use_ppcg", line):
            s.solver = PPCG; continue
        if _starts_with("coefficient_density", line):
            s.coefficient = CONDUCTIVITY; continue
        if _starts_with("coefficient_inverse_density", line):
            s.coefficient = RECIP_CONDUCTIVITY; continue

    nstates = 0
    for line in lines:
        if _starts_with("state", line):
            nstates = max(nstates, _atoi(_read_value(line, "state")))
    states = [State() for _ in range(nstates)]
    geoms = {"rectangle": 0, "circular": 1, "point": 2} if dims == 2 else \
            {"cuboid": 0, "spherical": 1, "point": 2}
    for line in lines:
        if not _starts_with("state", line):
            continue
        num = _atoi(_read_value(line, "state"))
        st = states[num - 1]
        if st.defined:
            raise DeckError(f"State number {num} defined twice.")
        st.density = _atof(_read_value(line, "density"))
        st.energy = _atof(_read_value(line, "energy"))
        if num > 1:
            st.x_min = _atof(_read_value(line, "xmin")) + s.dx / 100.0
            st.y_min = _atof(_read_value(line, "ymin")) + s.dy / 100.0
            st.x_max = _atof(_read_value(line, "xmax")) - s.dx / 100.0
            st.y_max = _atof(_read_value(line, "ymax")) - s.dy / 100.0
            if dims == 3:
                st.z_min = _atof(_read_value(line, "zmin")) + s.dz / 100.0
                st.z_max = _atof(_read_value(line, "zmax")) - s.dz / 100.0
            geo = _read_value(line, "geometry")
            if geo not in geoms:
                raise DeckError(f"unknown geometry '{geo}' for state {num}")
            st.geometry = geoms[geo]
            if st.geometry == 1:
                st.radius = _atof(_read_value(line, "radius"))
        st.defined = 1
    if not states:
        raise DeckError("deck defines no states")
    return s, states


def load_deck(path: str | Path, dims: int | None = None) -> tuple[Settings, list[State]]:
    text = Path(path).read_text()
    if dims is None:
        dims = 3 if re.search(r"^\s*z_cells", text, re.M) else 2
    return parse_deck(text, dims)


# ---------------------------------------------------------------- standard decks

LAYERED_2D = """*tea
state 1 density=100.0 energy=0.0001
state 2 density=0.1 energy=25.0 geometry=rectangle xmin=0.0 xmax=1.0 ymin=1.0 ymax=2.0
state 3 density=0.1 energy=0.1 geometry=rectangle xmin=1.0 xmax=6.0 ymin=1.0 ymax=2.0
state 4 density=0.1 energy=0.1 geometry=rectangle xmin=5.0 xmax=6.0 ymin=1.0 ymax=8.0
state 5 density=0.1 energy=0.1 geometry=rectangle xmin=5.0 xmax=10.0 ymin=7.0 ymax=8.0
xmin = 0.0
ymin = 0.0
xmax = 10.0
ymax = 10.0
x_cells = {n}
y_cells = {n}
{solver}
use_c_kernels
check_result
eps = {eps}
max_iters = {max_iters}
initial_timestep = {dt}
end_step = {steps}
end_time = 100.0
halo_depth = 2
num_chunks_per_rank = 1
ppcg_inner_steps = {inner}
epslim = {epslim}
presteps = {presteps}
{errswitch}
*endtea
"""

CUBOID_3D = """*tea
state 1 density=0.2 energy=1.0
state 2 density=1.0 energy=2.5 geometry=cuboid xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0 zmin=0.0 zmax=2.0
xmin = 0.0
ymin = 0.0
zmin = 0.0
xmax = 10.0
ymax = 10.0
zmax = 10.0
{solver}
use_c_kernels
check_result
x_cells = {n}
y_cells = {n}
z_cells = {n}
eps = {eps}
max_iters = {max_iters}
end_time = 100.0
end_step = {steps}
{dt}
halo_depth = 2
num_chunks_per_rank = 1
ppcg_inner_steps = {inner}
epslim = 0.0001
presteps = {presteps}
{coef}
*endtea
"""

_SOLVER_KEY = {CG: "use_cg", CHEBY: "use_chebyshev", PPCG: "use_ppcg", JACOBI: "use_jacobi"}


def layered_2d(n: int, steps: int, solver: int = CG, eps: float = 1e-15,
               max_iters: int = 10000, inner: int = 10, presteps: int = 30,
               dt: float = 0.004, epslim: float | None = None) -> str:
    """The reference's layered-rectangle deck (2d/tea.in) at n x n cells.
    With `epslim` the deck selects the reference's error switch
    (tl_use_errswitch): Chebyshev/PPCG start once the CG error is below
    epslim and more than 20 CG iterations have run."""
    return LAYERED_2D.format(n=n, steps=steps, solver=_SOLVER_KEY[solver], eps=eps,
                             max_iters=max_iters, inner=inner, presteps=presteps, dt=dt,
                             epslim=0.0001 if epslim is None else epslim,
                             errswitch="" if epslim is None else "errswitch")


def cuboid_3d(n: int, steps: int, solver: int = CG, eps: float = 1e-15,
              max_iters: int = 1000, inner: int = 10, presteps: int = 20,
              coefficient: int = CONDUCTIVITY, dt: float | None = None) -> str:
    """The reference's 3D cuboid deck (3d/tea.in) at n^3 cells (its timestep
    is the settings default unless `dt` is given)."""
    coef = "coefficient_inverse_density" if coefficient == RECIP_CONDUCTIVITY else ""
    return CUBOID_3D.format(n=n, steps=steps, solver=_SOLVER_KEY[solver], eps=eps,
                            max_iters=max_iters, inner=inner, presteps=presteps, coef=coef,
                            dt="" if dt is None else f"initial_timestep = {dt}")
