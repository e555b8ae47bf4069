"""Kernel-level cases shared by the oracle tests, the GPU parity tests and the
golden-fixture generator.

Each case names one reference entry point and lists its arguments in the
order of the omp3 binding (2d/c_kernels/omp3/c_kernels.h,
3d/c_kernels/omp3/c_kernels.h), which the oracle (oracle/oracle.h) and the
suite's C ABI (include/tealeaf_hip.h) share.  The fused Chebyshev / PPCG
sweeps of the suite take an extra output buffer (u_out / sd_out), which the
GPU runner supplies and maps back onto the in-place field.

Runners:
  run_host(dll, prefix, case)  -- oracle or reference omp3 (host arrays)
  run_gpu(case)                -- libtealeaf_hip.so (device arrays)
Both return {field name: array, scalar name: float}.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

# 2D / 3D reference State layouts (2d/settings.h:133-144, 3d/settings.h:131-145),
# used only to call the reference omp3 set_chunk_state directly.


class RefState2(ctypes.Structure):
    _fields_ = [("defined", ctypes.c_bool), ("density", ctypes.c_double),
                ("energy", ctypes.c_double), ("x_min", ctypes.c_double),
                ("y_min", ctypes.c_double), ("x_max", ctypes.c_double),
                ("y_max", ctypes.c_double), ("radius", ctypes.c_double),
                ("geometry", ctypes.c_int)]


class RefState3(ctypes.Structure):
    _fields_ = [("defined", ctypes.c_bool), ("density", ctypes.c_double),
                ("energy", ctypes.c_double), ("x_min", ctypes.c_double),
                ("y_min", ctypes.c_double), ("z_min", ctypes.c_double),
                ("x_max", ctypes.c_double), ("y_max", ctypes.c_double),
                ("z_max", ctypes.c_double), ("radius", ctypes.c_double),
                ("geometry", ctypes.c_int)]


# Argument tokens
def F(name):            # field (host array or device buffer)
    return ("F", name)


def D(v):               # double scalar in
    return ("D", float(v))


def I(v):               # int scalar in
    return ("I", int(v))


def B(v):               # bool in
    return ("B", bool(v))


def O(name):            # double scalar out (host)
    return ("O", name)


def H(arr):             # host array (neighbours, flags, buffers, states)
    return ("H", arr)


@dataclass
class Case:
    kernel: str                 # reference name, e.g. "cg_calc_w"
    dims: int
    x: int
    y: int
    z: int
    hd: int
    fields: dict                # name -> np.ndarray (inputs, also initial outputs)
    args: list                  # tokens, omp3 order (sizes included)
    outputs: list               # field names compared
    scalars: dict = field(default_factory=dict)   # out scalar name -> initial value
    gpu_name: str = ""          # entry point in the suite
    gpu_args: list = None       # tokens for the suite if they differ
    gpu_alias: dict = field(default_factory=dict)  # suite out buffer -> field name
    ref_args: list = None       # tokens for the reference if they differ
    tag: str = ""


def _ctype(tok, arrays, outs, keep):
    kind, v = tok
    if kind == "F":
        return ctypes.c_void_p(arrays[v].ctypes.data)
    if kind == "D":
        return ctypes.c_double(v)
    if kind == "I":
        return ctypes.c_int(v)
    if kind == "B":
        return ctypes.c_bool(v)
    if kind == "O":
        return ctypes.c_void_p(ctypes.addressof(outs[v]))
    if kind == "H":
        keep.append(v)
        if isinstance(v, np.ndarray):
            return ctypes.c_void_p(v.ctypes.data)
        return ctypes.c_void_p(ctypes.addressof(v))
    raise ValueError(tok)


def run_host(dll, prefix: str, case: Case, ref: bool = False) -> dict:
    arrays = {k: v.copy() for k, v in case.fields.items()}
    outs = {k: ctypes.c_double(v) for k, v in case.scalars.items()}
    keep = []
    toks = case.ref_args if (ref and case.ref_args is not None) else case.args
    cargs = [_ctype(t, arrays, outs, keep) for t in toks]
    fn = getattr(dll, prefix + case.kernel)
    fn.restype = ctypes.c_int
    fn(*cargs)
    res = {k: arrays[k] for k in case.outputs}
    res.update({k: o.value for k, o in outs.items()})
    return res


def run_gpu(case: Case) -> dict:
    from tealeaf_amd.abi import lib
    from tealeaf_amd.device import DeviceArray
    L = lib()
    toks = case.gpu_args if case.gpu_args is not None else case.args
    bufs = {k: DeviceArray.from_host(v) for k, v in case.fields.items()}
    for tok in toks:   # extra device outputs (u_out, sd_out) start as copies
        if tok[0] == "F" and tok[1] not in bufs:
            src = case.gpu_alias[tok[1]]
            bufs[tok[1]] = DeviceArray.from_host(case.fields[src])
    outs = {k: ctypes.c_double(v) for k, v in case.scalars.items()}
    keep = []
    cargs = []
    for kind, v in toks:
        if kind == "F":
            cargs.append(bufs[v].ptr)
        elif kind == "O":
            cargs.append(ctypes.addressof(outs[v]))
        elif kind == "H":
            keep.append(v)
            cargs.append(v.ctypes.data if isinstance(v, np.ndarray) else ctypes.addressof(v))
        else:
            cargs.append(v)
    name = case.gpu_name or ("tlhip%dd_%s" % (case.dims, case.kernel))
    L.call(name, *cargs)
    L.tlhip_synchronize()
    res = {}
    for k in case.outputs:
        if k == "buffer" and getattr(case, "gpu_host_buffer", False):
            res[k] = keep[0].copy()     # pack writes the caller's host buffer
            continue
        src = next((b for b, f in case.gpu_alias.items() if f == k), k)
        res[k] = bufs[src].download()
    res.update({k: o.value for k, o in outs.items()})
    return res


# ---------------------------------------------------------------- generators

def _rng(seed):
    return np.random.default_rng(seed)


def _grid(dims, nx, ny, nz, hd):
    x, y = nx + 2 * hd, ny + 2 * hd
    z = nz + 2 * hd if dims == 3 else 1
    return x, y, z, x * y * z


def _sz(dims, x, y, z):
    return [I(x), I(y)] + ([I(z)] if dims == 3 else [])


def make_cases(dims: int, nx: int, ny: int, nz: int = 1, hd: int = 2, seed: int = 0,
               coefficient: int = 1) -> list[Case]:
    """One case per reference kernel on a seeded random chunk."""
    x, y, z, n = _grid(dims, nx, ny, nz, hd)
    g = _rng(seed)

    def f(lo=-1.0, hi=1.0):
        return g.uniform(lo, hi, n)

    density = np.exp(g.uniform(np.log(0.1), np.log(100.0), n))
    energy = g.uniform(0.1, 25.0, n)
    kx, ky, kz = f(0.05, 40.0), f(0.05, 40.0), f(0.05, 40.0)
    base = dict(density=density, energy=energy, energy0=energy.copy(),
                u=f(), u0=f(), p=f(), r=f(), w=f(), sd=f(), kx=kx, ky=ky, kz=kz,
                volume=f(0.5, 2.0))
    S = _sz(dims, x, y, z)
    Sh = S + [I(hd)]
    d3 = dims == 3
    k3 = [F("kx"), F("ky")] + ([F("kz")] if d3 else [])
    rr = [D(0.3), D(0.7)] + ([D(0.45)] if d3 else [])
    nm = (lambda a, b: b) if d3 else (lambda a, b: a)   # 2D name, 3D name
    cases = []

    def add(kernel, flds, args, outputs, scalars=None, **kw):
        cases.append(Case(kernel, dims, x, y, z, hd, {k: base[k] for k in flds}, args,
                          outputs, scalars or {}, **kw))

    add(nm("cg_init", "cg_solver_init"), ["density", "energy", "u", "p", "r", "w"] + ["kx", "ky"] + (["kz"] if d3 else []),
        Sh + [I(coefficient)] + rr + [O("rro"), F("density"), F("energy"), F("u"), F("p"),
                                       F("r"), F("w")] + k3,
        ["u", "p", "r", "w", "kx", "ky"] + (["kz"] if d3 else []), {"rro": 0.25})
    add(nm("cg_calc_w", "cg_solver_calc_w"), ["p", "w", "kx", "ky"] + (["kz"] if d3 else []),
        Sh + [O("pw"), F("p"), F("w")] + k3, ["w"], {"pw": 0.5})
    add(nm("cg_calc_ur", "cg_solver_calc_ur"), ["u", "p", "r", "w"],
        Sh + [D(0.37), O("rrn"), F("u"), F("p"), F("r"), F("w")], ["u", "r"], {"rrn": 0.125})
    add(nm("cg_calc_p", "cg_solver_calc_p"), ["p", "r"], Sh + [D(0.81), F("p"), F("r")], ["p"])
    add(nm("cheby_init", "cheby_solver_init"), ["u", "u0", "p", "r", "w"] + ["kx", "ky"] + (["kz"] if d3 else []),
        Sh + [D(1.7), F("u"), F("u0"), F("p"), F("r"), F("w")] + k3, ["u", "p", "r", "w"])
    it_args = Sh + [D(0.93), D(0.41)]
    add(nm("cheby_iterate", "cheby_solver_iterate"), ["u", "u0", "p", "r", "w"] + ["kx", "ky"] + (["kz"] if d3 else []),
        it_args + [F("u"), F("u0"), F("p"), F("r"), F("w")] + k3, ["u", "p", "r", "w"],
        gpu_args=it_args + [F("u"), F("u_out"), F("u0"), F("p"), F("r"), F("w")] + k3,
        gpu_alias={"u_out": "u"})
    add(nm("jacobi_init", "jacobi_solver_init"), ["density", "energy", "u0", "u"] + ["kx", "ky"] + (["kz"] if d3 else []),
        Sh + [I(coefficient)] + rr + [F("density"), F("energy"), F("u0"), F("u")] + k3,
        ["u0", "u", "kx", "ky"] + (["kz"] if d3 else []))
    add(nm("jacobi_iterate", "jacobi_solver_iterate"), ["u0", "u", "r"] + ["kx", "ky"] + (["kz"] if d3 else []),
        Sh + [O("error")] + k3 + [F("u0"), F("u"), F("r")], ["u", "r"], {"error": 99.0})
    add("ppcg_init", ["r", "sd"], Sh + [D(2.3), F("r"), F("sd")], ["sd"])
    pi_args = Sh + [D(0.66), D(0.29), F("u"), F("r")] + k3
    add("ppcg_inner_iteration", ["u", "r", "sd"] + ["kx", "ky"] + (["kz"] if d3 else []),
        pi_args + [F("sd")], ["u", "r", "sd"],
        gpu_args=pi_args + [F("sd"), F("sd_out")], gpu_alias={"sd_out": "sd"})
    add(nm("copy_u", "solver_copy_u"), ["u0", "u"], Sh + [F("u0"), F("u")], ["u0"])
    add("calculate_residual", ["u", "u0", "r"] + ["kx", "ky"] + (["kz"] if d3 else []),
        Sh + [F("u"), F("u0"), F("r")] + k3, ["r"])
    add("calculate_2norm", ["r"], Sh + [F("r"), O("norm")], [], {"norm": 1.5})
    add(nm("finalise", "solver_finalise"), ["energy", "density", "u"],
        Sh + [F("energy"), F("density"), F("u")], ["energy"])
    add("store_energy", ["energy0", "energy"], S + [F("energy0"), F("energy")], ["energy"])
    add("field_summary", ["volume", "density", "energy0", "u"],
        Sh + [F("volume"), F("density"), F("energy0"), F("u"), O("vol"), O("mass"), O("ie"),
              O("temp")], [], {"vol": 1.0, "mass": 2.0, "ie": 3.0, "temp": 4.0})

    nfaces = 6 if d3 else 4
    for depth in (1, hd):
        nb = np.full(nfaces, -1, dtype=np.int32)
        if depth == 1:
            nb[1] = 7            # one internal face: left untouched
        flags = np.array([1, 1, 0, 1, 1, 1], dtype=np.bool_)
        add("local_halos", ["density", "energy0", "energy", "u", "p", "sd"],
            S + [I(depth), I(hd), H(nb), H(flags), F("density"), F("energy0"), F("energy"),
                 F("u"), F("p"), F("sd")],
            ["density", "energy0", "energy", "u", "p", "sd"], tag=f"depth{depth}")
    xi, yi, zi = nx, ny, nz
    for face in range(nfaces):
        for pack in (True, False):
            depth = min(2, hd) if face % 2 else 1    # the reference needs depth <= halo_depth
            if face < 2:
                blen = depth * yi * (zi if d3 else 1)
            elif face < 4:
                blen = depth * xi * (zi if d3 else 1)
            else:
                blen = depth * xi * yi
            buf = g.uniform(-1, 1, blen)
            c = Case("pack_or_unpack", dims, x, y, z, hd, {"u": base["u"], "buffer": buf},
                     S + [I(depth), I(hd), I(face), B(pack), F("u"), F("buffer")],
                     ["buffer"] if pack else ["u"], tag=f"face{face}_{'pack' if pack else 'unpack'}")
            # the suite's buffer is host memory
            c.gpu_args = S + [I(depth), I(hd), I(face), B(pack), F("u"), H(buf.copy())]
            c.gpu_host_buffer = True
            cases.append(c)
    return cases


def state_case(dims: int, nx: int, ny: int, nz: int = 1, hd: int = 2, seed: int = 3):
    """set_chunk_state with every geometry, on the vertices set_chunk_data makes."""
    from tealeaf_amd.deck import CState
    x, y, z, n = _grid(dims, nx, ny, nz, hd)
    g = _rng(seed)
    dx, dy, dz = 10.0 / nx, 10.0 / ny, 10.0 / max(nz, 1)
    vx = -hd * dx + dx * np.arange(x + 1)
    vy = -hd * dy + dy * np.arange(y + 1)
    vz = -hd * dz + dz * np.arange(z + 1)
    cx = 0.5 * (vx[:-1] + vx[1:])
    cy = 0.5 * (vy[:-1] + vy[1:])
    cz = 0.5 * (vz[:-1] + vz[1:])
    sts = [dict(density=100.0, energy=1e-4, geometry=0),
           dict(density=0.1, energy=25.0, geometry=0, x_min=1.0 + dx / 100, x_max=6.0 - dx / 100,
                y_min=2.0 + dy / 100, y_max=8.0 - dy / 100, z_min=0.5, z_max=7.5),
           dict(density=5.0, energy=3.0, geometry=1, x_min=cx[x // 2], y_min=cy[y // 2],
                z_min=cz[z // 2], radius=2.5),
           dict(density=7.0, energy=9.0, geometry=2, x_min=vx[hd + 1], y_min=vy[hd + 2],
                z_min=vz[min(hd + 1, z)])]
    ours = (CState * len(sts))()
    Ref = RefState3 if dims == 3 else RefState2
    refs = (Ref * len(sts))()
    for i, s in enumerate(sts):
        for k, v in s.items():
            setattr(ours[i], k, v)
            if hasattr(refs[i], k):
                setattr(refs[i], k, v)
        ours[i].defined = refs[i].defined = 1
    flds = dict(vertex_x=vx, vertex_y=vy, vertex_z=vz, cell_x=cx, cell_y=cy, cell_z=cz,
                density=g.uniform(0, 1, n), energy0=g.uniform(0, 1, n), u=g.uniform(0, 1, n))
    S = _sz(dims, x, y, z)
    geo = [F("vertex_x"), F("vertex_y")] + ([F("vertex_z")] if dims == 3 else []) + \
          [F("cell_x"), F("cell_y")] + ([F("cell_z")] if dims == 3 else [])
    tail = [F("density"), F("energy0"), F("u"), I(len(sts))]
    c = Case("set_chunk_state", dims, x, y, z, hd, flds, S + geo + tail + [H(ours)],
             ["density", "energy0", "u"])
    c.ref_args = S + geo + tail + [H(refs)]
    return c


def data_case(dims: int, nx: int, ny: int, nz: int = 1, hd: int = 2):
    """set_chunk_data on a chunk whose mesh starts at (3, 5, 2) cells in."""
    x, y, z, n = _grid(dims, nx, ny, nz, hd)
    m = x + y + z + 3
    flds = {k: np.zeros(n if k in ("volume", "x_area", "y_area", "z_area") else m)
            for k in ("cell_x", "cell_y", "cell_z", "vertex_x", "vertex_y", "vertex_z",
                      "volume", "x_area", "y_area", "z_area")}
    if dims == 2:
        args = [I(x), I(y), I(hd), I(3), I(5), D(0.0), D(-1.0), D(10.0 / nx), D(10.0 / ny),
                F("cell_x"), F("cell_y"), F("vertex_x"), F("vertex_y"), F("volume"),
                F("x_area"), F("y_area")]
        outs = ["cell_x", "cell_y", "vertex_x", "vertex_y", "volume", "x_area", "y_area"]
    else:
        args = [I(x), I(y), I(z), I(hd), I(3), I(5), I(2), D(0.0), D(-1.0), D(2.0),
                D(10.0 / nx), D(10.0 / ny), D(10.0 / nz), F("cell_x"), F("cell_y"),
                F("cell_z"), F("vertex_x"), F("vertex_y"), F("vertex_z"), F("volume"),
                F("x_area"), F("y_area"), F("z_area")]
        outs = list(flds)
    return Case("set_chunk_data", dims, x, y, z, hd, flds, args, outs)


def case_id(c: Case) -> str:
    return f"{c.dims}d-{c.kernel}{('-' + c.tag) if c.tag else ''}"


# Fixtures written by tests/golden/make_golden.py from the reference omp3
# kernels: for every golden case, its inputs ("<id>:in:<field>") and the
# reference outputs ("<id>:<field or scalar>").
from pathlib import Path as _Path

GOLDEN_FILE = _Path(__file__).resolve().parent / "golden" / "omp3_kernels.npz"


def golden_cases() -> list[Case]:
    return (make_cases(2, 6, 5, seed=101) + make_cases(3, 4, 3, 3, seed=102)
            + [state_case(2, 8, 6), state_case(3, 5, 4, 4)])


def with_golden_inputs(case: Case, data) -> Case:
    """Replace a case's inputs with the ones stored in the fixture."""
    cid = case_id(case)
    for k in list(case.fields):
        case.fields[k] = data[f"{cid}:in:{k}"].copy()
    return case
