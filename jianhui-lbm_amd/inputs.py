"""The reference's positional ``INPUT.dat`` parameter files.

Both solvers read their parameter file line by line: on a data line the leading
whitespace-separated tokens are the values and the rest of the line is a comment; header
lines are skipped whole.  The line maps below are the ones the reference's parsers consume
(SP:255-308 for the single-phase file, CG:371-451 for the colour-gradient file).  Grid sizes
are written as counts (the reference subtracts one on read, SP:313).

``write_*`` produce files the reference accepts; ``read_*`` accept the reference's own
sample files.  Each data line's comment is kept short because the reference reads the rest of
a line into a 128-byte buffer.
"""
from __future__ import annotations

from dataclasses import dataclass, field, fields
from typing import List, Optional, Tuple

# (name, number of values, type) per line; None = header line.
SP_LINES: List[Optional[Tuple[str, int, type]]] = [
    None,                              # 1 banner
    ("filename", 1, str),              # 2 geometry file
    ("n_xyz", 3, int),                 # 3 nx ny nz (counts)
    ("n_max", 1, int),                 # 4 last time step index (runs n_max+1 steps)
    ("reso", 1, float),                # 5 voxel size in micrometres
    ("in_BC", 1, int),                 # 6 pressure/velocity boundaries on/off
    ("g", 3, float),                   # 7 body force
    ("pre_x", 4, float),               # 8 pressure BC x: flag p flag p
    ("pre_y", 4, float),               # 9
    ("pre_z", 4, float),               # 10
    ("vel_x", 4, float),               # 11 velocity BC x
    ("vel_y", 4, float),               # 12
    ("vel_z", 4, float),               # 13
    ("in_vis", 1, float),              # 14 kinematic viscosity
    ("ini_v", 3, float),               # 15 initial velocity
    None,                              # 16 header
    ("wr_per", 1, int),                # 17 write permeability file
    ("PerDir", 1, int),                # 18 permeability direction 1..3
    ("freRe", 1, int),                 # 19 diagnostics interval
    ("Out_Mode", 1, int),              # 20 1 = stop at steady u_ave
    ("freVe", 1, int),                 # 21 velocity VTK interval (-1 off)
    ("freDe", 1, int),                 # 22 density VTK interval (-1 off)
    None,                              # 23 header
    ("lattice", 3, float),             # 24 lattice_v dx dt
    ("outputfile", 1, str),            # 25 output prefix
    ("Sub_BC", 1, int),                # 26 BC flavour
    ("loc_perm", 1, int),              # 27 local permeability on/off
    ("par_per", 3, int),               # 28 partial permeability box on/off per axis
    ("per_x", 2, int),                 # 29 box x: xp xn
    ("per_y", 2, int),                 # 30 box y: yp yn
    ("per_z", 2, int),                 # 31 box z: zp zn
    None,                              # 32 header
    ("fre_backup", 1, int),            # 33 checkpoint interval (-1 none, 0 end only)
    ("mode_backup_ini", 1, int),       # 34 restart from checkpoint label
    ("vel_sol", 1, int),               # 35 velocity export for solute
    None,                              # 36 header
    ("gperm", 1, int),                 # 37 group permeability interval
    ("n_gperm", 4, int),               # 38
    ("size_gperm", 4, int),            # 39
    ("c_gperm1", 3, int),              # 40
    ("c_gperm2", 3, int),              # 41
    ("c_gperm3", 3, int),              # 42
    ("c_gperm4", 3, int),              # 43
    None,                              # 44 header
    ("decbin", 1, int),                # 45 geometry format 0 ASCII / 1 binary
    None,                              # 46 header
    ("nnf_module", 1, int),            # 47 non-Newtonian on/off
    ("nnf_mn", 2, float),              # 48
]

CG_LINES: List[Optional[Tuple[str, int, type]]] = [
    None,                              # 1 banner
    ("filename", 1, str),              # 2 geometry
    ("filenamepsi", 1, str),           # 3 phase file
    ("n_xyz", 3, int),                 # 4
    ("n_max", 1, int),                 # 5
    ("reso", 1, float),                # 6
    ("in_BC", 1, int),                 # 7
    ("in_psi_BC", 1, int),             # 8
    ("g", 3, float),                   # 9
    ("pre_x", 4, float),               # 10
    ("pre_y", 4, float),               # 11
    ("pre_z", 4, float),               # 12
    ("vel_x", 4, float),               # 13
    ("vel_y", 4, float),               # 14
    ("vel_z", 4, float),               # 15
    ("psi_x", 2, int),                 # 16
    ("psi_y", 2, int),                 # 17
    ("psi_z", 2, int),                 # 18
    ("niu_l", 1, float),               # 19 viscosity of component A (psi = 1)
    ("niu_g", 1, float),               # 20 viscosity of component B (psi = -1)
    ("cos_theta", 1, float),           # 21 contact angle parameter (psi_solid)
    ("CapA", 1, float),                # 22 surface tension parameter
    ("ini_v", 3, float),               # 23
    ("Permeability", 1, float),        # 24 single-phase permeability (mD) for k_r
    None,                              # 25 header
    ("wr_per", 1, int),                # 26
    ("PerDir", 1, int),                # 27
    ("freRe", 1, int),                 # 28
    ("Out_Mode", 1, int),              # 29
    ("freVe", 1, int),                 # 30
    ("freDe", 1, int),                 # 31
    ("frePsi", 1, int),                # 32
    None,                              # 33 header
    ("lattice", 3, float),             # 34
    ("outputfile", 1, str),            # 35
    ("Sub_BC", 1, int),                # 36
    ("stab", 2, int),                  # 37 stabiliser on/off, switch-on step
    ("ini_Sat", 1, float),             # 38 <0: read the phase file
    ("ini_buf", 1, int),               # 39
    ("rel_perm_psi", 2, float),        # 40 psi thresholds for the phase fluxes
    ("par_per", 3, int),               # 41
    ("per_x", 2, int),                 # 42
    ("per_y", 2, int),                 # 43
    ("per_z", 2, int),                 # 44
    None,                              # 45 header
    ("fre_backup", 1, int),            # 46
    ("mode_backup_ini", 1, int),       # 47
    None,                              # 48 header
    ("decbin", 1, int),                # 49
    None,                              # 50 header
    ("pressure_change", 2, int),       # 51
    ("pre_chan1", 4, float),           # 52
    ("pre_chan2", 4, float),           # 53
    ("pre_chan3", 4, float),           # 54
    ("pre_chan4", 4, float),           # 55
    ("pre_chan5", 4, float),           # 56
    ("bodyforce_apply", 1, int),       # 57
    None,                              # 58 header
    ("pressure_change2", 3, int),      # 59
    ("pre_chan_s", 5, float),          # 60
    ("sat_cri_d", 1, float),           # 61
    None,                              # 62 header
    ("least_square", 2, int),          # 63
    None,                              # 64 header
    ("rel_perm_id", 2, int),           # 65
    ("rel_perm_ift", 2, float),        # 66
    ("input_dynamic", 2, int),         # 67
]


@dataclass
class SPInput:
    filename: str = "geometry.bin"
    n_xyz: Tuple[int, int, int] = (64, 64, 64)
    n_max: int = 1999
    reso: float = 1.0
    in_BC: int = 0
    g: Tuple[float, float, float] = (1e-6, 0.0, 0.0)
    pre_x: Tuple[float, ...] = (0, 1.0, 0, 1.0)
    pre_y: Tuple[float, ...] = (0, 1.0, 0, 1.0)
    pre_z: Tuple[float, ...] = (0, 1.0, 0, 1.0)
    vel_x: Tuple[float, ...] = (0, 0.0, 0, 0.0)
    vel_y: Tuple[float, ...] = (0, 0.0, 0, 0.0)
    vel_z: Tuple[float, ...] = (0, 0.0, 0, 0.0)
    in_vis: float = 1.0 / 6.0
    ini_v: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    wr_per: int = 1
    PerDir: int = 1
    freRe: int = 100
    Out_Mode: int = 0
    freVe: int = -1
    freDe: int = -1
    lattice: Tuple[float, float, float] = (0, 1.0, 1.0)
    outputfile: str = "./"
    Sub_BC: int = 2
    loc_perm: int = 0
    par_per: Tuple[int, int, int] = (0, 0, 0)
    per_x: Tuple[int, int] = (0, 0)
    per_y: Tuple[int, int] = (0, 0)
    per_z: Tuple[int, int] = (0, 0)
    fre_backup: int = -1
    mode_backup_ini: int = 0
    vel_sol: int = -1
    gperm: int = 0
    n_gperm: Tuple[int, ...] = (1, 1, 1, 1)
    size_gperm: Tuple[int, ...] = (8, 8, 8, 8)
    c_gperm1: Tuple[int, ...] = (0, 0, 0)
    c_gperm2: Tuple[int, ...] = (0, 0, 0)
    c_gperm3: Tuple[int, ...] = (0, 0, 0)
    c_gperm4: Tuple[int, ...] = (0, 0, 0)
    decbin: int = 1
    nnf_module: int = 0
    nnf_mn: Tuple[float, float] = (0.001, 0.25)


@dataclass
class CGInput:
    filename: str = "geometry.bin"
    filenamepsi: str = "phase.dat"
    n_xyz: Tuple[int, int, int] = (64, 64, 64)
    n_max: int = 1999
    reso: float = 1.0
    in_BC: int = 0
    in_psi_BC: int = 0
    g: Tuple[float, float, float] = (1e-5, 0.0, 0.0)
    pre_x: Tuple[float, ...] = (0, 1.0, 0, 1.0)
    pre_y: Tuple[float, ...] = (0, 1.0, 0, 1.0)
    pre_z: Tuple[float, ...] = (0, 1.0, 0, 1.0)
    vel_x: Tuple[float, ...] = (0, 0.0, 0, 0.0)
    vel_y: Tuple[float, ...] = (0, 0.0, 0, 0.0)
    vel_z: Tuple[float, ...] = (0, 0.0, 0, 0.0)
    psi_x: Tuple[int, int] = (0, 0)
    psi_y: Tuple[int, int] = (0, 0)
    psi_z: Tuple[int, int] = (0, 0)
    niu_l: float = 1.0 / 6.0
    niu_g: float = 1.0 / 6.0
    cos_theta: float = 0.0
    CapA: float = 0.01
    ini_v: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    Permeability: float = 1.0
    wr_per: int = 1
    PerDir: int = 1
    freRe: int = 100
    Out_Mode: int = 0
    freVe: int = -1
    freDe: int = -1
    frePsi: int = -1
    lattice: Tuple[float, float, float] = (0, 1.0, 1.0)
    outputfile: str = "./"
    Sub_BC: int = 2
    stab: Tuple[int, int] = (0, 0)
    ini_Sat: float = -1.0
    ini_buf: int = 0
    rel_perm_psi: Tuple[float, float] = (0.99, 0.92)
    par_per: Tuple[int, int, int] = (0, 0, 0)
    per_x: Tuple[int, int] = (0, 0)
    per_y: Tuple[int, int] = (0, 0)
    per_z: Tuple[int, int] = (0, 0)
    fre_backup: int = -1
    mode_backup_ini: int = 0
    decbin: int = 1
    pressure_change: Tuple[int, int] = (0, 2)
    pre_chan1: Tuple[float, ...] = (-1, 0.0, 0.0, 0.0)
    pre_chan2: Tuple[float, ...] = (-1, 0.0, 0.0, 0.0)
    pre_chan3: Tuple[float, ...] = (-1, 0.0, 0.0, 0.0)
    pre_chan4: Tuple[float, ...] = (-1, 0.0, 0.0, 0.0)
    pre_chan5: Tuple[float, ...] = (-1, 0.0, 0.0, 0.0)
    bodyforce_apply: int = 0
    pressure_change2: Tuple[int, int, int] = (0, 1, 2000)
    pre_chan_s: Tuple[float, ...] = (1.0, 1.0, 0.0, 1, 10)
    sat_cri_d: float = 1e-6
    least_square: Tuple[int, int] = (0, 10)
    rel_perm_id: Tuple[int, int] = (0, 5)
    rel_perm_ift: Tuple[float, float] = (0, 0.0)
    input_dynamic: Tuple[int, int] = (0, 1000)


def _fmt(v) -> str:
    # Integral values are written without a decimal point: several slots are C++ ints in the
    # reference's parser, and `fin >> int` would stop at a '.' and shift every later value.
    if isinstance(v, float):
        return str(int(v)) if v.is_integer() and abs(v) < 1e15 else repr(v)
    return str(v)


def _write(obj, lines, path: str, banner: str) -> None:
    out = []
    for spec in lines:
        if spec is None:
            out.append(banner if not out else "=====")
            continue
        name, nvals, _ = spec
        v = getattr(obj, name)
        vals = list(v) if isinstance(v, (tuple, list)) else [v]
        assert len(vals) == nvals, (name, vals)
        out.append(" ".join(_fmt(x) for x in vals) + "    :" + name)
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def _read(cls, lines, path: str):
    obj = cls()
    with open(path) as fh:
        raw = fh.read().split("\n")
    for lineno, spec in enumerate(lines):
        if spec is None:
            continue
        name, nvals, typ = spec
        toks = raw[lineno].split()
        if len(toks) < nvals:
            raise ValueError(f"{path}:{lineno + 1}: expected {nvals} value(s) for {name}")
        vals = []
        for t in toks[:nvals]:
            if typ is int:
                vals.append(int(float(t)))
            elif typ is float:
                vals.append(float(t))
            else:
                vals.append(t)
        setattr(obj, name, tuple(vals) if nvals > 1 else vals[0])
    return obj


def write_sp_input(inp: SPInput, path: str) -> None:
    _write(inp, SP_LINES, path, "=== SINGLE PHASE D3Q19 MRT INPUT ===")


def read_sp_input(path: str) -> SPInput:
    return _read(SPInput, SP_LINES, path)


def write_cg_input(inp: CGInput, path: str) -> None:
    _write(inp, CG_LINES, path, "=== COLOUR GRADIENT D3Q19 MRT INPUT ===")


def read_cg_input(path: str) -> CGInput:
    return _read(CGInput, CG_LINES, path)
