"""ctypes front end of the FP64 CPU oracle (oracle/liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, by ``__graft_entry__.smoke()`` as the checker and
by ``bench.py``'s cpu_baseline leg.  The product path (the HIP library) never imports it.

Arrays follow the reference's conventions: per-node arrays have ``count + 1`` rows and row 0
is padding; populations are ``[count + 1, 19]`` in the reference's own velocity ordering
(SP ordering for ``SPOracle``, CG ordering for ``CGOracle``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc is enough; no GPU toolchain)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "liblbm_oracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_number_fluid.restype = ctypes.c_int
        L.orc_sp_error.restype = ctypes.c_double
        L.orc_sp_error_ranks.restype = ctypes.c_double
        L.orc_cg_wet_count.restype = ctypes.c_double
        L.orc_invert19.restype = ctypes.c_int
        L.orc_st_mass.restype = ctypes.c_double
        _lib = L
    return _lib


class BC(ctypes.Structure):
    """Face boundaries in the reference's input order x+, x-, y+, y-, z+, z- (SP:262-267, 280)."""
    _fields_ = [("sub_bc", ctypes.c_int), ("pre", ctypes.c_int * 6), ("p", ctypes.c_double * 6),
                ("vel", ctypes.c_int * 6), ("v", ctypes.c_double * 6), ("no_quirks", ctypes.c_int)]


def make_bc(sub_bc=0, pre=None, vel=None, no_quirks=False) -> BC:
    """``pre`` / ``vel``: dicts face -> value, faces named "xp", "xn", "yp", "yn", "zp", "zn"."""
    faces = ["xp", "xn", "yp", "yn", "zp", "zn"]
    b = BC()
    b.sub_bc = int(sub_bc)
    b.no_quirks = int(bool(no_quirks))
    for name, val in (pre or {}).items():
        b.pre[faces.index(name)] = 1
        b.p[faces.index(name)] = float(val)
    for name, val in (vel or {}).items():
        b.vel[faces.index(name)] = 1
        b.v[faces.index(name)] = float(val)
    return b


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def lattice(which: int):
    """(e[19,3], w[19], opp[19]) in the SP (0) or CG (1) ordering."""
    e = np.zeros((19, 3), np.int32)
    w = np.zeros(19)
    opp = np.zeros(19, np.int32)
    lib().orc_lattice(ctypes.c_int(which), _i(e), _d(w), _i(opp))
    return e, w, opp


def moment_matrix(which: int) -> np.ndarray:
    M = np.zeros((19, 19))
    lib().orc_moment_matrix(ctypes.c_int(which), _d(M))
    return M


def invert19(A: np.ndarray) -> np.ndarray:
    A = np.ascontiguousarray(A, np.float64)
    out = np.zeros((19, 19))
    if lib().orc_invert19(_d(A), _d(out)) != 0:
        raise ValueError("singular matrix")
    return out


def number_fluid(solid: np.ndarray):
    """(count, idx[nx,ny,nz] 0=solid else 1-based, supinv[count+1]) for an [i][j][k] solid map."""
    solid = np.ascontiguousarray(solid, np.int32)
    nx, ny, nz = solid.shape
    idx = np.zeros(solid.shape, np.int32)
    count = lib().orc_number_fluid(nx, ny, nz, _i(solid), _i(idx), None)
    supinv = np.zeros(count + 1, np.int32)
    lib().orc_number_fluid(nx, ny, nz, _i(solid), _i(idx), _i(supinv))
    return count, idx, supinv


def slab_split(solid: np.ndarray, ranks: int) -> np.ndarray:
    """First x-plane of each of ``ranks`` ranks under the reference's split (SP:1085-1108)."""
    loc_por = np.ascontiguousarray((np.asarray(solid) == 0).reshape(solid.shape[0], -1).sum(axis=1), np.int32)
    disp = np.zeros(int(ranks), np.int32)
    lib().orc_slab_split(int(solid.shape[0]), _i(loc_por), int(ranks), _i(disp))
    return disp


def rank_ends(solid: np.ndarray, ranks: int) -> np.ndarray:
    """Exclusive end of each rank's range of the global (0-based, i-major) node numbering."""
    disp = slab_split(solid, ranks)
    per_plane = (np.asarray(solid) == 0).reshape(solid.shape[0], -1).sum(axis=1)
    first = np.concatenate([[0], np.cumsum(per_plane)])        # node id of each plane's first node
    bounds = list(disp[1:]) + [solid.shape[0]]
    return np.array([first[b] for b in bounds], np.int64)


def error_ranks(u: np.ndarray, u0: np.ndarray, ends) -> tuple:
    """Error (SP:3625-3704) over ranks owning the node ranges ending at ``ends``; u and u0 carry
    the padding row 0 (the oracle's layout)."""
    u = np.ascontiguousarray(u, np.float64)
    u0 = np.ascontiguousarray(u0, np.float64)
    ends = np.ascontiguousarray(ends, np.int64)
    umax = ctypes.c_double()
    ssum = ctypes.c_double()
    err = lib().orc_sp_error_ranks(int(ends.size), ends.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _d(u), _d(u0),
                                   ctypes.byref(umax), ctypes.byref(ssum))
    return err, umax.value, ssum.value


def sp_relaxation(nu: float):
    S = np.zeros(19)
    sv = ctypes.c_double()
    lib().orc_sp_relaxation(ctypes.c_double(nu), _d(S), ctypes.byref(sv))
    return S, sv.value


class SPOracle:
    """Single-phase state on the CPU, stepped exactly like the reference time loop."""

    def __init__(self, solid, nu, g, u0=(0.0, 0.0, 0.0), rho0=1.0, nthreads=1, bc=None, power_law=None):
        self.bc = bc
        self.power_law = None if power_law is None else np.ascontiguousarray(power_law, np.float64)
        self.solid = np.ascontiguousarray(solid, np.int32)
        self.nx, self.ny, self.nz = self.solid.shape
        self.count, self.idx, self.supinv = number_fluid(self.solid)
        self.nu = float(nu)
        self.g = np.ascontiguousarray(g, np.float64)
        self.S, self.s_v = sp_relaxation(self.nu)
        self.niu_vn = np.full(self.count + 1, self.nu)   # per-node viscosity of the power-law collision
        n1 = self.count + 1
        self.rho = np.zeros(n1)
        self.u = np.zeros((n1, 3))
        self.u0 = np.zeros((n1, 3))
        self.f = np.zeros((n1, 19))
        self.F = np.zeros((n1, 19))
        self.nthreads = int(nthreads)
        u0 = np.ascontiguousarray(u0, np.float64)
        lib().orc_sp_init(self.count, ctypes.c_double(rho0), _d(u0), _d(self.rho), _d(self.u), _d(self.f))
        self.steps = 0

    def load(self, f, u, rho):
        """Replace the state (per-node arrays without the padding row)."""
        self.f[1:] = f
        self.u[1:] = u
        self.rho[1:] = rho

    def run(self, nsteps: int):
        L = lib()
        if getattr(self, "power_law", None) is not None:
            bc = ctypes.byref(self.bc) if self.bc is not None else None
            rc = L.orc_sp_run_ex(self.nx, self.ny, self.nz, self.count, _i(self.idx), _i(self.supinv), _d(self.S),
                                 _d(self.g), bc, _d(self.power_law), _d(self.niu_vn), int(self.steps), _d(self.rho),
                                 _d(self.u), _d(self.u0), _d(self.f), _d(self.F), int(nsteps), self.nthreads)
            if rc != 0:
                raise ValueError("boundary kind not restated (velocity Sub_BC 2/3)")
            self.steps += nsteps
            return
        if getattr(self, "bc", None) is not None:
            rc = L.orc_sp_run_bc(self.nx, self.ny, self.nz, self.count, _i(self.idx), _i(self.supinv), _d(self.S),
                                 _d(self.g), ctypes.byref(self.bc), _d(self.rho), _d(self.u), _d(self.u0), _d(self.f),
                                 _d(self.F), int(nsteps), self.nthreads)
            if rc != 0:
                raise ValueError("boundary kind not restated (velocity Sub_BC 2/3)")
            self.steps += nsteps
            return
        L.orc_sp_run(self.nx, self.ny, self.nz, self.count, _i(self.idx), _i(self.supinv), _d(self.S),
                     _d(self.g), _d(self.rho), _d(self.u), _d(self.u0), _d(self.f), _d(self.F),
                     int(nsteps), self.nthreads)
        self.steps += nsteps

    def error(self, ranks: int = 1):
        """(err, u_max, sum of |u|) of Error (SP:3625-3704) as the reference computes it on
        ``ranks`` MPI ranks: each rank skips its own last node, err and u_max are the largest
        per-rank values, the speed sums add up."""
        return error_ranks(self.u, self.u0, rank_ends(self.solid, ranks))

    def flux(self, box=None):
        Q = np.zeros(3)
        b = None if box is None else (ctypes.c_int * 6)(*box)
        lib().orc_flux(self.ny, self.nz, self.count, _i(self.supinv), _d(self.u), b, _d(Q))
        return Q


class CGOracle:
    """Colour-gradient state on the CPU, stepped like the reference CG time loop."""

    def __init__(self, solid, psi_field, nu_l, nu_g, g, cap_a, cos_theta, u0=(0.0, 0.0, 0.0),
                 bodyforce_apply=0, rel_perm_psi=0.99, rel_perm_psi2=0.92):
        self.solid = np.ascontiguousarray(solid, np.int32)
        self.nx, self.ny, self.nz = self.solid.shape
        self.count, self.idx, self.supinv = number_fluid(self.solid)
        n1 = self.count + 1
        psi_node = np.zeros(n1)
        flat = np.clip(np.asarray(psi_field, np.float64).reshape(-1), -1.0, 1.0)
        psi_node[1:] = flat[self.supinv[1:]]
        self.visc = np.zeros(8)
        lib().orc_cg_visc_consts(ctypes.c_double(nu_l), ctypes.c_double(nu_g), _d(self.visc))
        self.nu_l, self.nu_g = float(nu_l), float(nu_g)
        self.params = _CGParams()
        for a in range(3):
            self.params.g[a] = float(g[a])
        self.params.cap_a = float(cap_a)
        self.params.psi_solid = float(cos_theta)
        for a in range(8):
            self.params.visc[a] = self.visc[a]
        self.params.bodyforce_apply = int(bodyforce_apply)
        self.params.rel_perm_psi = float(rel_perm_psi)
        self.params.rel_perm_psi2 = float(rel_perm_psi2)
        self.g = np.ascontiguousarray(g, np.float64)
        self.rho = np.zeros(n1)
        self.u = np.zeros((n1, 3))
        self.u0 = np.zeros((n1, 3))
        self.f = np.zeros((n1, 19))
        self.F = np.zeros((n1, 19))
        self.psi = np.zeros(n1)
        self.rho_r = np.zeros(n1)
        self.rho_b = np.zeros(n1)
        self.rhor = np.zeros(n1)
        self.rhob = np.zeros(n1)
        u0 = np.ascontiguousarray(u0, np.float64)
        lib().orc_cg_init(self.count, _d(u0), _d(psi_node), _d(self.rho), _d(self.u), _d(self.psi),
                          _d(self.rho_r), _d(self.rho_b), _d(self.rhor), _d(self.rhob), _d(self.f))
        self.steps = 0

    def run(self, nsteps: int):
        L = lib()
        for _ in range(int(nsteps)):
            L.orc_cg_collide(self.nx, self.ny, self.nz, self.count, _i(self.idx), _i(self.supinv),
                             ctypes.byref(self.params), _d(self.rho), _d(self.u), _d(self.f), _d(self.F),
                             _d(self.psi), _d(self.rho_r), _d(self.rho_b), _d(self.rhor), _d(self.rhob))
            L.orc_cg_macro(self.count, _d(self.g), _d(self.F), _d(self.f), _d(self.rho), _d(self.u),
                           _d(self.u0), _d(self.rho_r), _d(self.rho_b), _d(self.rhor), _d(self.rhob),
                           _d(self.psi))
            self.steps += 1

    def error(self, ranks: int = 1):
        """Error (CG:4979-5040, the single-phase routine's loop and rank reduction) on ``ranks`` ranks."""
        return error_ranks(self.u, self.u0, rank_ends(self.solid, ranks))

    def flux(self, box=None):
        Ql = np.zeros(3)
        Qg = np.zeros(3)
        b = None if box is None else (ctypes.c_int * 6)(*box)
        lib().orc_cg_flux(self.ny, self.nz, self.count, _i(self.supinv), _d(self.u), _d(self.psi), b,
                          ctypes.c_double(self.params.rel_perm_psi), ctypes.c_double(self.params.rel_perm_psi2),
                          _d(Ql), _d(Qg))
        return Ql, Qg

    def wet_count(self, box=None):
        b = None if box is None else (ctypes.c_int * 6)(*box)
        return lib().orc_cg_wet_count(self.ny, self.nz, self.count, _i(self.supinv), _d(self.psi), b)


class _CGParams(ctypes.Structure):
    _fields_ = [("g", ctypes.c_double * 3), ("cap_a", ctypes.c_double), ("psi_solid", ctypes.c_double),
                ("visc", ctypes.c_double * 8), ("bodyforce_apply", ctypes.c_int),
                ("rel_perm_psi", ctypes.c_double), ("rel_perm_psi2", ctypes.c_double)]


class STOracle:
    """Passive solute (D3Q7 BGK, rest weight w0; w0 = 0 is the reference's six-velocity scheme)
    in a frozen velocity field, stepped like the six-velocity program's time loop."""

    def __init__(self, solid, D, u_nodes, conc_field, w0=0.25, u_init=None, nthreads=1):
        self.solid = np.ascontiguousarray(solid, np.int32)
        self.nthreads = int(nthreads)
        self.nx, self.ny, self.nz = self.solid.shape
        self.count, self.idx, self.supinv = number_fluid(self.solid)
        self.w0 = float(w0)
        cs2 = (1.0 - self.w0) / 3.0
        self.tau = float(D) / cs2 + 0.5          # tau_s = D/c_s^2 + 1/2 (6V.cpp:1274)
        n1 = self.count + 1
        self.u = np.zeros((n1, 3))
        self.u[1:] = u_nodes
        self.C = np.zeros(n1)
        self.C[1:] = np.asarray(conc_field, np.float64).ravel()[self.supinv[1:]]
        self.g = np.zeros((n1, 7))
        self.G = np.zeros((n1, 7))
        ui = self.u if u_init is None else np.ascontiguousarray(u_init, np.float64)
        lib().orc_st_init(self.count, ctypes.c_double(self.w0), _d(self.C), _d(ui), _d(self.g))
        self.steps = 0

    def run(self, nsteps: int):
        lib().orc_st_step(self.nx, self.ny, self.nz, self.count, _i(self.idx), _i(self.supinv),
                          ctypes.c_double(self.w0), ctypes.c_double(self.tau), _d(self.u), _d(self.g), _d(self.G),
                          _d(self.C), int(nsteps), self.nthreads)
        self.steps += nsteps

    def mass(self) -> float:
        return lib().orc_st_mass(self.count, _d(self.C))
