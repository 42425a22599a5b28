"""Python mirror of the reference's single-phase operator interface over the C ABI.

The reference drives its hot path from ``main`` (SP:662-841): per step ``collision`` then
``comput_macro_variables``; every ``freRe`` steps ``Error`` and ``Comput_Perm``.  ``SinglePhase``
exposes the same operations with the same argument meaning; the arithmetic runs in the HIP
library (liblbm_mi355x.so).  There is no CPU fallback: without the library this module raises.
"""
from __future__ import annotations

import ctypes
import weakref
from typing import Optional, Sequence

import numpy as np

from . import _capi


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class SinglePhase:
    """Sparse D3Q19 MRT single-phase solver on one GPU.

    ``solid`` is the reference's geometry array, int32 ``[i][j][k]`` with 0 = pore.
    ``precision`` is "fp64" (parity build), "fp32" (shifted FP32) or "fp16" (FP16 storage).
    """

    def __init__(self, solid: np.ndarray, nu: float, g: Sequence[float], precision: str = "fp32",
                 rho0: float = 1.0, device: int = 0, fp16_scale: float = 0.0):
        L = _capi.lib()
        solid = np.ascontiguousarray(solid, np.int32)
        if solid.ndim != 3:
            raise ValueError("solid must be a 3-D [i][j][k] array")
        self.shape = solid.shape
        self.nu = float(nu)
        self.g = tuple(float(x) for x in g)
        self.precision = precision
        p = _capi.SpParams()
        p.nx, p.ny, p.nz = solid.shape
        p.nu = self.nu
        for a in range(3):
            p.g[a] = self.g[a]
        p.rho0 = float(rho0)
        p.precision = _capi.PRECISIONS[precision]
        p.fp16_scale = float(fp16_scale)
        p.device = int(device)
        h = ctypes.c_void_p()
        _capi.check(L.lbm_sp_create(ctypes.byref(p), solid.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                    ctypes.byref(h)), None)
        self._h = h
        n = ctypes.c_int64()
        _capi.check(L.lbm_sp_count(self._h, ctypes.byref(n)), self._h)
        self.count = int(n.value)
        self.porosity = self.count / float(np.prod(self.shape))
        self._perm_prev = np.full(3, np.nan)

    # -- lifecycle ---------------------------------------------------------------------------
    def launch(self, kernel: str = "sp_step") -> dict:
        """The step's launch settings (block, tile order; chosen by the autotuner on big boxes)."""
        return _capi.get_launch(self._h, kernel)

    def set_launch(self, block: int = 256, order: str = "runs8", kernel: str = "sp_step"):
        _capi.set_launch(self._h, kernel, block, order, "gather")

    def close(self):
        # a coupled solute uses this handle's device arrays: it goes first
        for st in list(getattr(self, "_solutes", ())):
            st.close()
        if getattr(self, "_h", None) is not None and self._h.value:
            _capi.lib().lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- faces ---------------------------------------------------------------------------------
    FACES = ("xp", "xn", "yp", "yn", "zp", "zn")

    def set_boundaries(self, sub_bc: int = 0, pre: Optional[dict] = None, vel: Optional[dict] = None):
        """Pressure faces ``pre`` {face: density} and velocity faces ``vel`` {face: normal velocity},
        faces "xp", "xn", "yp", "yn", "zp", "zn", with the reference's Sub_BC (0-3 for pressure,
        0-1 for velocity) -- boundary_pressure / boundary_velocity (SP:2751-3620).  Call before
        stepping; ``set_boundaries()`` with no faces removes them."""
        b = _capi.SpBC()
        b.sub_bc = int(sub_bc)
        for name, val in (pre or {}).items():
            b.pre[self.FACES.index(name)] = 1
            b.p[self.FACES.index(name)] = float(val)
        for name, val in (vel or {}).items():
            b.vel[self.FACES.index(name)] = 1
            b.v[self.FACES.index(name)] = float(val)
        _capi.check(_capi.lib().lbm_sp_set_boundaries(self._h, ctypes.byref(b)), self._h)
        self._pre = dict(pre or {})

    def set_power_law(self, m: float, n: float):
        """Power-law fluid nu = m |S|^(n-1) from the 12th step on -- collision_nnf (SP:2017-2700);
        m <= 0 switches it off.  Call before stepping."""
        _capi.check(_capi.lib().lbm_sp_set_power_law(self._h, float(m), float(n)), self._h)

    def viscosity(self) -> np.ndarray:
        """Per-node viscosity niu_vn[count]."""
        nu = np.empty(self.count)
        _capi.check(_capi.lib().lbm_sp_get_viscosity(self._h, _dp(nu)), self._h)
        return nu

    def pressure_gradient(self, PerDir: int = 1) -> float:
        """dp of Comput_Perm (SP:4600-4614): |p_+ - p_-| c_s^2 / L along PerDir (the INPUT's
        in_BC = 1 case), 0 without pressure faces."""
        pre = getattr(self, "_pre", {})
        ax = {1: 0, 2: 1, 3: 2}.get(PerDir, 0)
        a = "xyz"[ax]
        if not pre:
            return 0.0
        return abs(pre.get(a + "p", 1.0) - pre.get(a + "n", 1.0)) / 3.0 / self.shape[ax]

    # -- state ---------------------------------------------------------------------------------
    def init(self, rho: float = 1.0, u0: Sequence[float] = (0.0, 0.0, 0.0)):
        """f = feq(rho, u0) at every node -- the reference's init (SP:1282-1383)."""
        u = np.ascontiguousarray(u0, np.float64)
        _capi.check(_capi.lib().lbm_sp_init(self._h, float(rho), _dp(u)), self._h)

    def load_state(self, f: np.ndarray, u: np.ndarray, rho: np.ndarray):
        """Post-streaming state in the reference layout -- Backup_init (SP:4715-4828)."""
        f = np.ascontiguousarray(f, np.float64).reshape(self.count, 19)
        u = np.ascontiguousarray(u, np.float64).reshape(self.count, 3)
        rho = np.ascontiguousarray(rho, np.float64).reshape(self.count)
        _capi.check(_capi.lib().lbm_sp_load_state(self._h, _dp(f), _dp(u), _dp(rho)), self._h)

    def step(self, nsteps: int = 1):
        """nsteps x (collision + comput_macro_variables)."""
        _capi.check(_capi.lib().lbm_sp_step(self._h, int(nsteps)), self._h)

    def step_timed(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        _capi.check(_capi.lib().lbm_sp_step_timed(self._h, int(nsteps), ctypes.byref(ms)), self._h)
        return ms.value

    @property
    def steps_done(self) -> int:
        s = ctypes.c_int64()
        _capi.check(_capi.lib().lbm_sp_steps_done(self._h, ctypes.byref(s)), self._h)
        return int(s.value)

    def state(self):
        """(f[count,19], u[count,3], rho[count]) post-streaming, reference order and layout."""
        f = np.empty((self.count, 19))
        u = np.empty((self.count, 3))
        rho = np.empty(self.count)
        _capi.check(_capi.lib().lbm_sp_get_state(self._h, _dp(f), _dp(u), _dp(rho)), self._h)
        return f, u, rho

    def state_noalloc(self):
        """state() through lbm_sp_get_state_noalloc: the library's fixed device array, no allocation."""
        f = np.empty((self.count, 19))
        u = np.empty((self.count, 3))
        rho = np.empty(self.count)
        _capi.check(_capi.lib().lbm_sp_get_state_noalloc(self._h, _dp(f), _dp(u), _dp(rho)), self._h)
        return f, u, rho

    def dense(self):
        """(u[nx,ny,nz,3], rho[nx,ny,nz]) on the box, solid voxels u = 0 and rho = 1: the arrays
        the reference's field writers gather (output_velocity_b SP:4265-4390)."""
        u = np.empty(self.shape + (3,))
        rho = np.empty(self.shape)
        _capi.check(_capi.lib().lbm_sp_get_dense(self._h, _dp(u), _dp(rho)), self._h)
        return u, rho

    def velocity(self):
        u = np.empty((self.count, 3))
        _capi.check(_capi.lib().lbm_sp_get_state(self._h, None, _dp(u), None), self._h)
        return u

    # -- diagnostics ---------------------------------------------------------------------------
    def diagnostics(self, box: Optional[Sequence[int]] = None) -> dict:
        d = _capi.SpDiag()
        b = None
        if box is not None:
            b = (ctypes.c_int32 * 6)(*[int(x) for x in box])
        _capi.check(_capi.lib().lbm_sp_diagnostics(self._h, b, ctypes.byref(d)), self._h)
        return {"Q": np.array(d.Q[:]), "err": d.err, "umax": d.umax, "sum_speed": d.sum_speed,
                "nan_count": d.nan_count, "count_in_box": d.count_in_box}

    def box_volume(self, box=None) -> int:
        if box is None:
            return int(np.prod(self.shape))
        return (box[1] - box[0] + 1) * (box[3] - box[2] + 1) * (box[5] - box[4] + 1)

    def Error(self, porosity: Optional[float] = None, ranks: int = 1):
        """(error, u_max, u_ave) as Error() returns them (SP:3625-3704) in a run of the reference
        on ``ranks`` MPI ranks: each rank skips its own last node and the largest per-rank
        error and u_max are taken (the reference's one-rank run aborts; it runs on 2 or more)."""
        out = np.zeros(3)
        _capi.check(_capi.lib().lbm_sp_error_ranks(self._h, int(ranks), _dp(out)), self._h)
        por = self.porosity if porosity is None else porosity
        return float(out[0]), float(out[1]), float(out[2]) / (self.box_volume() * por)

    def Comput_Perm(self, PerDir: int = 1, box: Optional[Sequence[int]] = None, dp: float = 0.0):
        """(k[3], relative change of k[PerDir-1], u_ave) as Comput_Perm (SP:4565-4712).

        k_d = Q_d / V * nu / (g_d + dp) in lattice units; the reference prints k * reso^2 * 1000.
        The first call's relative change compares against nothing and is NaN.
        """
        d = self.diagnostics(box)
        V = self.box_volume(box)
        with np.errstate(divide="ignore", invalid="ignore"):
            k = d["Q"] / V * self.nu / (np.asarray(self.g) + dp)
        i = {1: 0, 2: 1, 3: 2}.get(PerDir, 0)
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = (k[i] - self._perm_prev[i]) / self._perm_prev[i]
        self._perm_prev = k.copy()
        por = d["count_in_box"] / V
        avg = d["Q"] / (V * por)
        return k, rel, float(np.sqrt(avg @ avg))

    def Comput_Perm_LOCAL(self, box: Optional[Sequence[int]] = None, dp: float = 0.0) -> np.ndarray:
        """k[3] as Comput_Perm_LOCAL (SP:4831-4883): the flux through the x-plane before the last
        (the last rank's Solid[nx_l - 2]) over the cross-section of the per_* box (the whole box
        if None) in each direction, times nu / (g_d + dp)."""
        nx, ny, nz = self.shape
        if nx < 2:
            raise ValueError("the local permeability needs at least two x-planes")
        d = self.diagnostics((nx - 2, nx - 2, 0, ny - 1, 0, nz - 1))
        b = (0, nx - 1, 0, ny - 1, 0, nz - 1) if box is None else box
        ex, ey, ez = b[1] - b[0] + 1, b[3] - b[2] + 1, b[5] - b[4] + 1
        with np.errstate(divide="ignore", invalid="ignore"):
            return d["Q"] / np.array([ey * ez, ex * ez, ex * ey], float) * self.nu / (np.asarray(self.g) + dp)

    def Comput_Grop_Perm(self, n: Sequence[int], size: Sequence[int], origins: Sequence[Sequence[int]],
                         PerDir: int = 1, dp: float = 0.0) -> list:
        """Comput_Grop_Perm (SP:4886-5090): for each cube grid g (n[g]^3 cubes of edge size[g]
        from origins[g], x then y then z), a list of (k along PerDir, porosity) per cube, with
        k = Q / size^3 * nu / (g + dp) -- the lines of the reference's Perm_<size>.output."""
        i = {1: 0, 2: 1, 3: 2}.get(PerDir, 0)
        out = []
        for m, L, c in zip(n, size, origins):
            vol = float(L) ** 3
            sums = self.cube_sums(c, L, m)
            out.append([(q[i] / vol * self.nu / (self.g[i] + dp), q[3] / vol) for q in sums])
        return out

    def cube_sums(self, origin: Sequence[int], size: int, m: int) -> np.ndarray:
        """(m^3, 4): per cube of the grid (index (ix*m + iy)*m + iz), the sum of u over its fluid
        nodes and their number, in one launch (lbm_sp_cube_sums)."""
        out = np.zeros((int(m) ** 3, 4))
        o = np.ascontiguousarray(origin, np.int32)
        _capi.check(_capi.lib().lbm_sp_cube_sums(self._h, _ip(o), int(size), int(m), _dp(out)), self._h)
        return out


class ColorGradient:
    """Sparse D3Q19 MRT colour-gradient two-component solver on one GPU (the reference's CG
    program, MP_3D/Multi_Component_CG_MPI_SPARSE_Software.cpp).

    ``psi_field`` for ``init`` is the per-voxel phase field ``[i][j][k]`` (the content of the
    reference's phase.dat); ``state()`` returns f in the CG program's velocity ordering.
    """

    def __init__(self, solid: np.ndarray, nu_l: float, nu_g: float, cos_theta: float, cap_a: float,
                 g: Sequence[float], precision: str = "fp32", bodyforce_apply: int = 0,
                 rel_perm_psi: float = 0.99, rel_perm_psi2: float = 0.92, rho0: float = 1.0, device: int = 0,
                 perm_dir: int = 1):
        L = _capi.lib()
        solid = np.ascontiguousarray(solid, np.int32)
        self.shape = solid.shape
        self.nu_l, self.nu_g = float(nu_l), float(nu_g)
        self.g = tuple(float(x) for x in g)
        self.precision = precision
        self.rel_perm_psi, self.rel_perm_psi2 = float(rel_perm_psi), float(rel_perm_psi2)
        p = _capi.CgParams()
        p.nx, p.ny, p.nz = solid.shape
        p.nu_l, p.nu_g = self.nu_l, self.nu_g
        p.cos_theta = float(cos_theta)
        p.cap_a = float(cap_a)
        for a in range(3):
            p.g[a] = self.g[a]
        p.bodyforce_apply = int(bodyforce_apply)
        p.rel_perm_psi, p.rel_perm_psi2 = self.rel_perm_psi, self.rel_perm_psi2
        p.rho0 = float(rho0)
        p.precision = _capi.PRECISIONS[precision]
        p.device = int(device)
        p.perm_dir = int(perm_dir)
        h = ctypes.c_void_p()
        _capi.check(L.lbm_cg_create(ctypes.byref(p), solid.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                    ctypes.byref(h)), None)
        self._h = h
        n = ctypes.c_int64()
        _capi.check(L.lbm_cg_count(self._h, ctypes.byref(n)), self._h)
        self.count = int(n.value)
        self.porosity = self.count / float(np.prod(self.shape))
        self._perm_prev = np.full(3, np.nan)

    def launch(self, kernel: str) -> dict:
        """Launch settings of "cg_collide" or "cg_colour" (chosen by the autotuner on big boxes)."""
        return _capi.get_launch(self._h, kernel)

    def set_launch(self, kernel: str, block: int = 256, order: str = "runs8", variant: str = "gather"):
        _capi.set_launch(self._h, kernel, block, order, variant)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _capi.lib().lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, psi_field: np.ndarray, u0: Sequence[float] = (0.0, 0.0, 0.0)):
        psi = np.ascontiguousarray(psi_field, np.float64)
        if psi.shape != self.shape:
            raise ValueError("phase field must have the box shape")
        u = np.ascontiguousarray(u0, np.float64)
        _capi.check(_capi.lib().lbm_cg_init(self._h, _dp(psi), _dp(u)), self._h)

    def load_state(self, f, u, rho, psi, rho_r, rho_b):
        arrs = [np.ascontiguousarray(x, np.float64) for x in (f, u, rho, psi, rho_r, rho_b)]
        _capi.check(_capi.lib().lbm_cg_load_state(self._h, *[_dp(a) for a in arrs]), self._h)

    def step(self, nsteps: int = 1):
        _capi.check(_capi.lib().lbm_cg_step(self._h, int(nsteps)), self._h)

    def step_timed(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        _capi.check(_capi.lib().lbm_cg_step_timed(self._h, int(nsteps), ctypes.byref(ms)), self._h)
        return ms.value

    def set_body_force(self, g: Sequence[float]):
        gg = np.ascontiguousarray(g, np.float64)
        _capi.check(_capi.lib().lbm_cg_set_body_force(self._h, _dp(gg)), self._h)
        self.g = tuple(float(x) for x in gg)

    @property
    def steps_done(self) -> int:
        s = ctypes.c_int64()
        _capi.check(_capi.lib().lbm_cg_steps_done(self._h, ctypes.byref(s)), self._h)
        return int(s.value)

    def colour(self):
        """(rho_r[count], rho_b[count]) after the last colour streaming."""
        rr, rb = np.empty(self.count), np.empty(self.count)
        _capi.check(_capi.lib().lbm_cg_get_state(self._h, None, None, None, None, _dp(rr), _dp(rb)), self._h)
        return rr, rb

    def state(self):
        """(f[count,19] CG ordering, u[count,3], rho, psi, rho_r, rho_b)."""
        f = np.empty((self.count, 19))
        u = np.empty((self.count, 3))
        rho, psi, rr, rb = (np.empty(self.count) for _ in range(4))
        _capi.check(_capi.lib().lbm_cg_get_state(self._h, _dp(f), _dp(u), _dp(rho), _dp(psi), _dp(rr), _dp(rb)),
                    self._h)
        return f, u, rho, psi, rr, rb

    def diagnostics(self, box: Optional[Sequence[int]] = None) -> dict:
        d = _capi.CgDiag()
        b = None if box is None else (ctypes.c_int32 * 6)(*[int(x) for x in box])
        _capi.check(_capi.lib().lbm_cg_diagnostics(self._h, b, ctypes.byref(d)), self._h)
        return {"Ql": np.array(d.Ql[:]), "Qg": np.array(d.Qg[:]), "err": d.err, "umax": d.umax,
                "sum_speed": d.sum_speed, "nan_count": d.nan_count, "wet_count": d.wet_count,
                "count_in_box": d.count_in_box, "sd_l": d.sd_l, "sd_g": d.sd_g}

    def box_volume(self, box=None) -> int:
        if box is None:
            return int(np.prod(self.shape))
        return (box[1] - box[0] + 1) * (box[3] - box[2] + 1) * (box[5] - box[4] + 1)

    def Comput_Perm(self, PerDir: int = 1, box=None, dp: float = 0.0):
        """(k_l[3], k_g[3], relative change of k_l[PerDir-1]) as CG's Comput_Perm (CG:6298-6480)."""
        d = self.diagnostics(box)
        V = self.box_volume(box)
        g = np.asarray(self.g)
        with np.errstate(divide="ignore", invalid="ignore"):
            kl = d["Ql"] / V * self.nu_l / (g + dp)
            kg = d["Qg"] / V * self.nu_g / (g + dp)
            i = {1: 0, 2: 1, 3: 2}.get(PerDir, 0)
            rel = (kl[i] - self._perm_prev[i]) / self._perm_prev[i]
        self._perm_prev = kl.copy()
        return kl, kg, rel

    def Comput_Saturation(self, box=None) -> float:
        """S_l = #(psi >= 0) / (V * porosity) over the box (CG:6604-6763)."""
        d = self.diagnostics(box)
        return d["wet_count"] / d["count_in_box"]


class Solute:
    """Passive solute on the fluid nodes: D3Q7 BGK advection-diffusion (the reference's
    six-velocity program, Single_Phase_3D/Solute_Transport_MPI_Software_Spars_6V.cpp, when
    ``w0`` = 0).

    Standalone (``solid`` given): the velocity is a frozen per-node field set with
    ``set_velocity`` (the reference reads it from the flow run's Velocity_for_solute files).
    Coupled (``flow`` given, a ``SinglePhase``): every ``step`` advances the flow one step and
    the solute one step with that step's velocity, sharing the flow's node list on the device.
    """

    def __init__(self, D: float, solid: Optional[np.ndarray] = None, flow: Optional[SinglePhase] = None,
                 w0: float = 0.25, precision: str = "fp64", device: int = 0):
        if (solid is None) == (flow is None):
            raise ValueError("give exactly one of solid (standalone) or flow (coupled)")
        L = _capi.lib()
        p = _capi.StParams()
        p.D = float(D)
        p.w0 = float(w0)
        p.precision = _capi.PRECISIONS[precision]
        p.device = int(device)
        h = ctypes.c_void_p()
        if flow is None:
            solid = np.ascontiguousarray(solid, np.int32)
            if solid.ndim != 3:
                raise ValueError("solid must be a 3-D [i][j][k] array")
            self.shape = solid.shape
            p.nx, p.ny, p.nz = solid.shape
            _capi.check(L.lbm_st_create(ctypes.byref(p), solid.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                        ctypes.byref(h)), None)
        else:
            self.shape = flow.shape
            _capi.check(L.lbm_st_create_coupled(ctypes.byref(p), flow._h, ctypes.byref(h)), None)
            if not hasattr(flow, "_solutes"):
                flow._solutes = weakref.WeakSet()
            flow._solutes.add(self)
        self._h = h
        self.flow = flow
        self.D, self.w0, self.precision = float(D), float(w0), precision
        self.cs2 = (1.0 - self.w0) / 3.0
        self.tau = self.D / self.cs2 + 0.5
        n = ctypes.c_int64()
        _capi.check(L.lbm_st_count(self._h, ctypes.byref(n)), self._h)
        self.count = int(n.value)

    def launch(self, kernel: str = "st_step") -> dict:
        """The solute step's launch settings (coupled: the fused flow + solute step)."""
        return _capi.get_launch(self._h, kernel)

    def set_launch(self, block: int = 256, order: str = "hw", kernel: str = "st_step"):
        _capi.set_launch(self._h, kernel, block, order, "gather")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _capi.lib().lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_velocity(self, u: np.ndarray):
        """Frozen velocity u[count,3] in node order -- Backup_input_v (ST6:3640-3665)."""
        u = np.ascontiguousarray(u, np.float64).reshape(self.count, 3)
        _capi.check(_capi.lib().lbm_st_set_velocity(self._h, _dp(u)), self._h)

    def init(self, conc_field: np.ndarray):
        """g = g_eq(C, u) from a per-voxel concentration ``[i][j][k]`` -- init (ST6:1249-1320)."""
        c = np.ascontiguousarray(conc_field, np.float64)
        if c.shape != tuple(self.shape):
            raise ValueError("concentration field must have the box shape")
        _capi.check(_capi.lib().lbm_st_init(self._h, _dp(c)), self._h)

    def load_state(self, g: np.ndarray):
        """Post-streaming g[count,7] (index 0 = rest) -- Backup_init (ST6:3716-3770)."""
        g = np.ascontiguousarray(g, np.float64).reshape(self.count, 7)
        _capi.check(_capi.lib().lbm_st_load_state(self._h, _dp(g)), self._h)

    def step(self, nsteps: int = 1):
        _capi.check(_capi.lib().lbm_st_step(self._h, int(nsteps)), self._h)

    def step_timed(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        _capi.check(_capi.lib().lbm_st_step_timed(self._h, int(nsteps), ctypes.byref(ms)), self._h)
        return ms.value

    def concentration(self):
        """C[count] post-streaming (the populations stay on the device)."""
        c = np.empty(self.count)
        _capi.check(_capi.lib().lbm_st_get_state(self._h, None, _dp(c)), self._h)
        return c

    def state(self):
        """(g[count,7], C[count]) post-streaming."""
        g = np.empty((self.count, 7))
        c = np.empty(self.count)
        _capi.check(_capi.lib().lbm_st_get_state(self._h, _dp(g), _dp(c)), self._h)
        return g, c

    def concentration(self) -> np.ndarray:
        c = np.empty(self.count)
        _capi.check(_capi.lib().lbm_st_get_state(self._h, None, _dp(c)), self._h)
        return c

    def diagnostics(self) -> dict:
        d = _capi.StDiag()
        _capi.check(_capi.lib().lbm_st_diagnostics(self._h, ctypes.byref(d)), self._h)
        return {"mass": d.mass, "cmin": d.cmin, "cmax": d.cmax, "moment1": np.array(d.moment1[:]),
                "moment2": np.array(d.moment2[:]), "nan_count": d.nan_count}
