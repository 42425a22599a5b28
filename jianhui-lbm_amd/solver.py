"""Python mirror of the reference's single-phase operator interface over the C ABI.

The reference drives its hot path from ``main`` (SP:662-841): per step ``collision`` then
``comput_macro_variables``; every ``freRe`` steps ``Error`` and ``Comput_Perm``.  ``SinglePhase``
exposes the same operations with the same argument meaning; the arithmetic runs in the HIP
library (liblbm_mi355x.so).  There is no CPU fallback: without the library this module raises.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _capi


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


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
    def close(self):
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

    def Error(self, porosity: Optional[float] = None):
        """(error, u_max, u_ave) as Error() returns them (SP:3625-3704)."""
        d = self.diagnostics()
        por = self.porosity if porosity is None else porosity
        return d["err"], d["umax"], d["sum_speed"] / (self.box_volume() * por)

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
