"""Continuous-batching permeability server over the C ABI (``lbm_srv_*``).

Many independent single-phase permeability solves -- the reference's sub-sample workflow of
one INPUT file and one solver run per sub-sample (CPP_SCRIPT/HPC_SUBDOMAIN_INPUT*.cpp) --
served from one device arena: requests are admitted by HBM footprint as they arrive, stepped
together, and retired when k changes by less than ``tol`` (relative) over ``check_interval``
steps.  The request loop runs on a native thread inside the library.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np

from . import _capi, geometry


def request_stream(nreq: int = 32, rate: float = 200.0, seed: int = 7):
    """The serving workload's seeded request stream: edge lengths uniform in 48..160 lu,
    porosities uniform in 0.15..0.45, thresholded Gaussian random fields of correlation length
    6 lu (request i from seed 1000 seed + i), Poisson arrivals at ``rate`` per second.
    Returns (samples, arrival times in seconds from the first, edges, porosities)."""
    rng = np.random.default_rng(seed)
    edges = rng.integers(48, 161, nreq)
    pors = rng.uniform(0.15, 0.45, nreq)
    gaps = rng.exponential(1.0 / rate, nreq)
    arrivals = np.cumsum(gaps) - gaps[0]
    samples = [geometry.grf_sample(int(e), 6.0, float(p), seed=seed * 1000 + i) for i, (e, p) in enumerate(zip(edges, pors))]
    return samples, arrivals, edges, pors


class PermeabilityServer:
    def __init__(self, capacity_nodes: int = 0, mem_fraction: float = 0.5, check_interval: int = 100,
                 tol: float = 1e-6, min_steps: int = 0, max_steps: int = 200000, precision: str = "fp32",
                 block: int = 0, device: int = 0, accel_depth: int = 0, accel_period: int = 0,
                 accel_warmup: int = 0, escalate_k: float = 0.0):
        p = _capi.SrvParams()
        p.capacity_nodes = int(capacity_nodes)
        p.mem_fraction = float(mem_fraction)
        p.check_interval = int(check_interval)
        p.tol = float(tol)
        p.min_steps = int(min_steps)
        p.max_steps = int(max_steps)
        p.precision = _capi.PRECISIONS[precision]
        p.block = int(block)
        p.device = int(device)
        p.accel_depth = int(accel_depth)
        p.accel_period = int(accel_period)
        p.accel_warmup = int(accel_warmup)
        p.escalate_k = float(escalate_k)
        h = ctypes.c_void_p()
        _capi.check(_capi.lib().lbm_srv_create(ctypes.byref(p), ctypes.byref(h)), None)
        self._h = h

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

    def submit(self, solid: np.ndarray, nu: float, g: Sequence[float], perm_dir: int = 1) -> int:
        solid = np.ascontiguousarray(solid, np.int32)
        r = _capi.SrvRequest()
        r.nx, r.ny, r.nz = solid.shape
        r.solid_ijk = solid.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        r.nu = float(nu)
        for a in range(3):
            r.g[a] = float(g[a])
        r.perm_dir = int(perm_dir)
        t = ctypes.c_int64()
        _capi.check(_capi.lib().lbm_srv_submit(self._h, ctypes.byref(r), ctypes.byref(t)), self._h)
        return int(t.value)

    def poll(self, max_results: int = 1024) -> List[dict]:
        buf = (_capi.SrvResult * max_results)()
        n = ctypes.c_int32()
        _capi.check(_capi.lib().lbm_srv_poll(self._h, buf, max_results, ctypes.byref(n)), self._h)
        out = []
        for r in buf[: n.value]:
            out.append({"ticket": r.ticket, "k": np.array(r.k[:]), "rel_change": r.rel_change, "steps": r.steps,
                        "fluid_nodes": r.fluid_nodes, "converged": r.converged,
                        "accelerations": r.accelerations,
                        "precision": {0: "fp64", 1: "fp32"}.get(r.precision, str(r.precision)), "t_submit": r.t_submit,
                        "t_admit": r.t_admit, "t_done": r.t_done})
        return out

    def wait_idle(self, timeout_s: float = 0.0) -> bool:
        rc = _capi.lib().lbm_srv_wait_idle(self._h, float(timeout_s))
        if rc == -16:
            return False
        _capi.check(rc, self._h)
        return True

    def stats(self) -> dict:
        s = _capi.SrvStats()
        _capi.check(_capi.lib().lbm_srv_get_stats(self._h, ctypes.byref(s)), self._h)
        return {k: getattr(s, k) for k, _ in _capi.SrvStats._fields_ if k != "pad"}

    def clock(self) -> float:
        return _capi.lib().lbm_srv_clock(self._h)
