"""ctypes binding of liblbm_mi355x.so (the C ABI in include/lbm_mi355x.h)."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "liblbm_mi355x.so")
HEADER = os.path.join(REPO_DIR, "include", "lbm_mi355x.h")

LBM_OK = 0
LBM_FP64, LBM_FP32, LBM_FP16 = 0, 1, 2
PRECISIONS = {"fp64": LBM_FP64, "fp32": LBM_FP32, "fp16": LBM_FP16}


class LBMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lbm error {code}: {msg}")
        self.code = code


class SpParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("nu", ctypes.c_double), ("g", ctypes.c_double * 3), ("rho0", ctypes.c_double),
                ("precision", ctypes.c_int32), ("fp16_scale", ctypes.c_double), ("device", ctypes.c_int32)]


class SpBC(ctypes.Structure):
    _fields_ = [("sub_bc", ctypes.c_int32), ("pre", ctypes.c_int32 * 6), ("p", ctypes.c_double * 6),
                ("vel", ctypes.c_int32 * 6), ("v", ctypes.c_double * 6)]


class SpDiag(ctypes.Structure):
    _fields_ = [("Q", ctypes.c_double * 3), ("err", ctypes.c_double), ("umax", ctypes.c_double),
                ("sum_speed", ctypes.c_double), ("nan_count", ctypes.c_double), ("count_in_box", ctypes.c_double)]


class CgParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("nu_l", ctypes.c_double), ("nu_g", ctypes.c_double), ("cos_theta", ctypes.c_double),
                ("cap_a", ctypes.c_double), ("g", ctypes.c_double * 3), ("bodyforce_apply", ctypes.c_int32),
                ("rel_perm_psi", ctypes.c_double), ("rel_perm_psi2", ctypes.c_double), ("rho0", ctypes.c_double),
                ("precision", ctypes.c_int32), ("device", ctypes.c_int32), ("perm_dir", ctypes.c_int32)]


class CgDiag(ctypes.Structure):
    _fields_ = [("Ql", ctypes.c_double * 3), ("Qg", ctypes.c_double * 3), ("err", ctypes.c_double),
                ("umax", ctypes.c_double), ("sum_speed", ctypes.c_double), ("nan_count", ctypes.c_double),
                ("wet_count", ctypes.c_double), ("count_in_box", ctypes.c_double), ("sd_l", ctypes.c_double),
                ("sd_g", ctypes.c_double)]


class StParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("D", ctypes.c_double), ("w0", ctypes.c_double), ("precision", ctypes.c_int32),
                ("device", ctypes.c_int32)]


class StDiag(ctypes.Structure):
    _fields_ = [("mass", ctypes.c_double), ("cmin", ctypes.c_double), ("cmax", ctypes.c_double),
                ("moment1", ctypes.c_double * 3), ("moment2", ctypes.c_double * 3),
                ("nan_count", ctypes.c_double)]


class Launch(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("order", ctypes.c_int32), ("variant", ctypes.c_int32),
                ("tuned", ctypes.c_int32), ("time_us", ctypes.c_double)]


KERNELS = {"sp_step": 1, "cg_collide": 2, "cg_colour": 3, "st_step": 4}
ORDERS = {0: "hw", 1: "runs8", 2: "slab"}
VARIANTS = {0: "gather", 1: "staged"}


def get_launch(h, kernel: str) -> dict:
    ls = Launch()
    check(lib().lbm_get_launch(h, KERNELS[kernel], ctypes.byref(ls)), h)
    return {"block": ls.block, "order": ORDERS.get(ls.order, ls.order), "variant": VARIANTS.get(ls.variant, ls.variant),
            "tuned": bool(ls.tuned), "time_us": round(ls.time_us, 2)}


def set_launch(h, kernel: str, block: int = 256, order: str = "runs8", variant: str = "gather"):
    ls = Launch()
    ls.block = int(block)
    ls.order = {v: k for k, v in ORDERS.items()}[order]
    ls.variant = {v: k for k, v in VARIANTS.items()}[variant]
    check(lib().lbm_set_launch(h, KERNELS[kernel], ctypes.byref(ls)), h)


class SrvParams(ctypes.Structure):
    _fields_ = [("capacity_nodes", ctypes.c_int64), ("mem_fraction", ctypes.c_double),
                ("check_interval", ctypes.c_int32), ("tol", ctypes.c_double), ("min_steps", ctypes.c_int32),
                ("max_steps", ctypes.c_int32), ("precision", ctypes.c_int32), ("block", ctypes.c_int32),
                ("device", ctypes.c_int32), ("accel_depth", ctypes.c_int32), ("accel_period", ctypes.c_int32),
                ("accel_warmup", ctypes.c_int32), ("escalate_k", ctypes.c_double)]


class SrvRequest(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("solid_ijk", ctypes.POINTER(ctypes.c_int32)), ("nu", ctypes.c_double), ("g", ctypes.c_double * 3),
                ("perm_dir", ctypes.c_int32)]


class SrvResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_int64), ("k", ctypes.c_double * 3), ("rel_change", ctypes.c_double),
                ("steps", ctypes.c_int64), ("fluid_nodes", ctypes.c_int64), ("converged", ctypes.c_int32),
                ("accelerations", ctypes.c_int32), ("precision", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double)]


class SrvStats(ctypes.Structure):
    _fields_ = [("pending", ctypes.c_int64), ("active", ctypes.c_int64), ("done", ctypes.c_int64),
                ("arena_nodes", ctypes.c_int64), ("arena_used", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("node_updates", ctypes.c_int64), ("block", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("busy_seconds", ctypes.c_double)]


_lib = None


def build(jobs: int = 8) -> str:
    """Compile the library (hipcc, gfx950) in place."""
    subprocess.run(["make", "-s", "-C", PKG_DIR, f"-j{jobs}"], check=True)
    return LIB_PATH


def declared_symbols():
    """Function names declared in include/lbm_mi355x.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(lbm_[a-z0-9_]+)\s*\(", text, re.M)))


def lib():
    """Load the built library; fails loudly if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with jianhui-lbm_amd/Makefile "
                               "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        L.lbm_version.restype = ctypes.c_char_p
        L.lbm_last_error.restype = ctypes.c_char_p
        L.lbm_last_error.argtypes = [vp]
        L.lbm_destroy.argtypes = [vp]
        L.lbm_destroy.restype = None
        L.lbm_geometry_wraps.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ip, ctypes.c_int32, ip]
        L.lbm_sp_create.argtypes = [ctypes.POINTER(SpParams), ip, ctypes.POINTER(vp)]
        L.lbm_sp_count.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.lbm_sp_init.argtypes = [vp, ctypes.c_double, dp]
        L.lbm_sp_set_boundaries.argtypes = [vp, ctypes.POINTER(SpBC)]
        L.lbm_sp_set_power_law.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.lbm_sp_get_viscosity.argtypes = [vp, dp]
        L.lbm_sp_load_state.argtypes = [vp, dp, dp, dp]
        L.lbm_sp_step.argtypes = [vp, ctypes.c_int32]
        L.lbm_sp_step_timed.argtypes = [vp, ctypes.c_int32, dp]
        L.lbm_sp_get_state.argtypes = [vp, dp, dp, dp]
        L.lbm_sp_get_state_noalloc.argtypes = [vp, dp, dp, dp]
        L.lbm_sp_get_dense.argtypes = [vp, dp, dp]
        L.lbm_sp_diagnostics.argtypes = [vp, ip, ctypes.POINTER(SpDiag)]
        L.lbm_sp_error_ranks.argtypes = [vp, ctypes.c_int32, dp]
        L.lbm_sp_cube_sums.argtypes = [vp, ip, ctypes.c_int32, ctypes.c_int32, dp]
        L.lbm_sp_steps_done.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.lbm_cg_create.argtypes = [ctypes.POINTER(CgParams), ip, ctypes.POINTER(vp)]
        L.lbm_cg_count.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.lbm_cg_init.argtypes = [vp, dp, dp]
        L.lbm_cg_load_state.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.lbm_cg_step.argtypes = [vp, ctypes.c_int32]
        L.lbm_cg_step_timed.argtypes = [vp, ctypes.c_int32, dp]
        L.lbm_cg_set_body_force.argtypes = [vp, dp]
        L.lbm_cg_get_state.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.lbm_cg_diagnostics.argtypes = [vp, ip, ctypes.POINTER(CgDiag)]
        L.lbm_cg_steps_done.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.lbm_st_create.argtypes = [ctypes.POINTER(StParams), ip, ctypes.POINTER(vp)]
        L.lbm_st_create_coupled.argtypes = [ctypes.POINTER(StParams), vp, ctypes.POINTER(vp)]
        L.lbm_st_count.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.lbm_st_set_velocity.argtypes = [vp, dp]
        L.lbm_st_init.argtypes = [vp, dp]
        L.lbm_st_load_state.argtypes = [vp, dp]
        L.lbm_st_step.argtypes = [vp, ctypes.c_int32]
        L.lbm_st_step_timed.argtypes = [vp, ctypes.c_int32, dp]
        L.lbm_st_get_state.argtypes = [vp, dp, dp]
        L.lbm_st_diagnostics.argtypes = [vp, ctypes.POINTER(StDiag)]
        L.lbm_get_launch.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(Launch)]
        L.lbm_set_launch.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(Launch)]
        L.lbm_srv_create.argtypes = [ctypes.POINTER(SrvParams), ctypes.POINTER(vp)]
        L.lbm_srv_submit.argtypes = [vp, ctypes.POINTER(SrvRequest), ctypes.POINTER(ctypes.c_int64)]
        L.lbm_srv_poll.argtypes = [vp, ctypes.POINTER(SrvResult), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.lbm_srv_wait_idle.argtypes = [vp, ctypes.c_double]
        L.lbm_srv_get_stats.argtypes = [vp, ctypes.POINTER(SrvStats)]
        L.lbm_srv_clock.argtypes = [vp]
        L.lbm_srv_clock.restype = ctypes.c_double
        _lib = L
    return _lib


def check(rc: int, handle=None):
    if rc != LBM_OK:
        msg = lib().lbm_last_error(handle)
        raise LBMError(rc, msg.decode() if msg else "")
