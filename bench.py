"""Benchmark of the single-phase sparse D3Q19 MRT step on one MI355X.

Workload (the configuration the headline metric is quoted on): a 256^3 periodic box of
overlapping solid spheres, radius 8 lu, porosity 0.35 (seed 2024 -> 5,871,208 fluid nodes),
tau = 1 (nu = 1/6), Guo body force g = (1e-6, 0, 0), FP32 shifted populations with FP64
device reductions.  A step is one fused pull-stream + MRT collide launch over every fluid
node (collision + comput_macro_variables of the reference, SP:662-684).

Prints one JSON line: fluid MLUPS (whole job), the HBM roofline of the step kernel, and the
FP64 CPU oracle timed on this host's cores on a bounded sample of the same workload.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "jianhui-lbm_amd")

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E, vendor figure
# algorithmic bytes per fluid-node update: 19 populations read + 19 written + 18 int32
# upstream indices read
B_ALG = {"fp32": 2 * 19 * 4 + 18 * 4, "fp64": 2 * 19 * 8 + 18 * 4, "fp16": 2 * 19 * 2 + 18 * 4}
DTYPE = {"fp32": "f32", "fp64": "f64", "fp16": "f16 storage / f32 compute"}


def load_package():
    if "jianhui_lbm_amd" in sys.modules:
        return sys.modules["jianhui_lbm_amd"]
    spec = importlib.util.spec_from_file_location("jianhui_lbm_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["jianhui_lbm_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def read_traffic(precision):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by
    the same counters on a copy kernel of known size), if a summary has been committed."""
    path = os.path.join(ROOT, "profiles", f"sp_step_{precision}_256cube_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(solid, nu, g, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    o = oracle.SPOracle(solid, nu, g, nthreads=threads)
    t0 = time.perf_counter()
    o.run(1)
    t1 = time.perf_counter()
    per = max(t1 - t0, 1e-3)
    n = max(1, int(budget_s / per))
    t0 = time.perf_counter()
    o.run(n)
    dt = time.perf_counter() - t0
    return {"value": round(o.count * n / dt / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"FP64 C oracle (OpenMP) on the same 256^3 geometry, {n} steps after 1 warm-up step, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp16"])
    ap.add_argument("--edge", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design only "
                                   "(x slabs with 5-population halos), no figure at several GPUs"}))
        sys.exit(2)

    pkg = load_package()
    edge = args.edge
    solid = pkg.geometry.sphere_pack((edge, edge, edge), 8.0, 0.35, seed=2024)
    nu, g = 1.0 / 6.0, (1e-6, 0.0, 0.0)
    s = pkg.SinglePhase(solid, nu, g, precision=args.precision, device=0)
    N = s.count
    s.step(args.warmup)
    s.diagnostics()                      # synchronises the handle stream
    t0 = time.perf_counter()
    kernel_ms = s.step_timed(args.steps)   # HIP events on the handle stream around the K launches
    wall = time.perf_counter() - t0
    diag = s.diagnostics()
    if diag["nan_count"] != 0:
        raise RuntimeError("non-finite velocity in the benchmark run")
    ms_per_step = wall * 1e3 / args.steps
    value = N * args.steps / wall / 1e6
    avg_launch_s = kernel_ms / 1e3 / args.steps
    achieved = B_ALG[args.precision] * N / avg_launch_s / 1e9
    traffic = read_traffic(args.precision)
    out = {
        "metric": "fluid MLUPS",
        "value": round(value, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": f"sp_d3q19_mrt_{edge}cube_spheres_r8_phi0.35", "fluid_nodes": N,
                   "box": [edge, edge, edge], "precision": args.precision, "tau": 1.0, "g": list(g),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "sp_step_kernel", "bytes_per_node": B_ALG[args.precision],
                     "avg_launch_us": round(avg_launch_s * 1e6, 2)},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(solid, nu, g, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
