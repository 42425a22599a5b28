"""Benchmark of the single-phase sparse D3Q19 MRT step on one MI355X.

Workload (the configuration the headline metric is quoted on): a 256^3 periodic box of
overlapping solid spheres, radius 8 lu, porosity 0.35 (seed 2024 -> 5,871,208 fluid nodes),
tau = 1 (nu = 1/6), Guo body force g = (1e-6, 0, 0), FP32 shifted populations with FP64
device reductions.  A step is one fused pull-stream + MRT collide launch over every fluid
node (collision + comput_macro_variables of the reference, SP:662-684).

Prints one JSON line: fluid MLUPS (whole job), the HBM roofline of the step kernel, and the
FP64 CPU oracle timed on this host's cores on a bounded sample of the same workload.

The same line carries the serving figures ("serving"): a seeded Poisson stream of single-phase
permeability requests on thresholded Gaussian-random-field subsamples (edge 48..160 lu,
correlation length 6 lu, porosity 0.15..0.45, FP64 reductions, stop when k changes by < 1e-6
relative over 100 steps) served by the continuous-batching server (FP32 arena), and the same stream served by
sequential plain FP32 solves;
permeabilities per hour, p50/p99 latency, and each served k against the sequential FP32 one.
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


def read_traffic(tag):
    """HBM bytes per launch (per step for the two-kernel colour-gradient step) measured with
    rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a copy kernel of
    known size), if a summary for this kernel and workload has been committed."""
    path = os.path.join(ROOT, "profiles", f"{tag}_traffic.md")
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


def serving_bench(pkg, nreq, rate, seed=7, max_steps=20000):
    """Server vs sequential plain solves on the same seeded request stream."""
    geo = pkg.geometry
    rng = np.random.default_rng(seed)
    edges = rng.integers(48, 161, nreq)
    pors = rng.uniform(0.15, 0.45, nreq)
    gaps = rng.exponential(1.0 / rate, nreq)
    arrivals = np.cumsum(gaps) - gaps[0]
    samples = [geo.grf_sample(int(e), 6.0, float(p), seed=seed * 1000 + i) for i, (e, p) in enumerate(zip(edges, pors))]
    nu, g, tol, interval = 1.0 / 6.0, (1e-6, 0.0, 0.0), 1e-6, 100
    nodes = [int(np.count_nonzero(s == 0)) for s in samples]

    def serve(precision):
        """The stream through the continuous-batching server with the given arena precision."""
        with pkg.PermeabilityServer(check_interval=interval, tol=tol, max_steps=max_steps, block=0,
                                    precision=precision) as srv:
            t0 = srv.clock()
            tickets = []
            for i in range(nreq):
                while srv.clock() - t0 < arrivals[i]:
                    time.sleep(min(0.002, max(0.0, arrivals[i] - (srv.clock() - t0))))
                tickets.append(srv.submit(samples[i], nu, g))
            srv.wait_idle()
            res = {r["ticket"]: r for r in srv.poll(4 * nreq)}
            st = srv.stats()
        rs = [res[t] for t in tickets]
        lat = np.array([r["t_done"] - r["t_submit"] for r in rs])
        span = max(r["t_done"] for r in rs) - min(r["t_submit"] for r in rs)
        return rs, lat, span, st

    runs = {p: serve(p) for p in ("fp32",)}

    # sequential plain solves, FIFO on the same arrival schedule
    seq_lat, seq_k = [], []
    t0 = time.perf_counter()
    for i in range(nreq):
        now = time.perf_counter() - t0
        if now < arrivals[i]:
            time.sleep(arrivals[i] - now)
        s = pkg.SinglePhase(samples[i], nu, g, precision="fp32")
        steps, k_prev = 0, None
        while True:
            s.step(interval)
            steps += interval
            k, _, _ = s.Comput_Perm(1)
            kd = k[0]
            if not np.isfinite(kd):
                break
            if k_prev is not None and steps >= 2 * interval and abs(kd - k_prev) <= tol * abs(kd):
                break
            if steps >= max_steps:
                break
            k_prev = kd
        s.close()
        seq_k.append(kd)
        seq_lat.append(time.perf_counter() - t0 - arrivals[i])
    seq_span = time.perf_counter() - t0
    seq_lat = np.array(seq_lat)
    k_seq = np.array(seq_k)
    ok = np.isfinite(k_seq) & (np.abs(k_seq) > 0)

    def summary(precision):
        rs, lat, span, st = runs[precision]
        k_srv = np.array([r["k"][0] for r in rs])
        dev = float(np.max(np.abs(k_srv[ok] - k_seq[ok]) / np.abs(k_seq[ok]))) if ok.any() else None
        return {
            "arena_precision": precision,
            "perms_per_hour": round(nreq / span * 3600.0, 1),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
            "speedup_vs_sequential": round(seq_span / span, 2),
            "converged": int(sum(1 for r in rs if r["converged"] == 1)),
            "max_rel_k_diff_vs_sequential_fp32": dev,
            "tile": st["block"], "server_device_busy_s": round(st["busy_seconds"], 3),
            "server_mlups": round(st["node_updates"] / max(st["busy_seconds"], 1e-9) / 1e6, 1),
        }

    head = summary("fp32")
    out = {"requests": nreq, "arrival_rate_per_s": rate, "seed": seed, "fluid_nodes_total": int(sum(nodes)),
           "stop_rule": f"|k_n - k_(n-{interval})| <= {tol} |k_n|, same for server and sequential solves"}
    out.update(head)
    out["sequential_plain_fp32"] = {
        "perms_per_hour": round(nreq / seq_span * 3600.0, 1),
        "p50_latency_s": round(float(np.percentile(seq_lat, 50)), 4),
        "p99_latency_s": round(float(np.percentile(seq_lat, 99)), 4)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "fp16"])
    ap.add_argument("--edge", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--workload", default="sp", choices=["sp", "cg"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design only "
                                   "(x slabs with 5-population halos), no figure at several GPUs"}))
        sys.exit(2)

    pkg = load_package()
    if args.serve_only:
        print(json.dumps({"serving": serving_bench(pkg, args.serve_requests, args.serve_rate)}))
        return
    if args.workload == "cg":
        print(json.dumps(cg_bench(pkg, args.steps, args.warmup, args.precision if args.precision != "fp16" else "fp32")))
        return
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
    traffic = read_traffic(f"sp_step_{args.precision}_256cube") if edge == 256 else None
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
    s.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(solid, nu, g, args.cpu_seconds)
    if not args.no_serve:
        out["serving"] = serving_bench(pkg, args.serve_requests, args.serve_rate)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
