"""Benchmark of the single-phase sparse D3Q19 MRT step on one MI355X.

Workload (the configuration the headline metric is quoted on): a 256^3 periodic box of
overlapping solid spheres, radius 8 lu, porosity 0.35 (seed 2024 -> 5,871,208 fluid nodes),
tau = 1 (nu = 1/6), Guo body force g = (1e-6, 0, 0), FP32 shifted populations with FP64
device reductions.  A step is one fused pull-stream + MRT collide launch over every fluid
node (collision + comput_macro_variables of the reference, SP:662-684).

Prints one JSON line: fluid MLUPS (whole job), the HBM roofline of the step kernel, the launch
settings the autotuner chose, and the reference program (FP64, under mpirun) and the FP64 C
oracle timed on this host's cores on a bounded sample of the same workload.

The default line also carries sub-records, each with its own roofline and launch settings:
"fp64" and "fp16" (the same workload in the other storage precisions), "accuracy" (the FP64,
FP32 and FP16 runs to steady state: k and velocity-field errors against FP64), "cg" (the
colour-gradient step on a 192^3 pack, its FP64 twin, the reference colour-gradient program as
CPU baseline and the FP32-vs-FP64 run over 20,000 steps) and "solute" (the coupled flow +
D3Q7 solute step on a 384^3 fibre pack).

The same line carries the serving figures ("serving"): a seeded Poisson stream of single-phase
permeability requests on thresholded Gaussian-random-field subsamples (edge 48..160 lu,
correlation length 6 lu, porosity 0.15..0.45, FP32 populations, FP64 reductions), arriving
faster than they can be served (200/s), through the continuous-batching server with Anderson
acceleration, and the same stream through sequential plain solves, one sample at a time:
stopped at the first check where k is within 1e-4 of the FP64 reference value (the least work
plain stepping can do at the server's accuracy), and by the reference's relative-change test
(1e-6 over 100 steps, and 1e-7), which miss that accuracy; permeabilities per hour, p50/p99
latency, and the largest error of each against FP64 reference values.
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


def cpu_baseline(solid, nu, g, budget_s, faces=None, power_law=None):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    bc = oracle.make_bc(faces["sub_bc"], faces.get("pre"), faces.get("vel")) if faces else None
    o = oracle.SPOracle(solid, nu, g, nthreads=threads, bc=bc, power_law=power_law)
    if power_law:
        o.run(11)    # the power law acts from the 12th step on
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


def reference_cpu_baseline(solid, nu, g, budget_s, faces=None, power_law=None):
    """The reference program itself (oracle/_ref/sp_ref, compiled from its own source by
    `make -C oracle ref`), under mpirun on the host's cores, on the same geometry: the time per
    step is the difference of two runs of different length, so start-up and geometry reading
    drop out.  None if the binary or mpirun is not there (the C port is used instead)."""
    import subprocess
    import tempfile
    binary = os.path.join(ROOT, "oracle", "_ref", "sp_ref")
    mpirun = "/opt/conda/bin/mpirun"
    if not (os.path.exists(binary) and os.path.exists(mpirun)) or faces or power_law:
        return None
    pkg = load_package()
    ranks = max(2, min(16, os.cpu_count() or 2))
    nx, ny, nz = solid.shape
    env = dict(os.environ, PATH=os.path.dirname(mpirun) + ":" + os.environ.get("PATH", ""))

    def run(n_max, wd):
        inp = pkg.inputs.SPInput(filename="geometry.bin", n_xyz=(nx, ny, nz), n_max=n_max, reso=1.0, g=tuple(g),
                                 in_vis=nu, freRe=1 << 30, outputfile="./", fre_backup=-1, decbin=1)
        pkg.inputs.write_sp_input(inp, os.path.join(wd, "INPUT.dat"))
        t0 = time.perf_counter()
        proc = subprocess.run([mpirun, "-np", str(ranks), binary, "INPUT.dat"], cwd=wd, env=env, capture_output=True,
                              text=True, timeout=900)
        if proc.returncode != 0:
            raise RuntimeError("reference run failed: " + proc.stdout[-500:] + proc.stderr[-500:])
        return time.perf_counter() - t0

    try:
        with tempfile.TemporaryDirectory() as wd:
            pkg.geometry.write_binary(os.path.join(wd, "geometry.bin"), solid)
            t_short = run(1, wd)                       # 2 steps (n = 0, 1)
            t_probe = run(3, wd)                       # 4 steps
            per = max((t_probe - t_short) / 2.0, 1e-3)
            extra = max(2, int(budget_s / per))
            t_long = run(1 + extra, wd)
    except (RuntimeError, subprocess.TimeoutExpired, OSError) as e:
        print(f"[bench] reference CPU run unavailable: {e}", file=sys.stderr, flush=True)
        return None
    per = (t_long - t_short) / extra
    count = int(np.count_nonzero(solid == 0))
    return {"value": round(count / per / 1e6, 3), "unit": "MLUPS", "cores": ranks, "kind": "reference",
            "sample": f"the reference program (SINGLE_PHASE_MPI_Software_Spars, FP64) under mpirun -np {ranks} on the "
                      f"same geometry: ({t_long:.1f} s for {2 + extra} steps - {t_short:.1f} s for 2 steps) / {extra}"}


def load_serving_truth(seed, nreq):
    """FP64 reference permeabilities of the serving stream (tests/golden/make_serving_truth.py)."""
    path = os.path.join(ROOT, "tests", "golden", "serving_stream_k.json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path))
    if d.get("seed") != seed or d.get("requests") < nreq:
        return None
    return np.array([x["k"] for x in d["samples"][:nreq]])


def serving_bench(pkg, nreq, rate, seed=7, max_steps=60000, quick=False):
    """Server vs sequential plain solves on the same seeded request stream (quick: the
    accelerated server alone, e.g. under a profiler)."""
    samples, arrivals, _, _ = pkg.request_stream(nreq, rate, seed)
    nu, g, interval = 1.0 / 6.0, (1e-6, 0.0, 0.0), 100
    nodes = [int(np.count_nonzero(s == 0)) for s in samples]
    truth = load_serving_truth(seed, nreq)

    def errors(k):
        if truth is None:
            return None
        return float(np.max(np.abs(np.asarray(k) - truth) / np.abs(truth)))

    def serve(accel):
        """The stream through the continuous-batching server (FP32 arena)."""
        kw = dict(accel_depth=8, accel_period=100, accel_warmup=200, escalate_k=1e-2) if accel else {}
        with pkg.PermeabilityServer(check_interval=interval, tol=1e-6, max_steps=max_steps, block=0, **kw) as srv:
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
        return {"perms_per_hour": round(nreq / span * 3600.0, 1),
                "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
                "max_rel_err_vs_fp64": errors([r["k"][0] for r in rs]),
                "converged": int(sum(1 for r in rs if r["converged"] == 1)),
                "escalated_to_fp64": int(sum(1 for r in rs if r["precision"] == "fp64")),
                "node_steps": int(sum(r["steps"] * r["fluid_nodes"] for r in rs)),
                "tile": st["block"], "device_busy_s": round(st["busy_seconds"], 3),
                "mlups": round(st["node_updates"] / max(st["busy_seconds"], 1e-9) / 1e6, 1)}, span

    def sequential(tol):
        """FIFO plain solves on the same arrival schedule, one sample on the GPU at a time,
        stopped by the reference's relative-change test over 100 steps."""
        lat, ks, steps_nodes = [], [], 0
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
                kd = s.Comput_Perm(1)[0][0]
                if not np.isfinite(kd):
                    break
                if k_prev is not None and steps >= 2 * interval and abs(kd - k_prev) <= tol * abs(kd):
                    break
                if steps >= max_steps:
                    break
                k_prev = kd
            s.close()
            ks.append(kd)
            steps_nodes += steps * nodes[i]
            lat.append(time.perf_counter() - t0 - arrivals[i])
        span = time.perf_counter() - t0
        lat = np.array(lat)
        return {"stop_tol": tol, "perms_per_hour": round(nreq / span * 3600.0, 1),
                "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
                "max_rel_err_vs_fp64": errors(ks), "node_steps": int(steps_nodes)}, span

    def sequential_oracle():
        """FIFO plain solves on the same arrival schedule, one sample at a time, each stopped at
        the first check (every 100 steps) at which its k is within 1e-4 of the FP64 reference
        value: the least work any stop rule for plain time stepping could do at that accuracy
        (plain k(t) approaches its limit monotonically on every sample of this stream, so the
        first entry into the band is final).  Samples whose reference k is below 1e-3 are
        solved in FP64: FP32 deviations cannot resolve their flux."""
        lat, ks, steps_nodes = [], [], 0
        t0 = time.perf_counter()
        for i in range(nreq):
            now = time.perf_counter() - t0
            if now < arrivals[i]:
                time.sleep(arrivals[i] - now)
            prec = "fp64" if abs(truth[i]) < 1e-3 else "fp32"
            s = pkg.SinglePhase(samples[i], nu, g, precision=prec)
            steps, kd = 0, float("nan")
            while steps < 400000:
                s.step(interval)
                steps += interval
                kd = s.Comput_Perm(1)[0][0]
                if abs(kd - truth[i]) <= 1e-4 * abs(truth[i]):
                    break
            s.close()
            ks.append(kd)
            steps_nodes += steps * nodes[i]
            lat.append(time.perf_counter() - t0 - arrivals[i])
        span = time.perf_counter() - t0
        lat = np.array(lat)
        return {"stop": "first check with |k - k_ref| <= 1e-4 k_ref (k_ref known)", "perms_per_hour": round(nreq / span * 3600.0, 1),
                "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
                "max_rel_err_vs_fp64": errors(ks), "node_steps": int(steps_nodes),
                "fp64_samples": int(sum(1 for t in truth if abs(t) < 1e-3))}, span

    served, span_srv = serve(True)
    if quick:
        return dict(requests=nreq, arrival_rate_per_s=rate, seed=seed, **served)
    batched, span_bat = serve(False)
    seq_ref, span_ref = sequential(1e-6)
    seq_acc, span_acc = sequential(1e-7)
    seq_orc, span_orc = sequential_oracle() if truth is not None else (None, None)
    out = {"requests": nreq, "arrival_rate_per_s": rate, "seed": seed, "fluid_nodes_total": int(sum(nodes)),
           "accuracy_target": "k within 1e-4 of the FP64 reference (tests/golden/serving_stream_k.json)"}
    out.update(served)
    out["server"] = ("FP32 arena, Anderson acceleration m=8 every 100 steps after 200, retire at 2 x |dk/k| <= 1e-6, "
                     "samples whose pores do not wrap round the box along x sent to FP64 (from rest) when their "
                     "connectivity test reports, samples with k < 1e-2 continued in FP64; samples numbered and "
                     "linked on the device at admission")
    # the same accuracy: every k within 1e-4 of the FP64 reference, which only the oracle-stopped
    # sequential solves reach (the 1e-7 and 1e-6 relative-change rules miss it)
    out["speedup_vs_sequential_same_accuracy"] = round(span_orc / span_srv, 2) if span_orc else None
    out["speedup_vs_sequential_tol_1e-7"] = round(span_acc / span_srv, 2)
    out["speedup_vs_sequential_reference_rule"] = round(span_ref / span_srv, 2)
    out["server_plain_stepping"] = batched
    out["sequential_plain_oracle_stop"] = seq_orc
    out["sequential_plain_tol_1e-7"] = seq_acc
    out["sequential_plain_tol_1e-6"] = seq_ref
    return out


def cg_reference_cpu_baseline(solid, psi, nu_l, nu_g, cos_theta, cap_a, g, budget_s):
    """The reference colour-gradient program (oracle/_ref/cg_ref, compiled from its own source
    by `make -C oracle ref`) under mpirun on the host's cores, on the same geometry and phase
    field (read from phase.dat, ini_Sat < 0), diagnostics off: the time per step is the
    difference of two runs of different length, so start-up and file reading drop out.  None if
    the binary or mpirun is not there."""
    import subprocess
    import tempfile
    binary = os.path.join(ROOT, "oracle", "_ref", "cg_ref")
    mpirun = "/opt/conda/bin/mpirun"
    if not (os.path.exists(binary) and os.path.exists(mpirun)):
        return None
    pkg = load_package()
    ranks = max(2, min(16, os.cpu_count() or 2))
    env = dict(os.environ, PATH=os.path.dirname(mpirun) + ":" + os.environ.get("PATH", ""))

    def run(n_max, wd):
        inp = pkg.inputs.CGInput(filename="geometry.bin", filenamepsi="phase.dat", n_xyz=tuple(solid.shape),
                                 n_max=n_max, reso=1.0, g=tuple(g), niu_l=nu_l, niu_g=nu_g, cos_theta=cos_theta,
                                 CapA=cap_a, Permeability=1.0, PerDir=1, freRe=1 << 30, outputfile="./",
                                 ini_Sat=-1.0, fre_backup=-1, decbin=1)
        pkg.inputs.write_cg_input(inp, os.path.join(wd, "INPUT.dat"))
        t0 = time.perf_counter()
        proc = subprocess.run([mpirun, "-np", str(ranks), binary, "INPUT.dat"], cwd=wd, env=env, capture_output=True,
                              text=True, timeout=900)
        if proc.returncode != 0:
            raise RuntimeError("reference run failed: " + proc.stdout[-500:] + proc.stderr[-500:])
        return time.perf_counter() - t0

    try:
        with tempfile.TemporaryDirectory() as wd:
            pkg.geometry.write_binary(os.path.join(wd, "geometry.bin"), solid)
            pkg.geometry.write_phase_ascii(os.path.join(wd, "phase.dat"), psi)
            t_short = run(1, wd)                       # 2 steps (n = 0, 1)
            t_probe = run(5, wd)                       # 6 steps
            per = max((t_probe - t_short) / 4.0, 1e-3)
            extra = max(2, min(2000, int(budget_s / per)))
            t_long = run(1 + extra, wd)
    except (RuntimeError, subprocess.TimeoutExpired, OSError) as e:
        print(f"[bench] reference CPU run unavailable: {e}", file=sys.stderr, flush=True)
        return None
    per = (t_long - t_short) / extra
    count = int(np.count_nonzero(solid == 0))
    return {"value": round(count / per / 1e6, 3), "unit": "MLUPS", "cores": ranks, "kind": "reference",
            "sample": f"the reference program (Multi_Component_CG_MPI_SPARSE_Software, FP64) under mpirun -np {ranks} "
                      f"on the same geometry and phase field: ({t_long:.1f} s for {2 + extra} steps - {t_short:.1f} s "
                      f"for 2 steps) / {extra}"}


def cg_bench(pkg, steps, warmup, precision, edge=192, cpu_s=15.0, no_cpu=False):
    """Colour-gradient workload: 192^3 sphere pack (r = 6, porosity 0.45), S_w = 0.5 by voxel
    (seeded), nu_l = nu_g = 1/6, CapA = 0.01, cos theta = 0, g = 1e-5."""
    geo = pkg.geometry
    solid = geo.sphere_pack((edge, edge, edge), 6.0, 0.45, seed=2025)
    psi = np.where(np.random.default_rng(2026).uniform(size=solid.shape) < 0.5, 1.0, -1.0)
    s = pkg.ColorGradient(solid, 1 / 6, 1 / 6, 0.0, 0.01, (1e-5, 0.0, 0.0), precision=precision)
    pin(s, "cg_collide")
    pin(s, "cg_colour")
    s.init(psi)
    s.step(warmup)
    s.diagnostics()
    t0 = time.perf_counter()
    ms = s.step_timed(steps)
    wall = time.perf_counter() - t0
    d = s.diagnostics()
    N = s.count
    b_alg = 2 * 21 * (8 if precision == "fp64" else 4) + 18 * 4   # f + rho_r + rho_b read and written, indices
    achieved = b_alg * N / (ms / 1e3 / steps) / 1e9
    out = {"metric": "fluid MLUPS (colour gradient)", "value": round(N * steps / wall / 1e6, 1), "unit": "MLUPS",
           "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(wall * 1e3 / steps, 5),
           "higher_is_better": True, "dtype": DTYPE[precision], "data": "synthetic",
           "config": {"workload": f"cg_d3q19_mrt_{edge}cube_spheres_r6_phi0.45_sw0.5", "fluid_nodes": N,
                      "precision": precision},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": read_traffic(f"cg_step_{precision}_{edge}cube"), "bytes_per_node": b_alg,
                        "kernels": "cg_collide_kernel + cg_colour_kernel per step",
                        "avg_step_us": round(ms * 1e3 / steps, 2)},
           "saturation": d["wet_count"] / d["count_in_box"], "nan_count": d["nan_count"],
           "launch": {"cg_collide": s.launch("cg_collide"), "cg_colour": s.launch("cg_colour")}, "cpu_baseline": None}
    s.close()
    if not no_cpu:
        out["cpu_baseline"] = cg_reference_cpu_baseline(solid, psi, 1 / 6, 1 / 6, 0.0, 0.01, (1e-5, 0.0, 0.0), cpu_s)
    return out


# D3Q7 solute, per node: 7 populations read + 7 written, 6 int32 upstream indices
B_ALG_ST = {"fp32": 2 * 7 * 4 + 6 * 4, "fp64": 2 * 7 * 8 + 6 * 4}


def solute_cpu_baseline(solid, nu, g, D, w0, budget_s, edge=128):
    """The FP64 C oracle, single-phase step then solute step, on a crop of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    crop = np.ascontiguousarray(solid[:edge, :edge, :edge])
    fo = oracle.SPOracle(crop, nu, g, nthreads=threads)
    conc = np.zeros(crop.shape)
    conc[: edge // 8] = 1.0
    so = oracle.STOracle(crop, D, fo.u[1:].copy(), conc, w0=w0, nthreads=threads)
    times = []
    for o in (fo, so):
        t0 = time.perf_counter()
        o.run(1)
        per = max(time.perf_counter() - t0, 1e-3)
        n = max(1, int(0.5 * budget_s / per))
        t0 = time.perf_counter()
        o.run(n)
        times.append(((time.perf_counter() - t0) / n, n))
    step_s = times[0][0] + times[1][0]
    return {"value": round(fo.count / step_s / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"FP64 C oracle (OpenMP), single-phase step + solute step on the {edge}^3 corner of the same "
                      f"fibre pack ({fo.count} fluid nodes): {times[0][1]} flow steps and {times[1][1]} solute steps "
                      f"after 1 warm-up step each"}


def solute_bench(pkg, steps, warmup, precision, edge=384, cpu_s=15.0, no_cpu=False):
    """Coupled flow + passive solute: a 384^3 periodic box of straight fibres (radius 4 lu,
    random orientations, porosity 0.8, seed 2027), nu = 1/6, g = (1.165e-4, 0, 0), D3Q7 solute
    with rest weight 1/4 and D = 0.008, started from a C = 1 slab over the first eighth of the
    box.  The pack's steady permeability is 11.44 lu^2 (single-phase run to a relative change
    below 1e-4 per 1000 steps), so the steady pore velocity is 0.0100 lu/step and
    Pe = u d / D = 10 with the fibre diameter d = 8.  A step is one launch that advances the
    flow and, with the velocity it has just computed, the solute."""
    geo = pkg.geometry
    solid = geo.fibre_pack((edge, edge, edge), 4.0, 0.8, seed=2027)
    nu, g, D, w0 = 1.0 / 6.0, (1.165e-4, 0.0, 0.0), 0.008, 0.25
    conc = np.zeros(solid.shape)
    conc[: edge // 8] = 1.0
    flow = pkg.SinglePhase(solid, nu, g, precision=precision)
    st = pkg.Solute(D, flow=flow, w0=w0, precision=precision)
    pin(st, "st_step")
    st.init(conc)
    N = flow.count
    st.step(warmup)
    m0 = st.diagnostics()["mass"]            # synchronises the shared stream
    t0 = time.perf_counter()
    ms = st.step_timed(steps)                # HIP events on the shared stream around 2K launches
    wall = time.perf_counter() - t0
    d = st.diagnostics()
    fd = flow.diagnostics()
    if d["nan_count"] or fd["nan_count"]:
        raise RuntimeError("non-finite state in the solute benchmark")
    b_step = B_ALG[precision] + B_ALG_ST[precision]
    achieved = b_step * N / (ms / 1e3 / steps) / 1e9
    u_pore = fd["Q"][0] / N
    out = {"metric": "fluid MLUPS (coupled flow + solute)", "value": round(N * steps / wall / 1e6, 1),
           "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": round(wall * 1e3 / steps, 5), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": DTYPE[precision], "data": "synthetic",
           "config": {"workload": f"sp_st_d3q19_d3q7_{edge}cube_fibres_r4_phi0.8_pe10", "fluid_nodes": N,
                      "porosity": round(N / solid.size, 5), "precision": precision, "nu": nu, "g": list(g), "D": D,
                      "w0": w0, "parallelism": "single GPU"},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": read_traffic(f"sp_st_step_{precision}_{edge}cube"), "bytes_per_node": b_step,
                        "kernels": "one sp_step_kernel launch per step with the solute step fused in",
                        "avg_step_us": round(ms * 1e3 / steps, 2)},
           "mass_rel_drift": abs(d["mass"] - m0) / m0, "pore_velocity_at_end": u_pore,
           "pe_at_end": u_pore * 8.0 / D, "launch": {"sp_step_fused": st.launch()}, "cpu_baseline": None}
    st.close()
    flow.close()
    # the solute kernel alone on the same pack with a frozen velocity field: the reference's
    # own solute program (velocity read from the flow run's files)
    fs = pkg.Solute(D, solid=solid, w0=w0, precision=precision)
    pin(fs, "st_step")
    rng = np.random.default_rng(2028)
    fs.set_velocity(rng.uniform(-0.01, 0.01, (fs.count, 3)))
    fs.init(conc)
    fs.step(warmup)
    fs.diagnostics()
    t0 = time.perf_counter()
    ms = fs.step_timed(steps)
    wall = time.perf_counter() - t0
    avg = ms / 1e3 / steps
    ach = B_ALG_ST[precision] * fs.count / avg / 1e9
    out["solute_kernel"] = {"workload": "frozen velocity, same pack", "value": round(fs.count * steps / wall / 1e6, 1),
                            "unit": "MLUPS", "kernel": "st_step_kernel", "bytes_per_node": B_ALG_ST[precision],
                            "avg_launch_us": round(avg * 1e6, 2), "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                            "frac": round(ach / HBM_PEAK_GBS, 4),
                            "traffic": read_traffic(f"st_step_{precision}_{edge}cube"), "launch": fs.launch()}
    fs.close()
    if not no_cpu:
        out["cpu_baseline"] = solute_cpu_baseline(solid, nu, g, D, w0, cpu_s)
    return out


def vtk_bench(pkg, edge=256):
    """The field writers on the single-phase workload: lbm_sp writes the velocity and density
    VTK files of its first step (the reference's legacy ASCII format), once with the host's
    threads and once with one thread (the reference formats on one rank)."""
    import subprocess
    import tempfile
    solid = pkg.geometry.sphere_pack((edge, edge, edge), 8.0, 0.35, seed=2024)
    out = {"workload": f"sp_d3q19_mrt_{edge}cube_spheres_r8_phi0.35 velocity + density VTK", "points": edge ** 3}
    with tempfile.TemporaryDirectory() as wd:
        pkg.geometry.write_binary(os.path.join(wd, "geometry.bin"), solid)
        inp = pkg.inputs.SPInput(filename="geometry.bin", n_xyz=(edge, edge, edge), n_max=0, reso=1.0, g=(1e-6, 0, 0),
                                 in_vis=1 / 6, freRe=1, freVe=1, freDe=1, outputfile="./", fre_backup=-1, decbin=1)
        pkg.inputs.write_sp_input(inp, os.path.join(wd, "INPUT.dat"))
        for label, threads in (("threads", max(1, min(16, os.cpu_count() or 1))), ("one_thread", 1)):
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            proc = subprocess.run([os.path.join(PKG_DIR, "lbm_sp"), "INPUT.dat", "--precision", "fp32"], cwd=wd, env=env,
                                  capture_output=True, text=True, timeout=900)
            if proc.returncode != 0:
                raise RuntimeError("lbm_sp failed: " + proc.stderr[-2000:])
            rec = {"threads": threads}
            for line in proc.stdout.splitlines():
                if line.startswith("[lbm_sp] velocity VTK") or line.startswith("[lbm_sp] density VTK"):
                    kind = line.split()[1]
                    parts = line.replace(",", "").split()
                    rec[f"{kind}_gather_s"] = float(parts[parts.index("gather") + 1])
                    rec[f"{kind}_text_s"] = float(parts[parts.index("text") + 1])
            rec["velocity_MB"] = round(os.path.getsize(os.path.join(wd, "LBM_velocity_Vector_0.vtk")) / 1e6, 1)
            rec["density_MB"] = round(os.path.getsize(os.path.join(wd, "LBM_Density_0.vtk")) / 1e6, 1)
            out[label] = rec
            print(f"[vtk] {label}: {rec}", file=sys.stderr, flush=True)
    return out


def cg_accuracy(pkg, steps=20000, edge=192, every=1000):
    """The colour-gradient workload for its full 20,000 steps in FP32 and in FP64: saturation,
    phase fluxes and relative permeabilities of the FP32 run against the FP64 run every 1000
    steps (interfaces move, so pointwise fields drift apart; these integral quantities are the
    ones the reference reports)."""
    geo = pkg.geometry
    solid = geo.sphere_pack((edge, edge, edge), 6.0, 0.45, seed=2025)
    psi = np.where(np.random.default_rng(2026).uniform(size=solid.shape) < 0.5, 1.0, -1.0)
    runs = {}
    for precision in ("fp64", "fp32"):
        s = pkg.ColorGradient(solid, 1 / 6, 1 / 6, 0.0, 0.01, (1e-5, 0.0, 0.0), precision=precision)
        s.init(psi)
        rows = []
        t0 = time.perf_counter()
        for n in range(every, steps + 1, every):
            s.step(every)
            kl, kg, _ = s.Comput_Perm(1)
            d = s.diagnostics()
            rows.append((n, s.Comput_Saturation(), kl[0], kg[0], d["nan_count"]))
        runs[precision] = {"rows": rows, "seconds": round(time.perf_counter() - t0, 2)}
        s.close()
        print(f"[cg accuracy] {precision}: {steps} steps in {runs[precision]['seconds']} s, S_w {rows[-1][1]:.6f}",
              file=sys.stderr, flush=True)
    a, b = np.array(runs["fp32"]["rows"]), np.array(runs["fp64"]["rows"])
    return {"workload": f"cg_d3q19_mrt_{edge}cube_spheres_r6_phi0.45_sw0.5", "steps": steps,
            "fp64_final": {"saturation": b[-1, 1], "k_wetting": b[-1, 2], "k_nonwetting": b[-1, 3]},
            "fp32_final": {"saturation": a[-1, 1], "k_wetting": a[-1, 2], "k_nonwetting": a[-1, 3]},
            "max_abs_saturation_diff": float(np.abs(a[:, 1] - b[:, 1]).max()),
            "max_rel_k_wetting_diff": float((np.abs(a[:, 2] - b[:, 2]) / np.abs(b[:, 2])).max()),
            "max_rel_k_nonwetting_diff": float((np.abs(a[:, 3] - b[:, 3]) / np.abs(b[:, 3])).max()),
            "final_rel_k_wetting_diff": float(abs(a[-1, 2] - b[-1, 2]) / abs(b[-1, 2])),
            "final_rel_k_nonwetting_diff": float(abs(a[-1, 3] - b[-1, 3]) / abs(b[-1, 3])),
            "nan_count": float(a[:, 4].max() + b[:, 4].max()),
            "seconds": {k: v["seconds"] for k, v in runs.items()}}


def accuracy_study(pkg, edge=256, max_steps=100000):
    """The single-phase workload run to steady state (|k_n - k_(n-100)| <= 1e-6 |k_n|, the
    reference's relative-change test at freRe = 100) in FP64, FP32 and FP16 storage; each
    variant's k and velocity field against the FP64 run's."""
    solid = pkg.geometry.sphere_pack((edge, edge, edge), 8.0, 0.35, seed=2024)
    nu, g, tol, interval = 1.0 / 6.0, (1e-6, 0.0, 0.0), 1e-6, 100
    res = {}
    for precision in ("fp64", "fp32", "fp16"):
        s = pkg.SinglePhase(solid, nu, g, precision=precision)
        steps, k_prev, settled = 0, None, False
        t0 = time.perf_counter()
        while steps < max_steps:
            s.step(interval)
            steps += interval
            k = s.Comput_Perm(1)[0][0]
            if k_prev is not None and steps >= 2 * interval and abs(k - k_prev) <= tol * abs(k):
                settled = True
                break
            k_prev = k
        res[precision] = {"k": k, "steps": steps, "settled": settled, "u": s.velocity(),
                          "seconds": round(time.perf_counter() - t0, 2)}
        s.close()
        print(f"[accuracy] {precision}: k {k:.10e} after {steps} steps (settled {settled})", file=sys.stderr, flush=True)
    ref = res["fp64"]
    out = {"workload": f"sp_d3q19_mrt_{edge}cube_spheres_r8_phi0.35", "stop_rule": "|k_n - k_(n-100)| <= 1e-6 |k_n|",
           "fp64": {"k_lattice": ref["k"], "steps": ref["steps"], "settled": ref["settled"], "seconds": ref["seconds"]}}
    for precision in ("fp32", "fp16"):
        r = res[precision]
        du = r["u"] - ref["u"]
        out[precision] = {"k_lattice": r["k"], "steps": r["steps"], "settled": r["settled"], "seconds": r["seconds"],
                          "k_rel_err_vs_fp64": abs(r["k"] - ref["k"]) / abs(ref["k"]),
                          "u_l2_rel_err_vs_fp64": float(np.sqrt((du * du).sum() / (ref["u"] * ref["u"]).sum()))}
    return out


SUB_STEPS, SUB_WARMUP = 500, 20     # timed and untimed steps of each sub-record
LAUNCH = {}                          # --launch settings: kernel -> (block, order, variant)


def pin(handle, kernel):
    """Apply a --launch setting (the profiler runs pin the settings the autotuner chose in a
    normal run: under --pmc every dispatch is serialised and its timings mean nothing)."""
    if kernel not in LAUNCH:
        return
    block, order, variant = LAUNCH[kernel]
    if kernel == "sp_step":
        handle.set_launch(block, order)
    elif kernel == "st_step":
        handle.set_launch(block, order)
    else:
        handle.set_launch(kernel, block, order, variant)


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def sp_record(pkg, solid, precision, steps, warmup, nu, g, faces=None, power_law=None):
    """The single-phase step on the given pack: MLUPS over the wall clock of `steps` steps, and
    the step kernel's roofline from HIP events on the handle stream around the same launches."""
    edge = solid.shape[0]
    s = pkg.SinglePhase(solid, nu, g, precision=precision, device=0)
    pin(s, "sp_step")
    if faces:
        s.set_boundaries(faces["sub_bc"], faces["pre"])
        s.init(1.0, (0.0, 0.0, 0.0))
    if power_law:
        s.set_power_law(*power_law)
        s.init(1.0, (0.0, 0.0, 0.0))
    N = s.count
    s.step(warmup)
    s.diagnostics()                        # synchronises the handle stream
    t0 = time.perf_counter()
    kernel_ms = s.step_timed(steps)        # HIP events on the handle stream around the K launches
    wall = time.perf_counter() - t0
    diag = s.diagnostics()
    if diag["nan_count"] != 0:
        raise RuntimeError(f"non-finite velocity in the {precision} benchmark run")
    avg_launch_s = kernel_ms / 1e3 / steps
    # the power-law step also reads and writes the node's viscosity
    b_node = B_ALG[precision] + (2 * (8 if precision == "fp64" else 4) if power_law else 0)
    achieved = b_node * N / avg_launch_s / 1e9
    tag = f"sp_step_{precision}_{edge}cube" + ("_xfaces" if faces else "") + ("_powerlaw" if power_law else "")
    rec = {"value": round(N * steps / wall / 1e6, 1), "unit": "MLUPS", "steps": steps, "warmup": warmup,
           "ms_per_step": round(wall * 1e3 / steps, 5), "dtype": DTYPE[precision], "fluid_nodes": N,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(tag),
                        "kernel": "sp_step_kernel", "bytes_per_node": b_node,
                        "avg_launch_us": round(avg_launch_s * 1e6, 2)},
           "launch": s.launch()}
    s.close()
    return rec


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
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--serve-quick", action="store_true", help="with --serve-only: the accelerated server leg alone")
    ap.add_argument("--workload", default="sp", choices=["sp", "cg", "solute"])
    ap.add_argument("--faces", action="store_true",
                    help="single-phase workload driven by pressure faces on x (Sub_BC 3) instead of a body force")
    ap.add_argument("--power-law", action="store_true",
                    help="single-phase workload as a shear-thinning power-law fluid (nu = m |S|^(n-1), n = 0.8)")
    ap.add_argument("--vtk", action="store_true", help="time the VTK field writers of the lbm_sp driver")
    ap.add_argument("--accuracy", action="store_true",
                    help="run the single-phase workload to steady state in FP64/FP32/FP16 and compare")
    ap.add_argument("--launch", action="append", default=[], metavar="KERNEL=BLOCK:ORDER:VARIANT",
                    help="fix a kernel's launch settings instead of the autotuner's choice (sp_step, cg_collide, "
                         "cg_colour, st_step; order hw|runs8|slab; variant gather|staged)")
    ap.add_argument("--no-sub", action="store_true",
                    help="the single-phase FP32 line alone, without the fp64 / fp16 / accuracy / cg / solute records")
    args = ap.parse_args()
    for item in args.launch:
        kernel, spec = item.split("=", 1)
        block, order, variant = (spec.split(":") + ["gather"])[:3]
        LAUNCH[kernel] = (int(block), order, variant)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design only "
                                   "(x slabs with 5-population halos), no figure at several GPUs"}))
        sys.exit(2)

    pkg = load_package()
    if args.serve_only:
        print(json.dumps({"serving": serving_bench(pkg, args.serve_requests, args.serve_rate, quick=args.serve_quick)}))
        return
    if args.accuracy and args.workload == "sp":
        print(json.dumps({"accuracy": accuracy_study(pkg, args.edge)}))
        return
    if args.vtk:
        print(json.dumps({"vtk": vtk_bench(pkg, args.edge)}))
        return
    if args.workload == "solute":
        prec = args.precision if args.precision != "fp16" else "fp32"
        print(json.dumps(solute_bench(pkg, args.steps, args.warmup, prec, cpu_s=args.cpu_seconds, no_cpu=args.no_cpu)))
        return
    if args.workload == "cg" and args.accuracy:
        print(json.dumps({"cg_accuracy": cg_accuracy(pkg)}))
        return
    if args.workload == "cg":
        print(json.dumps(cg_bench(pkg, args.steps, args.warmup, args.precision if args.precision != "fp16" else "fp32",
                                  cpu_s=args.cpu_seconds, no_cpu=args.no_cpu)))
        return
    edge = args.edge
    solid = pkg.geometry.sphere_pack((edge, edge, edge), 8.0, 0.35, seed=2024)
    nu, g = 1.0 / 6.0, (1e-6, 0.0, 0.0)
    faces = None
    if args.faces:
        # the same mean driving as g = 1e-6: dp = |p+ - p-| c_s^2 / L = 1e-6 over 256 lu
        g = (0.0, 0.0, 0.0)
        faces = {"sub_bc": 3, "pre": {"xp": 1.0, "xn": 1.0 + 3e-6 * edge}}
    power_law = None
    if args.power_law:
        # m chosen so that nu = 1/6 at |S| = 1e-7, the strain rate scale of this flow
        power_law = (1.0 / 6.0 * 1e-7 ** 0.2, 0.8)
    t_start = time.perf_counter()
    rec = sp_record(pkg, solid, args.precision, args.steps, args.warmup, nu, g, faces, power_law)
    out = {
        "metric": "fluid MLUPS",
        "value": rec["value"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rec["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic",
        "config": {"workload": f"sp_d3q19_mrt_{edge}cube_spheres_r8_phi0.35" + ("_xpressure_faces_subbc3" if faces else "")
                   + ("_power_law_n0.8" if power_law else ""), "power_law": power_law,
                   "fluid_nodes": rec["fluid_nodes"], "box": [edge, edge, edge], "precision": args.precision, "tau": 1.0,
                   "g": list(g), "faces": faces, "parallelism": "single GPU"},
        "roofline": rec["roofline"],
        "launch": rec["launch"],
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        ref = reference_cpu_baseline(solid, nu, g, args.cpu_seconds, faces, power_law)
        port = cpu_baseline(solid, nu, g, args.cpu_seconds, faces, power_law)
        out["cpu_baseline"] = ref if ref is not None else port
        if ref is not None:
            out["cpu_port"] = port   # the FP64 C restatement on the same cores, for comparison
    default_run = not (faces or power_law) and edge == 256 and args.precision == "fp32"
    if default_run and not args.no_sub:
        # the other storage precisions of the same workload, their accuracy at steady state,
        # then the colour-gradient and coupled-solute workloads
        for prec in ("fp64", "fp16"):
            out[prec] = sp_record(pkg, solid, prec, SUB_STEPS, SUB_WARMUP, nu, g)
            _progress(f"sp {prec}: {out[prec]['value']} MLUPS")
        out["accuracy"] = accuracy_study(pkg, edge)
        _progress("accuracy done")
        cg = cg_bench(pkg, SUB_STEPS, SUB_WARMUP, "fp32", cpu_s=args.cpu_seconds, no_cpu=args.no_cpu)
        cg["fp64"] = {k: v for k, v in cg_bench(pkg, SUB_STEPS, SUB_WARMUP, "fp64", no_cpu=True).items()
                      if k in ("value", "unit", "ms_per_step", "roofline", "launch", "steps", "warmup")}
        cg["accuracy"] = cg_accuracy(pkg)
        out["cg"] = cg
        _progress(f"cg: {cg['value']} MLUPS")
        out["solute"] = solute_bench(pkg, SUB_STEPS, SUB_WARMUP, "fp32", cpu_s=args.cpu_seconds, no_cpu=args.no_cpu)
        _progress(f"solute: {out['solute']['value']} MLUPS")
    if not args.no_serve:
        out["serving"] = serving_bench(pkg, args.serve_requests, args.serve_rate)
    out["bench_seconds"] = round(time.perf_counter() - t_start, 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
