"""Generate the golden fixtures from the reference solvers themselves.

Runs the reference's own single-phase and colour-gradient programs (built from their sources
by ``make -C oracle ref``) on small seeded porous boxes and stores the full-precision FP64
state they write through their checkpoint routine (SP:4507-4562, CG:7020-7090), plus the
text result files, as ``tests/golden/*.npz``.  The fixtures are data only: inputs (geometry,
phase field, parameters) and the reference's outputs.

The reference needs MPI and cannot run with one rank (SP:1226 sends to rank+1), so it runs
with ``mpirun -np 2``.  Its per-node FP64 state at two ranks equals the single-rank semantics
(periodic x) that the oracle and the HIP library implement.  Checkpoint label n holds the
state after n+1 time steps (the loop runs n = 0..n_max, SP:662), and record 0 of each rank
file is padding; ranks are concatenated in order to get the global i-major node order.

Usage (in the build container, where /root/reference exists):
    python tests/golden/make_ref_fixtures.py
"""
from __future__ import annotations

import json
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "jianhui-lbm_amd"))
import geometry  # noqa: E402
import inputs  # noqa: E402

REF_DIR = os.path.join(REPO, "oracle", "_ref")
MPI_BIN = "/opt/conda/bin"


def _build_ref():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)


def _run(binary, workdir, input_name="INPUT.dat", ranks=2):
    env = dict(os.environ)
    env["PATH"] = MPI_BIN + ":" + env.get("PATH", "")
    proc = subprocess.run(["mpirun", "-np", str(ranks), binary, input_name], cwd=workdir, env=env,
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    if proc.returncode != 0:
        raise RuntimeError(f"reference run failed in {workdir}:\n{proc.stdout[-3000:]}")
    return proc.stdout


def _run_until_quiet(binary, workdir, need, ranks=2, limit=120.0, quiet=3.0):
    """Run the reference until the files ``need`` exist and no file in workdir has changed for
    ``quiet`` seconds, then end the process group this call started.  For the high-velocity
    stop (SP:839): its counter U_max_ref is raised on rank 0 only, from the maximum gathered
    there (SP:712), so rank 0 alone leaves the time loop, writes its checkpoint (SP:843) and
    finishes, while the other ranks wait in the next step's exchange and never return."""
    env = dict(os.environ)
    env["PATH"] = MPI_BIN + ":" + env.get("PATH", "")
    proc = subprocess.Popen(["mpirun", "-np", str(ranks), binary, "INPUT.dat"], cwd=workdir, env=env,
                            stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, start_new_session=True)
    t0, last, last_change = time.time(), None, time.time()
    try:
        while time.time() - t0 < limit:
            if proc.poll() is not None:
                break
            state = sorted((f, os.path.getsize(os.path.join(workdir, f))) for f in os.listdir(workdir))
            if state != last:
                last, last_change = state, time.time()
            if all(os.path.exists(os.path.join(workdir, f)) for f in need) and time.time() - last_change > quiet:
                break
            time.sleep(0.2)
    finally:
        if proc.poll() is None:
            os.killpg(proc.pid, signal.SIGKILL)
        out = proc.communicate()[0]
    missing = [f for f in need if not os.path.exists(os.path.join(workdir, f))]
    if missing:
        raise RuntimeError(f"reference run in {workdir} did not write {missing}:\n{out[-3000:]}")
    return out


def _read_ckpt(workdir, name, label, width, ranks=2):
    parts = []
    for rank in range(ranks):
        path = os.path.join(workdir, f"LBM_checkpoint_{name}_{label}.{rank}.bin_input")
        a = np.fromfile(path, dtype="<f8")
        a = a.reshape(-1, width) if width > 1 else a
        parts.append(a[1:])       # record 0 is padding
    return np.concatenate(parts, axis=0)


def _rank_counts(workdir, label, ranks=2):
    """Fluid nodes of each rank: the record count of its density checkpoint, less the padding."""
    return [int(np.fromfile(os.path.join(workdir, f"LBM_checkpoint_rho_{label}.{r}.bin_input"), "<f8").size - 1)
            for r in range(ranks)]


def _text(workdir, fname):
    p = os.path.join(workdir, fname)
    return open(p).read() if os.path.exists(p) else ""


def _bc_fields(inp, bc):
    """Pressure / velocity boundary lines of the INPUT file (SP:260-267, 280) from
    {"sub_bc": s, "pre": {face: rho}, "vel": {face: v}}, faces "xp", "xn", "yp", ..."""
    inp.in_BC = 1
    inp.Sub_BC = int(bc["sub_bc"])
    for kind, attr in (("pre", "pre_"), ("vel", "vel_")):
        vals = bc.get(kind, {})
        for ax in "xyz":
            p, n = vals.get(ax + "p"), vals.get(ax + "n")
            setattr(inp, attr + ax, (1 if p is not None else 0, p if p is not None else (1.0 if kind == "pre" else 0.0),
                                     1 if n is not None else 0, n if n is not None else (1.0 if kind == "pre" else 0.0)))


def _field_outputs(wd, freRe, n_max, out):
    """The reference's VTK, profile and solute-velocity files of one run (SP:538-543, 818-832)."""
    out["vtk_geometry"] = np.array(_text(wd, "LBM_Geometry.vtk"))
    for n in range(0, n_max + 1, freRe):
        out[f"vtk_velocity_{n}"] = np.array(_text(wd, f"LBM_velocity_Vector_{n}.vtk"))
        out[f"vtk_density_{n}"] = np.array(_text(wd, f"LBM_Density_{n}.vtk"))
        out[f"profile_{n}"] = np.array(_text(wd, f"LBM_velocity_{n}.out"))
        if n > 0:
            parts = [np.fromfile(os.path.join(wd, f"Velocity_for_solute_{n}.{r}.bin"), dtype="<f8").reshape(-1, 3)
                     for r in range(2)]
            out[f"solute_velocity_{n}"] = np.concatenate(parts, axis=0)


def sp_case(name, shape, radius, por, seed, nu, g, u0, steps_list, freRe=10, box=None, bc=None, power_law=None,
            field_outputs=False, loc_perm=False, gperm=None, per_dir=1):
    solid = geometry.sphere_pack(shape, radius, por, seed)
    out = {"solid": solid.astype(np.int8)}
    params = dict(shape=list(shape), nu=nu, g=list(g), u0=list(u0), freRe=freRe, box=box,
                  steps=list(steps_list), radius=radius, porosity_target=por, seed=seed, bc=bc,
                  power_law=power_law, loc_perm=loc_perm, gperm=gperm, per_dir=per_dir)
    for steps in steps_list:
        wd = tempfile.mkdtemp(prefix="sp_fix_")
        try:
            geometry.write_binary(os.path.join(wd, "geometry.bin"), solid)
            inp = inputs.SPInput(filename="geometry.bin", n_xyz=tuple(shape), n_max=steps - 1, reso=1.0,
                                 g=tuple(g), in_vis=nu, ini_v=tuple(u0), wr_per=1, PerDir=per_dir, freRe=freRe,
                                 Out_Mode=0, outputfile="./", fre_backup=0, decbin=1)
            if bc is not None:
                _bc_fields(inp, bc)
            if power_law is not None:                 # nnf_module = 1, nnf_m nnf_n (SP:307-308)
                inp.nnf_module = 1
                inp.nnf_mn = tuple(power_law)
            if field_outputs:                         # freVe, freDe, vel_sol = freRe
                inp.freVe = inp.freDe = inp.vel_sol = freRe
            if loc_perm:                              # Comput_Perm_LOCAL (SP:714-715, 4831-4883)
                inp.loc_perm = 1
            if gperm is not None:                     # Comput_Grop_Perm (SP:718-719, 4886-5090)
                inp.gperm = gperm["every"]
                inp.n_gperm = tuple(gperm["n"])
                inp.size_gperm = tuple(gperm["size"])
                inp.c_gperm1, inp.c_gperm2, inp.c_gperm3, inp.c_gperm4 = (tuple(c) for c in gperm["c"])
            if box is not None:
                inp.par_per = (1, 1, 1)
                inp.per_x = (box[1], box[0])
                inp.per_y = (box[3], box[2])
                inp.per_z = (box[5], box[4])
            inputs.write_sp_input(inp, os.path.join(wd, "INPUT.dat"))
            _run(os.path.join(REF_DIR, "sp_ref"), wd)
            label = steps - 1
            out[f"f_{steps}"] = _read_ckpt(wd, "f", label, 19)
            out[f"u_{steps}"] = _read_ckpt(wd, "velocity", label, 3)
            out[f"rho_{steps}"] = _read_ckpt(wd, "rho", label, 1)
            params["rank_counts"] = _rank_counts(wd, label)
            out[f"perm_txt_{steps}"] = np.array(_text(wd, "Permeability_error_local_Perm.txt"))
            out[f"vel_txt_{steps}"] = np.array(_text(wd, "Velocity_ave_max.txt"))
            out[f"general_txt_{steps}"] = np.array(_text(wd, "General_Results.txt"))
            if loc_perm:
                out[f"results_txt_{steps}"] = np.array(_text(wd, "Results.txt"))
            if gperm is not None:
                for size in sorted(set(gperm["size"])):
                    out[f"gperm_{size}_{steps}"] = np.array(_text(wd, f"Perm_{size}.output"))
            if field_outputs:
                _field_outputs(wd, freRe, steps - 1, out)
        finally:
            shutil.rmtree(wd, ignore_errors=True)
    out["params"] = np.array(json.dumps(params))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "count =", out[f"rho_{steps_list[0]}"].shape[0])


def cg_case(name, shape, radius, por, seed, psi_mode, psi_seed, nu_l, nu_g, cos_theta, cap_a, g,
            steps_list, freRe=10, bodyforce_apply=0, rel_perm=(0.99, 0.92), per_dir=1):
    solid = geometry.sphere_pack(shape, radius, por, seed)
    rng = np.random.default_rng(psi_seed)
    if psi_mode == "binary":
        psi = np.where(rng.uniform(size=shape) < 0.5, 1.0, -1.0)
    else:   # values on a 1/64 grid so that the reference's float read is exact
        psi = np.round(rng.uniform(-1.0, 1.0, size=shape) * 64.0) / 64.0
    out = {"solid": solid.astype(np.int8), "psi_field": psi}
    params = dict(shape=list(shape), nu_l=nu_l, nu_g=nu_g, cos_theta=cos_theta, cap_a=cap_a, g=list(g),
                  freRe=freRe, steps=list(steps_list), bodyforce_apply=bodyforce_apply,
                  rel_perm_psi=list(rel_perm), radius=radius, porosity_target=por, seed=seed,
                  psi_mode=psi_mode, psi_seed=psi_seed, per_dir=per_dir)
    for steps in steps_list:
        wd = tempfile.mkdtemp(prefix="cg_fix_")
        try:
            geometry.write_binary(os.path.join(wd, "geometry.bin"), solid)
            geometry.write_phase_ascii(os.path.join(wd, "phase.dat"), psi)
            inp = inputs.CGInput(filename="geometry.bin", filenamepsi="phase.dat", n_xyz=tuple(shape),
                                 n_max=steps - 1, reso=1.0, g=tuple(g), niu_l=nu_l, niu_g=nu_g,
                                 cos_theta=cos_theta, CapA=cap_a, Permeability=1.0, wr_per=1, PerDir=per_dir,
                                 freRe=freRe, outputfile="./", stab=(0, 0), ini_Sat=-1.0,
                                 rel_perm_psi=tuple(rel_perm), fre_backup=0, decbin=1,
                                 bodyforce_apply=bodyforce_apply)
            inputs.write_cg_input(inp, os.path.join(wd, "INPUT.dat"))
            _run(os.path.join(REF_DIR, "cg_ref"), wd)
            label = steps - 1
            out[f"f_{steps}"] = _read_ckpt(wd, "f", label, 19)
            out[f"u_{steps}"] = _read_ckpt(wd, "velocity", label, 3)
            out[f"rho_{steps}"] = _read_ckpt(wd, "rho", label, 1)
            out[f"psi_{steps}"] = _read_ckpt(wd, "psi", label, 1)
            out[f"rhor_{steps}"] = _read_ckpt(wd, "rhor", label, 1)
            out[f"rhob_{steps}"] = _read_ckpt(wd, "rhob", label, 1)
            out[f"relperm1_txt_{steps}"] = np.array(_text(wd, "Relative_Permeability_Component1.txt"))
            out[f"relperm2_txt_{steps}"] = np.array(_text(wd, "Relative_Permeability_Component2.txt"))
            out[f"bodyforce_txt_{steps}"] = np.array(_text(wd, "bodyforce.txt"))
        finally:
            shutil.rmtree(wd, ignore_errors=True)
    out["params"] = np.array(json.dumps(params))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "count =", out[f"rho_{steps_list[0]}"].shape[0])


def _write_st_input(path, shape, n_max, niu_s, freRe, fre_backup, sol_ini_n, restart=0):
    """INPUT file of the six-velocity solute program, by line as its reader takes it
    (Solute_Transport_MPI_Software_Spars_6V.cpp:219-270); every boundary module off."""
    L = ["=== solute input (generated) ===",
         "geometry.bin :geometry", "phase.dat :initial concentration",
         f"{shape[0]} {shape[1]} {shape[2]} :nx ny nz", f"{n_max} :n_max", "1.0 :reso", "0 :in_BC",
         "0.0 0.0 :buoyancy", "0.0 :gravity", "0.0 0.0 0.0 :body force",
         "0 1.0 0 1.0 :p x", "0 1.0 0 1.0 :p y", "0 1.0 0 1.0 :p z",
         "0 0.0 0 0.0 :v x", "0 0.0 0 0.0 :v y", "0 0.0 0 0.0 :v z",
         "0 1.0 0 1.0 :c x", "0 1.0 0 1.0 :c y", "0 1.0 0 1.0 :c z",
         "0 0.0 0 0.0 :zf x", "0 0.0 0 0.0 :zf y", "0 0.0 0 0.0 :zf z",
         "0.1 :viscosity", f"{niu_s!r} :diffusion coefficient", "0.0 0.0 0.0 :initial velocity",
         "=== output ===", "0 :wr_per", "1 :PerDir", f"{freRe} :freRe", "0 :Out_Mode", "-1 :freVe",
         "-1 :freDe", "-1 :frePsi", "0 1000 :dispersion statistics", "=== advanced ===",
         "0 1.0 1.0 :lattice_v", "./ :output path", "2 :Sub_BC", "1 :Sub_BC_psi", "-1 :ini_psi",
         "0 0 0 :partial perm", "0 0 :x", "0 0 :y", "0 0 :z", "=== backup ===",
         f"{fre_backup} :fre_backup", f"{restart} :mode_backup_ini", f"{sol_ini_n} :velocity label",
         "2 2 6 :peaks", "=== geometry reading ===", "1 :decbin"]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def st_case(name, sp_fixture, niu_s, u_scale, c_seed, n_max, labels, freRe=5):
    """Six-velocity passive solute (the D3Q7 oracle with rest weight 0) on the geometry and
    the reference velocity field of an existing single-phase fixture, the field scaled up so
    that advection shows.

    The program reads the velocity per rank from Velocity_for_solute_<label>.<rank>.bin
    (6V.cpp:3640-3665), after its init has already formed the initial populations from the
    not yet loaded velocity array (6V.cpp:473-475, 1306-1310), i.e. from uninitialised memory.
    The fixture therefore starts it through its restart path (mode_backup_ini, 6V.cpp:3716-3770)
    from an initial state written here: g_q = C/6 (1 + 3 e_q.u), the same formula, stored in the
    fixture.  A first short run gives the ranks' node counts (from its checkpoint sizes)."""
    d = np.load(os.path.join(HERE, f"{sp_fixture}.npz"))
    solid = d["solid"].astype(np.int32)
    shape = solid.shape
    u = d["u_101"] * u_scale
    rng = np.random.default_rng(c_seed)
    conc = np.round(rng.uniform(0.0, 1.0, size=shape) * 64.0) / 64.0   # exact in the float read
    wd = tempfile.mkdtemp(prefix="st_fix_")
    try:
        geometry.write_binary(os.path.join(wd, "geometry.bin"), solid)
        geometry.write_phase_ascii(os.path.join(wd, "phase.dat"), conc)
        for r in range(2):
            np.zeros((u.shape[0], 3)).tofile(os.path.join(wd, f"Velocity_for_solute_0.{r}.bin"))
        _write_st_input(os.path.join(wd, "INPUT.dat"), shape, 1, niu_s, 1, 1, 0)
        _run(os.path.join(REF_DIR, "st6_ref"), wd)
        counts = [np.fromfile(os.path.join(wd, f"LBM_checkpoint_psi_1.{r}.bin_input"), "<f8").size - 1
                  for r in range(2)]
        assert sum(counts) == u.shape[0], (counts, u.shape)
        for f in os.listdir(wd):
            if f.startswith("LBM_checkpoint"):
                os.remove(os.path.join(wd, f))
        # initial state: node concentrations in the global order, populations from them
        flat = conc.ravel()
        pore = np.flatnonzero(solid.ravel() == 0)
        c0 = flat[pore]
        e = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
        g0 = np.zeros((u.shape[0], 7))
        for q in range(1, 7):
            g0[:, q] = 1.0 / 6.0 * c0 * (1 + 3 * (u @ e[q]))
        lo = 0
        for r in range(2):
            hi = lo + counts[r]
            np.ascontiguousarray(u[lo:hi]).tofile(os.path.join(wd, f"Velocity_for_solute_0.{r}.bin"))
            np.concatenate([np.zeros((1, 7)), g0[lo:hi]]).tofile(os.path.join(wd, f"LBM_checkpoint_fg_1.{r}.bin_input"))
            np.concatenate([[0.0], c0[lo:hi]]).tofile(os.path.join(wd, f"LBM_checkpoint_psi_1.{r}.bin_input"))
            lo = hi
        _write_st_input(os.path.join(wd, "INPUT.dat"), shape, n_max, niu_s, freRe, freRe, 0, restart=1)
        _run(os.path.join(REF_DIR, "st6_ref"), wd)
        out = {"solid": solid.astype(np.int8), "conc_field": conc, "u": u, "g_0": g0, "c_0": c0}
        for lab in labels:
            out[f"g_{lab + 1}"] = _read_ckpt(wd, "fg", lab, 7)
            out[f"c_{lab + 1}"] = _read_ckpt(wd, "psi", lab, 1)
    finally:
        shutil.rmtree(wd, ignore_errors=True)
    params = dict(shape=list(shape), niu_s=niu_s, u_scale=u_scale, c_seed=c_seed, sp_fixture=sp_fixture,
                  steps=[lab + 1 for lab in labels], rank_counts=counts)
    out["params"] = np.array(json.dumps(params))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "count =", u.shape[0], "ranks", counts)


def _write_sp(wd, solid, **kw):
    geometry.write_binary(os.path.join(wd, "geometry.bin"), solid)
    base = dict(filename="geometry.bin", n_xyz=tuple(solid.shape), reso=1.0, ini_v=(0.0, 0.0, 0.0), wr_per=1,
                PerDir=1, Out_Mode=0, outputfile="./", fre_backup=0, decbin=1)
    base.update(kw)
    inputs.write_sp_input(inputs.SPInput(**base), os.path.join(wd, "INPUT.dat"))


def stop_case(name, kind):
    """The time loop's three stop rules (SP:812-816, 838-839), each on a reference run.

    converged      Out_Mode = 1: the run ends when Comput_Perm's mean speed changes by less than
                   1e-7 relative between two checks; the velocity files are written for that
                   step and the step counter jumps to n_max + 1 (SP:812-816).
    high_velocity  a force that drives a box without solid past u_max = 10: after the fifth check
                   at or above it the run stops (SP:712, 839).  Uniform flow along x keeps every
                   population that crosses a rank's x-face positive, so the reference's ranks
                   stay exact (its exchange keeps only positive values, SP:1993-1998).  Only
                   rank 0 stops; the fixture keeps its files (see _run_until_quiet).
    nan            a restart from a checkpoint with one non-finite node inside each rank's slab:
                   Error's velocity change is NaN at the first check and the run stops (SP:838).
    """
    wd = tempfile.mkdtemp(prefix="sp_stop_")
    out = {}
    try:
        if kind == "converged":
            solid = geometry.sphere_pack((16, 12, 8), 3.0, 0.6, 7)
            params = dict(kind=kind, shape=list(solid.shape), nu=0.1, g=[1e-5, 0.0, 0.0], freRe=10, n_max=20000,
                          Out_Mode=1)
            _write_sp(wd, solid, n_max=20000, g=(1e-5, 0.0, 0.0), in_vis=0.1, freRe=10, Out_Mode=1)
            _run(os.path.join(REF_DIR, "sp_ref"), wd)
            vel = [f for f in os.listdir(wd) if f.startswith("LBM_velocity_Vector_")]
            assert len(vel) == 1, vel
            n_stop = int(vel[0][len("LBM_velocity_Vector_"):-4])
            params["n_stop"] = n_stop
            out["vtk_velocity"] = np.array(_text(wd, vel[0]))
            out["profile"] = np.array(_text(wd, f"LBM_velocity_{n_stop}.out"))
            label, ranks = 20000, 2
        elif kind == "high_velocity":
            solid = np.zeros((8, 6, 4), np.int32)
            params = dict(kind=kind, shape=list(solid.shape), nu=0.1, g=[0.25, 0.0, 0.0], freRe=1, n_max=80)
            _write_sp(wd, solid, n_max=80, g=(0.25, 0.0, 0.0), in_vis=0.1, freRe=1)
            label, ranks = 80, 1   # rank 0's files only
            _run_until_quiet(os.path.join(REF_DIR, "sp_ref"), wd,
                             [f"LBM_checkpoint_{k}_{label}.0.bin_input" for k in ("f", "velocity", "rho")])
        else:
            solid = geometry.sphere_pack((16, 12, 8), 3.0, 0.6, 7)
            params = dict(kind=kind, shape=list(solid.shape), nu=0.1, g=[1e-5, 0.0, 0.0], freRe=10, n_max=30,
                          start_label=9)
            _write_sp(wd, solid, n_max=9, g=(1e-5, 0.0, 0.0), in_vis=0.1, freRe=10)
            _run(os.path.join(REF_DIR, "sp_ref"), wd)
            counts = _rank_counts(wd, 9)
            # one node in the middle x-plane of each rank's slab gets non-finite populations
            start = {k: _read_ckpt(wd, k, 9, w) for k, w in (("f", 19), ("velocity", 3), ("rho", 1))}
            planes = np.flatnonzero(solid.ravel() == 0) // (solid.shape[1] * solid.shape[2])
            lo = 0
            for r, c in enumerate(counts):
                mid = (planes[lo] + planes[lo + c - 1]) // 2
                node = lo + int(np.flatnonzero(planes[lo:lo + c] == mid)[0])
                start["f"][node, :] = np.nan
                lo += c
            lo = 0
            for r, c in enumerate(counts):
                for k, w in (("f", 19), ("velocity", 3), ("rho", 1)):
                    a = start[k][lo:lo + c].reshape(c, w)
                    np.concatenate([np.zeros((1, w)), a]).tofile(os.path.join(wd, f"LBM_checkpoint_{k}_9.{r}.bin_input"))
                lo += c
            for f in os.listdir(wd):
                if f.endswith(".txt"):
                    os.remove(os.path.join(wd, f))
            out.update({"f_start": start["f"], "u_start": start["velocity"], "rho_start": start["rho"]})
            _write_sp(wd, solid, n_max=30, g=(1e-5, 0.0, 0.0), in_vis=0.1, freRe=10, mode_backup_ini=9)
            _run(os.path.join(REF_DIR, "sp_ref"), wd)
            label, ranks = 30, 2
        out["solid"] = solid.astype(np.int8)
        out["f_end"] = _read_ckpt(wd, "f", label, 19, ranks)
        out["u_end"] = _read_ckpt(wd, "velocity", label, 3, ranks)
        out["rho_end"] = _read_ckpt(wd, "rho", label, 1, ranks)
        params["label"] = label
        params["rank_counts"] = _rank_counts(wd, label, ranks)
        out["vel_txt"] = np.array(_text(wd, "Velocity_ave_max.txt"))
        out["perm_txt"] = np.array(_text(wd, "Permeability_error_local_Perm.txt"))
    finally:
        shutil.rmtree(wd, ignore_errors=True)
    out["params"] = np.array(json.dumps(params))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, params)


def c1_case(name, steps=2000, ranks=8):
    """The reference's own CPU case at full size: 64^3 spheres r = 6 to porosity 0.6 (seed
    12345, 157,141 fluid nodes), tau = 1, g = (1e-6, 0, 0), FP64, 2000 steps.  The state is too
    large to keep whole, so the fixture holds every 64th node's f, u and rho, the per-direction
    sums of f over all nodes, the sums of u and rho, and the printed permeability file."""
    solid = geometry.sphere_pack((64, 64, 64), 6.0, 0.6, 12345)
    wd = tempfile.mkdtemp(prefix="sp_c1_")
    try:
        geometry.write_binary(os.path.join(wd, "geometry.bin"), solid)
        inp = inputs.SPInput(filename="geometry.bin", n_xyz=(64, 64, 64), n_max=steps - 1, reso=1.0, g=(1e-6, 0.0, 0.0),
                             in_vis=1.0 / 6.0, ini_v=(0.0, 0.0, 0.0), wr_per=1, PerDir=1, freRe=100, Out_Mode=0,
                             outputfile="./", fre_backup=0, decbin=1)
        inputs.write_sp_input(inp, os.path.join(wd, "INPUT.dat"))
        _run(os.path.join(REF_DIR, "sp_ref"), wd, ranks=ranks)
        f = _read_ckpt(wd, "f", steps - 1, 19, ranks)
        counts = _rank_counts(wd, steps - 1, ranks)
        u = _read_ckpt(wd, "velocity", steps - 1, 3, ranks)
        rho = _read_ckpt(wd, "rho", steps - 1, 1, ranks)
        pick = np.arange(0, f.shape[0], 64)
        out = {"solid": solid.astype(np.int8), "pick": pick, "f_pick": f[pick], "u_pick": u[pick], "rho_pick": rho[pick],
               "f_sum": f.sum(0), "u_sum": u.sum(0), "rho_sum": np.array(rho.sum()),
               "perm_txt": np.array(_text(wd, "Permeability_error_local_Perm.txt")),
               "params": np.array(json.dumps(dict(shape=[64, 64, 64], nu=1.0 / 6.0, g=[1e-6, 0.0, 0.0], u0=[0, 0, 0],
                                                  steps=steps, radius=6.0, porosity_target=0.6, seed=12345,
                                                  count=int(f.shape[0]), ranks=ranks, rank_counts=counts)))}
    finally:
        shutil.rmtree(wd, ignore_errors=True)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "count =", f.shape[0])


def main(only=()):
    """Write every fixture, or only those whose names are given on the command line."""
    _build_ref()
    cases = [
        ("sp_mrt_16x12x8", lambda n: sp_case(n, (16, 12, 8), 3.0, 0.6, 7, nu=0.1, g=(1e-5, 2e-6, -3e-6),
                                              u0=(1e-3, -2e-3, 5e-4), steps_list=[1, 101])),
        # PerDir = 3: the permeability files report the z component
        ("sp_perdir3_12x10x16", lambda n: sp_case(n, (12, 10, 16), 2.5, 0.55, 15, nu=0.12, g=(1e-6, 0.0, 2e-5),
                                                   u0=(0.0, 1e-3, 0.0), steps_list=[31], loc_perm=True, per_dir=3)),
        ("sp_tau1_box_20x10x12", lambda n: sp_case(n, (20, 10, 12), 2.5, 0.5, 11, nu=1.0 / 6.0, g=(1e-6, 0.0, 0.0),
                                                    u0=(0.0, 0.0, 0.0), steps_list=[1, 51],
                                                    box=[4, 15, 1, 8, 2, 10])),
        ("cg_binary_16x12x8", lambda n: cg_case(n, (16, 12, 8), 3.0, 0.6, 5, "binary", 6, nu_l=0.1, nu_g=0.05,
                                                 cos_theta=0.3, cap_a=0.01, g=(1e-5, 0.0, 0.0), steps_list=[1, 51])),
        ("cg_mixed_gated_14x10x8", lambda n: cg_case(n, (14, 10, 8), 2.5, 0.55, 9, "uniform", 10, nu_l=1.0 / 6.0,
                                                      nu_g=1.0 / 6.0, cos_theta=0.0, cap_a=0.005,
                                                      g=(2e-5, 1e-5, 0.0), steps_list=[1, 21], bodyforce_apply=1)),
        # PerDir = 2: the relative-permeability files report the y components and y-face pressures
        ("cg_perdir2_14x12x8", lambda n: cg_case(n, (14, 12, 8), 2.5, 0.55, 13, "binary", 14, nu_l=0.1, nu_g=0.1,
                                                  cos_theta=0.2, cap_a=0.01, g=(0.0, 2e-5, 0.0), steps_list=[1, 31],
                                                  per_dir=2)),
        ("st6_16x12x8", lambda n: st_case(n, "sp_mrt_16x12x8", niu_s=0.02, u_scale=100.0, c_seed=21, n_max=20,
                                          labels=[5, 20])),
        # face boundaries (SP:2751-3620), one fixture per rule, u0 != 0 so that the x- face
        # quirk of Sub_BC 1 and 3 (u_ls[2] = u[nb][1]) shows
        ("sp_bc_p0_16x12x10", lambda n: sp_case(n, (16, 12, 10), 3.0, 0.6, 31, nu=0.1, g=(0.0, 0.0, 0.0),
                                                 u0=(2e-3, 1e-3, -1e-3), steps_list=[1, 2, 21],
                                                 bc={"sub_bc": 0, "pre": {"xp": 1.0, "xn": 1.01}})),
        ("sp_bc_p1_16x12x10", lambda n: sp_case(n, (16, 12, 10), 3.0, 0.6, 32, nu=0.1, g=(0.0, 1e-5, 0.0),
                                                 u0=(2e-3, 1e-3, -1e-3), steps_list=[1, 2, 21],
                                                 bc={"sub_bc": 1, "pre": {"xp": 0.995, "xn": 1.005, "yp": 1.0}})),
        ("sp_bc_p2_16x12x10", lambda n: sp_case(n, (16, 12, 10), 3.0, 0.6, 33, nu=0.1, g=(0.0, 0.0, 0.0),
                                                 u0=(2e-3, 1e-3, -1e-3), steps_list=[1, 2, 21],
                                                 bc={"sub_bc": 2, "pre": {"xp": 1.0, "xn": 1.01, "zn": 1.002}})),
        ("sp_bc_p3_16x12x10", lambda n: sp_case(n, (16, 12, 10), 3.0, 0.6, 34, nu=0.1, g=(0.0, 0.0, 1e-5),
                                                 u0=(2e-3, 1e-3, -1e-3), steps_list=[1, 2, 21],
                                                 bc={"sub_bc": 3, "pre": {"xp": 1.0, "xn": 1.01, "yn": 1.0,
                                                                          "zp": 0.999}})),
        ("sp_bc_v1_16x12x10", lambda n: sp_case(n, (16, 12, 10), 3.0, 0.6, 35, nu=0.1, g=(0.0, 0.0, 0.0),
                                                 u0=(2e-3, 1e-3, -1e-3), steps_list=[1, 2, 21],
                                                 bc={"sub_bc": 1, "pre": {"xp": 1.0}, "vel": {"xn": 0.01, "zp": 0.002}})),
        # power-law (non-Newtonian) collision, collision_nnf (SP:2017-2700): switched on from
        # loop step n = 11, so the 12-step state is the first that shows it
        ("sp_nnf_16x12x8", lambda n: sp_case(n, (16, 12, 8), 3.0, 0.6, 7, nu=0.1, g=(1e-4, 2e-5, 0.0),
                                              u0=(1e-3, -2e-3, 5e-4), steps_list=[1, 11, 12, 13, 40],
                                              power_law=(0.02, 0.8))),
        ("sp_nnf_bc_16x12x10", lambda n: sp_case(n, (16, 12, 10), 3.0, 0.6, 36, nu=0.1, g=(0.0, 0.0, 0.0),
                                                  u0=(2e-3, 1e-3, -1e-3), steps_list=[12, 30],
                                                  bc={"sub_bc": 3, "pre": {"xp": 1.0, "xn": 1.02}},
                                                  power_law=(0.05, 1.3))),
        # the field writers: legacy VTK (velocity, density, geometry), the velocity profile and
        # the solute-velocity files at n = 0, 10, 20
        ("sp_c1_64cube", lambda n: c1_case(n)),
        # the local permeability (loc_perm = 1): the flux through the plane x = NX - 1 over the
        # per_* cross-section, with a partial box and body force, and with pressure faces (dp != 0)
        ("sp_locperm_box_20x10x12", lambda n: sp_case(n, (20, 10, 12), 2.5, 0.5, 11, nu=1.0 / 6.0,
                                                       g=(1e-5, 2e-6, -1e-6), u0=(0.0, 0.0, 0.0), steps_list=[21],
                                                       box=[4, 15, 1, 8, 2, 10], loc_perm=True)),
        ("sp_locperm_bc_16x12x10", lambda n: sp_case(n, (16, 12, 10), 3.0, 0.6, 37, nu=0.1, g=(2e-5, 0.0, 0.0),
                                                      u0=(2e-3, 1e-3, -1e-3), steps_list=[21],
                                                      bc={"sub_bc": 3, "pre": {"xp": 1.0, "xn": 1.01}},
                                                      loc_perm=True)),
        # the group permeability (gperm > 0): four cube grids, the last partly outside the box
        ("sp_gperm_20x10x12", lambda n: sp_case(n, (20, 10, 12), 2.5, 0.5, 12, nu=1.0 / 6.0, g=(1e-5, 0.0, 0.0),
                                                 u0=(0.0, 0.0, 0.0), steps_list=[21],
                                                 gperm={"every": 10, "n": [2, 1, 3, 2], "size": [4, 6, 3, 5],
                                                        "c": [[0, 0, 0], [5, 2, 3], [1, 1, 1], [12, 4, 6]]})),
        ("sp_out_16x12x8", lambda n: sp_case(n, (16, 12, 8), 3.0, 0.6, 7, nu=0.1, g=(1e-5, 2e-6, -3e-6),
                                              u0=(1e-3, -2e-3, 5e-4), steps_list=[21], field_outputs=True)),
        # the time loop's stop rules
        ("sp_stop_converged_16x12x8", lambda n: stop_case(n, "converged")),
        ("sp_stop_highvel_8x6x4", lambda n: stop_case(n, "high_velocity")),
        ("sp_stop_nan_16x12x8", lambda n: stop_case(n, "nan")),
    ]
    for name, make in cases:
        if not only or name in only:
            make(name)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
