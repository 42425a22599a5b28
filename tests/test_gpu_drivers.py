"""The drop-in driver programs (lbm_sp, lbm_cg) on the reference's own input files.

Each case writes the INPUT file, geometry and phase file that produced a golden fixture,
runs the driver in a scratch directory, and compares its result files and checkpoint files
with what the reference wrote for the same input."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG_DIR, load_fixture, rel_err

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rows(text):
    return [[float(x) for x in ln.split()] for ln in text.strip().splitlines() if ln.strip()]


def _read_ckpt(path, width):
    a = np.fromfile(path, dtype="<f8")
    return (a.reshape(-1, width) if width > 1 else a)[1:]


def _run(exe, cwd, *args):
    proc = subprocess.run([os.path.join(PKG_DIR, exe), "INPUT.dat", *args], cwd=cwd, capture_output=True, text=True,
                          timeout=300)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr[-2000:]
    return proc.stdout


def _sp_case(pkg, tmp_path, name, n_max, **over):
    d, p = load_fixture(name)
    solid = d["solid"].astype(np.int32)
    pkg.geometry.write_binary(str(tmp_path / "geometry.bin"), solid)
    inp = pkg.inputs.SPInput(filename="geometry.bin", n_xyz=tuple(p["shape"]), n_max=n_max, reso=1.0,
                             g=tuple(p["g"]), in_vis=p["nu"], ini_v=tuple(p["u0"]), wr_per=1, PerDir=p.get("per_dir", 1),
                             freRe=p["freRe"], Out_Mode=0, outputfile="./", fre_backup=0, decbin=1)
    box = p["box"]
    if box is not None:
        inp.par_per = (1, 1, 1)
        inp.per_x, inp.per_y, inp.per_z = (box[1], box[0]), (box[3], box[2]), (box[5], box[4])
    if p.get("loc_perm"):
        inp.loc_perm = 1
    for k, v in over.items():
        setattr(inp, k, v)
    pkg.inputs.write_sp_input(inp, str(tmp_path / "INPUT.dat"))
    return d, p


@pytest.mark.parametrize("name", ["sp_mrt_16x12x8", "sp_tau1_box_20x10x12", "sp_perdir3_12x10x16"])
def test_lbm_sp_outputs_match_reference(pkg, tmp_path, name):
    _log(f"[gpu] lbm_sp on {name}")
    d, p = _sp_case(pkg, tmp_path, name, n_max=p_last(name) - 1)
    # the fixtures come from 2-rank runs of the reference: Error's columns as 2 ranks form them
    _run("lbm_sp", str(tmp_path), "--ranks", str(len(p["rank_counts"])))
    last = p["steps"][-1]
    got = _rows(open(tmp_path / "Permeability_error_local_Perm.txt").read())
    ref = _rows(str(d[f"perm_txt_{last}"]))
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert abs(g[0] - r[0]) <= 2e-5 * abs(r[0])             # k, 6 significant digits
        if np.isfinite(r[1]) and abs(r[1]) > 1e-3:
            assert abs(g[1] - r[1]) <= 1e-4 * abs(r[1]) + 1e-6   # relative change of k
        assert abs(g[2] - r[2]) <= 2e-5 * abs(r[2]) + 1e-12       # local k (0 with loc_perm off)
    vg = _rows(open(tmp_path / "Velocity_ave_max.txt").read())
    vr = _rows(str(d[f"vel_txt_{last}"]))
    assert len(vg) == len(vr)
    # u_ave (Comput_Perm's), u_max and Error's relative velocity change, both of the latter the
    # largest over the reference's ranks, each rank skipping its own last node (SP:3651-3685):
    # all three to the 6 digits printed
    for g, r in zip(vg, vr):
        for c in range(3):
            assert abs(g[c] - r[c]) <= 6e-6 * abs(r[c]), (c, g, r)
    gen = open(tmp_path / "General_Results.txt").read().split()
    ref_gen = str(d[f"general_txt_{last}"]).split()
    assert gen[0] == ref_gen[0] == "geometry.bin" and len(gen) == len(ref_gen)
    for a, b in zip(gen[1:], ref_gen[1:]):                  # k along x, y, z (inf where g_d = 0)
        assert a == b or abs(float(a) - float(b)) <= 2e-5 * abs(float(b)), (gen, ref_gen)
    # the final checkpoint (label n_max) is the reference's state after n_max + 1 steps
    f = _read_ckpt(tmp_path / f"LBM_checkpoint_f_{last - 1}.0.bin_input", 19)
    u = _read_ckpt(tmp_path / f"LBM_checkpoint_velocity_{last - 1}.0.bin_input", 3)
    assert rel_err(f, d[f"f_{last}"]) < 1e-13
    assert np.abs(u - d[f"u_{last}"]).max() < 1e-14


def p_last(name):
    return load_fixture(name)[1]["steps"][-1]


def test_lbm_sp_restart_from_its_own_checkpoint(pkg, tmp_path):
    """Backup at label 50 (51 steps), restart with mode_backup_ini = 50 for 50 more steps,
    compare with the reference's state after 101 steps."""
    _log("[gpu] lbm_sp restart")
    d, p = _sp_case(pkg, tmp_path, "sp_mrt_16x12x8", n_max=50)
    _run("lbm_sp", str(tmp_path))
    _sp_case(pkg, tmp_path, "sp_mrt_16x12x8", n_max=49, mode_backup_ini=50)
    _run("lbm_sp", str(tmp_path), "./")
    f = _read_ckpt(tmp_path / "LBM_checkpoint_f_49.0.bin_input", 19)
    assert rel_err(f, d["f_101"]) < 1e-13


def test_lbm_sp_fp32_precision_flag(pkg, tmp_path):
    d, p = _sp_case(pkg, tmp_path, "sp_mrt_16x12x8", n_max=100)
    _run("lbm_sp", str(tmp_path), "--precision", "fp32")
    f = _read_ckpt(tmp_path / "LBM_checkpoint_f_100.0.bin_input", 19)
    assert rel_err(f, d["f_101"]) < 1e-6


@pytest.mark.parametrize("name", ["cg_binary_16x12x8", "cg_mixed_gated_14x10x8", "cg_perdir2_14x12x8"])
def test_lbm_cg_outputs_match_reference(pkg, tmp_path, name):
    _log(f"[gpu] lbm_cg on {name}")
    d, p = load_fixture(name)
    last = p["steps"][-1]
    pkg.geometry.write_binary(str(tmp_path / "geometry.bin"), d["solid"].astype(np.int32))
    pkg.geometry.write_phase_ascii(str(tmp_path / "phase.dat"), d["psi_field"])
    inp = pkg.inputs.CGInput(filename="geometry.bin", filenamepsi="phase.dat", n_xyz=tuple(p["shape"]),
                             n_max=last - 1, reso=1.0, g=tuple(p["g"]), niu_l=p["nu_l"], niu_g=p["nu_g"],
                             cos_theta=p["cos_theta"], CapA=p["cap_a"], Permeability=1.0, wr_per=1,
                             PerDir=p.get("per_dir", 1),
                             freRe=p["freRe"], outputfile="./", stab=(0, 0), ini_Sat=-1.0,
                             rel_perm_psi=tuple(p["rel_perm_psi"]), fre_backup=0, decbin=1,
                             bodyforce_apply=p["bodyforce_apply"])
    inp.least_square = (0, 10)
    pkg.inputs.write_cg_input(inp, str(tmp_path / "INPUT.dat"))
    _run("lbm_cg", str(tmp_path))
    bg = _rows(open(tmp_path / "bodyforce.txt").read())
    br = _rows(str(d[f"bodyforce_txt_{last}"]))
    assert len(bg) == len(br)
    # S_l counts psi >= 0: nodes whose psi is zero up to round-off (exact colour ties, common
    # right after a +-1 start) may fall on either side; the first row's band is the number of
    # such nodes in the reference's own 1-step state, later rows allow two nodes
    n = d["psi_1"].size
    ties = int(np.count_nonzero(np.abs(d["psi_1"]) < 1e-12))
    for i, (g, r) in enumerate(zip(bg, br)):
        band = (ties if i == 0 else 2) / n + 2e-6
        assert abs(g[0] - r[0]) <= band
        assert abs(g[3] - r[3]) <= 2e-3 * abs(r[3]) + 0.5 and abs(g[4] - r[4]) <= 2e-3 * abs(r[4]) + 0.5
    # columns: k, saturation, the local k through the plane x = NX - 1 (Comput_Perm_LOCAL,
    # CG:6214-6290), ..., the PerDir face pressures
    for comp in (1, 2):
        kg = _rows(open(tmp_path / f"Relative_Permeability_Component{comp}.txt").read())
        kr = _rows(str(d[f"relperm{comp}_txt_{last}"]))
        assert len(kg) == len(kr)
        for g, r in zip(kg, kr):
            for c in (0, 2):
                assert abs(g[c] - r[c]) <= 1e-4 * abs(r[c]) + 1e-6 * max(1.0, abs(r[c])), (comp, c, g, r)
            if comp == 1:
                assert g[4:7] == r[4:7]
    f = _read_ckpt(tmp_path / f"LBM_checkpoint_f_{last - 1}.0.bin_input", 19)
    psi = _read_ckpt(tmp_path / f"LBM_checkpoint_psi_{last - 1}.0.bin_input", 1)
    assert rel_err(f, d[f"f_{last}"]) < 1e-12
    assert np.abs(psi - d[f"psi_{last}"]).max() < 1e-12


def _bc_over(bc):
    """INPUT fields of a fixture's faces (SP:260-267, 280)."""
    over = {"in_BC": 1, "Sub_BC": bc["sub_bc"]}
    for kind, attr, off in (("pre", "pre_", 1.0), ("vel", "vel_", 0.0)):
        vals = bc.get(kind, {})
        for ax in "xyz":
            pv, nv = vals.get(ax + "p"), vals.get(ax + "n")
            over[attr + ax] = (int(pv is not None), off if pv is None else pv, int(nv is not None),
                               off if nv is None else nv)
    return over


@pytest.mark.parametrize("name", ["sp_bc_p0_16x12x10", "sp_bc_p3_16x12x10", "sp_bc_v1_16x12x10"])
def test_lbm_sp_faces_match_reference(pkg, tmp_path, name):
    """The driver on an INPUT file with pressure / velocity faces: the reference's printed
    permeability (with its dp term) and its final checkpoint."""
    _log(f"[gpu] lbm_sp with faces on {name}")
    d0, p0 = load_fixture(name)
    d, p = _sp_case(pkg, tmp_path, name, n_max=p0["steps"][-1] - 1, **_bc_over(p0["bc"]))
    _run("lbm_sp", str(tmp_path))
    last = p["steps"][-1]
    got = _rows(open(tmp_path / "Permeability_error_local_Perm.txt").read())
    ref = _rows(str(d[f"perm_txt_{last}"]))
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        if np.isfinite(r[0]):
            assert abs(g[0] - r[0]) <= 2e-5 * abs(r[0])
    f = _read_ckpt(tmp_path / f"LBM_checkpoint_f_{last - 1}.0.bin_input", 19)
    assert rel_err(f, d[f"f_{last}"]) < 1e-13


@pytest.mark.parametrize("name", ["sp_nnf_16x12x8", "sp_nnf_bc_16x12x10"])
def test_lbm_sp_power_law_matches_reference(pkg, tmp_path, name):
    """nnf_module = 1 in the INPUT file: the reference's permeability output and checkpoint."""
    _log(f"[gpu] lbm_sp power law on {name}")
    d0, p0 = load_fixture(name)
    over = {"nnf_module": 1, "nnf_mn": tuple(p0["power_law"])}
    if p0.get("bc"):
        over.update(_bc_over(p0["bc"]))
    d, p = _sp_case(pkg, tmp_path, name, n_max=p0["steps"][-1] - 1, **over)
    _run("lbm_sp", str(tmp_path))
    last = p["steps"][-1]
    got = _rows(open(tmp_path / "Permeability_error_local_Perm.txt").read())
    ref = _rows(str(d[f"perm_txt_{last}"]))
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        if np.isfinite(r[0]):
            assert abs(g[0] - r[0]) <= 2e-5 * abs(r[0])
    f = _read_ckpt(tmp_path / f"LBM_checkpoint_f_{last - 1}.0.bin_input", 19)
    assert rel_err(f, d[f"f_{last}"]) < 1e-13


@pytest.mark.parametrize("name", ["sp_locperm_box_20x10x12", "sp_locperm_bc_16x12x10"])
def test_lbm_sp_local_permeability_matches_reference(pkg, tmp_path, name):
    """loc_perm = 1 (Comput_Perm_LOCAL, SP:4831-4883): the flux through the plane x = NX - 1
    over the per_* cross-section, printed in Results.txt (all three directions) and in the third
    column of Permeability_error_local_Perm.txt, as the reference printed it (6 digits)."""
    _log(f"[gpu] lbm_sp local permeability on {name}")
    d0, p0 = load_fixture(name)
    over = {"loc_perm": 1}
    if p0.get("bc"):
        over.update(_bc_over(p0["bc"]))
    d, p = _sp_case(pkg, tmp_path, name, n_max=p0["steps"][-1] - 1, **over)
    _run("lbm_sp", str(tmp_path))
    last = p["steps"][-1]

    def local_lines(text):
        return [[float(x) for x in ln.split(":")[1].split(",")] for ln in text.splitlines() if "LOCAL" in ln]

    got, ref = local_lines(open(tmp_path / "Results.txt").read()), local_lines(str(d[f"results_txt_{last}"]))
    assert len(got) == len(ref) >= 2 and any(abs(v) > 0 for row in ref for v in row)
    for g, r in zip(got, ref):
        for a in range(3):
            assert abs(g[a] - r[a]) <= 2e-5 * abs(r[a]) + 1e-9, (g, r)
    got_p = _rows(open(tmp_path / "Permeability_error_local_Perm.txt").read())
    ref_p = _rows(str(d[f"perm_txt_{last}"]))
    assert len(got_p) == len(ref_p)
    for g, r in zip(got_p, ref_p):
        assert abs(g[2] - r[2]) <= 2e-5 * abs(r[2])


def test_lbm_sp_group_permeability_matches_reference(pkg, tmp_path):
    """gperm > 0 (Comput_Grop_Perm, SP:4886-5090): for four cube grids (the last partly outside
    the box), the permeability along PerDir and the porosity of each cube in Perm_<size>.output,
    as the reference printed them (6 digits)."""
    _log("[gpu] lbm_sp group permeability")
    d0, p0 = load_fixture("sp_gperm_20x10x12")
    gp = p0["gperm"]
    over = {"gperm": gp["every"], "n_gperm": tuple(gp["n"]), "size_gperm": tuple(gp["size"])}
    for i, c in enumerate(gp["c"]):
        over[f"c_gperm{i + 1}"] = tuple(c)
    d, p = _sp_case(pkg, tmp_path, "sp_gperm_20x10x12", n_max=p0["steps"][-1] - 1, **over)
    _run("lbm_sp", str(tmp_path))
    last = p["steps"][-1]
    for size in sorted(set(gp["size"])):
        got = _rows(open(tmp_path / f"Perm_{size}.output").read())
        ref = _rows(str(d[f"gperm_{size}_{last}"]))
        assert len(got) == len(ref) > 0
        for g, r in zip(got, ref):
            assert abs(g[0] - r[0]) <= 2e-5 * abs(r[0]) + 1e-9, (size, g, r)
            assert abs(g[1] - r[1]) <= 1e-6, (size, g, r)


def test_driver_refuses_out_of_path_features(pkg, tmp_path):
    _sp_case(pkg, tmp_path, "sp_mrt_16x12x8", n_max=3, lattice=(1, 1.0, 2.0))
    proc = subprocess.run([os.path.join(PKG_DIR, "lbm_sp"), "INPUT.dat"], cwd=str(tmp_path), capture_output=True,
                          text=True, timeout=60)
    assert proc.returncode != 0 and "not part of this build" in proc.stderr


def test_lbm_sp_field_writers_match_reference_files(pkg, tmp_path):
    """freVe / freDe / vel_sol: the geometry, velocity and density VTK files, the velocity
    profile and the solute-velocity files the reference wrote in the same run, text for text."""
    _log("[gpu] lbm_sp field writers")
    d, p = _sp_case(pkg, tmp_path, "sp_out_16x12x8", n_max=20, freVe=10, freDe=10, vel_sol=10)
    _run("lbm_sp", str(tmp_path))
    assert open(tmp_path / "LBM_Geometry.vtk").read() == str(d["vtk_geometry"])
    for n in (0, 10, 20):
        for ours, key in ((f"LBM_velocity_Vector_{n}.vtk", f"vtk_velocity_{n}"), (f"LBM_Density_{n}.vtk", f"vtk_density_{n}"),
                          (f"LBM_velocity_{n}.out", f"profile_{n}")):
            got, ref = open(tmp_path / ours).read(), str(d[key])
            if got != ref:   # six printed digits: only a value within round-off of a rounding boundary may differ
                gl, rl = got.splitlines(), ref.splitlines()
                assert len(gl) == len(rl)
                diff = [(a, b) for a, b in zip(gl, rl) if a != b]
                assert len(diff) <= 2, diff[:4]
                for a, b in diff:
                    np.testing.assert_allclose(np.array(a.split(), float), np.array(b.split(), float), rtol=2e-6)
        if n > 0:
            u = np.fromfile(tmp_path / f"Velocity_for_solute_{n}.0.bin", dtype="<f8").reshape(-1, 3)
            assert np.abs(u - d[f"solute_velocity_{n}"]).max() < 1e-14


# ----------------------------------------------------------------------------- stop rules
def _stop_case(pkg, tmp_path, name, **over):
    d, p = load_fixture(name)
    solid = d["solid"].astype(np.int32)
    pkg.geometry.write_binary(str(tmp_path / "geometry.bin"), solid)
    kw = dict(filename="geometry.bin", n_xyz=tuple(p["shape"]), n_max=p["n_max"], reso=1.0, g=tuple(p["g"]),
              in_vis=p["nu"], ini_v=(0.0, 0.0, 0.0), wr_per=1, PerDir=1, freRe=p["freRe"],
              Out_Mode=p.get("Out_Mode", 0), outputfile="./", fre_backup=0, decbin=1)
    kw.update(over)
    pkg.inputs.write_sp_input(pkg.inputs.SPInput(**kw), str(tmp_path / "INPUT.dat"))
    return d, p


def _vel_rows_match(got_text, ref_text):
    got, ref = _rows(got_text), _rows(ref_text)
    assert len(got) == len(ref), (len(got), len(ref))
    for g, r in zip(got, ref):
        for c in range(3):
            if np.isnan(r[c]):
                assert np.isnan(g[c]), (c, g, r)
            else:
                assert abs(g[c] - r[c]) <= 6e-6 * abs(r[c]), (c, g, r)


def test_lbm_sp_converged_stop_matches_reference(pkg, tmp_path):
    """Out_Mode = 1 (SP:812-816): the run stops at the first check where Comput_Perm's mean
    speed changes by less than 1e-7 relative, writes the velocity files of that step, and the
    final checkpoint (label n_max) holds the state of that step; every printed line and the
    checkpoint match the reference run."""
    _log("[gpu] lbm_sp Out_Mode = 1 stop")
    d, p = _stop_case(pkg, tmp_path, "sp_stop_converged_16x12x8")
    _run("lbm_sp", str(tmp_path), "--ranks", "2")
    n_stop = p["n_stop"]
    _vel_rows_match(open(tmp_path / "Velocity_ave_max.txt").read(), str(d["vel_txt"]))
    vtk = sorted(f for f in os.listdir(tmp_path) if f.startswith("LBM_velocity_Vector_"))
    assert vtk == [f"LBM_velocity_Vector_{n_stop}.vtk"]
    for ours, key in ((vtk[0], "vtk_velocity"), (f"LBM_velocity_{n_stop}.out", "profile")):
        got, ref = open(tmp_path / ours).read().splitlines(), str(d[key]).splitlines()
        assert len(got) == len(ref)
        diff = [(a, b) for a, b in zip(got, ref) if a != b]
        assert len(diff) <= 2, diff[:4]    # six printed digits: a value at a rounding boundary
        for a, b in diff:
            np.testing.assert_allclose(np.array(a.split(), float), np.array(b.split(), float), rtol=2e-6)
    f = _read_ckpt(tmp_path / f"LBM_checkpoint_f_{p['label']}.0.bin_input", 19)
    u = _read_ckpt(tmp_path / f"LBM_checkpoint_velocity_{p['label']}.0.bin_input", 3)
    assert rel_err(f, d["f_end"]) < 1e-13
    assert np.abs(u - d["u_end"]).max() < 1e-14


def test_lbm_sp_high_velocity_stop_matches_reference(pkg, tmp_path):
    """u_max >= 10 at five checks (SP:712, 839): the run stops after the fifth and writes its
    final checkpoint.  Uniform flow in a box without solid reaches u = (n + 1/2) g; the printed
    lines (u_max, and Error's change 1/(n + 1/2)) and the checkpoint match the reference's rank 0,
    the one rank that leaves its loop there."""
    _log("[gpu] lbm_sp high-velocity stop")
    d, p = _stop_case(pkg, tmp_path, "sp_stop_highvel_8x6x4")
    out = _run("lbm_sp", str(tmp_path), "--ranks", "2")
    assert "PROGRAM STOP DUE TO HIGH VELOCITY" in out
    _vel_rows_match(open(tmp_path / "Velocity_ave_max.txt").read(), str(d["vel_txt"]))
    rows = _rows(str(d["vel_txt"]))
    assert sum(r[1] >= 10.0 for r in rows) == 5 and rows[-1][1] >= 10.0
    n0 = p["rank_counts"][0]
    f = _read_ckpt(tmp_path / f"LBM_checkpoint_f_{p['label']}.0.bin_input", 19)
    # populations up to |f| ~ 60 at u ~ 11: the reference's M^-1 literals, a few ulp off the exact
    # inverse, leave ~3e-13 after 44 such steps
    assert rel_err(f[:n0], d["f_end"]) < 1e-12


def test_lbm_sp_nan_stop_matches_reference(pkg, tmp_path):
    """A restart from a state with non-finite populations at one node of each reference rank's
    slab: the first check's velocity change is NaN and the run stops there (SP:838); u_max
    leaves the non-finite nodes out as the reference's comparison does (SP:3660), and the final
    checkpoint has NaN at the same nodes and the reference's values elsewhere."""
    _log("[gpu] lbm_sp NaN stop")
    d, p = _stop_case(pkg, tmp_path, "sp_stop_nan_16x12x8", mode_backup_ini=p_start(), fre_backup=0)
    lab = p["start_label"]
    for key, name, w in (("f_start", "f", 19), ("u_start", "velocity", 3), ("rho_start", "rho", 1)):
        a = d[key].reshape(-1, w)
        np.concatenate([np.zeros((1, w)), a]).tofile(tmp_path / f"LBM_checkpoint_{name}_{lab}.0.bin_input")
    out = _run("lbm_sp", str(tmp_path), "--ranks", "2")
    assert "PROGRAM STOP" in out
    _vel_rows_match(open(tmp_path / "Velocity_ave_max.txt").read(), str(d["vel_txt"]))
    f = _read_ckpt(tmp_path / f"LBM_checkpoint_f_{p['label']}.0.bin_input", 19)
    ref = d["f_end"]
    bad = ~np.isfinite(ref)
    assert bad.any() and np.array_equal(~np.isfinite(f), bad)
    assert np.abs(f[~bad] - ref[~bad]).max() < 1e-14


def p_start():
    return load_fixture("sp_stop_nan_16x12x8")[1]["start_label"]
