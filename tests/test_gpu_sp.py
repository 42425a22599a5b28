"""Single-phase HIP kernels against the reference's checkpoints and the FP64 CPU oracle.

Every test here runs the HIP library on the GPU through the C ABI (ctypes) and checks it with
the oracle or the reference's own fixture states, at sizes the oracle finishes in seconds.
Tolerances: FP64 within 1e-12 (norm-wise relative) of the oracle and the reference; FP32
permeability within 1e-4 relative of FP64 at steady state.
"""
import sys

import numpy as np
import pytest

from conftest import load_fixture, rel_err

pytestmark = pytest.mark.gpu

SP_FIXTURES = ["sp_mrt_16x12x8", "sp_tau1_box_20x10x12"]


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fixture_solver(pkg, name, precision="fp64"):
    d, p = load_fixture(name)
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision=precision)
    s.init(1.0, p["u0"])
    return s, d, p


@pytest.mark.parametrize("name", SP_FIXTURES)
def test_fp64_matches_reference_checkpoints(pkg, name):
    _log(f"[gpu] fp64 vs reference checkpoints {name}")
    s, d, p = _fixture_solver(pkg, name)
    assert s.count == d[f"rho_{p['steps'][0]}"].shape[0]
    done = 0
    for steps in p["steps"]:
        s.step(steps - done)
        done = steps
        f, u, rho = s.state()
        assert rel_err(f, d[f"f_{steps}"]) < 1e-13
        assert rel_err(rho, d[f"rho_{steps}"]) < 1e-13
        assert np.abs(u - d[f"u_{steps}"]).max() < 1e-14
        assert rel_err(u, d[f"u_{steps}"]) < 5e-10


@pytest.mark.parametrize("name", SP_FIXTURES)
def test_fp64_matches_oracle_step_by_step(pkg, oracle_mod, name):
    _log(f"[gpu] fp64 vs oracle {name}")
    s, d, p = _fixture_solver(pkg, name)
    o = oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"])
    for steps in (1, 2, 5, 30):
        s.step(steps - s.steps_done)
        o.run(steps - o.steps)
        f, u, rho = s.state()
        assert rel_err(f, o.f[1:]) < 1e-13
        assert np.abs(u - o.u[1:]).max() < 1e-14
        # diagnostics: Error() and the permeability flux, FP64 reductions on the device
        box = p["box"]
        diag = s.diagnostics(box)
        err, umax, ssum = o.error()
        Q = o.flux(box)
        assert abs(diag["err"] - err) <= 1e-9 * err
        # |u| carries the ~1e-17 absolute rounding of u = (sum e f + g/2)/rho
        assert abs(diag["umax"] - umax) <= 1e-14
        assert abs(diag["sum_speed"] - ssum) <= 1e-9 * ssum
        # components that nearly cancel over the box carry the summed ~1e-17 rounding of u
        np.testing.assert_allclose(diag["Q"], Q, rtol=1e-11, atol=1e-11 * ssum)
        assert diag["nan_count"] == 0


def test_restart_from_reference_checkpoint(pkg):
    """Backup_init path: load the reference's 1-step state, continue, compare at 101 steps."""
    _log("[gpu] restart from reference checkpoint")
    d, p = load_fixture("sp_mrt_16x12x8")
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp64")
    s.load_state(d["f_1"], d["u_1"], d["rho_1"])
    f, u, rho = s.state()
    np.testing.assert_array_equal(f, d["f_1"])
    s.step(100)
    f, u, rho = s.state()
    assert rel_err(f, d["f_101"]) < 1e-13
    assert np.abs(u - d["u_101"]).max() < 1e-14


def test_init_state_is_equilibrium(pkg, oracle_mod):
    _log("[gpu] init state")
    d, p = load_fixture("sp_mrt_16x12x8")
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp64")
    s.init(1.02, p["u0"])
    f, u, rho = s.state()
    o = oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"], rho0=1.02)
    np.testing.assert_array_equal(f, o.f[1:])
    np.testing.assert_array_equal(u, o.u[1:])
    np.testing.assert_array_equal(rho, o.rho[1:])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-5)])
def test_uniform_acceleration(pkg, precision, tol):
    _log(f"[gpu] uniform acceleration {precision}")
    solid = np.zeros((8, 6, 10), np.int32)
    g = np.array([1e-5, -2e-6, 3e-6])
    s = pkg.SinglePhase(solid, 0.1, g, precision=precision)
    for steps in (1, 10, 40):
        s.step(steps - s.steps_done)
        u = s.velocity()
        np.testing.assert_allclose(u, np.tile((steps + 0.5) * g, (s.count, 1)), rtol=tol, atol=1e-16)


def test_plane_poiseuille(pkg):
    _log("[gpu] plane Poiseuille")
    H, nu, g = 10, 0.1, 1e-6
    solid = np.zeros((4, H + 1, 3), np.int32)
    solid[:, 0, :] = 1
    y = np.arange(1, H + 1)
    exact = g / (2 * nu) * (y - 0.5) * (H + 0.5 - y)
    for precision, tol in (("fp64", 1e-9), ("fp32", 2e-5)):
        s = pkg.SinglePhase(solid, nu, (g, 0, 0), precision=precision)
        s.step(int(50 * H * H / nu))
        u = s.velocity()[:, 0].reshape(4, H, 3)
        np.testing.assert_allclose(u, np.broadcast_to(exact[None, :, None], u.shape), rtol=tol)


@pytest.mark.parametrize("shape,por", [((1, 1, 1), 1.0), ((5, 7, 9), 1.0), ((7, 9, 11), 0.7), ((16, 16, 17), 0.5)])
def test_edge_sizes_against_oracle(pkg, oracle_mod, shape, por):
    """Node counts below, at and across the block size; a single node; an all-fluid box."""
    _log(f"[gpu] edge size {shape} porosity {por}")
    if por >= 1.0:
        solid = np.zeros(shape, np.int32)
    else:
        solid = (np.random.default_rng(sum(shape)).uniform(size=shape) > por).astype(np.int32)
        solid.reshape(-1)[0] = 0
    g = (2e-5, 1e-5, -1e-5)
    s = pkg.SinglePhase(solid, 0.12, g, precision="fp64")
    s.init(1.0, (1e-3, 0, -2e-3))
    o = oracle_mod.SPOracle(solid, 0.12, g, (1e-3, 0, -2e-3))
    s.step(13)
    o.run(13)
    f, u, rho = s.state()
    assert f.shape[0] == o.count
    assert rel_err(f, o.f[1:]) < 1e-13
    assert np.abs(u - o.u[1:]).max() < 1e-14


def test_runs_are_bitwise_reproducible(pkg):
    _log("[gpu] reproducibility")
    d, p = load_fixture("sp_tau1_box_20x10x12")
    outs = []
    for _ in range(2):
        s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp32")
        s.step(37)
        outs.append((s.state(), s.diagnostics(p["box"])))
    for a, b in zip(outs[0][0], outs[1][0]):
        np.testing.assert_array_equal(a, b)
    for k in outs[0][1]:
        assert np.array_equal(np.asarray(outs[0][1][k]), np.asarray(outs[1][1][k])), k


def test_nan_is_reported(pkg):
    """A non-finite population is caught by the device check (the reference's NaN stop)."""
    _log("[gpu] nan detection")
    d, p = load_fixture("sp_mrt_16x12x8")
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp64")
    f = d["f_1"].copy()
    f[17, 3] = np.nan
    s.load_state(f, d["u_1"], d["rho_1"])
    s.step(2)
    assert s.diagnostics()["nan_count"] >= 1


@pytest.mark.parametrize("precision,k_tol", [("fp32", 1e-4), ("fp16", 2e-2)])
def test_reduced_precision_permeability_vs_fp64(pkg, precision, k_tol):
    """FP32 shifted populations (and FP16 storage) against the FP64 build on a sphere pack run
    towards steady state; the velocity L2 error and the permeability error are printed."""
    _log(f"[gpu] {precision} vs fp64 permeability")
    solid = pkg.geometry.sphere_pack((40, 40, 40), 5.0, 0.45, seed=21)
    runs = {}
    for prec in ("fp64", precision):
        s = pkg.SinglePhase(solid, 1.0 / 6.0, (1e-6, 0, 0), precision=prec)
        s.step(3000)
        k, _, _ = s.Comput_Perm(1)
        runs[prec] = (k[0], s.velocity())
    k64, u64 = runs["fp64"]
    kp, up = runs[precision]
    k_err = abs(kp - k64) / abs(k64)
    u_err = rel_err(up, u64)
    _log(f"[gpu] {precision}: permeability rel err {k_err:.3e}, velocity L2 rel err {u_err:.3e}")
    assert k_err < k_tol
    assert u_err < 10 * k_tol


def test_fp64_full_size_reference_case(pkg):
    """The reference's own case at full size (64^3 spheres, 157,141 fluid nodes, tau = 1,
    g = 1e-6, 2000 steps): every 64th node's state, the per-direction sums of f over all nodes
    and the permeability the reference printed every 100 steps."""
    _log("[gpu] fp64 full-size reference case 64^3 x 2000 steps")
    d, p = load_fixture("sp_c1_64cube")
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp64")
    assert s.count == p["count"]
    s.init(1.0, p["u0"])
    printed = [float(line.split()[0]) for line in str(d["perm_txt"]).strip().splitlines()]
    ks = []
    for n in range(0, p["steps"], 100):   # loop steps n = 0, 100, ..., 1900: after n + 1 steps
        s.step(n + 1 - s.steps_done)
        ks.append(s.Comput_Perm(1)[0][0] * 1000)
    s.step(p["steps"] - s.steps_done)
    f, u, rho = s.state()
    pick = d["pick"]
    # the reference's few-ulp M^-1 literals accumulate to ~1.4e-13 on f over 2000 steps
    assert np.abs(f[pick] - d["f_pick"]).max() < 5e-13
    assert np.abs(u[pick] - d["u_pick"]).max() < 2e-14
    np.testing.assert_allclose(f.sum(0), d["f_sum"], rtol=1e-12)
    np.testing.assert_allclose(u.sum(0), d["u_sum"], rtol=1e-8)
    assert len(printed) == len(ks)
    for a, b in zip(ks, printed):
        assert abs(a - b) <= 2e-5 * abs(b)    # six printed digits


def test_python_local_and_group_permeability_match_reference(pkg):
    """SinglePhase.Comput_Perm_LOCAL and Comput_Grop_Perm (SP:4831-5090) after the fixtures'
    21 steps against what the reference printed: Results.txt's LOCAL line and the
    Perm_<size>.output files (6 digits)."""
    d, p = load_fixture("sp_locperm_box_20x10x12")
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp64")
    s.init(1.0, p["u0"])
    s.step(p["steps"][-1])
    k = s.Comput_Perm_LOCAL(p["box"]) * 1000.0
    line = [ln for ln in str(d["results_txt_21"]).splitlines() if "LOCAL" in ln][-1]
    ref = [float(x) for x in line.split(":")[1].split(",")]
    np.testing.assert_allclose(k, ref, rtol=2e-5)

    d, p = load_fixture("sp_gperm_20x10x12")
    gp = p["gperm"]
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp64")
    s.init(1.0, p["u0"])
    s.step(p["steps"][-1])
    groups = s.Comput_Grop_Perm(gp["n"], gp["size"], gp["c"])
    last_of_size = {size: rows for size, rows in zip(gp["size"], groups)}   # a later grid rewrites the file
    for size, rows in last_of_size.items():
        ref = [[float(x) for x in ln.split()] for ln in str(d[f"gperm_{size}_21"]).strip().splitlines()]
        assert len(rows) == len(ref)
        for (kk, por), (rk, rp) in zip(rows, ref):
            assert abs(kk * 1000.0 - rk) <= 2e-5 * abs(rk) + 1e-9 and abs(por - rp) <= 1e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cube_sums_match_box_diagnostics(pkg, precision):
    """lbm_sp_cube_sums (all cubes of a grid in one launch) against one box diagnostic per cube:
    the same flux sums to round-off and the same node counts, for a grid that reaches past the
    box on every side, after pull steps and before the first step; the grid's node lists are
    kept, so a second call gives the same sums bit for bit."""
    _log(f"[gpu] cube sums {precision}")
    solid = pkg.geometry.sphere_pack((48, 40, 36), 4.0, 0.4, seed=11)
    s = pkg.SinglePhase(solid, 1.0 / 6.0, (1e-5, 2e-6, -1e-6), precision=precision)
    for steps in (0, 30):
        s.step(steps)
        for origin, size, m in (((-3, 2, -5), 9, 6), ((0, 0, 0), 12, 3), ((5, 7, 1), 1, 4)):
            got = s.cube_sums(origin, size, m)
            assert np.array_equal(got, s.cube_sums(origin, size, m))
            c = 0
            for ix in range(m):
                for iy in range(m):
                    for iz in range(m):
                        lo = [origin[0] + ix * size, origin[1] + iy * size, origin[2] + iz * size]
                        box = (lo[0], lo[0] + size - 1, lo[1], lo[1] + size - 1, lo[2], lo[2] + size - 1)
                        d = s.diagnostics(box)
                        assert got[c, 3] == d["count_in_box"], (origin, size, m, c)
                        scale = np.abs(d["Q"]).max() + 1e-300
                        assert np.abs(got[c, :3] - d["Q"]).max() <= 1e-12 * scale, (origin, size, m, c)
                        c += 1
    with pytest.raises(pkg.LBMError):
        s.cube_sums((0, 0, 0), 0, 2)
    with pytest.raises(pkg.LBMError):
        s.cube_sums((0, 0, 0), 4, 257)
    s.close()


@pytest.mark.parametrize("name", ["sp_mrt_16x12x8", "sp_tau1_box_20x10x12"])
def test_error_over_ranks_matches_reference_printed_values(pkg, oracle_mod, name):
    """lbm_sp_error_ranks with the reference's 2 ranks gives the u_max and velocity change its
    Velocity_ave_max.txt prints (6 digits), and the oracle's 2-rank Error to round-off."""
    _log(f"[gpu] Error over ranks {name}")
    s, d, p = _fixture_solver(pkg, name)
    o = oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"])
    last = p["steps"][-1]
    rows = [[float(x) for x in ln.split()] for ln in str(d[f"vel_txt_{last}"]).strip().splitlines()]
    for line_no, n in enumerate(range(0, last, p["freRe"])):
        s.step(n + 1 - s.steps_done)
        o.run(n + 1 - o.steps)
        err, umax, _ = s.Error(ranks=2)
        oerr, oumax, ossum = o.error(ranks=2)
        # |u| carries the ~1e-17 absolute rounding of u = (sum e f + g/2)/rho
        assert abs(err - oerr) <= 1e-9 * oerr and abs(umax - oumax) <= 1e-14
        assert abs(umax - rows[line_no][1]) <= 6e-6 * rows[line_no][1]
        assert abs(err - rows[line_no][2]) <= 6e-6 * rows[line_no][2]
        # one rank is the plain diagnostics
        dg = s.diagnostics()
        e1, um1, _ = s.Error(ranks=1)
        assert e1 == dg["err"] and um1 == dg["umax"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_error_closed_form_on_a_solid_free_box(pkg, precision):
    """No solid, uniform body force: u_n = (n + 1/2) g at every node after n steps, so Error's
    velocity change is 1/(n + 1/2) on any rank count and u_max = (n + 1/2)|g|."""
    _log(f"[gpu] Error closed form {precision}")
    solid = np.zeros((6, 5, 4), np.int32)
    g = np.array([3e-4, -1e-4, 2e-4])
    s = pkg.SinglePhase(solid, 0.1, g, precision=precision)
    s.init()
    tol = 1e-11 if precision == "fp64" else 1e-4   # FP32: u to ~1e-9 absolute
    for n in (2, 3, 7, 20):
        s.step(n - s.steps_done)
        for ranks in (1, 2, 3):
            err, umax, _ = s.Error(ranks=ranks)
            assert abs(err - 1.0 / (n + 0.5)) < tol * 1.0 / (n + 0.5), (n, ranks, err)
            assert abs(umax - (n + 0.5) * np.linalg.norm(g)) < tol * umax
    s.close()
