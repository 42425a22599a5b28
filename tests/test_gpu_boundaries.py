"""Pressure and velocity faces (SP:2751-3620) on the GPU against the reference's checkpoints
and the FP64 oracle.

The fixtures are reference runs with pressure faces under Sub_BC 0, 1, 2 and 3 and a
velocity face under Sub_BC 1 (tests/golden/sp_bc_*.npz, written by
tests/golden/make_ref_fixtures.py).  Tolerances as for the periodic step: FP64 f within 1e-13
(norm-wise relative) of the reference, u within 1e-14 absolute; FP32 within 2e-5 of FP64.
"""
import sys

import numpy as np
import pytest

from conftest import load_fixture, rel_err

pytestmark = pytest.mark.gpu

BC_FIXTURES = ["sp_bc_p0_16x12x10", "sp_bc_p1_16x12x10", "sp_bc_p2_16x12x10", "sp_bc_p3_16x12x10",
               "sp_bc_v1_16x12x10"]


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _solver(pkg, d, p, precision="fp64"):
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision=precision)
    bc = p["bc"]
    s.set_boundaries(bc["sub_bc"], bc.get("pre"), bc.get("vel"))
    s.init(1.0, p["u0"])
    return s


@pytest.mark.parametrize("name", BC_FIXTURES)
def test_fp64_faces_match_reference_checkpoints(pkg, name):
    _log(f"[gpu] faces fp64 vs reference {name}")
    d, p = load_fixture(name)
    s = _solver(pkg, d, p)
    for steps in p["steps"]:
        s.step(steps - s.steps_done)
        f, u, rho = s.state()
        assert rel_err(f, d[f"f_{steps}"]) < 1e-13
        assert rel_err(rho, d[f"rho_{steps}"]) < 1e-13
        assert np.abs(u - d[f"u_{steps}"]).max() < 1e-14


@pytest.mark.parametrize("name", ["sp_bc_p2_16x12x10", "sp_bc_p3_16x12x10", "sp_bc_v1_16x12x10"])
def test_fp64_faces_restart_from_reference_checkpoint(pkg, name):
    """load_state of the reference's 2-step state, then 19 steps: its 21-step state."""
    d, p = load_fixture(name)
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp64")
    bc = p["bc"]
    s.set_boundaries(bc["sub_bc"], bc.get("pre"), bc.get("vel"))
    s.load_state(d["f_2"], d["u_2"], d["rho_2"])
    s.step(19)
    f, u, rho = s.state()
    assert rel_err(f, d["f_21"]) < 1e-13
    assert np.abs(u - d["u_21"]).max() < 1e-14


@pytest.mark.parametrize("sub_bc", [0, 1, 2, 3])
def test_fp64_all_faces_match_oracle(pkg, oracle_mod, sub_bc):
    """Pressure on all six faces (edges and corners resolved by the reference's face order),
    step by step against the oracle."""
    _log(f"[gpu] six pressure faces, Sub_BC {sub_bc}, vs oracle")
    rng = np.random.default_rng(40 + sub_bc)
    solid = (rng.uniform(size=(14, 11, 9)) < 0.25).astype(np.int32)
    pre = {"xp": 1.0, "xn": 1.004, "yp": 0.999, "yn": 1.001, "zp": 1.002, "zn": 0.998}
    nu, g, u0 = 0.08, (1e-5, -2e-6, 3e-6), (1e-3, 2e-3, -1e-3)
    s = pkg.SinglePhase(solid, nu, g, precision="fp64")
    s.set_boundaries(sub_bc, pre)
    s.init(1.0, u0)
    o = oracle_mod.SPOracle(solid, nu, g, u0, bc=oracle_mod.make_bc(sub_bc, pre))
    for steps in (1, 2, 3, 10, 40):
        s.step(steps - s.steps_done)
        o.run(steps - o.steps)
        f, u, rho = s.state()
        assert rel_err(f, o.f[1:]) < 1e-13
        assert np.abs(u - o.u[1:]).max() < 1e-14
        d = s.diagnostics()
        np.testing.assert_allclose(d["Q"], o.u[1:].sum(0), rtol=1e-12, atol=1e-16)


def test_pressure_and_velocity_faces_together_match_oracle(pkg, oracle_mod):
    rng = np.random.default_rng(50)
    solid = (rng.uniform(size=(12, 10, 12)) < 0.2).astype(np.int32)
    pre, vel = {"xp": 1.0, "yn": 1.002}, {"xn": 0.01, "zp": -0.004, "yn": 0.003}
    s = pkg.SinglePhase(solid, 0.1, (0.0, 0.0, 0.0), precision="fp64")
    s.set_boundaries(1, pre, vel)
    s.init(1.0, (0.0, 0.0, 0.0))
    o = oracle_mod.SPOracle(solid, 0.1, (0.0, 0.0, 0.0), bc=oracle_mod.make_bc(1, pre, vel))
    s.step(30)
    o.run(30)
    f, u, rho = s.state()
    assert rel_err(f, o.f[1:]) < 1e-13


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-5), ("fp16", 1e-1)])
def test_reduced_precision_faces_close_to_fp64(pkg, precision, tol):
    d, p = load_fixture("sp_bc_p3_16x12x10")
    ref = _solver(pkg, d, p)
    s = _solver(pkg, d, p, precision)
    ref.step(300)
    s.step(300)
    u64, u = ref.velocity(), s.velocity()
    err = rel_err(u, u64)
    _log(f"[gpu] faces {precision} u rel err after 300 steps {err:.2e}")
    assert err < tol


def test_faces_refused_after_stepping_and_unsupported(pkg):
    d, p = load_fixture("sp_bc_p0_16x12x10")
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp64")
    with pytest.raises(pkg.LBMError):
        s.set_boundaries(2, vel={"xn": 0.01})       # velocity Sub_BC 2 is not supported
    with pytest.raises(pkg.LBMError):
        s.set_boundaries(0, pre={"xp": -1.0})
    s.set_boundaries(0, pre={"xp": 1.0, "xn": 1.01})
    s.step(2)
    with pytest.raises(pkg.LBMError):
        s.set_boundaries(0, pre={"xp": 1.0})
    s.init(1.0, p["u0"])                            # a new start accepts new faces
    s.set_boundaries()                              # and no faces: the periodic step again
    o = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp64")
    o.init(1.0, p["u0"])
    s.step(5)
    o.step(5)
    assert np.array_equal(s.state()[0], o.state()[0])


def test_pressure_gradient_matches_reference_permeability(pkg):
    """k = Q/V nu/(g + dp) with the reference's dp = |p+ - p-| c_s^2 / L (SP:4600-4614) gives
    the permeability the reference printed (6 significant digits, x 1000)."""
    d, p = load_fixture("sp_bc_p0_16x12x10")
    s = _solver(pkg, d, p)
    s.step(21)
    dgn = s.diagnostics()
    V = np.prod(p["shape"])
    k = dgn["Q"][0] / V * p["nu"] / (p["g"][0] + s.pressure_gradient(1))
    printed = float(str(d["perm_txt_21"]).strip().splitlines()[-1].split()[0])
    assert abs(k * 1000 - printed) / printed < 1e-5
