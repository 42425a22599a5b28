"""Power-law (non-Newtonian) collision on the GPU against the reference's checkpoints and the
FP64 oracle.

The reference switches its collision_nnf (SP:2017-2700) on from loop step n = 11: the 11-step
state is still Newtonian, the 12-step state is the first with per-node viscosities.  Fixtures:
tests/golden/sp_nnf_16x12x8.npz (shear-thinning, body force) and sp_nnf_bc_16x12x10.npz
(shear-thickening, pressure faces).  Tolerances: FP64 f within 1e-13 (norm-wise relative),
u within 1e-14; the per-node viscosity within 1e-12 of the oracle's (1e-5 absolute where the
strain rate is round-off).
"""
import sys

import numpy as np
import pytest

from conftest import load_fixture, rel_err

pytestmark = pytest.mark.gpu

NNF_FIXTURES = ["sp_nnf_16x12x8", "sp_nnf_bc_16x12x10"]


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _solver(pkg, d, p, precision="fp64"):
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision=precision)
    if p.get("bc"):
        bc = p["bc"]
        s.set_boundaries(bc["sub_bc"], bc.get("pre"), bc.get("vel"))
    s.set_power_law(*p["power_law"])
    s.init(1.0, p["u0"])
    return s


def _oracle(oracle_mod, d, p):
    bc = p.get("bc")
    b = oracle_mod.make_bc(bc["sub_bc"], bc.get("pre"), bc.get("vel")) if bc else None
    return oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"], bc=b, power_law=p["power_law"])


@pytest.mark.parametrize("name", NNF_FIXTURES)
def test_fp64_power_law_matches_reference_checkpoints(pkg, name):
    _log(f"[gpu] power law fp64 vs reference {name}")
    d, p = load_fixture(name)
    s = _solver(pkg, d, p)
    for steps in p["steps"]:
        s.step(steps - s.steps_done)
        f, u, rho = s.state()
        assert rel_err(f, d[f"f_{steps}"]) < 1e-13
        assert np.abs(u - d[f"u_{steps}"]).max() < 1e-14


@pytest.mark.parametrize("name", NNF_FIXTURES)
def test_fp64_viscosity_matches_oracle(pkg, oracle_mod, name):
    d, p = load_fixture(name)
    s = _solver(pkg, d, p)
    o = _oracle(oracle_mod, d, p)
    for steps in (5, 11, 12, 20, 60):
        s.step(steps - s.steps_done)
        o.run(steps - o.steps)
        nu = s.viscosity()
        # nodes whose strain rate is round-off (f = feq on equilibrium faces) get
        # nu = m |round-off|^(n-1): ~1e-6 for the shear-thickening case, with no digits in common
        np.testing.assert_allclose(nu, o.niu_vn[1:], rtol=1e-12, atol=1e-5)
        if steps <= 11:
            assert np.all(nu == p["nu"])
        assert rel_err(s.state()[0], o.f[1:]) < 1e-13


def test_fp32_power_law_close_to_fp64(pkg):
    d, p = load_fixture("sp_nnf_16x12x8")
    ref = _solver(pkg, d, p)
    s = _solver(pkg, d, p, "fp32")
    ref.step(400)
    s.step(400)
    err = rel_err(s.velocity(), ref.velocity())
    nerr = rel_err(s.viscosity(), ref.viscosity())
    _log(f"[gpu] power law fp32: u rel err {err:.2e}, viscosity rel err {nerr:.2e}")
    assert err < 1e-4 and nerr < 1e-4


def test_restart_starts_newtonian_again(pkg, oracle_mod):
    """load_state resets every node's viscosity to nu and the step count: the power law acts
    again from the 12th step after the restart."""
    d, p = load_fixture("sp_nnf_16x12x8")
    s = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp64")
    s.set_power_law(*p["power_law"])
    s.load_state(d["f_12"], d["u_12"], d["rho_12"])
    o = _oracle(oracle_mod, d, p)
    o.load(d["f_12"], d["u_12"], d["rho_12"])
    s.step(25)
    o.run(25)
    assert rel_err(s.state()[0], o.f[1:]) < 1e-13
    np.testing.assert_allclose(s.viscosity(), o.niu_vn[1:], rtol=1e-12, atol=1e-5)


def test_power_law_off_and_refusals(pkg):
    d, p = load_fixture("sp_nnf_16x12x8")
    solid = d["solid"].astype(np.int32)
    s = pkg.SinglePhase(solid, p["nu"], p["g"], precision="fp64")
    s.set_power_law(*p["power_law"])
    s.set_power_law(0.0, 1.0)                     # off again
    s.init(1.0, p["u0"])
    n = pkg.SinglePhase(solid, p["nu"], p["g"], precision="fp64")
    n.init(1.0, p["u0"])
    s.step(20)
    n.step(20)
    assert np.array_equal(s.state()[0], n.state()[0])
    with pytest.raises(pkg.LBMError):
        s.set_power_law(*p["power_law"])          # after stepping
    fresh = pkg.SinglePhase(solid, p["nu"], p["g"], precision="fp64")
    with pytest.raises(pkg.LBMError):
        fresh.set_power_law(0.02, float("nan"))
