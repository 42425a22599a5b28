"""Passive-solute (D3Q7) HIP kernels against the reference's six-velocity checkpoints and the
FP64 CPU oracle.

Every test runs the HIP library on the GPU through the C ABI and checks it with the oracle, the
reference program's own states (tests/golden/st6_16x12x8.npz, written by
Single_Phase_3D/Solute_Transport_MPI_Software_Spars_6V.cpp) or closed forms.  Tolerances: FP64
within 1e-14 absolute of the reference and the oracle (populations are O(0.1)); FP32 within
2e-6 relative; FP64 mass balance to 1e-12 relative.
"""
import sys

import numpy as np
import pytest

from conftest import load_fixture, rel_err

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _random_case(seed, shape=(18, 14, 12), porosity=0.7, umax=0.05):
    rng = np.random.default_rng(seed)
    solid = (rng.uniform(size=shape) > porosity).astype(np.int32)
    n = int(np.count_nonzero(solid == 0))
    u = rng.uniform(-umax, umax, (n, 3))
    conc = rng.uniform(0.0, 1.0, shape)
    return solid, u, conc


def test_fp64_w0_zero_matches_reference_checkpoints(pkg):
    """Rest weight 0 is the reference's six-velocity scheme: from its initial state the GPU
    follows its checkpoints at steps 6 and 21."""
    _log("[gpu] solute fp64 vs reference six-velocity checkpoints st6_16x12x8")
    d, p = load_fixture("st6_16x12x8")
    s = pkg.Solute(p["niu_s"], solid=d["solid"].astype(np.int32), w0=0.0, precision="fp64")
    assert s.count == d["g_0"].shape[0]
    s.set_velocity(d["u"])
    s.load_state(d["g_0"])
    g, c = s.state()
    assert np.array_equal(g, d["g_0"])          # a loaded state reads back unchanged
    done = 0
    for steps in p["steps"]:
        s.step(steps - done)
        done = steps
        g, c = s.state()
        assert np.abs(g - d[f"g_{steps}"]).max() < 1e-14
        assert np.abs(c - d[f"c_{steps}"]).max() < 1e-14


def test_fp64_init_is_reference_equilibrium(pkg):
    """init(C) gives g = C/6 (1 + 3 e.u) for w0 = 0: the reference's initial checkpoint."""
    d, p = load_fixture("st6_16x12x8")
    s = pkg.Solute(p["niu_s"], solid=d["solid"].astype(np.int32), w0=0.0, precision="fp64")
    s.set_velocity(d["u"])
    s.init(d["conc_field"])
    g, c = s.state()
    np.testing.assert_allclose(g, d["g_0"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(c, d["c_0"], rtol=0, atol=1e-15)


@pytest.mark.parametrize("w0", [0.0, 0.25])
def test_fp64_matches_oracle_step_by_step(pkg, oracle_mod, w0):
    _log(f"[gpu] solute fp64 vs oracle, w0={w0}")
    solid, u, conc = _random_case(11)
    D = 0.03
    s = pkg.Solute(D, solid=solid, w0=w0, precision="fp64")
    s.set_velocity(u)
    s.init(conc)
    o = oracle_mod.STOracle(solid, D, u, conc, w0=w0)
    assert s.count == o.count
    for steps in (1, 2, 7, 30):
        s.step(steps - o.steps)
        o.run(steps - o.steps)
        g, c = s.state()
        assert np.abs(g - o.g[1:]).max() < 1e-14
        assert np.abs(c - o.C[1:]).max() < 1e-14


def test_fp32_close_to_oracle(pkg, oracle_mod):
    solid, u, conc = _random_case(12)
    D, w0 = 0.05, 0.25
    s = pkg.Solute(D, solid=solid, w0=w0, precision="fp32")
    s.set_velocity(u)
    s.init(conc)
    o = oracle_mod.STOracle(solid, D, u, conc, w0=w0)
    s.step(100)
    o.run(100)
    g, c = s.state()
    err = rel_err(c, o.C[1:])
    _log(f"[gpu] solute fp32 C rel err after 100 steps {err:.2e}")
    assert err < 2e-6
    assert rel_err(g, o.g[1:]) < 2e-6


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_mass_conserved_with_walls(pkg, precision, tol):
    """Bounce-back at solid sites reflects solute: the FP64 sum of C over the fluid nodes is
    invariant (zero-flux walls, periodic box)."""
    solid, u, conc = _random_case(13, shape=(40, 36, 32), porosity=0.6)
    s = pkg.Solute(0.02, solid=solid, w0=0.25, precision=precision)
    s.set_velocity(u)
    s.init(conc)
    m0 = s.diagnostics()["mass"]
    s.step(500)
    d = s.diagnostics()
    drift = abs(d["mass"] - m0) / m0
    _log(f"[gpu] solute {precision} mass drift after 500 steps {drift:.2e}")
    assert d["nan_count"] == 0
    assert drift < tol


def test_diagnostics_match_state(pkg):
    solid, u, conc = _random_case(14)
    s = pkg.Solute(0.04, solid=solid, w0=0.25, precision="fp64")
    s.set_velocity(u)
    s.init(conc)
    for steps in (0, 9):
        s.step(steps)
        _, c = s.state()
        d = s.diagnostics()
        xyz = np.stack(np.nonzero(solid == 0), 1).astype(float)   # i-major, the node order
        np.testing.assert_allclose(d["mass"], c.sum(), rtol=1e-13)
        assert d["cmin"] == c.min() and d["cmax"] == c.max()
        np.testing.assert_allclose(d["moment1"], (xyz * c[:, None]).sum(0), rtol=1e-13)
        np.testing.assert_allclose(d["moment2"], (xyz ** 2 * c[:, None]).sum(0), rtol=1e-13)
        assert d["nan_count"] == 0


@pytest.mark.parametrize("w0", [0.0, 0.25])
def test_pulse_moments_closed_form(pkg, w0):
    """A unit pulse in an open box at equilibrium: the mean moves exactly u t, and at u = 0 the
    variance per axis grows by c_s^2 [(2 tau - 1) t - 2 tau (tau - 1)(1 - (1 - 1/tau)^t)]."""
    L, t, D = 48, 20, 0.05
    cs2 = (1 - w0) / 3
    tau = D / cs2 + 0.5
    solid = np.zeros((L, L, L), np.int32)
    conc = np.zeros((L, L, L))
    conc[L // 2, L // 2, L // 2] = 1.0
    for u in ((0.02, -0.01, 0.005), (0.0, 0.0, 0.0)):
        s = pkg.Solute(D, solid=solid, w0=w0, precision="fp64")
        s.set_velocity(np.tile(np.asarray(u), (L ** 3, 1)))
        s.init(conc)
        s.step(t)
        d = s.diagnostics()
        assert abs(d["mass"] - 1.0) < 1e-14
        mean = d["moment1"]
        np.testing.assert_allclose(mean - L // 2, np.asarray(u) * t, rtol=0, atol=1e-12)
        if not any(u):
            var = d["moment2"] - 2 * (L // 2) * mean + (L // 2) ** 2
            expect = cs2 * ((2 * tau - 1) * t - 2 * tau * (tau - 1) * (1 - (1 - 1 / tau) ** t))
            np.testing.assert_allclose(var, expect, rtol=1e-11)


def test_coupled_matches_oracle(pkg, oracle_mod):
    """Coupled: each step is one flow step and one solute step with the velocity that flow
    step collided with.  The oracle steps the single-phase and solute restatements the same
    way."""
    _log("[gpu] coupled flow + solute fp64 vs oracle")
    d, p = load_fixture("sp_mrt_16x12x8")
    solid = d["solid"].astype(np.int32)
    flow = pkg.SinglePhase(solid, p["nu"], p["g"], precision="fp64")
    flow.init(1.0, p["u0"])
    conc = np.random.default_rng(15).uniform(0, 1, solid.shape)
    D, w0 = 0.02, 0.25
    st = pkg.Solute(D, flow=flow, w0=w0, precision="fp64")
    st.init(conc)
    fo = oracle_mod.SPOracle(solid, p["nu"], p["g"], p["u0"])
    so = oracle_mod.STOracle(solid, D, fo.u[1:].copy(), conc, w0=w0)
    assert st.count == flow.count == so.count
    for steps in (1, 3, 10):
        st.step(steps - so.steps)
        while so.steps < steps:
            so.u[1:] = fo.u[1:]
            so.run(1)
            fo.run(1)
        g, c = st.state()
        f, u, rho = flow.state()
        assert rel_err(f, fo.f[1:]) < 1e-13
        assert np.abs(g - so.g[1:]).max() < 1e-14
        assert np.abs(c - so.C[1:]).max() < 1e-14
    assert flow.steps_done == 10
    flow.close()   # closes the attached solute first
    assert st._h is None


def test_coupled_refuses_mismatch_and_frozen_velocity(pkg):
    d, p = load_fixture("sp_mrt_16x12x8")
    flow = pkg.SinglePhase(d["solid"].astype(np.int32), p["nu"], p["g"], precision="fp32")
    with pytest.raises(pkg.LBMError):
        pkg.Solute(0.02, flow=flow, precision="fp64")
    st = pkg.Solute(0.02, flow=flow, precision="fp32")
    with pytest.raises(pkg.LBMError):
        st.set_velocity(np.zeros((st.count, 3)))
    with pytest.raises(pkg.LBMError):
        st.step(1)    # the solute has no state yet
    flow.init(1.0, p["u0"])
    st.init(np.ones(flow.shape))
    st.step(2)
    assert flow.steps_done == 2
    # streaming with bounce-back moves every population somewhere: the total is kept
    assert abs(st.diagnostics()["mass"] / st.count - 1.0) < 1e-6


def test_coupled_init_after_flow_steps(pkg, oracle_mod):
    """A solute attached to a flow that has already run starts from g_eq(C, u) with the
    velocity of the flow's current state."""
    d, p = load_fixture("sp_mrt_16x12x8")
    solid = d["solid"].astype(np.int32)
    flow = pkg.SinglePhase(solid, p["nu"], p["g"], precision="fp64")
    flow.init(1.0, p["u0"])
    flow.step(5)
    conc = np.random.default_rng(17).uniform(0, 1, solid.shape)
    st = pkg.Solute(0.03, flow=flow, w0=0.0, precision="fp64")
    st.init(conc)
    fo = oracle_mod.SPOracle(solid, p["nu"], p["g"], p["u0"])
    fo.run(5)
    so = oracle_mod.STOracle(solid, 0.03, fo.u[1:].copy(), conc, w0=0.0)
    g, _ = st.state()
    assert np.abs(g - so.g[1:]).max() < 1e-14
    for _ in range(4):
        so.u[1:] = fo.u[1:]
        so.run(1)
        fo.run(1)
    st.step(4)
    g, c = st.state()
    assert np.abs(g - so.g[1:]).max() < 1e-14
    assert flow.steps_done == 9


def test_step_before_init_refused(pkg):
    solid, u, conc = _random_case(16)
    s = pkg.Solute(0.02, solid=solid, precision="fp64")
    with pytest.raises(pkg.LBMError):
        s.step(1)
    with pytest.raises(pkg.LBMError):
        pkg.Solute(-1.0, solid=solid)
    with pytest.raises(pkg.LBMError):
        pkg.Solute(0.02, solid=solid, w0=1.0)
