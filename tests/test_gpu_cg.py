"""Colour-gradient HIP kernels against the reference's checkpoints and the FP64 CPU oracle."""
import sys

import numpy as np
import pytest

from conftest import load_fixture, rel_err

pytestmark = pytest.mark.gpu

CG_FIXTURES = ["cg_binary_16x12x8", "cg_mixed_gated_14x10x8"]


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cg_from_fixture(pkg, d, p, precision="fp64"):
    s = pkg.ColorGradient(d["solid"].astype(np.int32), p["nu_l"], p["nu_g"], p["cos_theta"], p["cap_a"], p["g"],
                          precision=precision, bodyforce_apply=p["bodyforce_apply"],
                          rel_perm_psi=p["rel_perm_psi"][0], rel_perm_psi2=p["rel_perm_psi"][1])
    s.init(d["psi_field"])
    return s


def _oracle_from_fixture(oracle_mod, d, p):
    return oracle_mod.CGOracle(d["solid"].astype(np.int32), d["psi_field"], p["nu_l"], p["nu_g"], p["g"], p["cap_a"],
                               p["cos_theta"], bodyforce_apply=p["bodyforce_apply"],
                               rel_perm_psi=p["rel_perm_psi"][0], rel_perm_psi2=p["rel_perm_psi"][1])


@pytest.mark.parametrize("name", CG_FIXTURES)
def test_fp64_matches_reference_checkpoints(pkg, name):
    _log(f"[gpu] cg fp64 vs reference checkpoints {name}")
    d, p = load_fixture(name)
    s = _cg_from_fixture(pkg, d, p)
    assert s.count == d[f"rho_{p['steps'][0]}"].shape[0]
    for steps in p["steps"]:
        s.step(steps - s.steps_done)
        f, u, rho, psi, rr, rb = s.state()
        assert rel_err(f, d[f"f_{steps}"]) < 1e-12
        assert rel_err(rho, d[f"rho_{steps}"]) < 1e-12
        assert np.abs(u - d[f"u_{steps}"]).max() < 1e-13
        assert np.abs(psi - d[f"psi_{steps}"]).max() < 1e-12
        assert np.abs(rr - d[f"rhor_{steps}"]).max() < 1e-12
        assert np.abs(rb - d[f"rhob_{steps}"]).max() < 1e-12


@pytest.mark.parametrize("name", CG_FIXTURES)
def test_fp64_matches_oracle_with_diagnostics(pkg, oracle_mod, name):
    _log(f"[gpu] cg fp64 vs oracle {name}")
    d, p = load_fixture(name)
    s = _cg_from_fixture(pkg, d, p)
    o = _oracle_from_fixture(oracle_mod, d, p)
    box = [2, d["solid"].shape[0] - 3, 1, d["solid"].shape[1] - 2, 0, d["solid"].shape[2] - 1]
    for steps in (1, 3, 12):
        s.step(steps - s.steps_done)
        o.run(steps - o.steps)
        f, u, rho, psi, rr, rb = s.state()
        assert rel_err(f, o.f[1:]) < 1e-12
        assert np.abs(psi - o.psi[1:]).max() < 1e-12
        for bx in (None, box):
            diag = s.diagnostics(bx)
            Ql, Qg = o.flux(bx)
            scale = np.abs(o.u[1:]).sum()
            # phase membership is decided by psi thresholds: a node at round-off distance from a
            # threshold may switch sides, which the tolerance below does not allow, so check
            psi_o = o.psi[1:]
            margin = min(np.abs(psi_o - p["rel_perm_psi"][0]).min(), np.abs(psi_o + p["rel_perm_psi"][1]).min(),
                         np.abs(psi_o).min())
            if margin > 1e-12:
                np.testing.assert_allclose(diag["Ql"], Ql, rtol=1e-10, atol=1e-11 * scale)
                np.testing.assert_allclose(diag["Qg"], Qg, rtol=1e-10, atol=1e-11 * scale)
                assert diag["wet_count"] == o.wet_count(bx)
        err, umax, ssum = o.error()
        diag = s.diagnostics()
        assert abs(diag["err"] - err) <= 1e-8 * err
        assert abs(diag["umax"] - umax) <= 1e-13
        assert diag["nan_count"] == 0


def test_colour_mass_is_conserved(pkg):
    """Recolouring moves colour between the two fields pairwise and colour bounce-back keeps
    it on the node, so sum rho_r and sum rho_b are each invariant."""
    _log("[gpu] cg colour conservation")
    solid = pkg.geometry.sphere_pack((24, 20, 16), 3.5, 0.55, seed=4)
    psi0 = np.where(np.random.default_rng(8).uniform(size=solid.shape) < 0.5, 1.0, -1.0)
    for precision, tol in (("fp64", 1e-12), ("fp32", 2e-6)):
        s = pkg.ColorGradient(solid, 0.1, 0.05, 0.4, 0.01, (1e-5, 0, 0), precision=precision)
        s.init(psi0)
        _, _, _, _, rr0, rb0 = s.state()
        s.step(200)
        _, _, rho, _, rr, rb = s.state()
        _log(f"[gpu] {precision}: colour mass drift over 200 steps r {abs(rr.sum() / rr0.sum() - 1):.2e} "
             f"b {abs(rb.sum() / rb0.sum() - 1):.2e}")
        assert abs(rr.sum() - rr0.sum()) < tol * rr0.sum()
        assert abs(rb.sum() - rb0.sum()) < tol * rb0.sum()
        assert abs(rho.sum() - s.count) < tol * s.count


def test_restart_from_reference_checkpoint(pkg):
    _log("[gpu] cg restart from reference checkpoint")
    d, p = load_fixture("cg_binary_16x12x8")
    s = pkg.ColorGradient(d["solid"].astype(np.int32), p["nu_l"], p["nu_g"], p["cos_theta"], p["cap_a"], p["g"],
                          precision="fp64")
    s.load_state(d["f_1"], d["u_1"], d["rho_1"], d["psi_1"], d["rhor_1"], d["rhob_1"])
    s.step(50)
    f, u, rho, psi, rr, rb = s.state()
    assert rel_err(f, d["f_51"]) < 1e-12
    assert np.abs(psi - d["psi_51"]).max() < 1e-12


def test_single_phase_limit(pkg):
    """psi = 1 everywhere with a wetting wall (cos theta = 1): no colour gradient, so the CG
    solver is the single-phase MRT with nu_l up to its incompressible equilibrium (u instead
    of rho u), i.e. the permeabilities agree to O(delta rho)."""
    _log("[gpu] cg single-phase limit")
    solid = pkg.geometry.sphere_pack((32, 32, 32), 4.0, 0.5, seed=12)
    cg = pkg.ColorGradient(solid, 0.1, 0.05, 1.0, 0.01, (1e-6, 0, 0), precision="fp64", rel_perm_psi=0.5)
    cg.init(np.ones(solid.shape))
    sp = pkg.SinglePhase(solid, 0.1, (1e-6, 0, 0), precision="fp64")
    cg.step(2000)
    sp.step(2000)
    kl, kg, _ = cg.Comput_Perm(1)
    k, _, _ = sp.Comput_Perm(1)
    assert abs(kl[0] - k[0]) < 1e-5 * abs(k[0])
    assert kg[0] == 0.0
    assert cg.Comput_Saturation() == 1.0


def test_fp32_vs_fp64_saturation_and_fluxes(pkg):
    _log("[gpu] cg fp32 vs fp64")
    solid = pkg.geometry.sphere_pack((48, 48, 48), 6.0, 0.45, seed=33)
    psi0 = np.where(np.random.default_rng(34).uniform(size=solid.shape) < 0.5, 1.0, -1.0)
    out = {}
    for prec in ("fp64", "fp32"):
        s = pkg.ColorGradient(solid, 1 / 6, 1 / 6, 0.0, 0.01, (1e-5, 0, 0), precision=prec)
        s.init(psi0)
        s.step(500)
        d = s.diagnostics()
        out[prec] = (s.Comput_Saturation(), d["Ql"][0] + d["Qg"][0], s.state()[1])
    sat_err = abs(out["fp32"][0] - out["fp64"][0])
    q_err = abs(out["fp32"][1] - out["fp64"][1]) / abs(out["fp64"][1])
    u_err = rel_err(out["fp32"][2], out["fp64"][2])
    _log(f"[gpu] cg fp32: saturation abs err {sat_err:.2e}, total flux rel err {q_err:.2e}, u L2 rel err {u_err:.2e}")
    assert sat_err < 2e-3
    assert q_err < 1e-2
