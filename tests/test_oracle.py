"""The FP64 CPU oracle, pinned against the reference's own outputs and against closed forms.

Golden fixtures (tests/golden/*.npz) hold checkpoint states written by the reference solvers
themselves (see tests/golden/make_ref_fixtures.py).  Everything here runs on the CPU.
"""
import numpy as np
import pytest

from conftest import load_fixture, rel_err

SP_FIXTURES = ["sp_mrt_16x12x8", "sp_tau1_box_20x10x12"]
CG_FIXTURES = ["cg_binary_16x12x8", "cg_mixed_gated_14x10x8"]


# ----------------------------------------------------------------------------- lattice & basis
@pytest.mark.parametrize("which", [0, 1])
def test_lattice_isotropy(oracle_mod, which):
    e, w, opp = oracle_mod.lattice(which)
    assert abs(w.sum() - 1.0) < 1e-15
    np.testing.assert_array_equal(e[opp], -e)
    # second and fourth moments of D3Q19: sum w e_a e_b = d_ab/3, sum w e_x^4 = 1/3, e_x^2 e_y^2 = 1/9
    np.testing.assert_allclose(np.einsum("q,qa,qb->ab", w, e, e), np.eye(3) / 3, atol=1e-15)
    assert abs(np.sum(w * e[:, 0] ** 4) - 1 / 3) < 1e-15
    assert abs(np.sum(w * e[:, 0] ** 2 * e[:, 1] ** 2) - 1 / 9) < 1e-15


def test_sp_basis_orthogonal_and_reference_rows(oracle_mod):
    M = oracle_mod.moment_matrix(0)
    G = M @ M.T
    np.testing.assert_allclose(G, np.diag(np.diag(G)), atol=0)
    np.testing.assert_array_equal(np.diag(G), [19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8])
    # rows spelled out in SP:56-75 for a few entries
    assert list(M[1, :8]) == [-30, -11, -11, -11, -11, -11, -11, 8]
    assert list(M[2, :8]) == [12, -4, -4, -4, -4, -4, -4, 1]
    assert list(M[16, 7:11]) == [1, -1, 1, -1] and list(M[18, 11:15]) == [-1, -1, 1, 1]


def test_cg_basis_reference_rows_and_inverse(oracle_mod):
    M = oracle_mod.moment_matrix(1)
    assert list(M[1]) == [-1] + [0] * 6 + [1] * 12
    assert list(M[2]) == [1] + [-2] * 6 + [1] * 12
    assert list(M[10]) == [0, -2, -2, 1, 1, 1, 1] + [1] * 8 + [-2] * 4
    Mi = oracle_mod.invert19(M)
    np.testing.assert_allclose(Mi @ M, np.eye(19), atol=1e-14)
    # M^-1 columns for rho: the lattice weights (so f(rho, u=0) = w rho)
    e, w, _ = oracle_mod.lattice(1)
    np.testing.assert_allclose(Mi[:, 0], w, atol=1e-16)


def test_cg_basis_is_block_change_of_sp_basis(oracle_mod):
    """M_CG = T M_SP with the small block T the device kernels use."""
    Msp = oracle_mod.moment_matrix(0)
    Mcg_cgorder = oracle_mod.moment_matrix(1)
    esp, _, _ = oracle_mod.lattice(0)
    ecg, _, _ = oracle_mod.lattice(1)
    perm = [[tuple(x) for x in esp].index(tuple(v)) for v in ecg]
    Mcg = np.zeros((19, 19))
    Mcg[:, perm] = Mcg_cgorder
    T = np.eye(19)
    T[1, :3] = [11 / 19, 1 / 19, 0]
    T[2, :3] = [1 / 19, 11 / 133, 2 / 7]
    for a, b in [(3, 4), (5, 6), (7, 8)]:
        T[b, a], T[b, b] = 2 / 5, 3 / 5
    for a, b in [(9, 10), (11, 12)]:
        T[b, a], T[b, b] = 1 / 3, 2 / 3
    np.testing.assert_allclose(T @ Msp, Mcg, atol=1e-13)


def test_closed_form_equilibrium_and_force_moments(oracle_mod):
    """M feq and M GuoF of the reference (SP:1386-1396, SP:1519-1593) equal the closed forms
    the device collision uses, for random rho, u, g."""
    e, w, _ = oracle_mod.lattice(0)
    M = oracle_mod.moment_matrix(0)
    rng = np.random.default_rng(3)
    for _ in range(20):
        rho = 1 + rng.uniform(-0.2, 0.2)
        u = rng.uniform(-0.2, 0.2, 3)
        g = rng.uniform(-1, 1, 3)
        eu = e @ u
        feq = w * rho * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * u @ u)
        guo = w * (((e - u) @ g) / (1 / 3) + eu * (e @ g) / (1 / 9))
        ux, uy, uz = u
        u2 = u @ u
        t9, t11 = 2 * ux * ux - uy * uy - uz * uz, uy * uy - uz * uz
        meq = [rho, rho * (-11 + 19 * u2), rho * (3 - 5.5 * u2), rho * ux, -2 / 3 * rho * ux, rho * uy,
               -2 / 3 * rho * uy, rho * uz, -2 / 3 * rho * uz, rho * t9, -0.5 * rho * t9, rho * t11,
               -0.5 * rho * t11, rho * ux * uy, rho * uy * uz, rho * ux * uz, 0, 0, 0]
        ug = u @ g
        f9, f11 = 2 * (2 * ux * g[0] - uy * g[1] - uz * g[2]), 2 * (uy * g[1] - uz * g[2])
        F = [0, 38 * ug, -11 * ug, g[0], -2 / 3 * g[0], g[1], -2 / 3 * g[1], g[2], -2 / 3 * g[2], f9, -0.5 * f9,
             f11, -0.5 * f11, ux * g[1] + uy * g[0], uy * g[2] + uz * g[1], ux * g[2] + uz * g[0], 0, 0, 0]
        np.testing.assert_allclose(M @ feq, meq, atol=1e-14)
        np.testing.assert_allclose(M @ guo, F, atol=1e-13)


def test_relaxation_rates(oracle_mod):
    S, sv = oracle_mod.sp_relaxation(1.0 / 6.0)
    assert sv == 1.0
    assert S[4] == 8 * (2 - 1) / (8 - 1)
    # TRT "magic" product of the two rates fixes the wall half-way between nodes
    S, sv = oracle_mod.sp_relaxation(0.05)
    sq = S[4]
    assert abs((1 / sv - 0.5) * (1 / sq - 0.5) - 3 / 16) < 1e-15
    assert S[0] == S[3] == S[5] == S[7] == 0.0


# ----------------------------------------------------------------------------- numbering
def test_number_fluid_is_i_major(oracle_mod):
    solid = np.ones((3, 4, 5), np.int32)
    pores = [(0, 0, 1), (0, 2, 0), (1, 0, 0), (2, 3, 4)]
    for p in pores:
        solid[p] = 0
    count, idx, supinv = oracle_mod.number_fluid(solid)
    assert count == 4
    assert [idx[p] for p in pores] == [1, 2, 3, 4]
    assert list(supinv[1:]) == [np.ravel_multi_index(p, solid.shape) for p in pores]


# ----------------------------------------------------------------------------- vs reference
@pytest.mark.parametrize("name", SP_FIXTURES)
def test_sp_oracle_matches_reference_checkpoints(oracle_mod, name):
    d, p = load_fixture(name)
    o = oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"])
    assert o.count == d[f"rho_{p['steps'][0]}"].shape[0]
    for s in p["steps"]:
        o.run(s - o.steps)
        assert rel_err(o.f[1:], d[f"f_{s}"]) < 1e-13
        assert rel_err(o.rho[1:], d[f"rho_{s}"]) < 1e-13
        # u = (sum e f + g/2)/rho cancels O(0.1) populations down to O(1e-6) velocities, so the
        # reference's own M^-1 literals (a few ulp off) show up at ~1e-16 absolute in u
        assert np.abs(o.u[1:] - d[f"u_{s}"]).max() < 1e-14
        assert rel_err(o.u[1:], d[f"u_{s}"]) < 5e-10


@pytest.mark.parametrize("name", SP_FIXTURES)
def test_sp_oracle_permeability_matches_reference_output(oracle_mod, name):
    """k printed every freRe steps (Permeability_error_local_Perm.txt, SP:760-775)."""
    d, p = load_fixture(name)
    last = p["steps"][-1]
    rows = [ln.split() for ln in str(d[f"perm_txt_{last}"]).strip().splitlines()]
    o = oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"])
    box = p["box"]
    V = np.prod(p["shape"]) if box is None else (box[1] - box[0] + 1) * (box[3] - box[2] + 1) * (box[5] - box[4] + 1)
    for r, row in enumerate(rows):
        n = r * p["freRe"]
        o.run(n + 1 - o.steps)            # label n = state after n+1 steps
        Q = o.flux(box)
        k = Q[0] / V * p["nu"] / p["g"][0] * 1.0 * 1.0 * 1000   # reso = 1
        assert abs(k - float(row[0])) <= 1e-5 * abs(k) + 1e-12


@pytest.mark.parametrize("name", CG_FIXTURES)
def test_cg_oracle_matches_reference_checkpoints(oracle_mod, name):
    d, p = load_fixture(name)
    o = oracle_mod.CGOracle(d["solid"].astype(np.int32), d["psi_field"], p["nu_l"], p["nu_g"], p["g"], p["cap_a"],
                            p["cos_theta"], bodyforce_apply=p["bodyforce_apply"],
                            rel_perm_psi=p["rel_perm_psi"][0], rel_perm_psi2=p["rel_perm_psi"][1])
    for s in p["steps"]:
        o.run(s - o.steps)
        assert rel_err(o.f[1:], d[f"f_{s}"]) < 1e-13
        assert rel_err(o.u[1:], d[f"u_{s}"]) < 1e-11
        assert rel_err(o.psi[1:], d[f"psi_{s}"]) < 1e-13
        assert rel_err(o.rho_r[1:], d[f"rhor_{s}"]) < 1e-13
        assert rel_err(o.rho_b[1:], d[f"rhob_{s}"]) < 1e-13


@pytest.mark.parametrize("name", CG_FIXTURES)
def test_cg_oracle_saturation_matches_reference_output(oracle_mod, name):
    """S_l printed every freRe steps in bodyforce.txt (CG:1115-1117)."""
    d, p = load_fixture(name)
    last = p["steps"][-1]
    rows = [ln.split() for ln in str(d[f"bodyforce_txt_{last}"]).strip().splitlines()]
    o = oracle_mod.CGOracle(d["solid"].astype(np.int32), d["psi_field"], p["nu_l"], p["nu_g"], p["g"], p["cap_a"],
                            p["cos_theta"], bodyforce_apply=p["bodyforce_apply"],
                            rel_perm_psi=p["rel_perm_psi"][0], rel_perm_psi2=p["rel_perm_psi"][1])
    for r, row in enumerate(rows):
        o.run(r * p["freRe"] + 1 - o.steps)
        # S_l counts psi >= 0; nodes with psi at round-off distance from 0 (exact ties of
        # rho_r and rho_b) may land on either side, so the reference value must lie within
        # the band those nodes allow
        psi = o.psi[1:]
        lo = np.count_nonzero(psi >= 1e-12) / o.count
        hi = np.count_nonzero(psi >= -1e-12) / o.count
        assert lo - 1e-6 <= float(row[0]) <= hi + 1e-6
        assert abs(o.wet_count() / o.count - float(row[0])) <= (hi - lo) + 1e-6


# ----------------------------------------------------------------------------- physics
def test_uniform_acceleration_without_solid(oracle_mod):
    """No solid: each collision adds g to the momentum and u carries +g/2, so u = (s+1/2) g."""
    solid = np.zeros((4, 3, 5), np.int32)
    g = np.array([1e-5, -2e-6, 3e-6])
    o = oracle_mod.SPOracle(solid, 0.1, g)
    for s in (1, 7, 20):
        o.run(s - o.steps)
        np.testing.assert_allclose(o.u[1:], np.tile((s + 0.5) * g, (o.count, 1)), rtol=1e-10, atol=1e-16)
        np.testing.assert_allclose(o.rho[1:], 1.0, atol=1e-14)


def test_mass_conservation_with_solids(oracle_mod):
    d, p = load_fixture("sp_mrt_16x12x8")
    o = oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"])
    m0 = o.rho[1:].sum()
    o.run(50)
    assert abs(o.rho[1:].sum() - m0) < 1e-12 * m0


@pytest.mark.parametrize("nu", [1.0 / 6.0, 0.05, 0.5])
def test_plane_poiseuille_exact_with_halfway_bounce_back(oracle_mod, nu):
    """Channel of H fluid layers between solid walls: the steady MRT solution with the
    reference's s_q = 8(2-s_v)/(8-s_v) is the parabola with walls half-way between nodes,
    u(y) = g/(2 nu) (y - 1/2)(H + 1/2 - y), for any viscosity."""
    H = 6
    solid = np.zeros((1, H + 1, 1), np.int32)
    solid[0, 0, 0] = 1
    g = 1e-6
    o = oracle_mod.SPOracle(solid, nu, (g, 0.0, 0.0))
    steps = int(60 * H * H / nu) + 2000
    o.run(steps)
    y = np.arange(1, H + 1)
    exact = g / (2 * nu) * (y - 0.5) * (H + 0.5 - y)
    np.testing.assert_allclose(o.u[1:, 0], exact, rtol=1e-9)
    # permeability of the slit per unit cell, k = (Q/V) nu / g, with Q the node sum of the
    # parabola: sum_{t=1/2..H-1/2} t (H - t) / 2 = (H^3/12 + H/24)
    k = o.flux()[0] / (H + 1) * nu / g
    assert abs(k - (H ** 3 / 12 + H / 24) / (H + 1)) < 1e-9 * k


def test_error_skips_last_node(oracle_mod):
    d, p = load_fixture("sp_mrt_16x12x8")
    o = oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"])
    o.run(3)
    err, umax, ssum = o.error()
    u, u0 = o.u[1:-1], o.u0[1:-1]
    ref = np.sqrt(((u - u0) ** 2).sum()) / (np.sqrt((u ** 2).sum()) + 1e-30)
    assert abs(err - ref) < 1e-12 * ref
    assert abs(umax - np.sqrt((u ** 2).sum(1)).max()) < 1e-18


# ----------------------------------------------------------------------------- passive solute
def test_st_oracle_w0_zero_matches_reference_six_velocity_program(oracle_mod):
    """Rest weight 0 is the reference's six-velocity scheme: from the same initial state the
    oracle follows the reference's checkpoints (6V.cpp) to round-off."""
    d, p = load_fixture("st6_16x12x8")
    o = oracle_mod.STOracle(d["solid"].astype(np.int32), p["niu_s"], d["u"], d["conc_field"], w0=0.0)
    np.testing.assert_allclose(o.g[1:], d["g_0"], rtol=0, atol=1e-15)   # same initial formula
    o.g[1:] = d["g_0"]
    o.C[1:] = d["c_0"]
    last = 0
    for steps in p["steps"]:
        o.run(steps - last)
        last = steps
        assert np.abs(o.g[1:] - d[f"g_{steps}"]).max() < 1e-14
        assert np.abs(o.C[1:] - d[f"c_{steps}"]).max() < 1e-14


def _pulse(oracle_mod, L, w0, D, u, solid=None):
    solid = np.zeros((L, L, L), np.int32) if solid is None else solid
    conc = np.zeros((L, L, L))
    conc[L // 2, L // 2, L // 2] = 1.0
    o = oracle_mod.STOracle(solid, D, np.tile(np.asarray(u, float), (int(np.count_nonzero(solid == 0)), 1)), conc,
                            w0=w0)
    xyz = np.stack(np.unravel_index(o.supinv[1:], solid.shape), 1).astype(float)
    return o, xyz


@pytest.mark.parametrize("w0", [0.0, 0.25, 0.5])
def test_st_pulse_moments_closed_form(oracle_mod, w0):
    """A unit pulse in an open box, started at equilibrium: the mean moves exactly u t, and with
    u = 0 the variance per axis grows by c_s^2 [(2 tau - 1) t - 2 tau (tau - 1)(1 - (1 - 1/tau)^t)]
    (the moment recursion of BGK streaming), i.e. 2 D t at long times with D = c_s^2 (tau - 1/2),
    c_s^2 = (1 - w0)/3.  A wrong sound speed, weight or equilibrium flux fails one of these."""
    L, t, D = 48, 20, 0.05
    cs2 = (1 - w0) / 3
    u = (0.02, -0.01, 0.005)
    o, xyz = _pulse(oracle_mod, L, w0, D, u)
    x0 = (xyz * o.C[1:, None]).sum(0)
    o.run(t)
    assert abs(o.mass() - 1.0) < 1e-14
    np.testing.assert_allclose((xyz * o.C[1:, None]).sum(0) - x0, np.asarray(u) * t, rtol=0, atol=1e-12)
    o, xyz = _pulse(oracle_mod, L, w0, D, (0.0, 0.0, 0.0))
    o.run(t)
    tau = D / cs2 + 0.5
    expect = cs2 * ((2 * tau - 1) * t - 2 * tau * (tau - 1) * (1 - (1 - 1 / tau) ** t))
    var = ((xyz - L // 2) ** 2 * o.C[1:, None]).sum(0)
    np.testing.assert_allclose(var, expect, rtol=1e-12)


def test_st_mass_conserved_with_walls(oracle_mod):
    """Bounce-back at solid sites reflects solute, so the total is invariant (zero-flux walls)."""
    solid = (np.random.default_rng(3).uniform(size=(18, 14, 12)) < 0.3).astype(np.int32)
    rng = np.random.default_rng(4)
    n = int(np.count_nonzero(solid == 0))
    o = oracle_mod.STOracle(solid, 0.03, rng.uniform(-0.05, 0.05, (n, 3)), rng.uniform(0, 1, solid.shape), w0=0.25)
    m0 = o.mass()
    o.run(200)
    assert abs(o.mass() - m0) < 1e-12 * m0


BC_FIXTURES = ["sp_bc_p0_16x12x10", "sp_bc_p1_16x12x10", "sp_bc_p2_16x12x10", "sp_bc_p3_16x12x10",
               "sp_bc_v1_16x12x10"]


def _bc_oracle(oracle_mod, d, p, **kw):
    bc = p["bc"]
    b = oracle_mod.make_bc(bc["sub_bc"], bc.get("pre"), bc.get("vel"), **kw)
    return oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"], bc=b)


@pytest.mark.parametrize("name", BC_FIXTURES)
def test_sp_boundaries_match_reference(oracle_mod, name):
    """Pressure (Sub_BC 0-3) and velocity (Sub_BC 1) faces: the oracle follows the reference's
    checkpoints after 1, 2 and 21 steps; without the faces it does not."""
    d, p = load_fixture(name)
    o = _bc_oracle(oracle_mod, d, p)
    per = oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"])
    for steps in p["steps"]:
        o.run(steps - o.steps)
        per.run(steps - per.steps)
        ref = d[f"f_{steps}"]
        assert np.abs(o.f[1:] - ref).max() < 1e-14
        assert np.abs(o.rho[1:] - d[f"rho_{steps}"]).max() < 1e-13
        assert np.abs(o.u[1:] - d[f"u_{steps}"]).max() < 1e-14
        assert np.abs(per.f[1:] - ref).max() > 1e-6


@pytest.mark.parametrize("name", ["sp_bc_p1_16x12x10", "sp_bc_p2_16x12x10", "sp_bc_p3_16x12x10"])
def test_sp_boundary_quirks_are_visible_in_reference_output(oracle_mod, name):
    """The two quirks kept from the reference (x- face u_ls[2] = u[nb][1] in Sub_BC 1 and 3; x
    faces with a solid inward neighbour under Sub_BC 2 keep their streamed populations): the
    oracle with the quirks corrected misses the reference's state, so the fixtures pin them."""
    d, p = load_fixture(name)
    o = _bc_oracle(oracle_mod, d, p, no_quirks=True)
    o.run(p["steps"][-1])
    assert np.abs(o.f[1:] - d[f"f_{p['steps'][-1]}"]).max() > 1e-9


@pytest.mark.parametrize("name", ["sp_nnf_16x12x8", "sp_nnf_bc_16x12x10"])
def test_sp_power_law_matches_reference(oracle_mod, name):
    """collision_nnf (SP:2017-2700): Newtonian up to loop step 10, then per-node viscosity from
    the strain rate.  The oracle follows the reference's checkpoints; the Newtonian oracle
    follows them up to 11 steps and not after."""
    d, p = load_fixture(name)
    bc = p.get("bc")
    b = oracle_mod.make_bc(bc["sub_bc"], bc.get("pre"), bc.get("vel")) if bc else None
    solid = d["solid"].astype(np.int32)
    o = oracle_mod.SPOracle(solid, p["nu"], p["g"], p["u0"], bc=b, power_law=p["power_law"])
    newt = oracle_mod.SPOracle(solid, p["nu"], p["g"], p["u0"], bc=b)
    for steps in p["steps"]:
        o.run(steps - o.steps)
        newt.run(steps - newt.steps)
        ref = d[f"f_{steps}"]
        assert np.abs(o.f[1:] - ref).max() < 1e-14
        assert np.abs(o.u[1:] - d[f"u_{steps}"]).max() < 1e-14
        if steps <= 11:
            assert np.abs(newt.f[1:] - ref).max() < 1e-14
        else:
            assert np.abs(newt.f[1:] - ref).max() > 1e-6


def test_sp_oracle_full_size_reference_case(oracle_mod):
    """The reference's own case at full size (64^3, 157,141 nodes, 2000 steps, 8 ranks): the
    oracle's state at every 64th node and its per-direction sums over all nodes."""
    d, p = load_fixture("sp_c1_64cube")
    o = oracle_mod.SPOracle(d["solid"].astype(np.int32), p["nu"], p["g"], p["u0"], nthreads=8)
    assert o.count == p["count"]
    o.run(p["steps"])
    pick = d["pick"]
    # 2000 steps accumulate the reference's few-ulp M^-1 literals into ~1.4e-13 on f (4e-13
    # relative) and ~6e-15 on u (|u| up to 1.6e-4); the sum of u cancels to 3e-9 relative
    assert np.abs(o.f[1:][pick] - d["f_pick"]).max() < 5e-13
    assert np.abs(o.u[1:][pick] - d["u_pick"]).max() < 2e-14
    np.testing.assert_allclose(o.f[1:].sum(0), d["f_sum"], rtol=1e-12)
    np.testing.assert_allclose(o.u[1:].sum(0), d["u_sum"], rtol=1e-8)


# ----------------------------------------------------------------------------- Error over ranks
RANK_FIXTURES = ["sp_mrt_16x12x8", "sp_tau1_box_20x10x12", "sp_perdir3_12x10x16", "sp_bc_p1_16x12x10",
                 "sp_bc_p3_16x12x10", "sp_nnf_16x12x8", "sp_locperm_box_20x10x12", "sp_gperm_20x10x12",
                 "sp_c1_64cube"]


@pytest.mark.parametrize("name", RANK_FIXTURES)
def test_slab_split_matches_reference_rank_counts(oracle_mod, name):
    """The reference's x-slab split (SP:1085-1108) gives each rank the node count that rank's
    checkpoint file holds (2 ranks; 8 for the 64^3 case)."""
    d, p = load_fixture(name)
    counts = p["rank_counts"]
    ends = oracle_mod.rank_ends(d["solid"].astype(np.int32), len(counts))
    assert list(np.diff(np.concatenate([[0], ends]))) == counts


def _vel_lines(text):
    return [[float(x) for x in line.split()] for line in text.strip().splitlines() if line.strip()]


def _bc_of(oracle_mod, p):
    bc = p.get("bc")
    return None if bc is None else oracle_mod.make_bc(bc["sub_bc"], bc.get("pre"), bc.get("vel"))


@pytest.mark.parametrize("name", ["sp_mrt_16x12x8", "sp_tau1_box_20x10x12", "sp_perdir3_12x10x16",
                                  "sp_bc_p3_16x12x10", "sp_nnf_16x12x8"])
def test_error_over_ranks_matches_reference_printed_values(oracle_mod, name):
    """Velocity_ave_max.txt prints, every freRe steps, u_max and Error's velocity change from
    the reference's 2 ranks (SP:781-782); both are reproduced from the oracle's u of the step
    and of the step before with the 2-rank split, to the 6 digits printed.  The one-rank value
    and a variant that forgets to skip each rank's last node are planted mistakes the printed
    values must reject."""
    d, p = load_fixture(name)
    solid = d["solid"].astype(np.int32)
    last = p["steps"][-1]
    printed = _vel_lines(str(d[f"vel_txt_{last}"]))
    o = oracle_mod.SPOracle(solid, p["nu"], p["g"], p["u0"], bc=_bc_of(oracle_mod, p),
                            power_law=p.get("power_law"))
    ends = oracle_mod.rank_ends(solid, 2)
    assert list(np.diff(np.concatenate([[0], ends]))) == p["rank_counts"]
    rejected_one_rank = rejected_no_skip = 0
    for line_no, n in enumerate(range(0, last, p["freRe"])):
        o.run(n + 1 - o.steps)          # loop step n ends with u after n + 1 steps (SP:662-684)
        err, umax, _ = o.error(ranks=2)
        row = printed[line_no]
        assert abs(umax - row[1]) <= 6e-6 * abs(row[1]), (n, umax, row)
        assert abs(err - row[2]) <= 6e-6 * abs(row[2]), (n, err, row)
        # planted mistakes, on the CPU: one rank; every rank's last node included
        err1, _, _ = o.error(ranks=1)
        rejected_one_rank += abs(err1 - row[2]) > 6e-6 * abs(row[2])
        u, u0 = o.u[1:], o.u0[1:]
        bad = max(np.sqrt(((u[a:b] - u0[a:b]) ** 2).sum()) / (np.sqrt((u[a:b] ** 2).sum()) + 1e-30)
                  for a, b in ((0, ends[0]), (ends[0], ends[1])))
        rejected_no_skip += abs(bad - row[2]) > 6e-6 * abs(row[2])
    assert len(printed) == line_no + 1
    assert rejected_one_rank > 0 and rejected_no_skip > 0


def test_error_closed_form_on_a_solid_free_box(oracle_mod):
    """Body force on a periodic box with no solid: every node has u_n = (n + 1/2) g after n steps
    (Guo forcing adds g to the momentum each step and the velocity carries half a step more),
    so Error's velocity change is exactly 1/(n + 1/2) for n >= 2, on any rank count, and
    u_max = (n + 1/2) |g|."""
    solid = np.zeros((6, 5, 4), np.int32)
    g = np.array([3e-4, -1e-4, 2e-4])
    o = oracle_mod.SPOracle(solid, 0.1, g)
    # u is a difference of populations ~0.05: FP64 round-off of ~1e-17 in u
    for n in (2, 3, 7, 20):
        o.run(n - o.steps)
        np.testing.assert_allclose(o.u[1:], np.tile((n + 0.5) * g, (o.count, 1)), rtol=0, atol=1e-15)
        for ranks in (1, 2, 3):
            err, umax, ssum = o.error(ranks=ranks)
            assert abs(err - 1.0 / (n + 0.5)) < 1e-11
            assert abs(umax - (n + 0.5) * np.linalg.norm(g)) < 1e-11 * umax
            # every rank skips one node
            assert abs(ssum - (o.count - ranks) * (n + 0.5) * np.linalg.norm(g)) < 1e-11 * ssum
