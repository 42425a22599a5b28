"""State transfers in bounded chunks (through a 16 MiB device buffer per handle).

Each case is sized to need several chunks, and checks a transfer against another path to the
same numbers: the dense fields against the per-node state scattered on the host, and
get_state -> load_state -> get_state round trips, which must be exact in FP64.
"""
import sys

import numpy as np
import pytest

from conftest import load_fixture, rel_err

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _c1_solid():
    d, _ = load_fixture("sp_c1_64cube")
    return d["solid"].astype(np.int32)


def test_dense_fields_match_scattered_state(pkg):
    # 96^3 voxels: 884,736 voxels, two chunks of the dense u + rho transfer
    _log("[gpu] staging: dense fields of a 96^3 pack")
    solid = pkg.geometry.sphere_pack((96, 96, 96), 6.0, 0.5, 31)
    s = pkg.SinglePhase(solid, 1.0 / 6.0, [1e-5, 0.0, 0.0], precision="fp32")
    s.init()
    s.step(4)
    _, u, rho = s.state()
    ud, rd = s.dense()
    pores = solid == 0
    assert np.array_equal(ud[pores], u)
    assert np.array_equal(rd[pores], rho)
    assert (ud[~pores] == 0.0).all() and (rd[~pores] == 1.0).all()
    s.close()


def test_sp_round_trip_exact_fp64(pkg):
    # 157,141 nodes: f + u + rho in two chunks
    _log("[gpu] staging: single-phase round trip")
    solid = _c1_solid()
    a = pkg.SinglePhase(solid, 1.0 / 6.0, [1e-5, 0.0, 0.0], precision="fp64")
    a.init()
    a.step(3)
    f, u, rho = a.state()
    b = pkg.SinglePhase(solid, 1.0 / 6.0, [1e-5, 0.0, 0.0], precision="fp64")
    b.load_state(f, u, rho)
    f2, u2, rho2 = b.state()
    assert np.array_equal(f, f2) and np.array_equal(u, u2) and np.array_equal(rho, rho2)
    # the restarted run collides its first step with the loaded rho (rho0 + drho rounded to a
    # double, as the reference's restart does) where the continuing run has drho itself: equal
    # to round-off, not bit for bit
    a.step(2)
    b.step(2)
    assert rel_err(a.state()[0], b.state()[0]) < 1e-14
    a.close()
    b.close()


def test_cg_round_trip_exact_fp64(pkg):
    # 157,141 nodes: 26 doubles per node in two chunks; the phase field of init in one
    _log("[gpu] staging: colour-gradient round trip")
    solid = _c1_solid()
    rng = np.random.default_rng(5)
    psi = np.where(rng.random(solid.shape) < 0.5, 1.0, -1.0)
    args = (solid, 1.0 / 6.0, 1.0 / 6.0, 0.0, 0.01, [1e-5, 0.0, 0.0])
    a = pkg.ColorGradient(*args, precision="fp64")
    a.init(psi)
    st0 = a.state()
    assert np.array_equal(st0[3], psi[solid == 0])
    a.step(3)
    st = a.state()
    b = pkg.ColorGradient(*args, precision="fp64")
    b.load_state(*st)
    st2 = b.state()
    for x, y in zip(st, st2):
        assert np.array_equal(x, y)
    a.step(2)
    b.step(2)
    for x, y in zip(a.state(), b.state()):
        assert rel_err(x, y) < 1e-14
    a.close()
    b.close()


def test_solute_round_trip_exact_fp64(pkg):
    # 157,141 nodes: g in two chunks
    _log("[gpu] staging: solute round trip")
    solid = _c1_solid()
    conc = np.random.default_rng(6).random(solid.shape)
    a = pkg.Solute(0.05, solid=solid, precision="fp64")
    a.init(conc)
    a.step(3)
    g, c = a.state()
    assert np.allclose(c, g.sum(axis=1), rtol=0, atol=1e-14)
    b = pkg.Solute(0.05, solid=solid, precision="fp64")
    b.load_state(g)
    g2, c2 = b.state()
    assert np.array_equal(g, g2) and np.array_equal(c, c2)
    a.close()
    b.close()
