"""Launch settings (lbm_get_launch / lbm_set_launch) and the autotuner.

Block size, tile order and the colour-streaming variant change how the work is dealt to the
GPU, never the arithmetic: every setting must give the same state bit for bit.  Boxes of 2^20
or more fluid nodes are tuned at create; smaller ones keep the defaults.
"""
import itertools
import sys

import numpy as np
import pytest

from conftest import load_fixture

pytestmark = pytest.mark.gpu

BLOCKS = (128, 256, 512)
ORDERS = ("hw", "runs8", "slab")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _c1_solid():
    d, _ = load_fixture("sp_c1_64cube")
    return d["solid"].astype(np.int32)


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp16"])
def test_single_phase_settings_give_identical_states(pkg, precision):
    _log(f"[gpu] launch settings, single phase {precision}")
    solid = _c1_solid()
    ref = None
    for block, order in itertools.product(BLOCKS, ORDERS):
        s = pkg.SinglePhase(solid, 1.0 / 6.0, [1e-5, 0, 0], precision=precision)
        assert s.launch() == {"block": 256, "order": "runs8", "variant": "gather", "tuned": False, "time_us": 0.0}
        s.set_launch(block, order)
        s.step(25)
        f = s.state()[0]
        if ref is None:
            ref = f
        else:
            assert np.array_equal(f, ref), (block, order)
        s.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_colour_gradient_settings_give_identical_states(pkg, precision):
    _log(f"[gpu] launch settings, colour gradient {precision}")
    solid = pkg.geometry.sphere_pack((48, 40, 36), 4.0, 0.5, seed=3)
    psi = np.where(np.random.default_rng(4).uniform(size=solid.shape) < 0.5, 1.0, -1.0)
    variants = ("gather", "staged")
    ref = None
    cases = [(b, o, "gather", "gather") for b, o in itertools.product(BLOCKS, ORDERS)]
    cases += [(256, o, cv, v) for cv, v, o in itertools.product(("gather", "staged"), variants, ORDERS)]
    for block, order, collide_variant, variant in cases:
        s = pkg.ColorGradient(solid, 0.1, 0.05, 0.3, 0.01, (1e-5, 0, 0), precision=precision)
        # the staged kernels run 512-node blocks
        s.set_launch("cg_collide", 512 if collide_variant == "staged" else block, order, collide_variant)
        s.set_launch("cg_colour", 512 if variant == "staged" else 256, order, variant)
        s.init(psi)
        s.step(20)
        st = s.state()
        if ref is None:
            ref = st
        else:
            for a, b in zip(st, ref):
                assert np.array_equal(a, b), (block, order, collide_variant, variant)
        s.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("shape", [(128, 12, 12), (64, 24, 24)])
def test_staged_kernels_on_small_planes(pkg, precision, shape):
    """Boxes of few nodes per x-plane, whose 512-node blocks stage 1024 (128x12x12) and 1536
    (64x24x24) states rather than 2560: the staged kernels give the gathers' bits."""
    _log(f"[gpu] staged colour gradient, {precision}, {shape}")
    solid = pkg.geometry.sphere_pack(shape, 4.0, 0.5, seed=3)
    psi = np.where(np.random.default_rng(4).uniform(size=solid.shape) < 0.5, 1.0, -1.0)
    states = []
    for variant in ("gather", "staged"):
        s = pkg.ColorGradient(solid, 0.1, 0.05, 0.3, 0.01, (1e-5, 0, 0), precision=precision)
        block = 512 if variant == "staged" else 256
        s.set_launch("cg_collide", block, "runs8", variant)
        s.set_launch("cg_colour", block, "slab", variant)
        s.init(psi)
        s.step(20)
        states.append(s.state())
        s.close()
    for a, b in zip(*states):
        assert np.array_equal(a, b)


def test_solute_settings_give_identical_states(pkg):
    _log("[gpu] launch settings, solute")
    solid = _c1_solid()
    conc = np.random.default_rng(9).random(solid.shape)
    u = np.random.default_rng(10).uniform(-0.01, 0.01, (int((solid == 0).sum()), 3))
    ref = ref_c = None
    for block, order in itertools.product(BLOCKS, ORDERS):
        st = pkg.Solute(0.02, solid=solid, precision="fp32")
        st.set_launch(block, order)
        st.set_velocity(u)
        st.init(conc)
        st.step(15)
        g = st.state()[0]
        st.close()
        # coupled: the fused flow + solute launch
        flow = pkg.SinglePhase(solid, 1.0 / 6.0, [1e-4, 0, 0], precision="fp32")
        cs = pkg.Solute(0.02, flow=flow, precision="fp32")
        cs.set_launch(block, order)
        cs.init(conc)
        cs.step(15)
        c = cs.concentration()
        cs.close()
        flow.close()
        if ref is None:
            ref, ref_c = g, c
        else:
            assert np.array_equal(g, ref) and np.array_equal(c, ref_c), (block, order)


def test_invalid_settings_are_refused(pkg):
    from jianhui_lbm_amd import _capi
    solid = _c1_solid()[:16, :16, :16].copy()
    s = pkg.SinglePhase(solid, 0.1, [1e-5, 0, 0], precision="fp32")
    for bad in (dict(block=64), dict(block=256, order="slab", variant="staged")):
        with pytest.raises(_capi.LBMError):
            _capi.set_launch(s._h, "sp_step", bad.get("block", 256), bad.get("order", "runs8"),
                             bad.get("variant", "gather"))
    with pytest.raises(_capi.LBMError):
        _capi.set_launch(s._h, "cg_colour", 256, "slab", "gather")
    s.close()
    c = pkg.ColorGradient(solid, 0.1, 0.1, 0.0, 0.01, (1e-5, 0, 0), precision="fp64")
    with pytest.raises(_capi.LBMError):
        c.set_launch("cg_colour", 256, "slab", "staged")   # staging is per 512-node block
    with pytest.raises(_capi.LBMError):
        c.set_launch("cg_colour", 512, "slab", "gather")
    with pytest.raises(_capi.LBMError):
        c.set_launch("cg_collide", 256, "slab", "staged")   # staging is per 512-node block
    c.close()


def test_large_boxes_are_tuned(pkg):
    _log("[gpu] autotuner on a 2^20+ node box")
    solid = pkg.geometry.sphere_pack((128, 128, 128), 6.0, 0.6, seed=5)
    assert int((solid == 0).sum()) >= 1 << 20
    s = pkg.SinglePhase(solid, 1.0 / 6.0, [1e-6, 0, 0], precision="fp32")
    ls = s.launch()
    assert ls["tuned"] and ls["time_us"] > 0 and ls["block"] in BLOCKS and ls["order"] in ORDERS
    s.close()
