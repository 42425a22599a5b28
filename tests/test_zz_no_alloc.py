"""lbm_sp_get_state_noalloc against lbm_sp_get_state, through the C ABI.

The no-allocation entry point streams the state through the library's fixed device array; its
output must equal the allocating entry point's bit for bit.  Two calls in a row on two handles,
the second on a smaller input: the first is large enough to need several chunks of the array
(a 64^3 sphere pack, 157,141 nodes against 91,136 nodes per chunk), the second one chunk, so the
second call also reads a array whose tail still holds the first call's data.
"""
import ctypes
import sys

import numpy as np
import pytest

from conftest import load_fixture

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc(s):
    """Call the C entry point directly (and the allocating one) on the same handle."""
    from jianhui_lbm_amd import _capi
    L = _capi.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    out = []
    for fn in (L.lbm_sp_get_state, L.lbm_sp_get_state_noalloc):
        f = np.full((s.count, 19), np.nan)
        u = np.full((s.count, 3), np.nan)
        rho = np.full(s.count, np.nan)
        _capi.check(fn(s._h, f.ctypes.data_as(dp), u.ctypes.data_as(dp), rho.ctypes.data_as(dp)), s._h)
        out.append((f, u, rho))
    # one output only, the others NULL
    u = np.full((s.count, 3), np.nan)
    _capi.check(L.lbm_sp_get_state_noalloc(s._h, None, u.ctypes.data_as(dp), None), s._h)
    out.append((None, u, None))
    return out


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_state_equals_get_state(pkg, precision):
    d, p = load_fixture("sp_c1_64cube")
    big = d["solid"].astype(np.int32)
    d2, p2 = load_fixture("sp_mrt_16x12x8")
    small = d2["solid"].astype(np.int32)
    for solid, nu, g, steps in ((big, p["nu"], p["g"], 7), (small, p2["nu"], p2["g"], 5)):
        _log(f"[gpu] noalloc get_state {precision} box {solid.shape}")
        s = pkg.SinglePhase(solid, nu, g, precision=precision)
        s.init(1.0, [0.0, 0.0, 0.0])
        s.step(steps)
        (f, u, rho), (f2, u2, rho2), (_, u3, _) = _noalloc(s)
        assert np.isfinite(f).all()
        assert np.array_equal(f, f2)
        assert np.array_equal(u, u2)
        assert np.array_equal(rho, rho2)
        assert np.array_equal(u, u3)
        s.close()


def test_noalloc_rejects_wrong_handle(pkg):
    from jianhui_lbm_amd import _capi
    L = _capi.lib()
    assert L.lbm_sp_get_state_noalloc(None, None, None, None) == -22
