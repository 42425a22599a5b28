"""Continuous-batching permeability server: results independent of batch mates, admission by
HBM footprint, and agreement with a plain single-sample solve."""
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _samples(pkg):
    geo = pkg.geometry
    return [geo.grf_sample(24, 3.0, 0.4, seed=101),
            geo.grf_sample(32, 3.0, 0.35, seed=102),
            geo.sphere_pack((28, 20, 36), 3.0, 0.5, seed=103),
            geo.grf_sample(20, 2.5, 0.45, seed=104),
            geo.sphere_pack((40, 24, 24), 4.0, 0.55, seed=105)]


def _run(srv, reqs):
    tickets = [srv.submit(*r) for r in reqs]
    assert srv.wait_idle(120.0)
    res = {r["ticket"]: r for r in srv.poll()}
    return [res[t] for t in tickets]


def test_result_is_the_same_alone_and_in_the_mix(pkg):
    _log("[gpu] serve: alone vs mixed")
    samples = _samples(pkg)
    reqs = [(s, 1.0 / 6.0 + 0.01 * i, (1e-5, 0.0, 0.0)) for i, s in enumerate(samples)]
    alone = []
    for r in reqs:
        with pkg.PermeabilityServer(capacity_nodes=400_000, check_interval=50, tol=1e-5, block=256) as srv:
            alone.append(_run(srv, [r])[0])
    with pkg.PermeabilityServer(capacity_nodes=400_000, check_interval=50, tol=1e-5, block=256) as srv:
        mixed = _run(srv, reqs)
    for a, m in zip(alone, mixed):
        assert a["converged"] == 1 and m["converged"] == 1
        assert a["steps"] == m["steps"]
        np.testing.assert_array_equal(a["k"], m["k"])


def test_staggered_arrivals_and_footprint_admission(pkg):
    """An arena that holds only ~two samples at a time: later requests wait for memory, are
    admitted as earlier ones retire, and still give the same answers."""
    _log("[gpu] serve: footprint admission")
    samples = _samples(pkg)
    reqs = [(s, 1.0 / 6.0, (1e-5, 0.0, 0.0)) for s in samples]
    with pkg.PermeabilityServer(capacity_nodes=400_000, check_interval=50, tol=1e-5, block=256) as srv:
        ref = _run(srv, reqs)
    biggest = max(int(np.count_nonzero(s == 0)) for s in samples)
    cap = 2 * ((biggest + 255) // 256 * 256)
    with pkg.PermeabilityServer(capacity_nodes=cap, check_interval=50, tol=1e-5, block=256) as srv:
        tickets = []
        saw_pending = False
        for (s, nu, g) in reqs:
            tickets.append(srv.submit(s, nu, g))
            st = srv.stats()
            saw_pending |= st["pending"] > 0
            assert st["arena_used"] <= st["arena_nodes"]
            time.sleep(0.01)
        assert srv.wait_idle(120.0)
        got = {r["ticket"]: r for r in srv.poll()}
        st = srv.stats()
    assert st["done"] == len(reqs)
    for t, r in zip(tickets, ref):
        np.testing.assert_array_equal(got[t]["k"], r["k"])
        assert got[t]["t_admit"] >= got[t]["t_submit"] and got[t]["t_done"] >= got[t]["t_admit"]


def test_matches_plain_solver(pkg):
    _log("[gpu] serve: vs plain solve")
    s = pkg.geometry.grf_sample(32, 3.0, 0.35, seed=102)
    with pkg.PermeabilityServer(capacity_nodes=200_000, check_interval=100, tol=1e-6, block=256) as srv:
        r = _run(srv, [(s, 1.0 / 6.0, (1e-5, 0.0, 0.0))])[0]
    assert r["converged"] == 1
    sp = pkg.SinglePhase(s, 1.0 / 6.0, (1e-5, 0.0, 0.0), precision="fp32")
    sp.step(r["steps"])
    k, _, _ = sp.Comput_Perm(1)
    assert abs(r["k"][0] - k[0]) <= 1e-6 * abs(k[0])
    # and the converged value is the FP64 answer to the stated accuracy
    sp64 = pkg.SinglePhase(s, 1.0 / 6.0, (1e-5, 0.0, 0.0), precision="fp64")
    sp64.step(r["steps"] + 2000)
    k64, _, _ = sp64.Comput_Perm(1)
    _log(f"[gpu] serve k {r['k'][0]:.8e} after {r['steps']} steps, fp64 long run {k64[0]:.8e}")
    assert abs(r["k"][0] - k64[0]) <= 1e-4 * abs(k64[0])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_matches_plain_solver_with_short_and_degenerate_axes(pkg, precision):
    """The arena's 16-bit upstream deltas wrap modulo the sample's node count: a box one voxel
    thick in x (every x link is a self link) and a box with a short z axis give the plain
    solver's answer."""
    _log(f"[gpu] serve: degenerate axes, {precision}")
    rng = np.random.default_rng(31)
    thin = (rng.uniform(size=(1, 24, 30)) < 0.3).astype(np.int32)
    short = (rng.uniform(size=(20, 18, 3)) < 0.25).astype(np.int32)
    for s in (thin, short):
        with pkg.PermeabilityServer(capacity_nodes=100_000, check_interval=50, tol=1e-6, block=256,
                                    precision=precision) as srv:
            r = _run(srv, [(s, 1.0 / 6.0, (0.0, 1e-5, 0.0), 2)])[0]
        assert r["converged"] == 1
        sp = pkg.SinglePhase(s, 1.0 / 6.0, (0.0, 1e-5, 0.0), precision=precision)
        sp.step(r["steps"])
        k, _, _ = sp.Comput_Perm(2)
        assert abs(r["k"][1] - k[1]) <= 1e-6 * abs(k[1])


AA = dict(accel_depth=5, accel_period=100, accel_warmup=200)


def test_accelerated_result_is_the_same_alone_and_in_the_mix(pkg):
    """Anderson acceleration is per sample: its dot products are reduced over the sample's own
    tiles in tile order, so the accelerated result does not depend on the batch mates either."""
    _log("[gpu] serve: accelerated, alone vs mixed")
    samples = _samples(pkg)
    reqs = [(s, 1.0 / 6.0, (1e-5, 0.0, 0.0)) for s in samples]
    alone = []
    for r in reqs:
        with pkg.PermeabilityServer(capacity_nodes=400_000, check_interval=50, tol=1e-6, block=256, **AA) as srv:
            alone.append(_run(srv, [r])[0])
    with pkg.PermeabilityServer(capacity_nodes=400_000, check_interval=50, tol=1e-6, block=256, **AA) as srv:
        mixed = _run(srv, reqs)
    for a, m in zip(alone, mixed):
        assert a["converged"] == 1 and m["converged"] == 1
        assert a["steps"] == m["steps"] and a["accelerations"] == m["accelerations"]
        np.testing.assert_array_equal(a["k"], m["k"])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_acceleration_reaches_the_fp64_answer_in_fewer_steps(pkg, precision):
    """A low-porosity sample whose slow pressure mode drags plain stepping: accelerated, the
    served k is within 1e-5 of a long FP64 run after far fewer steps."""
    _log(f"[gpu] serve: acceleration accuracy, {precision}")
    s = pkg.geometry.grf_sample(56, 5.0, 0.2, seed=91)
    with pkg.PermeabilityServer(capacity_nodes=200_000, check_interval=100, tol=1e-7, block=256,
                                precision=precision, accel_depth=5, accel_period=200, accel_warmup=500) as srv:
        r = _run(srv, [(s, 1.0 / 6.0, (1e-6, 0.0, 0.0))])[0]
    assert r["converged"] == 1 and r["accelerations"] > 0
    sp64 = pkg.SinglePhase(s, 1.0 / 6.0, (1e-6, 0.0, 0.0), precision="fp64")
    sp64.step(40000)
    k64 = sp64.Comput_Perm(1)[0][0]
    sp64.step(200)
    k64b = sp64.Comput_Perm(1)[0][0]
    _log(f"[gpu] accelerated k {r['k'][0]:.10e} after {r['steps']} steps ({r['accelerations']} accelerations); "
         f"FP64 after 40000 steps {k64:.10e} (still moving {abs(k64b - k64) / k64:.1e} per 200 steps), "
         f"rel diff {abs(r['k'][0] - k64) / k64:.2e}")
    assert r["steps"] < 20000
    assert abs(r["k"][0] - k64) <= 1e-5 * abs(k64)


def test_escalated_sample_continues_in_fp64(pkg):
    """A threshold above every k sends each FP32 sample to the internal FP64 server after the
    warm-up; it continues from its FP32 state and settles on the FP64 server's answer (to the
    FP32 rounding carried over), marked fp64, with the FP32 steps counted."""
    _log("[gpu] serve: precision escalation")
    samples = _samples(pkg)[:3]
    reqs = [(s, 1.0 / 6.0, (1e-5, 0.0, 0.0)) for s in samples]
    kw = dict(capacity_nodes=400_000, check_interval=50, tol=1e-7, block=256, **AA)
    with pkg.PermeabilityServer(escalate_k=1e3, **kw) as srv:
        esc = _run(srv, reqs)
        st = srv.stats()
    with pkg.PermeabilityServer(precision="fp64", **kw) as srv:
        ref = _run(srv, reqs)
    assert st["done"] == len(reqs) and st["pending"] == 0 and st["active"] == 0
    for e, r in zip(esc, ref):
        assert e["precision"] == "fp64" and r["precision"] == "fp64"
        assert e["converged"] == 1 and e["steps"] > AA["accel_warmup"]
        assert abs(e["k"][0] - r["k"][0]) <= 1e-5 * abs(r["k"][0])
        assert e["t_done"] >= e["t_admit"] >= e["t_submit"]


def test_sample_without_throughflow_goes_to_fp64_at_admission(pkg):
    """With escalation on, a sample whose pores do not wrap round the box along the force
    direction (no throughflow: its k is the lattice's spurious-current residue) is sent to the
    FP64 server (128-node tiles) when it is prepared.  Its result is that of an FP64 server with
    the same tiles, bit for bit and step for step, with no FP32 steps before it; a sample that
    wraps stays in the FP32 arena."""
    _log("[gpu] serve: no-throughflow sample routed to fp64")
    geo = pkg.geometry
    open_s = geo.grf_sample(24, 3.0, 0.4, seed=101)
    shut = open_s.copy()
    shut[0] = 1                                      # a solid x-plane: no path wraps along x
    assert geo.pores_wrap(open_s, 0) and not geo.pores_wrap(shut, 0) and geo.pores_wrap(shut, 1)
    reqs = [(shut, 1.0 / 6.0, (1e-5, 0.0, 0.0)), (open_s, 1.0 / 6.0, (1e-5, 0.0, 0.0))]
    kw = dict(capacity_nodes=400_000, check_interval=50, tol=1e-6, max_steps=20000, block=256, **AA)
    with pkg.PermeabilityServer(escalate_k=1e-9, **kw) as srv:
        got = _run(srv, reqs)
        st = srv.stats()
    # the internal FP64 server cuts samples into 128-node tiles (tile order fixes the flux sums)
    with pkg.PermeabilityServer(precision="fp64", **{**kw, "block": 128}) as srv:
        ref = _run(srv, reqs[:1])[0]
    _log(f"[gpu] routed k {got[0]['k'][0]:.3e} after {got[0]['steps']} steps; open sample k {got[1]['k'][0]:.3e}")
    assert st["done"] == 2 and st["pending"] == 0 and st["active"] == 0
    assert got[0]["precision"] == "fp64" and got[1]["precision"] == "fp32"
    assert got[0]["steps"] == ref["steps"] and got[0]["converged"] == ref["converged"]
    np.testing.assert_array_equal(got[0]["k"], ref["k"])
    # what remains of the decaying transient when the relative-change rule retires it
    assert abs(got[0]["k"][0]) < 1e-2 * abs(got[1]["k"][0])


def test_rejects_bad_requests(pkg):
    with pkg.PermeabilityServer(capacity_nodes=50_000, block=256) as srv:
        with pytest.raises(pkg.LBMError):
            srv.submit(np.ones((8, 8, 8), np.int32), 0.1, (1e-5, 0, 0))       # no pore voxels
        with pytest.raises(pkg.LBMError):
            srv.submit(np.zeros((8, 8, 8), np.int32), 0.1, (0.0, 1e-5, 0))    # no force along x
        with pytest.raises(pkg.LBMError):
            srv.submit(np.zeros((64, 64, 64), np.int32), 0.1, (1e-5, 0, 0))   # larger than the arena
    with pytest.raises(pkg.LBMError):
        pkg.PermeabilityServer(capacity_nodes=50_000, block=256, precision="fp16")
    with pytest.raises(pkg.LBMError):
        pkg.PermeabilityServer(capacity_nodes=50_000, block=256, accel_depth=9)
    with pytest.raises(pkg.LBMError):
        pkg.PermeabilityServer(capacity_nodes=50_000, block=256, check_interval=100, accel_depth=3, accel_period=150)


def test_fp64_arena_matches_the_oracle(pkg, oracle_mod):
    """The FP64 arena's k on a 32^3 Gaussian-random-field sample against the CPU oracle stepped
    the same number of steps (Comput_Perm, SP:4617-4677), to 1e-12."""
    _log("[gpu] serve: fp64 arena vs oracle")
    s = pkg.geometry.grf_sample(32, 3.0, 0.35, seed=102)
    nu, g = 1.0 / 6.0, (1e-5, 0.0, 0.0)
    with pkg.PermeabilityServer(capacity_nodes=200_000, check_interval=100, tol=1e-6, block=256,
                                precision="fp64") as srv:
        r = _run(srv, [(s, nu, g)])[0]
    assert r["converged"] == 1 and r["precision"] == "fp64"
    o = oracle_mod.SPOracle(s.astype(np.int32), nu, g, nthreads=max(1, min(16, os.cpu_count() or 1)))
    o.run(r["steps"])
    k = o.flux()[0] / s.size * nu / g[0]
    _log(f"[gpu] fp64 arena k {r['k'][0]:.15e}, oracle {k:.15e} after {r['steps']} steps")
    assert abs(r["k"][0] - k) <= 1e-12 * abs(k)


def test_diverging_request_is_isolated(pkg):
    """A request whose body force makes it blow up retires with converged = -1 (non-finite
    velocity) while its batch mates run on: their results are bit for bit those they get
    alone.  (The reference stops the whole run on a NaN, SP:838-839.)"""
    _log("[gpu] serve: diverging request isolated")
    samples = _samples(pkg)[:3]
    good = [(s, 1.0 / 6.0, (1e-5, 0.0, 0.0)) for s in samples]
    bad = (samples[1], 1.0 / 6.0, (50.0, 0.0, 0.0))
    kw = dict(capacity_nodes=400_000, check_interval=100, tol=1e-6, block=256, **AA)
    with pkg.PermeabilityServer(**kw) as srv:
        alone = _run(srv, good)
    with pkg.PermeabilityServer(**kw) as srv:
        mixed = _run(srv, [good[0], bad, good[1], good[2]])
    assert mixed[1]["converged"] == -1
    for a, m in zip(alone, [mixed[0], mixed[2], mixed[3]]):
        assert a["converged"] == 1 and m["converged"] == 1
        assert a["steps"] == m["steps"]
        np.testing.assert_array_equal(a["k"], m["k"])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_wide_sample_matches_plain_solver(pkg, precision):
    """x-planes of more than 32,767 pores put some links' upstream nodes beyond a 16-bit delta:
    such a sample runs on 32-bit deltas (its own launch).  Its k is the plain solver's, and a
    narrow sample sharing the arena with it gives the same k as alone."""
    _log(f"[gpu] serve: wide sample, {precision}")
    rng = np.random.default_rng(77)
    wide = (rng.uniform(size=(4, 200, 200)) < 0.1).astype(np.int32)
    assert np.count_nonzero(wide[0] == 0) > 32767
    narrow = pkg.geometry.grf_sample(24, 3.0, 0.4, seed=101)
    kw = dict(capacity_nodes=400_000, check_interval=50, tol=1e-6, block=256, precision=precision)
    with pkg.PermeabilityServer(**kw) as srv:
        alone = _run(srv, [(narrow, 1.0 / 6.0, (1e-5, 0.0, 0.0))])[0]
    with pkg.PermeabilityServer(**kw) as srv:
        r, n = _run(srv, [(wide, 1.0 / 6.0, (1e-5, 0.0, 0.0)), (narrow, 1.0 / 6.0, (1e-5, 0.0, 0.0))])
    assert r["converged"] == 1 and n["converged"] == 1
    np.testing.assert_array_equal(n["k"], alone["k"])
    sp = pkg.SinglePhase(wide, 1.0 / 6.0, (1e-5, 0.0, 0.0), precision=precision)
    sp.step(r["steps"])
    k, _, _ = sp.Comput_Perm(1)
    sp.close()
    _log(f"[gpu] wide sample k {r['k'][0]:.10e} after {r['steps']} steps, plain solver {k[0]:.10e}")
    # two FP64 kernels (arena and plain solver) differ by round-off only
    assert abs(r["k"][0] - k[0]) <= (1e-11 if precision == "fp64" else 1e-6) * abs(k[0])


def test_autotuned_block(pkg):
    with pkg.PermeabilityServer(capacity_nodes=2_000_000, block=0) as srv:
        assert srv.stats()["block"] in (128, 256, 512)
