"""Host-side logic on the CPU: file formats, generators, and the C-ABI library's exports.

No GPU work happens here: the library is loaded and its symbol table checked, nothing that
touches HIP is called."""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT


def test_library_exports_every_declared_symbol(pkg):
    lib_path = pkg._capi.LIB_PATH
    if not os.path.exists(lib_path):
        pytest.skip("library not built (run __graft_entry__.build())")
    declared = pkg.declared_symbols()
    assert "lbm_sp_step" in declared and "lbm_sp_create" in declared
    L = ctypes.CDLL(lib_path)
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, f"declared but not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(declared) <= exported


def test_library_version_string(pkg):
    if not os.path.exists(pkg._capi.LIB_PATH):
        pytest.skip("library not built")
    v = pkg.lib().lbm_version().decode()
    assert "gfx950" in v


def test_binary_geometry_roundtrip(tmp_path, pkg):
    geo = pkg.geometry
    solid = (np.random.default_rng(1).uniform(size=(5, 4, 3)) < 0.3).astype(np.int32)
    p = tmp_path / "g.bin"
    geo.write_binary(str(p), solid)
    raw = np.fromfile(p, dtype="<i4")
    # [i][j][k] C order, z fastest (SP:1058-1065)
    assert raw[1] == solid[0, 0, 1] and raw[3] == solid[0, 1, 0] and raw[12] == solid[1, 0, 0]
    np.testing.assert_array_equal(geo.read_binary(str(p), solid.shape), solid)


def test_ascii_geometry_is_x_fastest_and_skips_other_tokens(tmp_path, pkg):
    geo = pkg.geometry
    solid = (np.random.default_rng(2).uniform(size=(4, 3, 2)) < 0.4).astype(np.int32)
    p = tmp_path / "g.dat"
    geo.write_ascii(str(p), solid)
    toks = open(p).read().split()
    assert int(toks[1]) == solid[1, 0, 0] and int(toks[4]) == solid[0, 1, 0]
    # the reference's reader skips tokens that are neither 0 nor 1 (SP:1030-1034)
    with open(p) as fh:
        text = fh.read()
    (tmp_path / "g2.dat").write_text("2 7 " + text.replace(" ", " 5 ", 3))
    np.testing.assert_array_equal(geo.read_ascii(str(tmp_path / "g2.dat"), solid.shape), solid)


def test_input_files_roundtrip_and_reference_samples(tmp_path, pkg):
    inp = pkg.inputs
    a = inp.SPInput(n_xyz=(64, 32, 16), n_max=1999, g=(1e-6, 0.0, 0.0), in_vis=1 / 6, freRe=100,
                    par_per=(1, 0, 1), per_x=(40, 3), outputfile="./out_")
    p = tmp_path / "INPUT.dat"
    inp.write_sp_input(a, str(p))
    b = inp.read_sp_input(str(p))
    assert b == a
    lines = open(p).read().splitlines()
    assert lines[2].split()[:3] == ["64", "32", "16"] and lines[13].split()[0] == repr(1 / 6)
    c = inp.CGInput(niu_l=0.1, niu_g=0.05, CapA=0.01, cos_theta=0.3)
    inp.write_cg_input(c, str(tmp_path / "CG.dat"))
    assert inp.read_cg_input(str(tmp_path / "CG.dat")) == c
    # integral values are written without a decimal point (int slots of the reference parser)
    assert open(tmp_path / "CG.dat").read().splitlines()[9].split()[:4] == ["0", "1", "0", "1"]


def test_sphere_pack_hits_porosity_and_is_seeded(pkg):
    geo = pkg.geometry
    a = geo.sphere_pack((32, 32, 32), 4.0, 0.6, seed=5)
    b = geo.sphere_pack((32, 32, 32), 4.0, 0.6, seed=5)
    np.testing.assert_array_equal(a, b)
    por = geo.porosity(a)
    assert 0.55 < por <= 0.6
    assert not np.array_equal(a, geo.sphere_pack((32, 32, 32), 4.0, 0.6, seed=6))


def _fibres_brute_force(shape, radius, porosity, seed):
    """Every voxel against every fibre, in-plane displacements by the periodic minimum image
    (no windows): the same fibre set as geometry.fibre_pack, built the slow way."""
    L = np.asarray(shape, float)
    rng = np.random.default_rng(seed)
    solid = np.zeros(shape, bool)
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), -1).astype(float)
    while 1.0 - solid.mean() > porosity:
        c = rng.uniform(0, 1, 3) * L
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        a = int(np.argmax(np.abs(v)))
        dk = idx[..., a] - c[a]
        dk -= L[a] * np.round(dk / L[a])
        d = np.zeros(shape + (3,))
        for x in range(3):
            if x == a:
                continue
            px = c[x] + v[x] * dk / v[a]
            d[..., x] = idx[..., x] - px
            d[..., x] -= L[x] * np.round(d[..., x] / L[x])
        t = d @ v
        solid |= (d * d).sum(-1) - t * t <= radius ** 2
    return solid.astype(np.int32)


def test_fibre_pack_matches_brute_force(pkg):
    for shape, r, phi, seed in (((40, 36, 32), 3.0, 0.85, 5), ((24, 30, 28), 2.0, 0.7, 9)):
        fast = pkg.geometry.fibre_pack(shape, r, phi, seed=seed)
        slow = _fibres_brute_force(shape, r, phi, seed)
        assert 1.0 - fast.mean() <= phi
        # boundary voxels may flip by round-off in the distance test; none in these cases
        assert np.array_equal(fast, slow)


def test_grf_sample_porosity(pkg):
    s = pkg.geometry.grf_sample(40, 6.0, 0.3, seed=3)
    assert abs(pkg.geometry.porosity(s) - 0.3) < 1e-3


def test_golden_fixtures_are_data_only():
    d = os.path.join(ROOT, "tests", "golden")
    for name in os.listdir(d):
        if name.endswith(".npz"):
            with np.load(os.path.join(d, name), allow_pickle=False) as z:
                assert "solid" in z.files and "params" in z.files


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: asked for more it prints one JSON line saying so and exits
    with status 2, before it loads the library or touches a device."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
