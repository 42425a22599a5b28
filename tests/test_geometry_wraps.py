"""lbm_geometry_wraps (host code, no device): does a cluster of pore voxels, joined by the 18
D3Q19 links with the box periodic on every axis, wrap round the box along an axis?

The library decides it with a union-find that carries each node's displacement from its set's
root.  It is checked here against a breadth-first search that labels every voxel with its
unwrapped coordinate and reports a wrap when a voxel is reached at two different positions,
and against cases whose answer is fixed by construction: a 6-connected channel, a closed
cavity, a staircase that only the edge-diagonal links join (wraps), and one joined only by body
diagonals (not D3Q19 links: does not wrap).
"""
import json
import os
from collections import deque

import numpy as np
import pytest

from conftest import ROOT, load_package

LINKS = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)] + \
        [(a, b, 0) for a in (1, -1) for b in (1, -1)] + \
        [(a, 0, b) for a in (1, -1) for b in (1, -1)] + \
        [(0, a, b) for a in (1, -1) for b in (1, -1)]


def bfs_wraps(solid, axis):
    shape = solid.shape
    unset = np.iinfo(np.int64).min
    pos = np.full(shape, unset, dtype=np.int64)
    for start in map(tuple, np.argwhere(solid == 0)):
        if pos[start] != unset:
            continue
        pos[start] = start[axis]
        todo = deque([start])
        while todo:
            v = todo.popleft()
            for e in LINKS:
                w = tuple((v[c] + e[c]) % shape[c] for c in range(3))
                if solid[w]:
                    continue
                p = pos[v] + e[axis]
                if pos[w] == unset:
                    pos[w] = p
                    todo.append(w)
                elif pos[w] != p:
                    return True
    return False


@pytest.fixture(scope="module")
def geo():
    return load_package().geometry


def test_channel_and_cavity(geo):
    s = np.ones((8, 6, 5), dtype=np.int32)
    s[:, 2, 3] = 0                                   # straight channel along x
    assert [geo.pores_wrap(s, a) for a in range(3)] == [True, False, False]
    c = np.ones((8, 6, 5), dtype=np.int32)
    c[2:6, 1:4, 1:4] = 0                             # closed cavity
    assert [geo.pores_wrap(c, a) for a in range(3)] == [False, False, False]
    assert all(geo.pores_wrap(np.zeros((4, 5, 6), dtype=np.int32), a) for a in range(3))


def test_edge_diagonals_join_body_diagonals_do_not(geo):
    n = 5
    s = np.ones((n, n, 4), dtype=np.int32)
    for i in range(n):
        s[i, i, 1] = 0                               # (1,1,0) staircase: D3Q19 links, wraps in x and y
    assert [geo.pores_wrap(s, a) for a in range(3)] == [True, True, False]
    b = np.ones((n, n, n), dtype=np.int32)
    for i in range(n):
        b[i, i, i] = 0                               # (1,1,1) staircase: no D3Q19 link joins it
    assert [geo.pores_wrap(b, a) for a in range(3)] == [False, False, False]


def test_random_boxes_match_breadth_first_search(geo):
    """Porosities 0.09-0.22, around the 18-neighbour site-percolation threshold, so that boxes
    wrap along all, some or none of the axes."""
    seen = set()
    for seed in range(10):
        rng = np.random.default_rng(300 + seed)
        shape = tuple(int(x) for x in rng.integers(6, 13, 3))
        solid = (rng.uniform(size=shape) < rng.uniform(0.78, 0.91)).astype(np.int32)
        if not (solid == 0).any():
            solid[0, 0, 0] = 0
        for axis in range(3):
            w = bfs_wraps(solid, axis)
            assert geo.pores_wrap(solid, axis) == w, (seed, shape, axis)
            seen.add(w)
    assert seen == {True, False}


def test_serving_stream_sample_without_throughflow(geo):
    """Sample 30 of the seed-7 serving stream (113^3, porosity 0.177) has a reference k of
    2.4e-7, five decades below the rest of the stream: its pores wrap along y and z but not
    along x, the direction of the body force."""
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "serving_stream_k.json")))
    rng = np.random.default_rng(7)
    edges = rng.integers(48, 161, 32)
    pors = rng.uniform(0.15, 0.45, 32)
    s = geo.grf_sample(int(edges[30]), 6.0, float(pors[30]), seed=7 * 1000 + 30)
    assert ref["samples"][30]["k"] < 1e-6
    assert [geo.pores_wrap(s, a) for a in range(3)] == [False, True, True]
    s19 = geo.grf_sample(int(edges[19]), 6.0, float(pors[19]), seed=7 * 1000 + 19)
    assert geo.pores_wrap(s19, 0)                    # the next-lowest k of the stream, 3.1e-2


def test_rejects_bad_axis(geo):
    with pytest.raises(Exception):
        geo.pores_wrap(np.zeros((3, 3, 3), dtype=np.int32), 3)
