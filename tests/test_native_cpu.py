"""CPU-side checks of the HIP library: it loads, exports every entry point
the public header declares, and its host BVH8 builder produces a structure
whose traversal (by the oracle, exact arithmetic) equals brute force."""
import glob
import os
import re

import numpy as np
import pytest

import oracle
from kazen_proto_amd import native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names.update(re.findall(r"\b(kp_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes signature table covers the whole header
    assert set(names) <= set(native.SIGNATURES), sorted(set(names) - set(native.SIGNATURES))


def _decode(nodes):
    n = nodes.reshape(-1, 128)
    p = n[:, 0:12].copy().view(np.float32)
    e = n[:, 12:15].astype(np.int8).astype(np.int32)
    imask = n[:, 15]
    base = n[:, 16:24].copy().view(np.uint32)
    meta = n[:, 24:32]
    # planes are binary16 integers 0..2047 (exact)
    q = n[:, 32:128].copy().view(np.float16).reshape(-1, 6, 8).astype(np.float64)
    assert np.all(q == np.round(q)) and q.min() >= 0 and q.max() <= 2047
    sc = np.ldexp(1.0, e)
    lo = p[:, :, None].astype(np.float64) + q[:, 0:3] * sc[:, :, None]
    hi = p[:, :, None].astype(np.float64) + q[:, 3:6] * sc[:, :, None]
    return lo, hi, imask, base, meta


def test_bvh8_structure_invariants():
    tris = scenes.small_test_scene(seed=4, n_random=3000)
    nodes, t12, depth = native.bvh8_build_host(tris, leaf_size=4)
    lo, hi, imask, base, meta = _decode(nodes)
    t12 = t12.reshape(-1, 12)
    # every triangle stored exactly once, with its caller index
    prim = t12[:, 3].view(np.uint32)
    assert np.array_equal(np.sort(prim), np.arange(tris.shape[0], dtype=np.uint32))
    verts = t12[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(-1, 3, 3)
    assert np.array_equal(verts.reshape(-1, 9), tris[prim])
    # Walk every root-to-leaf path: each triangle must lie inside the
    # intersection of all slot boxes above it (quantised boxes of a child need
    # not nest inside the parent's quantised box; the geometry must).
    seen_nodes = np.zeros(lo.shape[0], np.int32)
    seen_nodes[0] = 1
    stack = [(0, np.full(3, -np.inf), np.full(3, np.inf))]
    n_leaf_tris = 0
    while stack:
        i, clo, chi = stack.pop()
        rank = 0
        for s in range(8):
            m = meta[i, s]
            if m == 0:
                continue
            blo, bhi = np.maximum(clo, lo[i, :, s]), np.minimum(chi, hi[i, :, s])
            if (imask[i] >> s) & 1:
                c = base[i, 0] + rank
                rank += 1
                seen_nodes[c] += 1
                stack.append((c, blo, bhi))
            else:
                first, cnt = int(base[i, 1]) + int(m & 31), int(m >> 5)
                assert 1 <= cnt <= 4
                v = verts[first:first + cnt].reshape(-1, 3)
                assert np.all(v >= blo) and np.all(v <= bhi)
                n_leaf_tris += int(cnt)
    assert np.all(seen_nodes == 1)   # a tree: every node has one parent
    assert n_leaf_tris == tris.shape[0]
    assert depth >= 2


@pytest.mark.parametrize("leaf", [1, 2, 4])
def test_bvh8_traversal_matches_bruteforce(leaf):
    g = np.load(os.path.join(ROOT, "tests", "golden", "trace_small.npz"))
    nodes, t12, _ = native.bvh8_build_host(g["tris"], leaf_size=leaf)
    h8, nv, nt = oracle.bvh8_closest(nodes, t12, g["rays"])
    assert np.array_equal(h8.view(np.uint32), g["hits"].view(np.uint32))
    assert nv > 0 and nt > 0


def test_oracle_reproduces_golden_fixture():
    g = np.load(os.path.join(ROOT, "tests", "golden", "trace_small.npz"))
    assert np.array_equal(oracle.closest_bruteforce(g["tris"], g["rays"]).view(np.uint32),
                          g["hits"].view(np.uint32))
    assert np.array_equal(oracle.Bvh(g["tris"]).closest(g["rays"]).view(np.uint32),
                          g["hits"].view(np.uint32))


def test_host_build_rejects_bad_input():
    with pytest.raises(native.KpError):
        native.bvh8_build_host(np.zeros((0, 9), np.float32))
    bad = scenes.small_test_scene()
    bad[3, 4] = np.nan
    with pytest.raises(native.KpError):
        native.bvh8_build_host(bad)


def test_single_triangle_build():
    tri = np.array([[-1, 0, 1, 2, 0, 1, -1, 0, -2]], np.float32)
    nodes, t12, depth = native.bvh8_build_host(tri)
    r = np.array([[0, 1, 0, 0, 0, -1, 0, np.inf]], np.float32)
    h, _, _ = oracle.bvh8_closest(nodes, t12, r)
    assert h[0, 0] == 1.0 and h[0, 1].view(np.uint32) == 0
