"""Structure of the BVH8 the builder emits (csrc/bvh_build.cpp: binned-SAH
BVH2 collapsed by the cost-optimal plan), checked on the host records: every
triangle sits in exactly one leaf, leaves hold at most leaf_size triangles
and tile their node's triangle range, inner children are numbered as the
traversal computes them (child_base + rank of the slot), every node is
reached exactly once, and no node but the root has fewer than two children
(the plan always splits a wide node's subtree into at least two roots).
Hits stay equal to brute force, the layout the traversal kernels and the
host traversal read (csrc/kp_layout.h) being the same.  Replaces the
acceleration-structure build behind Accelerator::build (reference
src/core/accel.h:26-32, src/core/accel.cpp:185-346)."""
import numpy as np
import pytest

import oracle
from kazen_proto_amd import native
from test_instancing_gpu import random_rays, random_tris


def mismatches(a, b):
    return int(np.count_nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=1)))


def parse(nodes):
    rec = nodes.reshape(-1, 128)
    words = rec.view(np.uint32).reshape(-1, 32)
    imask = rec[:, 15].astype(np.uint32)
    child_base = words[:, 4]
    prim_base = words[:, 5]
    meta = rec[:, 24:32].astype(np.uint32)
    return imask, child_base, prim_base, meta


def check_structure(nodes, tris12, n_tris, leaf_size, width):
    imask, child_base, prim_base, meta = parse(nodes)
    n_nodes = imask.shape[0]
    covered = np.zeros(n_tris, np.int32)
    reached = np.zeros(n_nodes, np.int32)
    reached[0] = 1
    for i in range(n_nodes):
        used = int(np.count_nonzero(meta[i]))
        assert 1 <= used <= width
        if i > 0:
            assert used >= 2, "node %d has one child" % i
        rank = 0
        leaves = []
        for s in range(8):
            m = int(meta[i, s])
            inner = (int(imask[i]) >> s) & 1
            if m == 0:
                assert not inner
            elif inner:
                assert m == 1
                c = int(child_base[i]) + rank
                rank += 1
                assert 0 < c < n_nodes
                reached[c] += 1
            else:
                count, off = m >> 5, m & 31
                assert 1 <= count <= leaf_size
                leaves.append((off, count))
        # leaf ranges tile [prim_base, prim_base + total) in slot order
        pos = 0
        for off, count in leaves:
            assert off == pos
            pos += count
        first = int(prim_base[i])
        covered[first:first + pos] += 1
    assert np.all(reached == 1)
    assert np.all(covered == 1)
    prims = tris12.reshape(-1, 12)[:, 3].view(np.uint32)
    assert np.array_equal(np.sort(prims), np.arange(n_tris, dtype=np.uint32))


@pytest.mark.parametrize("n_tris", [1, 2, 4, 5, 9, 33, 3000])
@pytest.mark.parametrize("leaf", [1, 4])
def test_every_triangle_in_one_leaf(n_tris, leaf):
    rng = np.random.default_rng(100 + n_tris + leaf)
    tris = random_tris(rng, n_tris, spread=1.0, size=0.1)
    nodes, t12, _ = native.bvh8_build_host(tris, leaf_size=leaf)
    check_structure(nodes, t12, n_tris, leaf, 8)
    rays = random_rays(rng, 4000, box=1.3)
    h, _, _ = oracle.bvh8_closest(nodes, t12, rays)
    assert mismatches(h, oracle.closest_bruteforce(tris, rays)) == 0


@pytest.mark.parametrize("width", [2, 3, 5, 8])
def test_structure_at_every_node_width(width):
    rng = np.random.default_rng(7 + width)
    tris = random_tris(rng, 5000, spread=1.0, size=0.05)
    nodes, t12, _ = native.bvh8_build_host(tris, node_width=width)
    check_structure(nodes, t12, 5000, 4, width)


def test_coincident_triangles():
    # every centroid equal: no SAH split exists, the BVH2 splits by index
    one = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    tris = np.repeat(one, 200, axis=0)
    nodes, t12, _ = native.bvh8_build_host(tris)
    check_structure(nodes, t12, 200, 4, 8)
    rays = np.zeros((3, 8), np.float32)
    rays[:, 0:3] = [0.25, 0.25, -1.0]
    rays[:, 4:7] = [0.0, 0.0, 1.0]
    rays[:, 7] = np.inf
    h, _, _ = oracle.bvh8_closest(nodes, t12, rays)
    # equal distances: the lowest primitive wins
    assert np.all(h[:, 1].view(np.uint32) == 0) and np.allclose(h[:, 0], 1.0)


def test_collapse_flattens_small_subtrees():
    # the cost-optimal plan leaves few half-empty nodes: on a uniform soup
    # the average node holds close to the full width (6.35 of 8 here, 6.48
    # on the 1,148,576-triangle intersection microbench)
    rng = np.random.default_rng(3)
    tris = random_tris(rng, 20000, spread=1.0, size=0.03)
    nodes, _, _ = native.bvh8_build_host(tris)
    _, _, _, meta = parse(nodes)
    per_node = np.count_nonzero(meta, axis=1)
    assert per_node.mean() > 5.0
