"""A BVH deep enough that the traversal stack leaves LDS.

120 clusters of 8 triangles at x = 2^-i (i = 0..119), each triangle filling
the half-square y + z <= -1/4 of [-1/4, 0]^2, and a target triangle at
x = 2.  A ray along the x axis from x = -1 enters every cluster's
box (the boxes touch y = z = 0) and misses every cluster's triangles, so it
reaches the target only after visiting the whole tree, nearest first: each
level leaves its other inner children pending.  The kernels keep 8 pending
groups per lane in LDS and spill the rest to HBM; the CPU test checks, by
following the kernel's traversal order over the built BVH8, that this scene
needs more than 8; the GPU test checks the hits against the oracle."""
import struct

import numpy as np
import pytest

import oracle
from kazen_proto_amd import native

LDS_STACK = 8


def deep_scene(levels=120, per=8):
    tris = []
    s = 0.25
    for i in range(levels):
        x = 2.0 ** -i
        for k in range(per):
            xx = x * (1 + 0.01 * k)
            tris.append([xx, -s, -s, xx, 0.0, -s, xx, -s, 0.0])
    tris.append([2.0, -5, -5, 2.0, 5, -5, 2.0, 0, 5])
    return np.asarray(tris, np.float32)


def deep_rays(n_random=20000, seed=4):
    rng = np.random.default_rng(seed)
    axis = np.zeros((64, 8), np.float32)
    axis[:, 0] = -1.0 - rng.random(64)
    axis[:, 1] = -rng.random(64) * 1e-38 * (rng.random(64) < 0.5)
    axis[:, 2] = -rng.random(64) * 1e-38 * (rng.random(64) < 0.5)
    axis[:, 4] = 1.0
    axis[:, 7] = np.inf
    r = np.zeros((n_random, 8), np.float32)
    r[:, :3] = rng.uniform(-1.5, 2.5, (n_random, 3)) * [1, 0.5, 0.5]
    r[:, 4:7] = rng.normal(size=(n_random, 3))
    r[:, 7] = np.inf
    return np.concatenate([axis, r])


def _decode(nodes):
    out = []
    for i in range(nodes.size // 128):
        b = nodes[128 * i:128 * (i + 1)].tobytes()
        px, py, pz = struct.unpack("<3f", b[0:12])
        ex, ey, ez, imask = struct.unpack("<3bB", b[12:16])
        cb = struct.unpack("<I", b[16:20])[0]
        meta = list(b[24:32])
        q = np.frombuffer(b[32:128], dtype=np.float16).astype(np.float64).reshape(6, 8)
        lo = np.stack([px + q[0] * 2.0 ** ex, py + q[1] * 2.0 ** ey, pz + q[2] * 2.0 ** ez], 1)
        hi = np.stack([px + q[3] * 2.0 ** ex, py + q[4] * 2.0 ** ey, pz + q[5] * 2.0 ** ez], 1)
        out.append((imask, cb, meta, lo, hi))
    return out


def pending_groups(nodes, o, d):
    """Most pending groups a lane holds tracing (o, d) to the end, visiting
    the nearest hit inner child of each node first and keeping the others as
    one group (the kernels' order), with no closest-hit culling."""
    dec = _decode(nodes)
    inv = 1.0 / np.where(d == 0, 1e-300, d)
    stack, most, cur = [], 0, 0
    while True:
        if cur is not None:
            imask, cb, meta, lo, hi = dec[cur]
            hs = []
            for s in range(8):
                if meta[s] == 0 or not (imask >> s) & 1:
                    continue
                t0, t1 = (lo[s] - o) * inv, (hi[s] - o) * inv
                tn, tf = np.max(np.minimum(t0, t1)), np.min(np.maximum(t0, t1))
                if tn <= tf and tf >= 0:
                    hs.append((max(tn, 0.0), s))
            hs.sort()
            cur = None
            if hs:
                def child(s):
                    return cb + bin(imask & ((1 << s) - 1)).count("1")
                cur = child(hs[0][1])
                if len(hs) > 1:
                    stack.append([child(s) for _, s in hs[1:]])
                    most = max(most, len(stack))
        if cur is None:
            while stack and not stack[-1]:
                stack.pop()
            if not stack:
                return most
            cur = stack[-1].pop(0)


def test_deep_scene_needs_more_than_the_lds_stack():
    t = deep_scene()
    nodes, _, depth = native.bvh8_build_host(t)
    groups = pending_groups(nodes, np.array([-1.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0]))
    print("BVH8 depth %d, pending groups along the axis %d" % (depth, groups))
    assert depth >= 12 and groups > LDS_STACK
    # the host traversal (no stack limit of its kind) agrees with brute force
    rays = deep_rays(2000)
    got = native.HostBvh8(t).closest(rays)
    ref = oracle.closest_bruteforce(t, rays)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.all(got[:64, 1].view(np.uint32) == t.shape[0] - 1)   # the axis rays reach the target


@pytest.mark.gpu
def test_deep_bvh_spills_and_stays_exact():
    t = deep_scene()
    rays = deep_rays()
    ref = oracle.closest_bruteforce(t, rays)
    ctx = native.Context(0)
    ctx.mesh_add_soup(t)
    ctx.bvh_build()
    got = ctx.intersect_closest(rays)
    ctx.synchronize()   # a stack overflow or step limit would be reported here
    mm = int(np.count_nonzero(np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1)))
    print("deep BVH: %d rays, %d mismatches" % (rays.shape[0], mm), flush=True)
    assert mm == 0
    occ = ctx.occluded(rays)
    assert np.array_equal(occ.astype(bool), ref[:, 1].view(np.uint32) != 0xFFFFFFFF)
    ctx.close()
    # the same geometry as a prototype placed twice: the two-level kernels
    ctx = native.Context(0)
    b = ctx.blas_add_soup(t)
    eye = np.hstack([np.eye(3), np.zeros((3, 1))]).astype(np.float32)
    shifted = eye.copy()
    shifted[1, 3] = 20.0
    ctx.instance_add(b, eye)
    ctx.instance_add(b, shifted)
    ctx.bvh_build(instances=native.INSTANCES_TWO_LEVEL)
    got2 = ctx.intersect_closest(rays)
    ctx.synchronize()
    _, _, w2o = ctx.instance_get(1)
    ref2 = oracle.instanced_closest([t], [(0, eye), (0, w2o)], rays)
    ctx.close()
    assert np.array_equal(got2.view(np.uint32), ref2.view(np.uint32))
