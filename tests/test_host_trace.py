"""Host (CPU) traversal of the product's BVH8, on host-only contexts: the
thread-safe per-ray path that stands in for the reference's const
Hitable::intersect called from concurrent TBB tiles (src/core/hitable.h:12-33,
src/core/renderer.cpp:24-26).  Its hits must equal the oracle's brute force
bit for bit (one level) and oracle.instanced_closest (two levels); the GPU
tests check the same path against the kernels."""
import threading

import numpy as np
import pytest

import oracle
from kazen_proto_amd import native, scenes
from test_instancing_gpu import EYE, MISS, mixed_scene, random_rays, random_tris


def mismatches(a, b):
    return int(np.count_nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=1)))


def soup_and_rays(seed, n_tris=6000, n_rays=30000):
    rng = np.random.default_rng(seed)
    tris = random_tris(rng, n_tris, spread=1.0, size=0.08)
    rays = random_rays(rng, n_rays, box=1.3)
    rays[0, 0] = np.nan
    rays[1, 4:7] = 0.0
    rays[2, 3], rays[2, 7] = 5.0, 1.0
    rays[3, 4] = np.inf
    return tris, rays


@pytest.mark.parametrize("leaf", [1, 4])
def test_blob_closest_equals_bruteforce(leaf):
    tris, rays = soup_and_rays(5 + leaf)
    b = native.HostBvh8(tris, leaf_size=leaf)
    got = b.closest(rays)
    ref = oracle.closest_bruteforce(tris, rays)
    assert mismatches(got, ref) == 0
    assert np.all(got[:4, 1].view(np.uint32) == MISS)
    occ = b.occluded(rays)
    assert np.array_equal(occ.astype(bool), ref[:, 1].view(np.uint32) != MISS)


@pytest.mark.parametrize("width", [2, 3, 4, 6, 7])
def test_narrow_nodes_give_the_same_hits(width):
    # node width (children per node) changes the tree, never the hits; the
    # oracle's traversal of the same records agrees too
    tris, rays = soup_and_rays(40 + width)
    b = native.HostBvh8(tris, node_width=width)
    ref = oracle.closest_bruteforce(tris, rays)
    assert mismatches(b.closest(rays), ref) == 0
    nodes, t12, depth = native.bvh8_build_host(tris, node_width=width)
    wide, _, wide_depth = native.bvh8_build_host(tris)
    assert nodes.size > wide.size and depth >= wide_depth
    meta = nodes.reshape(-1, 128)[:, 24:32]
    assert np.max(np.count_nonzero(meta, axis=1)) <= width
    h, _, _ = oracle.bvh8_closest(nodes, t12, rays)
    assert mismatches(h, ref) == 0


def test_node_width_out_of_range_is_refused():
    tris, _ = soup_and_rays(3, n_tris=100, n_rays=10)
    for w in (1, 9):   # 1 (the device tuner's value) needs a device
        with pytest.raises(native.KpError):
            native.HostBvh8(tris, node_width=w)


def test_closed_mesh_edges_and_vertices_do_not_leak():
    # rays through every vertex and edge midpoint of a closed sphere from its
    # centre: each must hit (watertight), and the hit equals brute force
    tris = scenes.uv_sphere(24, 12, radius=1.0).reshape(-1, 9).astype(np.float32)
    v = tris.reshape(-1, 3, 3)
    pts = np.concatenate([v.reshape(-1, 3), 0.5 * (v[:, 0] + v[:, 1]), 0.5 * (v[:, 1] + v[:, 2])])
    rays = np.zeros((pts.shape[0], 8), np.float32)
    rays[:, 4:7] = pts
    rays[:, 7] = np.inf
    b = native.HostBvh8(tris)
    got = b.closest(rays)
    assert np.all(got[:, 1].view(np.uint32) != MISS)
    assert mismatches(got, oracle.closest_bruteforce(tris, rays)) == 0


def test_host_context_one_level_and_gpu_calls_refused():
    tris, rays = soup_and_rays(21)
    ctx = native.Context(native.DEVICE_HOST)
    ctx.mesh_add_soup(tris[:2000])
    ctx.mesh_add_soup(tris[2000:])
    ctx.bvh_build()
    got = ctx.intersect_closest_host(rays)
    assert mismatches(got, oracle.closest_bruteforce(tris, rays)) == 0
    assert ctx.prim_locate(2500) == (0, 1, 500)
    with pytest.raises(native.KpError) as e:
        ctx.intersect_closest(rays)
    assert e.value.code == 2
    with pytest.raises(native.KpError):
        ctx.synchronize()
    ctx.close()


def test_host_query_before_build_is_an_error():
    ctx = native.Context(native.DEVICE_HOST)
    with pytest.raises(native.KpError) as e:
        ctx.intersect_closest_host(np.zeros((1, 8), np.float32))
    assert "before kp_bvh_build" in str(e.value)
    ctx.close()


@pytest.mark.parametrize("with_world,leaf", [(True, 4), (False, 4), (True, 1)])
def test_host_context_two_level_equals_oracle(with_world, leaf):
    rng, world, protos, placements = mixed_scene(31 + leaf, with_world=with_world)
    ctx = native.Context(native.DEVICE_HOST)
    if world is not None:
        ctx.mesh_add_soup(world)
    ids = [ctx.blas_add_soup(p) for p in protos]
    for pi, o2w in placements:
        ctx.instance_add(ids[pi], o2w)
    ctx.bvh_build(leaf_size=leaf, instances=native.INSTANCES_TWO_LEVEL)
    assert ctx.stats()["two_level"] == 1
    rays = random_rays(rng, 12000, aim=6.0)
    meshes = ([world] if world is not None else []) + list(protos)
    off = 1 if world is not None else 0
    inst = [(0, EYE)] if world is not None else []
    for k, (pi, _) in enumerate(placements):
        b, _, w2o = ctx.instance_get(k)
        inst.append((pi + off, w2o))
    ref = oracle.instanced_closest(meshes, inst, rays)
    got = ctx.intersect_closest_host(rays)
    assert mismatches(got, ref) == 0
    occ = ctx.occluded_host(rays)
    assert np.array_equal(occ.astype(bool), ref[:, 1].view(np.uint32) != MISS)
    ctx.close()


def test_concurrent_host_queries_are_identical():
    # 16 threads query one context at once (ctypes drops the GIL in the call)
    tris, rays = soup_and_rays(41, n_rays=8000)
    ctx = native.Context(native.DEVICE_HOST)
    ctx.mesh_add_soup(tris)
    ctx.bvh_build()
    want = ctx.intersect_closest_host(rays)
    out = [None] * 16
    errs = []

    def work(k):
        try:
            r = np.roll(rays, 97 * k, axis=0)
            acc = np.empty_like(want)
            for i in range(0, r.shape[0], 500):   # many small calls, as per-ray callers make
                acc[i:i + 500] = ctx.intersect_closest_host(r[i:i + 500])
            out[k] = np.roll(acc, -97 * k, axis=0)
        except Exception as e:   # pragma: no cover - reported below
            errs.append(e)

    ts = [threading.Thread(target=work, args=(k,)) for k in range(16)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs
    for k in range(16):
        assert mismatches(out[k], want) == 0
    ctx.close()


def test_guiding_checkpoint_needs_a_network():
    ctx = native.Context(native.DEVICE_HOST)
    with pytest.raises(native.KpError) as e:
        ctx.guiding_save()
    assert e.value.code == 2
    with pytest.raises(native.KpError):
        ctx.guiding_load(b"KPGC" + b"\0" * 100)   # too short
    ctx.close()
