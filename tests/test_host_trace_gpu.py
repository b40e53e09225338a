"""The host traversal (kp_intersect_closest_host / kp_occluded_host, the
thread-safe per-ray path) against the GPU kernels on the same context: hits
bit-identical on the intersection microbench and on a two-level instanced
scene."""
import threading

import numpy as np
import pytest

from kazen_proto_amd import native, scenes
from test_instancing_gpu import build, mixed_scene, random_rays

pytestmark = pytest.mark.gpu


def mismatches(a, b):
    return int(np.count_nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=1)))


def host_threaded(ctx, rays, threads=16):
    out = np.empty((rays.shape[0], 4), np.float32)
    parts = np.array_split(np.arange(rays.shape[0]), threads)

    def work(ix):
        out[ix] = ctx.intersect_closest_host(rays[ix])

    ts = [threading.Thread(target=work, args=(ix,)) for ix in parts]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return out


def test_host_equals_gpu_on_microbench():
    tris = scenes.microbench_triangles()
    prim, inc = scenes.microbench_rays()
    ctx = native.Context(0)
    ctx.mesh_add_soup(tris)
    ctx.bvh_build()
    for name, rays in (("primary", prim[:200_000]), ("incoherent", inc[:400_000])):
        rays = np.ascontiguousarray(rays)
        gpu = ctx.intersect_closest(rays)
        host = host_threaded(ctx, rays)
        mm = mismatches(host, gpu)
        print("%s: %d rays, %d mismatches" % (name, rays.shape[0], mm), flush=True)
        assert mm == 0
    r = np.ascontiguousarray(inc[:100_000]).copy()
    r[:, 7] = 0.5
    assert np.array_equal(ctx.occluded_host(r), ctx.occluded(r))
    ctx.close()


@pytest.mark.parametrize("with_world", [True, False])
def test_host_equals_gpu_two_level(with_world):
    rng, world, protos, placements = mixed_scene(51, n_inst=200, with_world=with_world)
    ctx = build(world, protos, placements)
    assert ctx.stats()["two_level"] == 1
    rays = random_rays(rng, 50000, aim=6.0)
    gpu = ctx.intersect_closest(rays)
    host = host_threaded(ctx, rays)
    assert mismatches(host, gpu) == 0
    assert np.array_equal(ctx.occluded_host(rays), ctx.occluded(rays))
    ctx.close()
