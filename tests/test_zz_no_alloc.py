"""kp_intersect_closest_noalloc (static device scratch, no allocation in the
call) against kp_intersect_closest on the same device buffers: bit-identical
hits on two calls in a row, the second with a smaller input, one- and
two-level BVHs."""
import numpy as np
import pytest

from kazen_proto_amd import native, scenes
from test_instancing_gpu import build, mixed_scene, random_rays

pytestmark = pytest.mark.gpu


def mismatches(a, b):
    return int(np.count_nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=1)))


def check_two_calls(ctx, rays_np, n_small):
    rays = native.DeviceBuffer.from_array(ctx, rays_np)
    ref = ctx.intersect_closest(rays)
    ctx.synchronize()
    want = ref.numpy()
    out = native.DeviceBuffer(ctx, (rays_np.shape[0], 4), np.float32)
    ctx.intersect_closest_noalloc(rays, out)
    ctx.synchronize()
    got = out.numpy()
    print("noalloc call 1: %d rays, %d mismatches" % (rays_np.shape[0], mismatches(got, want)), flush=True)
    assert mismatches(got, want) == 0
    # second call, smaller input, into a cleared output: the first n_small
    # hits are written, the rest stay untouched
    out.upload(np.full((rays_np.shape[0], 4), -1.0, np.float32))
    ctx.intersect_closest_noalloc(rays, out, n=n_small)
    ctx.synchronize()
    got2 = out.numpy()
    print("noalloc call 2: %d rays, %d mismatches" % (n_small, mismatches(got2[:n_small], want[:n_small])),
          flush=True)
    assert mismatches(got2[:n_small], want[:n_small]) == 0
    assert np.all(got2[n_small:] == -1.0)
    for b in (rays, ref, out):
        b.free()


def test_noalloc_matches_closest_one_level():
    tris = scenes.microbench_triangles()
    _, inc = scenes.microbench_rays()
    ctx = native.Context(0)
    ctx.mesh_add_soup(tris)
    ctx.bvh_build()
    check_two_calls(ctx, np.ascontiguousarray(inc[:1_000_000]), 123_457)
    ctx.close()


def test_noalloc_matches_closest_two_level():
    rng, world, protos, placements = mixed_scene(61, n_inst=300)
    ctx = build(world, protos, placements)
    assert ctx.stats()["two_level"] == 1
    check_two_calls(ctx, random_rays(rng, 60_000, aim=6.0), 5_001)
    ctx.close()


def test_noalloc_argument_errors():
    ctx = native.Context(0)
    rays = native.DeviceBuffer(ctx, (4, 8), np.float32)
    out = native.DeviceBuffer(ctx, (4, 4), np.float32)
    with pytest.raises(native.KpError) as e:   # before kp_bvh_build
        ctx.intersect_closest_noalloc(rays, out)
    assert e.value.code == 2
    ctx.mesh_add_soup(scenes.uv_sphere(8, 4).reshape(-1, 9))
    ctx.bvh_build()
    assert native.lib().kp_intersect_closest_noalloc(ctx.handle, None, 4, out.ptr) == 1   # KP_ERR_INVALID
    assert native.lib().kp_intersect_closest_noalloc(ctx.handle, rays.ptr, 0, out.ptr) == 0
    rays.free()
    out.free()
    ctx.close()
