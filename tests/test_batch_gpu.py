"""Batched scenes in one workload (kp_batch_*).

Every path of a batch draws the same PCG32 stream, camera ray and BVH hits as
in a render of its scene alone, and the films fold samples in the same order
with the same Kahan compensation, so each scene's film must equal the
single-scene render bit for bit: eagerly, through the replayed pass graph,
for any pass size (passes that split a pixel's samples included), with scenes
of different depths, and after autotuning."""
import numpy as np
import pytest

from kazen_proto_amd import native, scenes

pytestmark = pytest.mark.gpu


def _scenes():
    return [
        scenes.cornell_box(width=40, height=32, spp=8, sphere_lon=16, sphere_lat=8),
        scenes.materials_scene(width=24, height=24, spp=4),
        scenes.terrain_scene(grid=64, n_spheres=6, sphere_lon=12, sphere_lat=6, width=32, height=24, spp=6,
                             max_depth=5, seed=21),
        scenes.furnace_box(albedo=0.5, width=16, height=16, spp=3, max_depth=3),
        scenes.guiding_scene(width=24, height=24, spp=5, max_depth=10, n_spheres=8),
    ]


def _alone(s):
    c = native.Context(0)
    c.load_scene(s)
    img = c.render(native.make_render_desc(s))
    c.close()
    return img


def _batch(owner, ss):
    ctxs = []
    for s in ss:
        c = native.Context(0)
        c.load_scene(s)
        ctxs.append(c)
    b = native.Batch(owner, ctxs)
    for c in ctxs:
        c.close()   # the batch holds its own copies
    return b


def _check(b, ss, alone, tag):
    for i, s in enumerate(ss):
        img = b.film_read(i, s.width, s.height)
        diff = int(np.count_nonzero(img.view(np.uint32) != alone[i].view(np.uint32)))
        assert diff == 0, "%s: scene %d (%s) differs in %d values" % (tag, i, s.name, diff)


@pytest.fixture(scope="module")
def setup():
    ss = _scenes()
    alone = [_alone(s) for s in ss]
    owner = native.Context(0)
    b = _batch(owner, ss)
    descs = [native.make_render_desc(s) for s in ss]
    yield ss, alone, owner, b, descs
    b.close()
    owner.close()


@pytest.mark.parametrize("use_graph,pass_paths", [(0, 1 << 21), (1, 1 << 21), (0, 777), (1, 3000)])
def test_batch_films_equal_single_scene_renders(setup, use_graph, pass_paths):
    ss, alone, owner, b, descs = setup
    b.set_tuning(use_graph=use_graph, pass_paths=pass_paths)
    b.render(descs)
    _check(b, ss, alone, "graph %d pass %d" % (use_graph, pass_paths))
    st = b.stats()
    work = sum(s.width * s.height * s.spp for s in ss)
    assert st["paths"] == work
    assert st["passes"] == (work + pass_paths - 1) // pass_paths
    assert st["extension_rays"] >= work and st["shadow_rays"] > 0


def test_graph_replays_match_with_eager_work_between(setup):
    ss, alone, owner, b, descs = setup
    other = scenes.cornell_box(width=16, height=16, spp=2, sphere_lon=8, sphere_lat=4)
    owner.load_scene(other)
    other_ref = owner.render(native.make_render_desc(other))
    b.set_tuning(use_graph=1, pass_paths=1024)
    for rep in range(3):
        b.render(descs)
        _check(b, ss, alone, "replay %d" % rep)
        # eager work on the same stream between replays
        again = owner.render(native.make_render_desc(other))
        assert np.array_equal(again.view(np.uint32), other_ref.view(np.uint32))


def test_autotune_keeps_results(setup):
    ss, alone, owner, b, descs = setup
    t = b.autotune(descs, passes=2)
    d = t.as_dict()
    print("autotuned:", d)
    assert d["pass_paths"] in (1 << 21, 1 << 22, 1 << 23, 1 << 24, 1 << 25)
    assert 1 <= d["shade_blocks_per_cu"] <= 16
    assert d["refill_below"] in (16, 32, 40, 48, 56)
    b.render(descs)
    _check(b, ss, alone, "autotuned")


def test_batch_rejects_guided_and_bad_input(setup):
    ss, alone, owner, b, descs = setup
    bad = [native.make_render_desc(s) for s in ss]
    bad[0].flags = native.RENDER_GUIDED
    with pytest.raises(native.KpError):
        b.render(bad)
    with pytest.raises(native.KpError):
        b.set_tuning(pass_paths=16)
    with pytest.raises(native.KpError):
        b.set_tuning(pass_paths=1 << 26)   # at most 2^25 paths per batch pass
    with pytest.raises(native.KpError):
        b.film_read(len(ss), 4, 4)


def test_batch_with_baked_instances_and_two_level_rejected():
    """A scene whose instances were baked into its one-level BVH batches like
    any other (films bit-identical to its render alone); a two-level scene is
    rejected by kp_batch_create."""
    inst = scenes.terrain_scene(grid=64, n_spheres=6, sphere_lon=12, sphere_lat=6, width=32, height=24, spp=4,
                                max_depth=5, seed=23, instanced=True)
    ss = [inst, scenes.cornell_box(width=24, height=24, spp=4, sphere_lon=16, sphere_lat=8)]
    alone = [_alone(s) for s in ss]
    owner = native.Context(0)
    b = _batch(owner, ss)
    b.render([native.make_render_desc(s) for s in ss])
    _check(b, ss, alone, "baked instances")
    b.close()
    two = native.Context(0)
    two.load_scene(inst, instances=native.INSTANCES_TWO_LEVEL)
    assert two.stats()["two_level"] == 1
    with pytest.raises(native.KpError):
        native.Batch(owner, [two])
    two.close()
    owner.close()


def test_batch_of_large_scenes_matches_renders_alone():
    """Scenes large enough that copying them into the merged arrays takes a
    while (≈ 1.1 M triangles in all, 26 MB of BVH8 nodes and triangles): the
    merged BVHs must be complete and rebased before any traversal runs, and
    the scene contexts may be closed as soon as kp_batch_create returns.
    Every film must equal the scene's render alone bit for bit and the batch
    must trace exactly the rays of the single-scene renders."""
    ss = [scenes.terrain_scene(grid=g, n_spheres=8, sphere_lon=16, sphere_lat=8, width=48, height=32, spp=4,
                               max_depth=6, seed=40 + i) for i, g in enumerate((512, 384, 256, 320))]
    alone, rays_alone = [], 0
    for s in ss:
        c = native.Context(0)
        c.load_scene(s)
        alone.append(c.render(native.make_render_desc(s)))
        st = c.render_stats()
        rays_alone += st["extension_rays"] + st["shadow_rays"]
        c.close()
    owner = native.Context(0)
    b = _batch(owner, ss)
    b.render([native.make_render_desc(s) for s in ss])
    _check(b, ss, alone, "large scenes")
    st = b.stats()
    assert st["extension_rays"] + st["shadow_rays"] == rays_alone
    b.close()
    owner.close()


def test_batch_workload_32_uneven_scenes():
    """The batched workload itself: 32 terrain scenes (seeds 100-131, 1e4-2e6
    triangles, 128^2-1024^2, 4-64 spp, depth 8) rendered as one autotuned
    batch through the replayed pass graph.  Every film equals the scene's
    render alone bit for bit and the batch traces the same rays; then, at one
    sample per pixel, 512 pixels of each of four scenes are traced path by
    path by the oracle (Renderer::render's per-pixel paths,
    src/core/renderer.cpp:18-90) and must agree."""
    import os

    import oracle
    ss = scenes.batch_scenes(32, seed=100)
    print("batch workload: %d scenes, %d triangles" % (len(ss), sum(s.n_tris for s in ss)), flush=True)
    owner = native.Context(0)
    ctxs, alone, rays_alone = [], [], 0
    for s in ss:
        c = native.Context(0)
        c.load_scene(s)
        alone.append(c.render(native.make_render_desc(s)))
        st = c.render_stats()
        rays_alone += st["extension_rays"] + st["shadow_rays"]
        c.release_buffers()
        ctxs.append(c)
    print("rendered alone: %d rays" % rays_alone, flush=True)
    b = native.Batch(owner, ctxs)
    descs = [native.make_render_desc(s) for s in ss]
    tuned = b.autotune(descs, passes=2).as_dict()
    b.render(descs)
    st = b.stats()
    print("batch: %s, %d rays" % (tuned, st["extension_rays"] + st["shadow_rays"]), flush=True)
    assert st["extension_rays"] + st["shadow_rays"] == rays_alone
    _check(b, ss, alone, "32-scene workload")
    # one sample per pixel against the oracle, path by path
    d1 = [native.make_render_desc(s, spp=1) for s in ss]
    b.render(d1)
    rng = np.random.default_rng(5)
    order = np.argsort([s.n_tris for s in ss])
    picks = [int(order[0]), int(order[1]), int(order[len(ss) // 2]), int(order[-1])]
    threads = min(16, os.cpu_count() or 1)
    agree = total = 0
    for i in picks:
        s = ss[i]
        img = b.film_read(i, s.width, s.height).reshape(-1, 3).astype(np.float64)
        px = rng.choice(s.width * s.height, 512, replace=False)
        d = oracle.RenderDesc()
        d.width, d.height, d.spp, d.max_depth, d.rr_depth, d.seed = s.width, s.height, 1, s.max_depth, s.rr_depth, s.seed
        d.spp_offset = 0
        d.eye[:], d.lookat[:], d.up[:] = s.eye, s.lookat, s.up
        d.tan_half_fov = s.tan_half_fov
        ref = oracle.scene_from(s).render_pixels(d, px, threads=threads)
        ok = np.all(np.abs(img[px] - ref) <= 2e-3 * np.abs(ref) + 1e-5, axis=1)
        print("scene %d (%d tris): %d / %d paths agree" % (i, s.n_tris, ok.sum(), ok.size), flush=True)
        agree += int(ok.sum())
        total += ok.size
    b.close()
    for c in ctxs:
        c.close()
    owner.close()
    assert agree >= 0.999 * total
