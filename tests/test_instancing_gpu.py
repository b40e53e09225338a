"""Instancing on the GPU against the oracle.

Two-level traversal: closest hits must equal oracle.instanced_closest bit for
bit, with the world-to-object matrices the library traverses with
(kp_instance_get); occlusion must equal "the oracle finds a hit".  Baked
instances: hits must equal the brute force over the placements' triangles
transformed as the library documents (FP64 ((m0 x + m1 y) + m2 z) + m3,
rounded to FP32).  Renders with instances (either way) must match the render
of the same scene with the instances added as meshes."""
import numpy as np
import pytest

import oracle
from kazen_proto_amd import native, scenes

pytestmark = pytest.mark.gpu
EYE = np.hstack([np.eye(3), np.zeros((3, 1))]).astype(np.float32)
MISS = 0xFFFFFFFF


def random_tris(rng, n, center=(0, 0, 0), spread=1.0, size=0.3):
    c = rng.uniform(-spread, spread, (n, 1, 3)) + np.asarray(center)
    return (c + rng.uniform(-size, size, (n, 3, 3))).reshape(n, 9).astype(np.float32)


def random_o2w(rng, spread=6.0, mirror=False):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    rot = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                    [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                    [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    s = rng.uniform(0.4, 1.6, 3)
    if mirror:
        s[0] = -s[0]
    return np.hstack([rot @ np.diag(s), rng.uniform(-spread, spread, (3, 1))]).astype(np.float32)


def random_rays(rng, n, box=8.0, finite_frac=0.3, aim=None):
    """Origins uniform in the box; directions uniform, or toward points
    uniform in [-aim, aim]^3 when aim is given."""
    o = rng.uniform(-box, box, (n, 3))
    d = rng.normal(size=(n, 3)) if aim is None else rng.uniform(-aim, aim, (n, 3)) - o
    r = np.zeros((n, 8), np.float32)
    r[:, :3], r[:, 4:7] = o, d
    r[:, 3] = 0.0
    r[:, 7] = np.inf
    k = int(n * finite_frac)
    r[:k, 3] = rng.uniform(0.0, 1.0, k)
    r[:k, 7] = r[:k, 3] + rng.uniform(0.5, 6.0, k)
    return r


def build(world, protos, placements, leaf=4, mode=native.INSTANCES_TWO_LEVEL):
    ctx = native.Context(0)
    if world is not None:
        ctx.mesh_add_soup(world)
    ids = [ctx.blas_add_soup(p) for p in protos]
    for pi, o2w in placements:
        ctx.instance_add(ids[pi], o2w)
    ctx.bvh_build(leaf_size=leaf, instances=mode)
    return ctx


def baked_soup(world, protos, placements):
    """Placements transformed as kp_bvh_build bakes them: per coordinate
    ((m0 x + m1 y) + m2 z) + m3 in float64, rounded to float32."""
    out = [] if world is None else [np.asarray(world, np.float32)]
    for pi, o2w in placements:
        m = np.asarray(o2w, np.float32).astype(np.float64)
        v = np.asarray(protos[pi], np.float32).reshape(-1, 3).astype(np.float64)
        w = np.empty_like(v)
        for a in range(3):
            w[:, a] = ((m[a, 0] * v[:, 0] + m[a, 1] * v[:, 1]) + m[a, 2] * v[:, 2]) + m[a, 3]
        out.append(w.astype(np.float32).reshape(-1, 9))
    return np.concatenate(out)


def oracle_hits(ctx, world, protos, placements, rays):
    meshes = ([world] if world is not None else []) + list(protos)
    off = 1 if world is not None else 0
    inst = [(0, EYE)] if world is not None else []
    for k, (pi, _) in enumerate(placements):
        b, _, w2o = ctx.instance_get(k)
        assert b == pi
        inst.append((pi + off, w2o))
    return oracle.instanced_closest(meshes, inst, rays)


def mismatches(a, b):
    return int(np.count_nonzero(np.any(a.view(np.uint32) != b.view(np.uint32), axis=1)))


def mixed_scene(seed, n_inst=60, with_world=True):
    rng = np.random.default_rng(seed)
    world = random_tris(rng, 400, spread=7.0, size=0.5) if with_world else None
    protos = [scenes.uv_sphere(16, 8, radius=0.8), random_tris(rng, 150, spread=0.8, size=0.25),
              scenes.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0))]
    placements = [(k % 3, random_o2w(rng, mirror=(k % 7 == 3))) for k in range(n_inst)]
    return rng, world, protos, placements


@pytest.mark.parametrize("with_world,leaf", [(True, 4), (False, 4), (True, 1)])
def test_closest_and_occluded_match_oracle(with_world, leaf):
    rng, world, protos, placements = mixed_scene(11 + leaf, with_world=with_world)
    ctx = build(world, protos, placements, leaf)
    st = ctx.stats()
    assert st["two_level"] == 1
    n_world = 0 if world is None else world.shape[0]
    assert st["n_tris"] == n_world + sum(protos[p].shape[0] for p, _ in placements)
    rays = random_rays(rng, 20000, aim=6.0)
    # degenerate rays miss: NaN origin, zero direction, empty interval, infinite direction
    rays[0, 0] = np.nan
    rays[1, 4:7] = 0.0
    rays[2, 3], rays[2, 7] = 5.0, 1.0
    rays[3, 4] = np.inf
    ref = oracle_hits(ctx, world, protos, placements, rays)
    got = ctx.intersect_closest(rays)
    assert mismatches(got, ref) == 0
    assert np.all(got[:4, 1].view(np.uint32) == MISS)
    hit = ref[:, 1].view(np.uint32) != MISS
    assert hit.mean() > 0.2
    occ = ctx.occluded(rays)
    assert np.array_equal(occ.astype(bool), hit)
    got2, nodes, tris = ctx.intersect_closest(rays, counted=True)
    assert mismatches(got2, ref) == 0 and nodes > rays.shape[0] and tris > 0
    # every hit primitive maps back to its mesh or instance and local triangle
    firsts = np.cumsum([0] + [protos[p].shape[0] for p, _ in placements])[:-1] + n_world
    for prim in np.unique(got[hit, 1].view(np.uint32))[::7]:
        kind, index, local = ctx.prim_locate(int(prim))
        if prim < n_world:
            assert (kind, index, local) == (0, 0, prim)
        else:
            k = int(np.searchsorted(firsts, prim, side="right") - 1)
            assert (kind, index, local) == (1, k, prim - firsts[k])
    with pytest.raises(native.KpError):
        ctx.prim_locate(st["n_tris"])
    ctx.close()


@pytest.mark.parametrize("mode", [native.INSTANCES_AUTO, native.INSTANCES_BAKE])
def test_baked_instances_match_brute_force(mode):
    rng, world, protos, placements = mixed_scene(21)
    ctx = build(world, protos, placements, mode=mode)
    st = ctx.stats()
    assert st["two_level"] == 0
    soup = baked_soup(world, protos, placements)
    assert st["n_tris"] == soup.shape[0]
    rays = random_rays(rng, 20000, aim=6.0)
    ref = oracle.closest_bruteforce(soup, rays)
    assert mismatches(ctx.intersect_closest(rays), ref) == 0
    assert np.array_equal(ctx.occluded(rays).astype(bool), ref[:, 1].view(np.uint32) != MISS)
    ctx.close()


def test_device_pointer_path_and_coincident_instances():
    rng = np.random.default_rng(5)
    proto = random_tris(rng, 80, spread=1.0)
    placements = [(0, EYE), (0, EYE), (0, random_o2w(rng, spread=2.0))]
    ctx = build(None, [proto], placements)
    rays = random_rays(rng, 8000, box=3.0, finite_frac=0.0, aim=0.8)
    ref = oracle_hits(ctx, None, [proto], placements, rays)
    db = native.DeviceBuffer.from_array(ctx, rays)
    got = ctx.intersect_closest(db).numpy()
    assert mismatches(got, ref) == 0
    p = got[:, 1].view(np.uint32)
    # the first of two coincident copies wins every tie
    assert not np.any((p >= 80) & (p < 160))
    assert np.count_nonzero(p < 80) > 1000
    ctx.close()


def test_many_instances_deep_top_level():
    rng = np.random.default_rng(7)
    proto = scenes.uv_sphere(6, 4, radius=0.15)
    placements = []
    for k in range(3000):
        m = np.eye(3) * rng.uniform(0.5, 1.5)
        placements.append((0, np.hstack([m, rng.uniform(-8, 8, (3, 1))]).astype(np.float32)))
    ctx = build(None, [proto], placements)
    rays = random_rays(rng, 4000, box=9.0, aim=8.0)
    ref = oracle_hits(ctx, None, [proto], placements, rays)
    got = ctx.intersect_closest(rays)
    assert mismatches(got, ref) == 0
    assert np.array_equal(ctx.occluded(rays).astype(bool), ref[:, 1].view(np.uint32) != MISS)
    ctx.close()


def test_instancing_errors():
    ctx = native.Context(0)
    proto = scenes.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0))
    b = ctx.blas_add_soup(proto)
    with pytest.raises(native.KpError):
        ctx.instance_add(b + 1, EYE)
    sing = EYE.copy()
    sing[2, 2] = 0.0
    with pytest.raises(native.KpError):
        ctx.instance_add(b, sing)
    bad = EYE.copy()
    bad[0, 3] = np.nan
    with pytest.raises(native.KpError):
        ctx.instance_add(b, bad)
    ctx.instance_add(b, EYE)
    ctx.bvh_build()
    ctx.close()


def instanced_room(spp=32, size=48):
    """The Cornell box's five walls as meshes; its ceiling light as one
    instanced prototype placed twice and its objects as instanced spheres
    (glossy and diffuse, scaled and rotated)."""
    s = scenes.cornell_box(width=size, height=size, spp=spp)
    lamp, diffuse = s.mesh_material[5], s.mesh_material[6]
    s.meshes, s.mesh_material = s.meshes[:5], s.mesh_material[:5]
    glossy = len(s.materials)
    s.materials.append(scenes.material(scenes.CONDUCTOR, dist=scenes.GGX, F0=(0.9, 0.7, 0.5), alpha=0.2))
    light = s.add_prototype(scenes.quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)), lamp)   # faces -y
    ball = s.add_prototype(scenes.uv_sphere(24, 12), glossy)
    dull = s.add_prototype(scenes.uv_sphere(12, 6), diffuse)
    for dx in (-0.4, 0.4):
        s.add_instance(light, [[0.12, 0, 0, dx], [0, 1, 0, 0.999], [0, 0, 0.12, 0]])
    c, sn = np.cos(0.7), np.sin(0.7)
    s.add_instance(ball, [[0.35 * c, 0, 0.35 * sn, -0.4], [0, 0.35, 0, -0.65], [-0.35 * sn, 0, 0.35 * c, -0.2]])
    s.add_instance(ball, [[0.3, 0, 0, 0.45], [0, 0.3, 0, -0.7], [0, 0, 0.3, 0.2]])
    s.add_instance(dull, [[0.25, 0, 0, 0.0], [0, 0.2, 0, -0.8], [0, 0, 0.25, 0.45]])
    return s


def flattened(s):
    f = scenes.Scene(meshes=[], mesh_material=[], materials=s.materials, eye=s.eye, lookat=s.lookat, up=s.up,
                     fov_deg=s.fov_deg, width=s.width, height=s.height, spp=s.spp, max_depth=s.max_depth,
                     rr_depth=s.rr_depth, seed=s.seed)
    for m, mi in s.placed_meshes():
        f.add_mesh(m, mi)
    return f


def render(s, spp, mode=native.INSTANCES_AUTO):
    ctx = native.Context(0)
    ctx.load_scene(s, instances=mode)
    assert ctx.stats()["two_level"] == (1 if mode == native.INSTANCES_TWO_LEVEL and s.instances else 0)
    ctx.film_reset(s.width, s.height)
    ctx.render_accumulate(native.make_render_desc(s, spp=spp))
    img = ctx.film_read(s.width, s.height)
    ctx.close()
    return img


@pytest.mark.parametrize("mode", [native.INSTANCES_TWO_LEVEL, native.INSTANCES_AUTO])
def test_instanced_render_matches_flattened(mode):
    s = instanced_room()
    a = render(s, 64, mode)
    b = render(flattened(s), 64)
    assert np.isfinite(a).all() and a.mean() > 0.01
    # same sampling protocol and nearly identical hits: the images agree
    # pixel by pixel except where a path's hit differs by rounding
    close = np.abs(a - b) <= 1e-3 * np.maximum(np.abs(b), 1e-2)
    assert close.mean() > 0.9
    assert abs(a.mean() - b.mean()) < 0.01 * b.mean()


def test_instanced_render_matches_oracle_statistically():
    s = instanced_room(spp=16, size=24)
    img = render(s, 256, native.INSTANCES_TWO_LEVEL)
    f = flattened(s)
    sc = oracle.scene_from(f)
    d = oracle.RenderDesc()
    d.width, d.height, d.spp, d.max_depth, d.rr_depth, d.seed = f.width, f.height, 256, f.max_depth, f.rr_depth, 99
    d.eye[:], d.lookat[:], d.up[:] = f.eye, f.lookat, f.up
    d.tan_half_fov = f.tan_half_fov
    ref = sc.render(d)
    # independent seeds: compare means and a noise-level relMSE
    assert abs(img.mean() - ref.mean()) < 0.05 * ref.mean()
    rel = float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))
    assert rel < 0.1


def test_instanced_terrain_matches_flattened():
    """The terrain workload with its spheres as one prototype placed with a
    material per instance renders the same image as with the spheres baked."""
    kw = dict(grid=129, n_spheres=24, sphere_lon=16, sphere_lat=8, tiles=2, width=40, height=40, spp=16,
              max_depth=6)
    a = render(scenes.terrain_scene(instanced=True, **kw), 16, native.INSTANCES_TWO_LEVEL)
    b = render(scenes.terrain_scene(**kw), 16)
    assert np.isfinite(a).all() and a.mean() > 0.01
    close = np.abs(a - b) <= 1e-3 * np.maximum(np.abs(b), 1e-2)
    assert close.mean() > 0.9
    assert abs(a.mean() - b.mean()) < 0.02 * b.mean()
