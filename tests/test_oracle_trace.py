"""Oracle intersection: watertight test KATs, closed forms, and BVH == brute
force.  Pins the oracle that the GPU parity tests rely on."""
import numpy as np
import pytest

import oracle
from kazen_proto_amd import scenes


def ray(o, d, tmin=0.0, tmax=np.inf):
    return np.array([*o, tmin, *d, tmax], np.float32)


def test_embree_single_triangle_kat():
    # tests/libusage_tests/embree_tests.cpp:19-62: triangle (-1,0,1),(2,0,1),(-1,0,-2),
    # ray from (0,1,0) along -y must hit.  Closed form: t = 1, u = v = 1/3.
    hit, t, u, v = oracle.intersect_triangle(ray((0, 1, 0), (0, -1, 0), 0.0, np.finfo(np.float32).max),
                                             (-1, 0, 1), (2, 0, 1), (-1, 0, -2))
    assert hit
    assert t == 1.0
    assert abs(u - 1 / 3) < 1e-6 and abs(v - 1 / 3) < 1e-6


def test_reference_intersection_test_miss_case():
    # tests/intersection_test.cpp:38-47: triangle (0,0,0),(2,0,0),(0,2,0) moved to
    # z=-10; ray (2,1.5,2) along -z passes outside (x+y = 3.5 > 2).
    hit, *_ = oracle.intersect_triangle(ray((2, 1.5, 2), (0, 0, -1)), (0, 0, -10), (2, 0, -10), (0, 2, -10))
    assert not hit
    hit, t, u, v = oracle.intersect_triangle(ray((0.5, 0.5, 2), (0, 0, -1)), (0, 0, -10), (2, 0, -10), (0, 2, -10))
    assert hit and t == 12.0 and u == 0.25 and v == 0.25


def test_two_sided_and_interval():
    tri = ((-1, -1, 0), (1, -1, 0), (0, 1, 0))
    assert oracle.intersect_triangle(ray((0, 0, 1), (0, 0, -1)), *tri)[0]
    assert oracle.intersect_triangle(ray((0, 0, -1), (0, 0, 1)), *tri)[0]        # back face
    assert not oracle.intersect_triangle(ray((0, 0, 1), (0, 0, -1), 0, 0.5), *tri)[0]
    assert oracle.intersect_triangle(ray((0, 0, 1), (0, 0, -1), 0, 1.0), *tri)[0]  # t == tmax counts
    assert not oracle.intersect_triangle(ray((0, 0, 1), (0, 0, -1), 1.5, 9), *tri)[0]
    assert not oracle.intersect_triangle(ray((0, 0, 1), (0, 0, 1)), *tri)[0]       # behind
    assert not oracle.intersect_triangle(ray((0, 0, 1), (1, 0, 0)), *tri)[0]       # parallel
    assert not oracle.intersect_triangle(ray((0, 0, 1), (0, 0, -1)), (0, 0, 0), (1, 1, 0), (2, 2, 0))[0]


def test_barycentrics_against_fp64():
    rng = np.random.default_rng(5)
    for _ in range(300):
        v = rng.normal(size=(3, 3))
        b = rng.dirichlet((1, 1, 1))
        p = b @ v
        o = p + rng.normal(size=3) * 2
        d = p - o
        hit, t, u, w = oracle.intersect_triangle(ray(o, d), *v.astype(np.float32))
        if not hit:
            continue
        assert abs(t - 1.0) < 1e-4
        assert abs(u - b[1]) < 1e-4 and abs(w - b[2]) < 1e-4


def _closed_box_tris():
    # unit cube, 12 triangles sharing edges and vertices
    c = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, cc, d in faces:
        tris.append(np.concatenate([c[a], c[b], c[cc]]))
        tris.append(np.concatenate([c[a], c[cc], c[d]]))
    return np.array(tris, np.float32)


def test_watertight_no_leak_through_edges_and_vertices():
    # Rays from inside a closed mesh aimed exactly at its edges and vertices must
    # always hit (no gaps between adjacent triangles).
    tris = _closed_box_tris()
    targets = []
    t = np.linspace(-1, 1, 33)
    for a in t:
        for fixed in ((None, 1, 1), (None, -1, 1), (1, None, -1), (-1, None, 1), (1, 1, None), (-1, -1, None)):
            targets.append([a if f is None else f for f in fixed])
    targets += [[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)]
    rng = np.random.default_rng(2)
    rays = []
    for o in rng.uniform(-0.9, 0.9, size=(40, 3)):
        for p in targets:
            rays.append(ray(o, np.asarray(p) - o))
    rays = np.array(rays, np.float32)
    hits = oracle.closest_bruteforce(tris, rays)
    prim = hits[:, 1].view(np.uint32)
    assert (prim != 0xFFFFFFFF).all()
    assert np.all(np.abs(hits[:, 0] - 1.0) < 1e-5)


def test_tie_break_lowest_primitive():
    tri = np.array([[-1, -1, 0, 1, -1, 0, 0, 1, 0]], np.float32)
    tris = np.concatenate([tri, tri, tri])   # three coincident triangles
    h = oracle.closest_bruteforce(tris, ray((0, 0, 1), (0, 0, -1))[None])
    assert h[0, 1].view(np.uint32) == 0
    bvh = oracle.Bvh(tris)
    assert bvh.closest(ray((0, 0, 1), (0, 0, -1))[None])[0, 1].view(np.uint32) == 0
    assert oracle.count_near_ties(tris, ray((0, 0, 1), (0, 0, -1))[None]) == 1


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_bvh_matches_bruteforce(seed):
    tris = scenes.small_test_scene(seed=seed + 3)
    rng = np.random.default_rng(seed)
    rays = np.concatenate([scenes.incoherent_rays(20000, rng, box=1.5),
                           scenes.pinhole_rays((0, 0.3, 3), (0, 0, 0), (0, 1, 0), 50, 64, 64)])
    bf = oracle.closest_bruteforce(tris, rays)
    bvh = oracle.Bvh(tris)
    assert np.array_equal(bvh.closest(rays).view(np.uint32), bf.view(np.uint32))
    occ = bvh.occluded(rays)
    assert np.array_equal(occ.astype(bool), bf[:, 1].view(np.uint32) != 0xFFFFFFFF)


def test_bvh_empty_and_degenerate_rays():
    tris = scenes.small_test_scene()
    bvh = oracle.Bvh(tris)
    bad = np.array([ray((0, 0, 3), (0, 0, 0)),                 # zero direction
                    ray((np.nan, 0, 3), (0, 0, -1)),           # NaN origin
                    ray((0, 0, 3), (0, 0, -1), 5.0, 1.0)],     # tmin > tmax
                   np.float32)
    h = bvh.closest(bad)
    assert (h[:, 1].view(np.uint32) == 0xFFFFFFFF).all()
    assert np.array_equal(h.view(np.uint32), oracle.closest_bruteforce(tris, bad).view(np.uint32))


def test_bvh_near_tie_count_equals_bruteforce():
    # a grid of shared edges and a duplicated patch: rays through edges and
    # through the duplicates tie; the BVH count must equal the brute force
    rng = np.random.default_rng(12)
    g = np.linspace(-1, 1, 9)
    tris = []
    for i in range(8):
        for j in range(8):
            a, b, c, d = (g[i], g[j], 0), (g[i + 1], g[j], 0), (g[i + 1], g[j + 1], 0), (g[i], g[j + 1], 0)
            tris += [a + b + c, a + c + d]
    tris = np.asarray(tris, np.float32)
    tris = np.concatenate([tris, tris[:10], tris[20:30] + np.float32([0, 0, 1e-7] * 3)])
    o = np.zeros((4000, 3), np.float32)
    o[:, :2] = rng.uniform(-1, 1, (4000, 2))
    o[:2000, :2] = np.round(o[:2000, :2] * 4) / 4        # many through edges and vertices
    o[:, 2] = 2.0
    rays = np.zeros((4000, 8), np.float32)
    rays[:, :3] = o
    rays[:, 6] = -1.0
    rays[:, 7] = np.inf
    bvh = oracle.Bvh(tris)
    hits = bvh.closest(rays)
    n = bvh.count_near_ties(rays, hits)
    assert n == oracle.count_near_ties(tris, rays)
    assert n > 1000
