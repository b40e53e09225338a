"""Pins the oracle integrator: furnace closed forms (NEE + BSDF sampling with
MIS must count every emission exactly once), Russian roulette unbiasedness,
and resumable sample indexing."""
import numpy as np
import pytest

import oracle
from kazen_proto_amd import scenes


def oracle_scene(s):
    return oracle.scene_from(s)


def desc(s, spp=None, spp_offset=0):
    d = oracle.RenderDesc()
    d.width, d.height = s.width, s.height
    d.spp = s.spp if spp is None else spp
    d.max_depth, d.rr_depth, d.seed = s.max_depth, s.rr_depth, s.seed
    d.spp_offset = spp_offset
    d.eye[:], d.lookat[:], d.up[:] = s.eye, s.lookat, s.up
    d.tan_half_fov = s.tan_half_fov
    return d


@pytest.mark.parametrize("albedo,depth,rr", [(0.5, 4, 64), (0.8, 3, 64), (0.7, 6, 2)])
def test_furnace_closed_form(albedo, depth, rr):
    s = scenes.furnace_box(albedo=albedo, Le=1.0, max_depth=depth, rr_depth=rr, spp=64)
    img = oracle_scene(s).render(desc(s))
    expect = scenes.furnace_expected(albedo, 1.0, depth)
    pix = img[..., 0].ravel()
    err = pix.std() / np.sqrt(pix.size)
    assert abs(pix.mean() - expect) < 5 * err + 1e-9, (pix.mean(), expect, err)
    assert np.allclose(img[..., 0], img[..., 1]) and np.allclose(img[..., 0], img[..., 2])


def test_white_furnace_albedo_one():
    # with roulette off and albedo 1 every vertex adds exactly Le: the
    # camera-visible emission is exact and each later vertex is an MIS
    # combination whose mean is Le
    s = scenes.furnace_box(albedo=1.0, Le=1.0, max_depth=3, rr_depth=64, spp=64)
    img = oracle_scene(s).render(desc(s))[..., 0].ravel()
    err = img.std() / np.sqrt(img.size)
    assert abs(img.mean() - 4.0) < 5 * err
    assert np.all(np.abs(img - 4.0) < 1.0)


def test_resume_matches_single_render():
    s = scenes.cornell_box(width=16, height=16, spp=8)
    sc = oracle_scene(s)
    full = sc.render(desc(s, spp=8))
    a = sc.render(desc(s, spp=4, spp_offset=0))
    b = sc.render(desc(s, spp=4, spp_offset=4))
    assert np.allclose(full, 0.5 * (a + b), rtol=1e-12, atol=1e-12)


def test_cornell_energy_is_finite_and_lit():
    s = scenes.cornell_box(width=24, height=24, spp=16)
    img = oracle_scene(s).render(desc(s))
    assert np.isfinite(img).all() and (img >= 0).all()
    assert img.mean() > 0.05
    # red wall on the left, green on the right
    left, right = img[:, 1:4].mean((0, 1)), img[:, -4:-1].mean((0, 1))
    assert left[0] > left[1] and right[1] > right[0]


def random_guide_params(seed, out_scale=3.0):
    """Glorot-uniform network parameters in the canonical order, output layer
    scaled by out_scale so the mixture is far from uniform."""
    rng = np.random.default_rng(seed)
    dims = [(48, 64), (64, 64), (64, 64), (64, 64), (64, 32)]
    parts = []
    for li, (din, dout) in enumerate(dims):
        lim = np.sqrt(6.0 / (din + dout)) * (out_scale if li == 4 else 1.0)
        parts.append(rng.uniform(-lim, lim, din * dout))
        parts.append(rng.uniform(-0.5, 0.5, dout) * (out_scale if li == 4 else 0.1))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == 17696
    return p


@pytest.mark.parametrize("albedo,depth,rr,gdepth", [(0.5, 4, 64, 0), (0.7, 5, 2, 0), (0.7, 5, 64, 2)])
def test_guided_furnace_closed_form(albedo, depth, rr, gdepth):
    # One-sample MIS between the BSDF and an arbitrary (here random, sharply
    # peaked) network density must stay unbiased, and next-event estimation
    # must weight against the same combined density: any slip in either
    # (a missing (1 - frac), the BSDF-only pdf in the power heuristic) moves
    # the mean away from the closed form.
    s = scenes.furnace_box(albedo=albedo, Le=1.0, max_depth=depth, rr_depth=rr, spp=128)
    lo, hi = s.soup().reshape(-1, 3).min(0), s.soup().reshape(-1, 3).max(0)
    # gdepth > 0: only the first gdepth vertices are guided, deeper ones sample
    # the BSDF alone; the mean must not move either
    g = oracle.make_guide(random_guide_params(7), lo, hi, bsdf_fraction=0.5, max_depth=gdepth)
    img = oracle_scene(s).render(desc(s), guide=g)
    expect = scenes.furnace_expected(albedo, 1.0, depth)
    pix = img[..., 0].ravel()
    err = pix.std() / np.sqrt(pix.size)
    assert abs(pix.mean() - expect) < 5 * err + 1e-9, (pix.mean(), expect, err)


def test_guided_network_changes_paths_not_the_mean():
    # The guide must actually be sampled (images differ per path from BSDF-only
    # sampling with the same 12-draw protocol) while the mean is unchanged.
    s = scenes.cornell_box(width=16, height=16, spp=64, sphere_lon=16, sphere_lat=8)
    lo, hi = s.soup().reshape(-1, 3).min(0), s.soup().reshape(-1, 3).max(0)
    p = random_guide_params(11)
    sc = oracle_scene(s)
    a = sc.render(desc(s), guide=oracle.make_guide(p, lo, hi, sample=False))
    b = sc.render(desc(s), guide=oracle.make_guide(p, lo, hi, sample=True))
    plain = sc.render(desc(s, spp=256))
    assert np.mean(np.abs(a - b)) > 1e-3

    def mean_err(img):
        # pixels are independent: the spread of pixel values bounds the
        # standard error of the image mean from above
        v = img.mean(2).ravel()
        return v.mean(), v.std() / np.sqrt(v.size)

    mp, ep = mean_err(plain)
    for img in (a, b):
        m, e = mean_err(img)
        assert abs(m - mp) < 4 * np.hypot(e, ep), (m, mp, e, ep)


@pytest.mark.parametrize("rr", [64, 2])
def test_composite_furnace_closed_form(rr):
    # a composite of two Lambert lobes (weights 0.3 and 0.4, albedos 1 and
    # 0.5) scatters like Lambert(0.5): the furnace closed form holds whatever
    # lobe-pick probabilities the throughput gives; per-channel weights make
    # the channels differ (albedo 0.5 / 0.5 / 0.35)
    s = scenes.furnace_box(albedo=0.5, Le=1.0, max_depth=4, rr_depth=rr, spp=64)
    s.materials[0] = scenes.composite(
        [((0.3, 0.3, 0.2), scenes.material(scenes.LAMBERT, albedo=(1, 1, 1))),
         ((0.4, 0.4, 0.3), scenes.material(scenes.LAMBERT, albedo=(0.5, 0.5, 0.5)))], Le=(1, 1, 1))
    img = oracle_scene(s).render(desc(s))
    for ch, alb in enumerate((0.5, 0.5, 0.35)):
        pix = img[..., ch].ravel()
        err = pix.std() / np.sqrt(pix.size)
        expect = scenes.furnace_expected(alb, 1.0, 4)
        assert abs(pix.mean() - expect) < 5 * err + 1e-9, (ch, pix.mean(), expect, err)


def test_geometric_shading_normals_change_nothing():
    # vertex normals equal to the geometric normal: the interpolated normal is
    # the geometric one and the shading point does not move
    s = scenes.furnace_box(albedo=0.6, Le=1.0, max_depth=3, spp=16)
    plain = oracle_scene(s).render(desc(s))
    t = s.meshes[0].reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    s.mesh_normals = [np.repeat(n[:, None, :], 3, axis=1).reshape(-1, 9).astype(np.float32)]
    smooth = oracle_scene(s).render(desc(s))
    assert np.allclose(plain, smooth, rtol=1e-5, atol=1e-6)


def test_shading_point_offset_closed_form():
    # triangle (0,0,0), (1,0,0), (0,1,0) with every vertex normal
    # n = (a, 0, c): at the centroid, (P - v_i).n = a/3, -2a/3, a/3, so only
    # v1's offset is projected and P' = P + (1/3)(2a/3) n = P + (2a/9) n
    tri = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)
    for a in (0.3, 0.6):
        n = np.array([a, 0.0, np.sqrt(1 - a * a)], np.float32)
        P, ns = oracle.shading_point(tri, np.tile(n, 3), 1 / 3, 1 / 3)
        want = np.array([1 / 3, 1 / 3, 0.0]) + (2 * a / 9) * n
        assert np.allclose(P, want, atol=1e-6), (P, want)
        assert np.allclose(ns, n, atol=1e-6)
    # one shared normal n: P' = P + sum_i b_i max(0, -(P - v_i).n) n, so with
    # n = (-0.5, 0, sqrt(.75)) at (u, v) = (0.2, 0.3) vertices 0 and 2 lift P
    # by 0.1 each, weighted 0.5 and 0.3
    n = np.array([-0.5, 0.0, np.sqrt(0.75)], np.float32)
    P, _ = oracle.shading_point(tri, np.tile(n, 3), 0.2, 0.3)
    assert np.allclose(P, np.array([0.2, 0.3, 0.0]) + 0.08 * n, atol=1e-6), P
    # a vertex under P lifts it by its own weight: at vertex 1 exactly, P' = v1
    P, ns = oracle.shading_point(tri, np.array([0, 0, 1, 0.6, 0, 0.8, 0, 0, 1], np.float32), 1.0, 0.0)
    assert np.allclose(P, [1, 0, 0], atol=1e-6)
    # the interpolated normal at the edge midpoint of v0 and v1
    _, ns = oracle.shading_point(tri, np.array([0, 0, 1, 0.6, 0, 0.8, 0, 0, 1], np.float32), 0.5, 0.0)
    m = np.array([0.3, 0, 0.9])
    assert np.allclose(ns, m / np.linalg.norm(m), atol=1e-6)
