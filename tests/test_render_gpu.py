"""GPU wavefront integrator against the oracle.

Three levels: (1) closed-form furnace radiance; (2) per-path agreement at one
sample per pixel (same sampling protocol and seeds: the GPU path radiance
must equal the oracle's up to float rounding for nearly every pixel);
(3) statistics: the GPU image's relMSE against an FP64 high-spp oracle
render must match the error level of an independent oracle render at the
same spp (relMSE = mean (I - R)^2 / (R^2 + 1e-2))."""
import os

import numpy as np
import pytest

import oracle
from kazen_proto_amd import native, scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPU_THREADS = min(16, os.cpu_count() or 1)


def rel_mse(img, ref):
    return float(np.mean((img.astype(np.float64) - ref) ** 2 / (ref ** 2 + 1e-2)))


def oracle_scene(s):
    return oracle.scene_from(s)


def odesc(s, spp=None, seed=None, spp_offset=0):
    d = oracle.RenderDesc()
    d.width, d.height = s.width, s.height
    d.spp = s.spp if spp is None else spp
    d.max_depth, d.rr_depth = s.max_depth, s.rr_depth
    d.seed = s.seed if seed is None else seed
    d.spp_offset = spp_offset
    d.eye[:], d.lookat[:], d.up[:] = s.eye, s.lookat, s.up
    d.tan_half_fov = s.tan_half_fov
    return d


def gdesc(s, spp=None, seed=None, spp_offset=0):
    d = native.make_render_desc(s, spp=spp, spp_offset=spp_offset)
    if seed is not None:
        d.seed = seed
    return d


def gpu_ctx(s):
    c = native.Context(0)
    c.load_scene(s)
    return c


@pytest.mark.parametrize("albedo,depth,rr", [(0.5, 4, 64), (0.7, 6, 2), (1.0, 3, 64)])
def test_furnace_closed_form(albedo, depth, rr):
    s = scenes.furnace_box(albedo=albedo, Le=1.0, max_depth=depth, rr_depth=rr, spp=64,
                           width=32, height=32)
    c = gpu_ctx(s)
    img = c.render(gdesc(s))
    expect = scenes.furnace_expected(albedo, 1.0, depth)
    pix = img[..., 0].astype(np.float64).ravel()
    err = pix.std() / np.sqrt(pix.size)
    print("furnace albedo %.2f depth %d: %.5f expected %.5f (+-%.5f)" % (albedo, depth, pix.mean(),
                                                                         expect, err))
    assert abs(pix.mean() - expect) < 5 * err + 1e-6
    c.close()


def _per_path(s):
    c = gpu_ctx(s)
    g = c.render(gdesc(s, spp=1)).astype(np.float64)
    o = oracle_scene(s).render(odesc(s, spp=1), threads=CPU_THREADS)
    c.close()
    close = np.all(np.abs(g - o) <= 2e-3 * np.abs(o) + 1e-5, axis=2)
    return float(close.mean()), g, o


def test_cornell_per_path_agreement():
    s = scenes.cornell_box(width=64, height=64)
    frac, g, o = _per_path(s)
    print("cornell: %.4f of pixels agree path by path" % frac)
    assert frac >= 0.999


def test_all_closures_per_path_agreement():
    s = scenes.materials_scene(width=48, height=48)
    frac, g, o = _per_path(s)
    print("materials: %.4f of pixels agree path by path" % frac)
    assert frac >= 0.999


def test_cornell_matches_fp64_oracle_image():
    ref = np.load(os.path.join(ROOT, "tests", "golden", "cornell64_ref.npz"))
    s = scenes.cornell_box(width=64, height=64)
    c = gpu_ctx(s)
    img = c.render(gdesc(s, spp=256, seed=3000))
    st = c.render_stats()
    c.close()
    e = rel_mse(img, ref["ref"])
    print("cornell 256 spp relMSE %.3e (independent oracle 256 spp: %.3e); %.1f Mpaths/s"
          % (e, float(ref["noise256"]), st["paths"] / st["seconds"] / 1e6))
    assert e < 1.5 * float(ref["noise256"])
    # image mean within 3 sigma of the oracle's
    assert abs(img.mean() - ref["ref"].mean()) < 0.01 * ref["ref"].mean()


def test_cornell_pixels_within_confidence_intervals_of_fp64_oracle():
    # Per-pixel check against the FP64 16,384-spp oracle image: 16 independent
    # 256-spp GPU renders give each pixel's mean (4,096 spp) and standard
    # error; the reference's own error is sigma / sqrt(16384).  With t = 15
    # degrees of freedom, |z| > 4 has probability 0.1 % per pixel channel.
    ref = np.load(os.path.join(ROOT, "tests", "golden", "cornell64_ref.npz"))["ref"]
    s = scenes.cornell_box(width=64, height=64)
    c = gpu_ctx(s)
    K, spp = 16, 256
    imgs = np.stack([c.render(gdesc(s, spp=spp, seed=5000 + k)).astype(np.float64) for k in range(K)])
    c.close()
    mean = imgs.mean(0)
    sd = imgs.std(0, ddof=1)   # of one 256-spp image
    se = sd * np.sqrt(1.0 / K + spp / 16384.0)
    flat = se == 0.0           # pixels with no variance (the light seen directly)
    assert np.allclose(mean[flat], ref[flat], rtol=1e-4, atol=1e-6)
    z = (mean[~flat] - ref[~flat]) / se[~flat]
    inside = np.mean(np.abs(z) <= 4.0)
    print("cornell 64x64: %d pixel channels, %.4f within 4 sigma, mean z^2 %.3f, mean z %.3f"
          % (z.size, inside, np.mean(z * z), np.mean(z)))
    assert inside >= 0.995
    assert 0.7 < np.mean(z * z) < 1.5     # t(15) variance 1.15
    assert abs(np.mean(z)) < 0.1          # no bias of the whole image


def test_all_closures_match_oracle_statistics():
    s = scenes.materials_scene(width=32, height=32)
    osc = oracle_scene(s)
    ref = osc.render(odesc(s, spp=4096, seed=900), threads=CPU_THREADS)
    indep = osc.render(odesc(s, spp=256, seed=901), threads=CPU_THREADS)
    c = gpu_ctx(s)
    img = c.render(gdesc(s, spp=256, seed=902))
    c.close()
    e, e0 = rel_mse(img, ref), rel_mse(indep, ref)
    print("materials 256 spp relMSE gpu %.3e oracle %.3e" % (e, e0))
    assert e < 1.6 * e0


def test_composite_anisotropic_gloss_and_smooth_normals_per_path():
    # composite closures, anisotropic microfacets, gloss and shading normals
    # with the shading-point offset, path by path against the oracle
    s = scenes.closures_scene(width=48, height=48)
    frac, g, o = _per_path(s)
    print("closures: %.4f of pixels agree path by path" % frac)
    assert frac >= 0.999
    # the scene's features change the image: flat normals / single lobes differ
    flat = scenes.closures_scene(width=48, height=48)
    flat.mesh_normals = [None] * len(flat.meshes)
    fo = oracle_scene(flat).render(odesc(flat, spp=1), threads=CPU_THREADS)
    assert np.mean(np.abs(fo - o) > 1e-3) > 0.02


def test_composite_anisotropic_gloss_and_smooth_normals_statistics():
    s = scenes.closures_scene(width=32, height=32)
    osc = oracle_scene(s)
    ref = osc.render(odesc(s, spp=4096, seed=910), threads=CPU_THREADS)
    indep = osc.render(odesc(s, spp=256, seed=911), threads=CPU_THREADS)
    c = gpu_ctx(s)
    img = c.render(gdesc(s, spp=256, seed=912))
    c.close()
    e, e0 = rel_mse(img, ref), rel_mse(indep, ref)
    print("closures 256 spp relMSE gpu %.3e oracle %.3e" % (e, e0))
    assert e < 1.6 * e0


def _small_terrain():
    return scenes.terrain_scene(grid=128, n_spheres=16, sphere_lon=16, sphere_lat=8, width=32, height=32,
                                spp=1, max_depth=8)


def test_terrain_per_path_agreement():
    frac, g, o = _per_path(_small_terrain())
    print("terrain: %.4f of pixels agree path by path" % frac)
    assert frac >= 0.999


def test_terrain_matches_oracle_statistics():
    s = _small_terrain()
    osc = oracle_scene(s)
    ref = osc.render(odesc(s, spp=4096, seed=700), threads=CPU_THREADS)
    indep = osc.render(odesc(s, spp=256, seed=701), threads=CPU_THREADS)
    c = gpu_ctx(s)
    img = c.render(gdesc(s, spp=256, seed=702))
    c.close()
    e, e0 = rel_mse(img, ref), rel_mse(indep, ref)
    print("terrain 256 spp relMSE gpu %.3e oracle %.3e" % (e, e0))
    assert e < 1.6 * e0


def test_ingested_xml_scene_per_path_agreement(tmp_path):
    # a scene in the reference's XML / OBJ format, through the loader
    from kazen_proto_amd import scene_xml
    from test_scene_xml import write_scene
    s = scene_xml.load_xml_scene(write_scene(str(tmp_path)), sphere_lon=16, sphere_lat=8)
    s.spp, s.max_depth, s.rr_depth, s.seed = 1, 5, 3, 2
    frac, g, o = _per_path(s)
    print("ingested xml scene: %.4f of pixels agree path by path" % frac)
    assert frac >= 0.999
    assert scene_xml.orient_film(g, s).shape == g.shape


def test_film_accumulate_resumes_bit_for_bit_samples():
    s = scenes.cornell_box(width=32, height=32, spp=8)
    c = gpu_ctx(s)
    full = c.render(gdesc(s, spp=8))
    c.film_reset(32, 32)
    c.render_accumulate(gdesc(s, spp=4, spp_offset=0))
    c.render_accumulate(gdesc(s, spp=4, spp_offset=4))
    two = c.film_read(32, 32)
    c.close()
    assert np.allclose(full, two, rtol=1e-5, atol=1e-7)


def test_render_state_errors():
    c = native.Context(0)
    s = scenes.cornell_box(width=8, height=8, spp=1)
    with pytest.raises(native.KpError):
        c.render(gdesc(s))          # no scene
    c.mesh_add_soup(s.soup())
    c.bvh_build()
    with pytest.raises(native.KpError):
        c.render(gdesc(s))          # no materials
    c.close()


def test_pass_size_and_buffer_release_do_not_change_the_film():
    """Paths are independent and every pixel folds its samples in sample order,
    so the film is bit-identical whatever the pass size (here 2^16 paths, i.e.
    several passes, against one pass of all 81,920 paths), also after the
    cached path state has been released and reallocated."""
    s = scenes.cornell_box(width=64, height=40, spp=32)
    c = gpu_ctx(s)
    one = c.render(gdesc(s))
    c.release_buffers()
    c.set_pass_paths(1 << 16)
    many = c.render(gdesc(s))
    assert c.render_stats()["passes"] > 1
    c.release_buffers()
    again = c.render(gdesc(s))
    with pytest.raises(native.KpError):
        c.set_pass_paths(100)
    c.close()
    assert np.array_equal(one.view(np.uint32), many.view(np.uint32))
    assert np.array_equal(one.view(np.uint32), again.view(np.uint32))
