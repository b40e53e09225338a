"""The pyzen-shaped Python surface (kazen_proto_amd/pyzen_api.py): scene
building through the reference's method names, PFM film output, progressive
callbacks, and (GPU) that its renders equal the library's own."""
import math
import os

import numpy as np
import pytest

from kazen_proto_amd import native, pyzen_api, scenes

I4 = np.eye(4)


def cornell_via_api():
    sc = pyzen_api.Scene()
    sc.set_film(32, 24, 32, 24, "")
    sc.set_camera((0, 0, 3.4), (0, 0, 0), (0, 1, 0), 1.0, 1000.0, 45.0)
    for name, cs, shader in (("white", (0.75, 0.75, 0.75), "diffuse"), ("red", (0.75, 0.1, 0.1), "matte"),
                             ("lamp", (15, 15, 15), "emitter"), ("chrome", (0.9, 0.9, 0.9), "mirror")):
        sc.begin_shader_group(name)
        sc.load_oso_shader("surface", shader, "layer1", "")
        sc.set_shader_param_vec3f("Cs", cs)
        sc.end_shader_group()
    sc.add_quad(I4, (0, -1, 0), (0, 1, 0), (0, 0, 1), 1, 1, "floor", "white")
    sc.add_quad(I4, (0, 0, -1), (0, 0, 1), (0, 1, 0), 1, 1, "back", "white")
    sc.add_quad(I4, (-1, 0, 0), (1, 0, 0), (0, 1, 0), 1, 1, "left", "red")
    sc.add_quad(I4, (0, 0.999, 0), (0, -1, 0), (0, 0, 1), 0.2, 0.2, "light", "lamp", True)
    w = np.eye(4)
    w[:3, 3] = (0.3, -0.6, -0.2)
    sc.add_sphere(w, (0, 0, 0), 0.35, "ball", "chrome")
    sc.add_triangle(I4, (-0.9, -0.99, 0.5), (-0.5, -0.99, 0.5), (-0.7, -0.5, 0.5), "tri", "white")
    vs = [(0.5, -0.99, 0.6), (0.9, -0.99, 0.6), (0.7, -0.6, 0.6)]
    sc.add_mesh(I4, vs, [], [], [(0, 1, 2)], "mesh", "red")
    sc.build_bvh()
    return sc


def test_scene_building_and_conventions():
    sc = cornell_via_api()
    s = sc.description
    assert (s.width, s.height) == (32, 24)
    assert math.isclose(math.tan(math.radians(s.fov_deg) / 2), math.tan(math.radians(22.5)) * 24 / 32)
    assert len(s.materials) == 4 and not s.warnings
    light = s.meshes[3].reshape(-1, 3, 3)
    n = np.cross(light[:, 1] - light[:, 0], light[:, 2] - light[:, 0])
    assert np.all(n[:, 1] < 0)                     # the light faces its dir (down)
    ball = s.meshes[4].reshape(-1, 3)
    assert np.allclose(np.linalg.norm(ball - (0.3, -0.6, -0.2), axis=1), 0.35, atol=1e-5)
    with pytest.raises(NotImplementedError):
        sc.set_integrator(pyzen_api.IntegratorType.WhittedIntegrator)
    sc.set_integrator(pyzen_api.IntegratorType.PathIntegrator)
    sc.set_accelerator(pyzen_api.AcceleratorType.HIP)


def test_pfm_round_trip(tmp_path):
    img = np.random.default_rng(1).random((5, 7, 3)).astype(np.float32)
    p = str(tmp_path / "x.pfm")
    pyzen_api.write_pfm(p, img)
    assert np.array_equal(pyzen_api.read_pfm(p), img)


@pytest.mark.gpu
def test_renderer_matches_library_and_calls_back(tmp_path):
    sc = cornell_via_api()

    class Cb(pyzen_api.RenderCallback):
        def __init__(self):
            self.calls = []

        def on_tile_end(self, film, tile_id):
            self.calls.append((tile_id, film.samples))

    cb = Cb()
    out = str(tmp_path / "cbox.pfm")
    r = pyzen_api.Renderer(16, 0, cb, max_depth=5, rr_depth=3, seed=4, chunk_spp=4)
    img = r.render(sc, out)
    assert cb.calls == [(0, 4), (1, 8), (2, 12), (3, 16)]
    assert np.array_equal(pyzen_api.read_pfm(out), img)
    s = sc.description
    ctx = native.Context(0)
    ctx.load_scene(s)
    ref = ctx.render(native.make_render_desc(s, spp=16))
    ctx.close()
    # 4 chunks of 4 samples fold into the film in sample order: the same film
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    px = r.render(sc, 10, 12)
    assert np.allclose(px, img[12, 10])


def test_reference_binding_test_runs():
    # the body of the reference's tests/binding_tests/scene_test.py with the
    # import of this package's pyzen layout
    import kazen_proto_amd.pyzen as pyzen
    scene = pyzen.api.Scene()
    mat = pyzen.mat.Mat4f()
    verts = [pyzen.vec.Vec3f(0, 1, 2)]
    normals = [pyzen.vec.Vec3f(0, 1, 2)]
    uvs = [pyzen.vec.Vec2f(0, 1)]
    faces = [pyzen.vec.Vec3i(0, 1, 2)]
    assert scene.add_mesh(mat, verts, normals, uvs, faces, 'test', False) is None
    # the face names vertices that do not exist: building reports it
    with pytest.raises(ValueError):
        scene.build_bvh()


def test_vec_and_mat_types():
    import kazen_proto_amd.pyzen as pyzen
    V, M = pyzen.vec.Vec3f, pyzen.mat
    a, b = V(1, 2, 3), V(0.5)
    assert (a + b)[2] == 3.5 and (a * 2)[1] == 4 and (-a)[0] == -1 and a.dot(V(1, 0, 0)) == 1
    assert np.allclose((V(1, 0, 0).cross(V(0, 1, 0))).v, (0, 0, 1))
    assert math.isclose(V(3, 4, 0).length(), 5.0) and math.isclose(V(3, 4, 0).normalized().length(), 1.0, rel_tol=1e-6)
    t = M.translate3f(V(1, 2, 3))
    r = M.rotate3f(pyzen.vec.Vec4f(0, 1, 0, math.pi / 2))
    s = M.scale3f(V(2, 2, 2))
    p = (t * r * s) * V(1, 0, 0)              # scale, rotate (x -> -z), translate
    assert np.allclose(p.v, (1, 2, 1), atol=1e-6)
    assert np.allclose((t.inverse() * t).m, np.eye(4), atol=1e-6)
    assert pyzen.vec.create_vec3f([1, 2, 3]) == V(1, 2, 3)
    with pytest.raises(IndexError):
        a[3]


def test_add_mesh_normals_light_flag_and_named_build():
    sc = pyzen_api.Scene()
    sc.set_film(8, 8, 8, 8, "")
    for name, shader, cs in (("white", "diffuse", (0.5, 0.5, 0.5)), ("lamp", "emitter", (4, 4, 4))):
        sc.begin_shader_group(name)
        sc.load_oso_shader("surface", shader, "layer1", "")
        sc.set_shader_param_vec3f("Cs", cs)
        sc.end_shader_group()
    vs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]
    ns = [(0, 0, 1), (0, 0, 1), (0, 0.6, 0.8), (0, 0.6, 0.8)]
    w = np.eye(4)
    w[:3, :3] = np.diag([2.0, 1.0, 1.0])       # non-uniform scale: normals take the inverse transpose
    sc.add_mesh(w, vs, ns, [], [(0, 1, 2), (0, 2, 3)], "smooth", "white")
    sc.add_mesh(I4, vs, [], [], [(0, 1, 2)], "lamp1", "lamp", True)
    sc.add_mesh(I4, vs, [], [], [(0, 1, 2)], "lamp2", "lamp", False)
    s = sc.description
    assert s.mesh_normals[0] is not None and s.mesh_normals[1] is None
    assert np.allclose(np.linalg.norm(s.mesh_normals[0].reshape(-1, 3), axis=1), 1.0)
    assert any("lamp2" in x and "is_light=False" in x for x in s.warnings)
    sc.build_bvh(["smooth", "lamp1"])
    assert len(s.meshes) == 2 and s.mesh_normals[0] is not None
    with pytest.raises(ValueError):
        sc.build_bvh(["nope"])


@pytest.mark.gpu
def test_renderer_matches_oracle_path_by_path():
    # the pyzen scene rendered through pyzen_api.Renderer at one sample per
    # pixel equals the oracle's paths (not only the library's own render)
    import oracle
    sc = cornell_via_api()
    r = pyzen_api.Renderer(1, 0, None, max_depth=5, rr_depth=3, seed=9)
    img = r.render(sc, "").astype(np.float64)
    s = sc.description
    d = oracle.RenderDesc()
    d.width, d.height, d.spp, d.max_depth, d.rr_depth, d.seed = s.width, s.height, 1, 5, 3, 9
    d.eye[:], d.lookat[:], d.up[:] = s.eye, s.lookat, s.up
    d.tan_half_fov = s.tan_half_fov
    ref = oracle.scene_from(s).render(d)
    close = np.all(np.abs(img - ref) <= 2e-3 * np.abs(ref) + 1e-5, axis=2)
    print("pyzen cornell: %.4f of pixels agree path by path" % close.mean())
    assert close.mean() >= 0.99   # 768 pixels: at most 7 float-rounding divergences
