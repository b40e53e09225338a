"""Python surface in the shape of the reference's `pyzen.api` module.

Mirrors the class and method names of the reference's pybind11 module
(src/binding/api.cpp:83-236; Scene methods src/core/scene.h:40-200) over the
HIP library's ctypes binding, so harness scripts written against pyzen keep
working: Scene (parse_from_file, set_film, set_camera, set_accelerator,
set_integrator, add_sphere / add_triangle / add_quad / add_mesh, build_bvh,
shader groups and parameters), RenderCallback (on_tile_end), Renderer (render
to a file, render a scene, debug render of one pixel), and the
AcceleratorType / IntegratorType enums.

Differences: only the path integrator exists (the others raise); OSL shader
groups are reduced to the integrator's closure table by shader name, as in
scene_xml; the film is written as PFM (the reference writes through OIIO);
one progress callback per chunk of samples, with the whole film as the tile;
a mesh emits (and is sampled as an area light) exactly when its shader group
emits, whatever is_light says (a mismatch is listed in the scene's
warnings).  The vector and matrix types of pyzen.vec / pyzen.mat are in
pyzen_types; `import kazen_proto_amd.pyzen as pyzen` gives all three.
"""
import enum
import math
import os

import numpy as np

from . import native, scene_xml, scenes


class AcceleratorType(enum.IntEnum):
    BVH = 0
    Embree = 1
    HIP = 2


class IntegratorType(enum.IntEnum):
    NormalIntegrator = 0
    AmbientOcclusionIntegrator = 1
    WhittedIntegrator = 2
    PathMatIntegrator = 3
    PathEmsIntegrator = 4
    PathIntegrator = 5


def write_pfm(path, img):
    """Colour PFM (little-endian, rows bottom to top)."""
    img = np.ascontiguousarray(img, np.float32)
    h, w = img.shape[:2]
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        f.write(np.ascontiguousarray(img[::-1]).astype("<f4").tobytes())


def read_pfm(path):
    with open(path, "rb") as f:
        kind = f.readline().strip()
        w, h = (int(x) for x in f.readline().split())
        scale = float(f.readline())
        ch = 3 if kind == b"PF" else 1
        data = np.frombuffer(f.read(), dtype="<f4" if scale < 0 else ">f4")
    return data.reshape(h, w, ch)[::-1].astype(np.float32)


class Film:
    """The film handed to RenderCallback.on_tile_end: the current mean image."""

    def __init__(self, width, height, filename):
        self.width, self.height, self.filename = width, height, filename
        self.image = np.zeros((height, width, 3), np.float32)
        self.samples = 0

    def write_tiles(self):
        if self.filename:
            write_pfm(self.filename, self.image)


def _matrix(world):
    if world is None:
        return None
    m = world.as_array() if hasattr(world, "as_array") else np.asarray(world, np.float64)
    return m.reshape(4, 4)


def _apply(world, tris):
    """Row-major 4x4 object-to-world matrix applied to [n, 9] vertices."""
    m = _matrix(world)
    if m is None:
        return np.asarray(tris, np.float32)
    v = np.asarray(tris, np.float64).reshape(-1, 3)
    v = v @ m[:3, :3].T + m[:3, 3]
    return v.reshape(-1, 9).astype(np.float32)


def _apply_normals(world, n9):
    """Normals [n, 9] under the inverse transpose of the matrix, unit length."""
    m = _matrix(world)
    v = np.asarray(n9, np.float64).reshape(-1, 3)
    if m is not None:
        v = v @ np.linalg.inv(m[:3, :3])
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).reshape(-1, 9).astype(np.float32)


def _rows(seq, n):
    """A list of pyzen vectors (or nested sequences / arrays) -> [k, n] array."""
    if hasattr(seq, "__len__") and len(seq) and hasattr(seq[0], "v"):
        return np.asarray([x.v for x in seq], np.float64).reshape(-1, n)
    return np.asarray(seq, np.float64).reshape(-1, n)


class Scene:
    """pyzen.api.Scene over the HIP path integrator."""

    def __init__(self):
        self._s = scenes.Scene(name="pyzen")
        self._s.flip_x = False
        self._s.warnings = []
        self._groups = {}
        self._group = None
        self._group_params = {}
        self._group_shader = None
        self._mat_of_group = {}
        self._film_path = ""
        self.accelerator = AcceleratorType.HIP
        self.integrator = IntegratorType.PathIntegrator
        self._built = False
        self._names = []      # per mesh of the description: its name
        self._lights = []     # per mesh: is_light as given

    # ---- files and settings
    def parse_from_file(self, file_path):
        self._s = scene_xml.load_xml_scene(str(file_path))
        self._names = [""] * len(self._s.meshes)
        self._lights = [any(x != 0.0 for x in self._s.materials[m]["Le"]) for m in self._s.mesh_material]
        self._built = False

    def set_film(self, w, h, xres=None, yres=None, f=""):
        self._s.width, self._s.height = int(w), int(h)
        self._film_path = f

    def set_camera(self, p, l, u, near_plane=1.0, far_plane=1000.0, fov=60.0):
        self._s.eye, self._s.lookat, self._s.up = tuple(map(float, p)), tuple(map(float, l)), tuple(map(float, u))
        # reference fov is horizontal (src/core/camera.h:68-96)
        t = math.tan(math.radians(fov) * 0.5) * self._s.height / self._s.width
        self._s.fov_deg = math.degrees(2.0 * math.atan(t))

    def set_accelerator(self, t):
        self.accelerator = AcceleratorType(t)

    def set_integrator(self, t):
        t = IntegratorType(t)
        if t != IntegratorType.PathIntegrator:
            raise NotImplementedError("only the path integrator runs on the HIP path (%s requested)" % t.name)
        self.integrator = t

    # ---- shader groups (reduced to the closure table)
    def set_shader_search_path(self, path):
        self._search = str(path)

    def begin_shader_group(self, name):
        self._group, self._group_shader, self._group_params = name, None, {}

    def load_oso_shader(self, type_, name, layer="", lib_path=""):
        if self._group_shader is None:
            self._group_shader = name
        return True

    def connect_shader(self, *args):
        pass

    def _param(self, name, value):
        self._group_params[name] = value

    set_shader_param_bool = set_shader_param_float = set_shader_param_int = _param

    def set_shader_param_vec3f(self, name, value):
        self._param(name, tuple(float(x) for x in value))

    set_shader_param_vec4f = set_shader_param_vec3f

    def set_shader_param_str(self, name, value):
        self._param(name, str(value))

    set_shader_param_ustr = set_shader_param_str

    def end_shader_group(self):
        if self._group is not None:
            name = self._group_shader or ""
            osl = os.path.join(getattr(self, "_search", ""), name + ".osl")
            m = None
            if getattr(self, "_search", None) and os.path.exists(osl):
                try:
                    m = scene_xml.shader_closure(osl, self._group_params, self._s.warnings, self._group)
                except ValueError as e:
                    self._s.warnings.append("%s: %s; mapped by shader name" % (self._group, e))
            if m is None:
                m = scene_xml._closure(name, self._group_params, self._s.warnings, self._group)
            self._groups[self._group] = m
        self._group = None

    def _material(self, shader_name):
        if shader_name not in self._mat_of_group:
            m = self._groups.get(shader_name)
            if m is None:
                self._s.warnings.append("unknown shader group %r mapped to Lambert 0.5" % shader_name)
                m = scenes.material(scenes.LAMBERT, albedo=(0.5, 0.5, 0.5))
            self._mat_of_group[shader_name] = len(self._s.materials)
            self._s.materials.append(m)
        return self._mat_of_group[shader_name]

    # ---- geometry
    def add_sphere(self, world, p, r, name="", shader_name="", is_light=False):
        t = scenes.uv_sphere(64, 32, center=tuple(p), radius=float(r))
        self._s.add_mesh(_apply(world, t), self._material(shader_name))
        self._names.append(str(name))
        self._lights.append(bool(is_light))
        self._built = False

    def add_triangle(self, world, a, b, c, name="", shader_name="", is_light=False):
        t = np.asarray([tuple(a) + tuple(b) + tuple(c)], np.float32)
        self._s.add_mesh(_apply(world, t), self._material(shader_name))
        self._names.append(str(name))
        self._lights.append(bool(is_light))
        self._built = False

    def add_quad(self, world, c, d, u, w, h, name="", shader_name="", is_light=False):
        c, d, u = (np.asarray(x, np.float64) for x in (c, d, u))
        hv = np.cross(u, d)
        hv = hv / np.linalg.norm(hv) * float(w)
        vv = u / np.linalg.norm(u) * float(h)
        p = [c - hv - vv, c + hv - vv, c + hv + vv, c - hv + vv]
        t = scenes.quad(*p)
        if np.dot(np.cross(p[1] - p[0], p[2] - p[0]), d) < 0:
            t = scenes.quad(p[0], p[3], p[2], p[1])
        self._s.add_mesh(_apply(world, t), self._material(shader_name))
        self._names.append(str(name))
        self._lights.append(bool(is_light))
        self._built = False

    def add_mesh(self, world, vs, ns, ts, idx, name="", shader_name="", is_light=False):
        """TriangleMesh of vertices vs, optional vertex normals ns (one per
        vertex: smooth shading), uvs ts (unused by the integrator) and faces
        idx (src/core/scene.h:144-160).  Faces that name missing vertices are
        an error when the BVH is built, as the reference's accelerator would
        read past the vertex array."""
        vs = _rows(vs, 3)
        idx = _rows(idx, 3).astype(np.int64)
        ns = _rows(ns, 3) if ns is not None and len(ns) else None
        shader = shader_name if isinstance(shader_name, str) else ""
        if idx.size and (idx.min() < 0 or idx.max() >= vs.shape[0]):
            self._pending_error = "add_mesh %r: a face names a vertex beyond the %d given" % (name, vs.shape[0])
            return None
        t = np.concatenate([vs[idx[:, 0]], vs[idx[:, 1]], vs[idx[:, 2]]], axis=1)
        normals = None
        if ns is not None and ns.shape[0] == vs.shape[0]:
            normals = _apply_normals(world, np.concatenate([ns[idx[:, 0]], ns[idx[:, 1]], ns[idx[:, 2]]], axis=1))
        mi = self._material(shader)
        emits = any(x != 0.0 for x in self._s.materials[mi]["Le"])
        if bool(is_light) != emits:
            self._s.warnings.append("mesh %r: is_light=%s but its shader group %s" % (
                name, bool(is_light), "emits" if emits else "emits nothing"))
        self._s.add_mesh(_apply(world, t), mi, normals=normals)
        self._names.append(str(name))
        self._lights.append(bool(is_light))
        self._built = False
        return None

    def build_bvh(self, names=None):
        """Accelerator::build(names) (src/core/accel.h:32): with names, only
        the meshes of those names stay in the scene."""
        err = getattr(self, "_pending_error", None)
        if err:
            raise ValueError(err)
        if names:
            keep = set(names)
            unknown = keep - set(self._names)
            if unknown:
                raise ValueError("build_bvh: unknown mesh names %s" % sorted(unknown))
            sel = [i for i, n in enumerate(self._names) if n in keep]
            s = self._s
            s.meshes = [s.meshes[i] for i in sel]
            s.mesh_material = [s.mesh_material[i] for i in sel]
            s.mesh_normals = [s.mesh_normals[i] if i < len(s.mesh_normals) else None for i in sel]
            self._names = [self._names[i] for i in sel]
            self._lights = [self._lights[i] for i in sel]
        self._built = True

    @property
    def description(self):
        """The scenes.Scene this object has built up."""
        return self._s


class RenderCallback:
    """Subclass and override on_tile_end(film, tile_id)."""

    def on_tile_end(self, film, tile_id):
        pass


class Renderer:
    """pyzen.api.Renderer: Renderer(), Renderer(nsamples, nthreads, callback)."""

    def __init__(self, nsamples=16, nthreads=0, callback=None, max_depth=8, rr_depth=3, seed=0, chunk_spp=None):
        self.nsamples = int(nsamples)
        self.callback = callback
        self.max_depth, self.rr_depth, self.seed = max_depth, rr_depth, seed
        self.chunk_spp = chunk_spp

    def _desc(self, s, spp, offset):
        s.max_depth, s.rr_depth, s.seed = self.max_depth, self.rr_depth, self.seed
        return native.make_render_desc(s, spp=spp, spp_offset=offset)

    def render(self, scene, output_or_x="", y=None):
        """render(file, output) / render(scene, output): the film (written to
        output as PFM if given); render(scene, x, y): the radiance of one
        pixel (the reference's debug render)."""
        if isinstance(scene, (str, os.PathLike)):
            sc = Scene()
            sc.parse_from_file(scene)
            scene = sc
        s = scene.description
        ctx = native.Context(0)
        try:
            ctx.load_scene(s)
            if y is not None:
                x = int(output_or_x)
                ctx.film_reset(s.width, s.height)
                ctx.render_accumulate(self._desc(s, self.nsamples, 0))
                img = scene_xml.orient_film(ctx.film_read(s.width, s.height), s)
                return img[int(y), x].copy()
            output = output_or_x or scene._film_path
            film = Film(s.width, s.height, output)
            ctx.film_reset(s.width, s.height)
            chunk = self.chunk_spp or self.nsamples
            done, tile = 0, 0
            while done < self.nsamples:
                n = min(chunk, self.nsamples - done)
                ctx.render_accumulate(self._desc(s, n, done))
                done += n
                film.image = scene_xml.orient_film(ctx.film_read(s.width, s.height), s)
                film.samples = done
                if self.callback is not None:
                    self.callback.on_tile_end(film, tile)
                tile += 1
            film.write_tiles()
            return film.image
        finally:
            ctx.close()
