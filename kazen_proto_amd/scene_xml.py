"""Scene ingest: the reference's XML scene description and OBJ meshes.

Reads the format of kazen_proto's scene files (parser: reference
src/core/scene.cpp:319-712; attribute encoding "type value ...",
scene.cpp:120-175; shapes Sphere / Triangle / Quad / Mesh with their
attribute names, src/core/shape.cpp:132-144,292-312,464-479; OBJ loading,
shape.cpp:719-790) into a scenes.Scene for the HIP integrator.

OSL shader groups are mapped onto the integrator's closure table by shader
name (the reference evaluates them with OSL, which this build does not have):

  diffuse, matte, lambert      -> Lambert(Cs)
  mirror                       -> GGX conductor, alpha 1e-3, F0 = Cs
  dielectric                   -> GGX dielectric, alpha 1e-3, eta
  area, emitter, emissive      -> emitter, Le = Cs
  rough                        -> Beckmann conductor, alpha, F0 = 0.04 (kd or Cs)
                                  (the dielectric Fresnel of eta 1.5 at normal
                                  incidence; a diffuse layer, if the shader has
                                  one, is dropped)
  anything else                -> Lambert(Cs or 0.5), listed in Scene.warnings

The camera's fov is horizontal (reference src/core/camera.h:68-96: the same
cotangent on both axes, the film's y range scaled by 1/aspect); a negative x
in the camera's `scale` mirrors the image horizontally (flip_x)."""
import math
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import scenes


def _parse_value(text):
    """'float3 1 2 3' -> (type, value)."""
    parts = text.split()
    if not parts:
        raise ValueError("empty attribute")
    t, rest = parts[0], parts[1:]
    if t == "bool":
        return t, rest[0] == "true"
    if t == "int":
        return t, int(rest[0])
    if t in ("float", "double"):
        return t, float(rest[0])
    if t in ("float3", "float4", "func_translate", "func_scale", "func_rotate"):
        return t, tuple(float(x) for x in rest)
    if t in ("string", "ustring"):
        return t, " ".join(rest)
    raise ValueError("unknown attribute type %r" % t)


def _attrs(node):
    out = {}
    for k, v in node.attrib.items():
        try:
            out[k] = _parse_value(v)[1]
        except ValueError:
            out[k] = v
    return out


def load_obj(path):
    """Triangles [n, 9] float32 of an OBJ file (v / f records; polygons are
    fanned; 1-based and negative indices)."""
    verts, tris = [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                verts.append((float(p[1]), float(p[2]), float(p[3])))
            elif p[0] == "f":
                idx = []
                for tok in p[1:]:
                    i = int(tok.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(verts) + i)
                for j in range(1, len(idx) - 1):
                    tris.append(verts[idx[0]] + verts[idx[j]] + verts[idx[j + 1]])
    return np.asarray(tris, np.float32).reshape(-1, 9)


def _rotation(axis_angle):
    """Rotation matrix; the attribute holds (axis x, y, z, angle in radians)
    as Hitable::rotate / Transform::rotate read it (reference
    src/core/hitable.h:65-68, src/core/transform.h:40-45)."""
    ax = np.asarray(axis_angle[:3], np.float64)
    ang = float(axis_angle[3])
    n = np.linalg.norm(ax)
    if n == 0.0:
        return np.eye(3)
    x, y, z = ax / n
    c, s = math.cos(ang), math.sin(ang)
    C = 1.0 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def _transform(tris, a, order):
    """Applies the object's translate / rotate / scale attributes.  Each one
    right-multiplies the object-to-world matrix in attribute order
    (reference src/core/scene.cpp:235-262), so the last one acts first."""
    v = tris.reshape(-1, 3).astype(np.float64)
    for name in reversed(order):
        if name == "scale":
            v = v * np.asarray(a["scale"][:3], np.float64)
        elif name == "rotate":
            v = v @ _rotation(a["rotate"]).T
        elif name == "translate":
            v = v + np.asarray(a["translate"][:3], np.float64)
    return v.reshape(-1, 9).astype(np.float32)


def _closure(shader, params, warnings, group):
    cs = params.get("Cs", params.get("kd", (0.5, 0.5, 0.5)))
    if isinstance(cs, float):
        cs = (cs, cs, cs)
    if shader in ("diffuse", "matte", "lambert"):
        return scenes.material(scenes.LAMBERT, albedo=cs)
    if shader == "mirror":
        return scenes.material(scenes.CONDUCTOR, scenes.GGX, F0=cs, alpha=1e-3)
    if shader == "dielectric":
        return scenes.material(scenes.DIELECTRIC, scenes.GGX, alpha=1e-3, eta=params.get("eta", 1.5))
    if shader in ("area", "emitter", "emissive"):
        return scenes.material(scenes.LAMBERT, albedo=(0, 0, 0), Le=cs)
    if shader == "rough":
        f0 = tuple(0.04 * c for c in cs)
        if "Cs" in params:
            warnings.append("%s: the diffuse layer of 'rough' is not represented" % group)
        return scenes.material(scenes.CONDUCTOR, scenes.BECKMANN, F0=f0, alpha=params.get("alpha", 0.28))
    warnings.append("%s: shader %r mapped to Lambert" % (group, shader))
    return scenes.material(scenes.LAMBERT, albedo=cs)


def load_xml_scene(path, sphere_lon=64, sphere_lat=32):
    """Parses a reference scene file; returns a scenes.Scene with extra
    attributes flip_x (bool) and warnings (list of str)."""
    root = ET.parse(path).getroot()
    base = os.path.dirname(os.path.abspath(path))
    s = scenes.Scene(name=os.path.splitext(os.path.basename(path))[0])
    s.flip_x = False
    s.warnings = []
    groups = {}
    film = root.find("Film")
    if film is not None:
        a = _attrs(film)
        s.width, s.height = int(a.get("width", 800)), int(a.get("height", 600))
    cam = root.find("Camera")
    hfov = 60.0
    if cam is not None:
        a = _attrs(cam)
        s.eye = tuple(a.get("position", (0.0, 0.0, 0.0))[:3])
        s.lookat = tuple(a.get("lookat", (0.0, 0.0, -1.0))[:3])
        s.up = tuple(a.get("up", (0.0, 1.0, 0.0))[:3])
        hfov = float(a.get("fov", 60.0))
        sc = a.get("scale")
        s.flip_x = bool(sc is not None and sc[0] < 0)
    # vertical field of view with the same image extent (tan v = tan h / aspect)
    t = math.tan(math.radians(hfov) * 0.5) * s.height / s.width
    s.fov_deg = math.degrees(2.0 * math.atan(t))
    mats = root.find("Materials")
    if mats is not None:
        current = None
        for node in mats:
            if node.tag == "ShaderGroupBegin":
                current = node.attrib.get("name")
                groups[current] = None
            elif node.tag == "Shader" and current is not None and groups[current] is None:
                params = {}
                for p in node.findall("Parameter"):
                    for k, v in p.attrib.items():
                        try:
                            params[k] = _parse_value(v)[1]
                        except ValueError:
                            pass
                groups[current] = _closure(node.attrib.get("name", ""), params, s.warnings, current)
            elif node.tag == "ShaderGroupEnd":
                current = None
    mat_index = {}

    def material_of(name):
        if name not in mat_index:
            m = groups.get(name)
            if m is None:
                s.warnings.append("unknown shader group %r mapped to Lambert 0.5" % name)
                m = scenes.material(scenes.LAMBERT, albedo=(0.5, 0.5, 0.5))
            mat_index[name] = len(s.materials)
            s.materials.append(m)
        return mat_index[name]

    objs = root.find("Objects")
    for node in (objs if objs is not None else []):
        a = _attrs(node)
        shader = a.get("shader_name", "")
        if node.tag == "Mesh":
            tris = load_obj(os.path.join(base, a["filename"]))
        elif node.tag == "Triangle":
            tris = np.asarray([a["verta"][:3] + a["vertb"][:3] + a["vertc"][:3]], np.float32)
        elif node.tag == "Quad":
            c = np.asarray(a["center"][:3], np.float64)
            d = np.asarray(a["dir"][:3], np.float64)
            v = np.asarray(a.get("vertical_vec", (0.0, 1.0, 0.0))[:3], np.float64)
            h = np.asarray(a["horizontal_vec"][:3], np.float64) if "horizontal_vec" in a else np.cross(v, d)
            hw, hh = float(a.get("half_width", 1.0)), float(a.get("half_height", 1.0))
            h = h / np.linalg.norm(h) * hw
            v = v / np.linalg.norm(v) * hh
            p = [c - h - v, c + h - v, c + h + v, c - h + v]
            tris = scenes.quad(*p)
            # wind so the geometric normal faces dir (the emitting side)
            n = np.cross(p[1] - p[0], p[2] - p[0])
            if np.dot(n, d) < 0:
                tris = scenes.quad(p[0], p[3], p[2], p[1])
        elif node.tag == "Sphere":
            cr = a["center_n_radius"]
            tris = scenes.uv_sphere(sphere_lon, sphere_lat, center=cr[:3], radius=cr[3])
        else:
            continue
        order = [k for k in node.attrib if k in ("translate", "rotate", "scale")]
        s.add_mesh(_transform(np.asarray(tris, np.float32), a, order), material_of(shader))
    return s


def orient_film(img, scene):
    """The image as the reference's camera lays it out (mirrored when the
    camera's scale has a negative x)."""
    return img[:, ::-1] if getattr(scene, "flip_x", False) else img
