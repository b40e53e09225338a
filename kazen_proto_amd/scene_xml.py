"""Scene ingest: the reference's XML scene description and OBJ meshes.

Reads the format of kazen_proto's scene files (parser: reference
src/core/scene.cpp:319-712; attribute encoding "type value ...",
scene.cpp:120-175; shapes Sphere / Triangle / Quad / Mesh with their
attribute names, src/core/shape.cpp:132-144,292-312,464-479; OBJ loading,
shape.cpp:719-790) into a scenes.Scene for the HIP integrator.

OSL shader groups become the integrator's closures.  The reference compiles
each group's shader (`<scene dir>/<shader name>.osl`) with OSL and turns the
closure tree of `Ci` into lobes (process_closure, src/shading/bsdf.cpp:
184-328).  OSL is not in this build; read_osl() reads the shader's declared
parameters with their defaults and evaluates its last `Ci = ...;` statement
as a weighted sum of closures (numbers, parameters, color(...), products and
sums, parentheses):

  diffuse(N)                          -> Lambert lobe
  microfacet(dist, N, a, eta, refr),
  microfacet(dist, N, ax, ay, eta, refr)
                                      -> refr 0: KpGloss lobe; 1, 2: KpGlass
                                         (bsdf.cpp:227-254)
  reflection(N, eta), refraction(N, eta)
                                      -> smooth (alpha 1e-3) gloss / dielectric
  kp_mirror()                         -> smooth conductor, F0 = 1
  kp_dielectric(int_ior, ext_ior)     -> smooth dielectric, eta = int / ext
  emission(), kp_emitter(albedo)      -> emitted radiance (the weight); the
                                         emitter's albedo term returns f = albedo
                                         with pdf 1 and samples no direction
                                         (bsdfs.cpp:340-353), so it adds no lobe
  anything else                       -> dropped, listed in Scene.warnings

The parameters a <Shader> node sets override the declared defaults; names
the shader does not declare are ignored (with a warning).  When the shader
file is missing, the group is mapped by shader name: diffuse / matte /
lambert -> Lambert(Cs), mirror -> smooth conductor of reflectance Cs,
dielectric -> smooth dielectric (eta), area / emitter / emissive -> emitter
(Le = Cs), rough -> Beckmann gloss (alpha, eta 1.5) weighted by kd or Cs, and
anything else -> Lambert(Cs or 0.5) with a warning.

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


def load_obj(path, with_normals=False):
    """Triangles [n, 9] float32 of an OBJ file (v / vn / f records; polygons
    are fanned; 1-based and negative indices).  with_normals: also the
    per-triangle vertex normals [n, 9] from the faces' v//vn references
    (None unless every face vertex names a normal), as TriangleMesh loads
    them (src/core/shape.cpp:719-790)."""
    verts, norms, tris, tnorm = [], [], [], []
    all_n = True
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "v":
                verts.append((float(p[1]), float(p[2]), float(p[3])))
            elif p[0] == "vn":
                norms.append((float(p[1]), float(p[2]), float(p[3])))
            elif p[0] == "f":
                idx, nid = [], []
                for tok in p[1:]:
                    parts = tok.split("/")
                    i = int(parts[0])
                    idx.append(i - 1 if i > 0 else len(verts) + i)
                    k = int(parts[2]) if len(parts) >= 3 and parts[2] else 0
                    k = k - 1 if k > 0 else len(norms) + k if k < 0 else -1
                    nid.append(k if 0 <= k < len(norms) else None)
                for j in range(1, len(idx) - 1):
                    tris.append(verts[idx[0]] + verts[idx[j]] + verts[idx[j + 1]])
                    c = (nid[0], nid[j], nid[j + 1])
                    if None in c:
                        all_n = False
                        tnorm.append((0.0,) * 9)
                    else:
                        tnorm.append(norms[c[0]] + norms[c[1]] + norms[c[2]])
    t = np.asarray(tris, np.float32).reshape(-1, 9)
    if not with_normals:
        return t
    return t, (np.asarray(tnorm, np.float32).reshape(-1, 9) if all_n and tris else None)


# ------------------------------------------------------------ OSL shaders --
def _strip_comments(text):
    import re
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _tokens(text):
    import re
    return re.findall(r'"[^"]*"|[A-Za-z_][A-Za-z_0-9]*|\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|[-+*/(),=;]',
                      text)


def _split_top(s, sep=","):
    depth, cur, out = 0, "", []
    for ch in s:
        if ch in "([":
            depth += 1
        elif ch in ")]":
            depth -= 1
        if ch == sep and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return out


def read_osl(path):
    """(shader name, {param: default value}, Ci expression text or None) of an
    OSL shader file; values are floats, RGB tuples or strings."""
    import re
    text = _strip_comments(open(path).read())
    text = re.sub(r"\[\[.*?\]\]", " ", text, flags=re.S)
    m = re.search(r"\b(?:surface|shader)\s+([A-Za-z_][A-Za-z_0-9]*)\s*\(", text)
    if not m:
        raise ValueError("%s: no shader declaration" % path)
    name = m.group(1)
    depth, i = 1, m.end()
    while depth and i < len(text):
        depth += {"(": 1, ")": -1}.get(text[i], 0)
        i += 1
    params = {}
    for decl in _split_top(text[m.end():i - 1]):
        d = decl.split("=", 1)
        words = d[0].split()
        if len(words) < 2 or len(d) < 2:
            continue
        try:
            params[words[-1]] = _Expr(_tokens(d[1]), {}).value()
        except ValueError:
            pass
    body = text[text.index("{", i):]
    ci = None
    for stmt in body.split(";"):
        mm = re.match(r"\s*[{}]?\s*Ci\s*=(.*)$", stmt, flags=re.S)
        if mm:
            ci = mm.group(1).strip()
    return name, params, ci


class _Expr:
    """Evaluates an OSL expression of numbers, parameters, color(...),
    closure calls, *, +, - and parentheses into a float, an RGB tuple or a
    closure sum [(weight RGB, closure name, args)]."""

    def __init__(self, toks, env):
        self.t, self.i, self.env = toks, 0, env

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else None

    def take(self, want=None):
        tok = self.peek()
        if tok is None or (want is not None and tok != want):
            raise ValueError("OSL expression: expected %r, found %r" % (want, tok))
        self.i += 1
        return tok

    def value(self):
        v = self.expr()
        if self.peek() not in (None, ";"):
            raise ValueError("OSL expression: trailing %r" % self.peek())
        return v

    @staticmethod
    def _rgb(v):
        return v if isinstance(v, tuple) else (float(v),) * 3

    def expr(self):
        v = self.term()
        while self.peek() in ("+", "-"):
            op = self.take()
            w = self.term()
            if isinstance(v, list) or isinstance(w, list):
                if op == "-" or not (isinstance(v, list) and isinstance(w, list)):
                    raise ValueError("OSL expression: unsupported closure arithmetic")
                v = v + w
            else:
                a, b = self._rgb(v), self._rgb(w)
                r = tuple(x + y if op == "+" else x - y for x, y in zip(a, b))
                v = r if isinstance(v, tuple) or isinstance(w, tuple) else r[0]
        return v

    def term(self):
        v = self.factor()
        while self.peek() in ("*", "/"):
            op = self.take()
            w = self.factor()
            if isinstance(v, list) and isinstance(w, list):
                raise ValueError("OSL expression: closure times closure")
            if op == "/":
                if isinstance(w, list):
                    raise ValueError("OSL expression: division by a closure")
                w = tuple(1.0 / x for x in self._rgb(w)) if isinstance(w, tuple) else 1.0 / w
            if isinstance(v, list) or isinstance(w, list):
                cl, s = (v, w) if isinstance(v, list) else (w, v)
                s = self._rgb(s)
                v = [(tuple(a * b for a, b in zip(s, wt)), n, args) for wt, n, args in cl]
            else:
                a, b = self._rgb(v), self._rgb(w)
                r = tuple(x * y for x, y in zip(a, b))
                v = r if isinstance(v, tuple) or isinstance(w, tuple) else r[0]
        return v

    def factor(self):
        tok = self.take()
        if tok == "-":
            v = self.factor()
            if isinstance(v, list):
                raise ValueError("OSL expression: negated closure")
            return tuple(-x for x in v) if isinstance(v, tuple) else -v
        if tok == "(":
            v = self.expr()
            self.take(")")
            return v
        if tok[0] == '"':
            return tok[1:-1]
        if tok[0].isdigit() or tok[0] == ".":
            return float(tok)
        if self.peek() == "(":
            self.take("(")
            args = []
            while self.peek() != ")":
                args.append(self.expr())
                if self.peek() == ",":
                    self.take(",")
            self.take(")")
            if tok in ("color", "normal", "vector", "point"):
                if len(args) == 1:
                    return (float(args[0]),) * 3
                return tuple(float(a) for a in args)
            return [((1.0, 1.0, 1.0), tok, args)]
        if tok in self.env:
            return self.env[tok]
        if tok in ("N", "Ng", "I", "P"):
            return tok
        raise ValueError("OSL expression: unknown name %r" % tok)


def _lobes_of(closures, group, warnings):
    """Closure sum -> (lobes [(weight, material dict)], Le)."""
    lobes, Le = [], [0.0, 0.0, 0.0]
    for w, name, args in closures:
        num = [a for a in args if isinstance(a, float)]
        dist = scenes.GGX
        if args and isinstance(args[0], str) and args[0] not in ("N", "Ng"):
            dist = scenes.BECKMANN if args[0] in ("beckmann", "default") else scenes.GGX
        if name == "diffuse":
            lobes.append((w, scenes.material(scenes.LAMBERT, albedo=(1, 1, 1))))
        elif name == "microfacet" and len(num) in (3, 4):
            ax = num[0]
            ay = num[1] if len(num) == 4 else num[0]
            eta, refr = num[-2], int(num[-1])
            t = scenes.GLOSS if refr == 0 else scenes.DIELECTRIC
            lobes.append((w, scenes.material(t, dist, alpha=max(ax, 1e-4), alpha_y=max(ay, 1e-4), eta=eta)))
        elif name == "reflection" and num:
            lobes.append((w, scenes.material(scenes.GLOSS, scenes.GGX, alpha=1e-3, eta=num[0])))
        elif name == "refraction" and num:
            lobes.append((w, scenes.material(scenes.DIELECTRIC, scenes.GGX, alpha=1e-3, eta=num[0])))
        elif name == "kp_mirror":
            lobes.append((w, scenes.material(scenes.CONDUCTOR, scenes.GGX, F0=(1, 1, 1), alpha=1e-3)))
        elif name == "kp_dielectric" and len(num) == 2:
            lobes.append((w, scenes.material(scenes.DIELECTRIC, scenes.GGX, alpha=1e-3, eta=num[0] / num[1])))
        elif name in ("emission", "kp_emitter"):
            Le = [a + b for a, b in zip(Le, w)]
        else:
            warnings.append("%s: closure %r is not supported and was dropped" % (group, name))
    if len(lobes) > 8:
        warnings.append("%s: %d lobes, only the first 8 kept" % (group, len(lobes)))
        lobes = lobes[:8]
    return lobes, tuple(Le)


def _material_of_lobes(lobes, Le):
    if not lobes:
        return scenes.material(scenes.LAMBERT, albedo=(0, 0, 0), Le=Le)
    if len(lobes) == 1 and lobes[0][1]["type"] == scenes.LAMBERT:
        # a weighted Lambert lobe is Lambert with the weight as albedo
        m = dict(lobes[0][1])
        m["albedo"] = tuple(float(x) for x in lobes[0][0])
        m["Le"] = tuple(float(x) for x in Le)
        return m
    return scenes.composite(lobes, Le=Le)


def shader_closure(path, params, warnings, group):
    """Material of a shader group whose shader file is `path`: the declared
    defaults overridden by the group's parameters, Ci evaluated to lobes."""
    _, defaults, ci = read_osl(path)
    env = dict(defaults)
    for k, v in params.items():
        if k in defaults:
            env[k] = tuple(float(x) for x in v[:3]) if isinstance(v, tuple) else v
        else:
            warnings.append("%s: parameter %r is not declared by the shader (ignored)" % (group, k))
    if ci is None:
        warnings.append("%s: the shader sets no Ci; mapped to black" % group)
        return scenes.material(scenes.LAMBERT, albedo=(0, 0, 0))
    cl = _Expr(_tokens(ci), env).value()
    if not isinstance(cl, list):
        raise ValueError("%s: Ci is not a closure" % group)
    return _material_of_lobes(*_lobes_of(cl, group, warnings))


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


def _transform_normals(n9, a, order):
    """Normals under the same attribute chain: the inverse transpose of each
    step (scale divides, rotate rotates, translate leaves them)."""
    v = n9.reshape(-1, 3).astype(np.float64)
    for name in reversed(order):
        if name == "scale":
            v = v / np.asarray(a["scale"][:3], np.float64)
        elif name == "rotate":
            v = v @ _rotation(a["rotate"]).T
    v = v / np.linalg.norm(v, axis=1, keepdims=True)
    return v.reshape(-1, 9).astype(np.float32)


def _closure(shader, params, warnings, group):
    cs = params.get("Cs", params.get("kd", (0.5, 0.5, 0.5)))
    if isinstance(cs, float):
        cs = (cs, cs, cs)
    if shader in ("diffuse", "matte", "lambert"):
        return scenes.material(scenes.LAMBERT, albedo=cs)
    if shader == "mirror":
        return scenes.composite([(cs, scenes.material(scenes.CONDUCTOR, scenes.GGX, F0=(1, 1, 1), alpha=1e-3))])
    if shader == "dielectric":
        return scenes.material(scenes.DIELECTRIC, scenes.GGX, alpha=1e-3, eta=params.get("eta", 1.5))
    if shader in ("area", "emitter", "emissive"):
        return scenes.material(scenes.LAMBERT, albedo=(0, 0, 0), Le=cs)
    if shader == "rough":
        warnings.append("%s: no rough.osl beside the scene; mapped to a Beckmann gloss lobe" % group)
        return scenes.composite([(cs, scenes.material(scenes.GLOSS, scenes.BECKMANN, alpha=params.get("alpha", 0.28),
                                                      eta=params.get("eta", 1.5)))])
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
                sname = node.attrib.get("name", "")
                osl = os.path.join(base, sname + ".osl")
                if os.path.exists(osl):
                    try:
                        groups[current] = shader_closure(osl, params, s.warnings, current)
                    except ValueError as e:   # shader code outside the subset read here
                        s.warnings.append("%s: %s; mapped by shader name" % (current, e))
                        groups[current] = _closure(sname, params, s.warnings, current)
                else:
                    groups[current] = _closure(sname, params, s.warnings, current)
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
        normals = None
        if node.tag == "Mesh":
            tris, normals = load_obj(os.path.join(base, a["filename"]), with_normals=True)
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
        if normals is not None:
            normals = _transform_normals(normals, a, order)
        s.add_mesh(_transform(np.asarray(tris, np.float32), a, order), material_of(shader), normals=normals)
    return s


def orient_film(img, scene):
    """The image as the reference's camera lays it out (mirrored when the
    camera's scale has a negative x)."""
    return img[:, ::-1] if getattr(scene, "flip_x", False) else img
