"""Python bindings (ctypes) of the CPU oracle.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module; the product
(kazen_proto_amd) never does. The C++ sources next to this file restate the
reference algorithm and cite it file:line.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_p = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int


class Material(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("dist", ctypes.c_int32),
                ("albedo", ctypes.c_float * 3), ("F0", ctypes.c_float * 3),
                ("alpha", ctypes.c_float), ("eta", ctypes.c_float),
                ("Le", ctypes.c_float * 3), ("alpha_y", ctypes.c_float)]


class Lobe(ctypes.Structure):
    _fields_ = [("weight", ctypes.c_float * 3), ("pad", ctypes.c_float), ("lobe", Material)]


MAX_LOBES = 8


class RenderDesc(ctypes.Structure):
    _fields_ = [("width", _u32), ("height", _u32), ("spp", _u32), ("max_depth", _u32),
                ("rr_depth", _u32), ("seed", _u32), ("spp_offset", _u32), ("flags", _u32),
                ("eye", ctypes.c_float * 3), ("lookat", ctypes.c_float * 3),
                ("up", ctypes.c_float * 3), ("tan_half_fov", ctypes.c_float)]


class GuideCfg(ctypes.Structure):
    _fields_ = [("params", ctypes.c_void_p), ("bmin", ctypes.c_float * 3), ("binv", ctypes.c_float * 3),
                ("bsdf_fraction", ctypes.c_float), ("min_alpha", ctypes.c_float), ("sample", _i32),
                ("max_depth", ctypes.c_uint32)]


def make_guide(params, lo, hi, bsdf_fraction=0.5, min_alpha=0.1, sample=True, max_depth=0):
    """Guide settings for the guided oracle paths; the box inverse is formed
    in FP32 as the HIP library forms it (1 / (hi - lo), 1 where the extent is 0)."""
    params = np.ascontiguousarray(params, np.float32)
    lo = np.asarray(lo, np.float32)
    ext = np.asarray(hi, np.float32) - lo
    inv = np.where(ext > 0, np.float32(1.0) / np.where(ext > 0, ext, np.float32(1.0)), np.float32(1.0))
    g = GuideCfg()
    g.params = params.ctypes.data
    g.bmin[:] = [float(x) for x in lo]
    g.binv[:] = [float(x) for x in inv.astype(np.float32)]
    g.bsdf_fraction = bsdf_fraction
    g.min_alpha = min_alpha
    g.sample = 1 if sample else 0
    g.max_depth = max_depth
    g._keep = params   # the parameters must outlive the calls
    return g


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "or_pcg32_u32": (None, [_u64, _u64, _p, _i32]),
            "or_pcg32_f32": (None, [_u64, _u64, _p, _i32]),
            "or_intersect_triangle": (_i32, [_p, _p, _p, _p, _p, _p, _p]),
            "or_closest_bruteforce": (None, [_p, _u32, _p, _u32, _p, _i32]),
            "or_count_near_ties": (_u32, [_p, _u32, _p, _u32, _i32]),
            "or_bvh_build": (_p, [_p, _u32]),
            "or_bvh_free": (None, [_p]),
            "or_bvh_node_count": (_u32, [_p]),
            "or_bvh_closest": (None, [_p, _p, _u32, _p, _i32]),
            "or_bvh_occluded": (None, [_p, _p, _u32, _p, _i32]),
            "or_bvh_count_near_ties": (_u32, [_p, _p, _u32, _p, _i32]),
            "or_bvh_closest_timed": (ctypes.c_double, [_p, _p, _u32, _p, _i32]),
            "or_hardware_threads": (_i32, []),
            "or_bvh8_closest": (None, [_p, _u64, _p, _u64, _p, _u32, _p, _p, _p, _i32]),
            "or_bsdf_eval": (None, [_p, _p, _p, _p, _p]),
            "or_bsdf_sample": (_i32, [_p, _p, _p, _p, _p, _p]),
            "or_mf_D": (ctypes.c_double, [_i32, ctypes.c_float, ctypes.c_float, _p]),
            "or_mf_lambda": (ctypes.c_double, [_i32, ctypes.c_float, ctypes.c_float, _p]),
            "or_mf_sample_vndf": (None, [_i32, ctypes.c_float, ctypes.c_float, _p, _p, _p]),
            "or_fresnel_dielectric": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
            "or_shading_point": (None, [_p, _p, ctypes.c_float, ctypes.c_float, _p, _p]),
            "or_erfinv": (ctypes.c_double, [ctypes.c_double]),
            "or_bsdf_eval_batch": (None, [_p, _p, _p, _i32, _p, _p]),
            "or_bsdf_sample_batch": (None, [_p, _p, _p, _i32, _p, _p, _p, _p]),
            "or_mf_eval_batch": (None, [_i32, ctypes.c_float, ctypes.c_float, _p, _i32, _p, _p]),
            "or_composite_eval_batch": (None, [_p, _p, _u32, _p, _p, _p, _i32, _p, _p]),
            "or_composite_sample_batch": (None, [_p, _p, _u32, _p, _p, _p, _i32, _p, _p, _p, _p]),
            "or_scene_set_composite": (_i32, [_p, _u32, _p, _u32, _p]),
            "or_scene_set_normals": (_i32, [_p, _p, _p]),
            "or_mlp_param_count": (_i32, []),
            "or_mlp_encode": (None, [_p, _p]),
            "or_mlp_forward": (None, [_p, _i32, _p, _i32, _p]),
            "or_vmf_pdf": (ctypes.c_double, [_p, _p]),
            "or_vmf_sample": (None, [_p, _p, _p, _p]),
            "or_mlp_grad": (ctypes.c_double, [_p, _i32, _p, _p, _p, _i32, _p]),
            "or_adam_step": (None, [_p, _p, _p, _p, _i32, _i32, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float]),
            "or_scene_create": (_p, [_p, _u32, _p, _p, _p, _u32]),
            "or_scene_free": (None, [_p]),
            "or_render": (_i32, [_p, _p, _p, _p, _i32]),
            "or_render_pixels": (_i32, [_p, _p, _p, _u32, _p, _i32]),
            "or_integrate_rays": (_i32, [_p, _p, _u32, _p, _p, _i32]),
            "or_render_guided": (_i32, [_p, _p, _p, _p, _p, _i32]),
            "or_instanced_closest": (None, [_p, _p, _p, _p, _p, _p, _u32, _p, _u32, _p, _i32]),
            "or_integrate_rays_guided": (_i32, [_p, _p, _u32, _p, _p, _p, _i32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def pcg32_u32(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    lib().or_pcg32_u32(initstate, initseq, _ptr(out), n)
    return out


def pcg32_f32(initstate, initseq, n):
    out = np.zeros(n, np.float32)
    lib().or_pcg32_f32(initstate, initseq, _ptr(out), n)
    return out


def intersect_triangle(ray8, v0, v1, v2):
    ray8 = np.ascontiguousarray(ray8, np.float32)
    v = [np.ascontiguousarray(x, np.float32) for x in (v0, v1, v2)]
    t = ctypes.c_float()
    u = ctypes.c_float()
    w = ctypes.c_float()
    hit = lib().or_intersect_triangle(_ptr(ray8), _ptr(v[0]), _ptr(v[1]), _ptr(v[2]),
                                      ctypes.byref(t), ctypes.byref(u), ctypes.byref(w))
    return bool(hit), t.value, u.value, w.value


def _hits_view(h):
    """float32 [n,4] hit array -> (t, prim, u, v)."""
    return h[:, 0], h[:, 1].view(np.uint32), h[:, 2], h[:, 3]


def closest_bruteforce(tris, rays, threads=0):
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    hits = np.zeros((rays.shape[0], 4), np.float32)
    lib().or_closest_bruteforce(_ptr(tris), tris.shape[0], _ptr(rays), rays.shape[0], _ptr(hits),
                                threads)
    return hits


def count_near_ties(tris, rays, threads=0):
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    return int(lib().or_count_near_ties(_ptr(tris), tris.shape[0], _ptr(rays), rays.shape[0],
                                        threads))


class Bvh:
    """The oracle's own binary SAH BVH (independent of the product's BVH8)."""

    def __init__(self, tris):
        self._tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self._h = lib().or_bvh_build(_ptr(self._tris), self._tris.shape[0])

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_bvh_free(self._h)
            self._h = None

    @property
    def node_count(self):
        return int(lib().or_bvh_node_count(self._h))

    def closest(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), np.float32)
        lib().or_bvh_closest(self._h, _ptr(rays), rays.shape[0], _ptr(hits), threads)
        return hits

    def closest_timed(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), np.float32)
        sec = lib().or_bvh_closest_timed(self._h, _ptr(rays), rays.shape[0], _ptr(hits), threads)
        return hits, sec

    def count_near_ties(self, rays, hits, threads=0):
        """Rays whose closest hit has a second primitive within 1 ulp of t."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.ascontiguousarray(hits, np.float32).reshape(-1, 4)
        return int(lib().or_bvh_count_near_ties(self._h, _ptr(rays), rays.shape[0], _ptr(hits), threads))

    def occluded(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.uint8)
        lib().or_bvh_occluded(self._h, _ptr(rays), rays.shape[0], _ptr(out), threads)
        return out


def hardware_threads():
    return int(lib().or_hardware_threads())


def bvh8_closest(nodes, tris12, rays, threads=0):
    """Traverse the product's BVH8 records on the host (exact decode)."""
    nodes = np.ascontiguousarray(nodes, np.uint8)
    tris12 = np.ascontiguousarray(tris12, np.float32)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    hits = np.zeros((rays.shape[0], 4), np.float32)
    nv = ctypes.c_uint64()
    nt = ctypes.c_uint64()
    lib().or_bvh8_closest(_ptr(nodes), nodes.size // 128, _ptr(tris12), tris12.size // 12,
                          _ptr(rays), rays.shape[0], _ptr(hits), ctypes.byref(nv),
                          ctypes.byref(nt), threads)
    return hits, int(nv.value), int(nt.value)


def material(type_, dist=0, albedo=(0, 0, 0), F0=(0, 0, 0), alpha=0.5, eta=1.5, Le=(0, 0, 0), alpha_y=0.0):
    m = Material()
    m.alpha_y = alpha_y
    m.type = type_
    m.dist = dist
    m.albedo[:] = albedo
    m.F0[:] = F0
    m.alpha = alpha
    m.eta = eta
    m.Le[:] = Le
    return m


def bsdf_eval(mat, wo, wi):
    wo = np.ascontiguousarray(wo, np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    f = np.zeros(3, np.float32)
    pdf = ctypes.c_float()
    lib().or_bsdf_eval(ctypes.byref(mat), _ptr(wo), _ptr(wi), _ptr(f), ctypes.byref(pdf))
    return f, pdf.value


def bsdf_sample(mat, wo, u3):
    wo = np.ascontiguousarray(wo, np.float32)
    u3 = np.ascontiguousarray(u3, np.float32)
    wi = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    pdf = ctypes.c_float()
    ok = lib().or_bsdf_sample(ctypes.byref(mat), _ptr(wo), _ptr(u3), _ptr(wi), _ptr(f),
                              ctypes.byref(pdf))
    return bool(ok), wi, f, pdf.value


def bsdf_eval_batch(mat, wo, wi):
    wo = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, np.float32), np.shape(wi)), np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    n = wi.shape[0]
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().or_bsdf_eval_batch(ctypes.byref(mat), _ptr(wo), _ptr(wi), n, _ptr(f), _ptr(pdf))
    return f, pdf


def bsdf_sample_batch(mat, wo, u3):
    u3 = np.ascontiguousarray(u3, np.float32)
    n = u3.shape[0]
    wo = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, np.float32), (n, 3)), np.float32)
    wi = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    ok = np.zeros(n, np.int32)
    lib().or_bsdf_sample_batch(ctypes.byref(mat), _ptr(wo), _ptr(u3), n, _ptr(wi), _ptr(f),
                               _ptr(pdf), _ptr(ok))
    return ok.astype(bool), wi, f, pdf


def lobes(pairs):
    """ctypes array of Lobe from [(weight rgb, Material), ...]."""
    arr = (Lobe * len(pairs))()
    for i, (w, m) in enumerate(pairs):
        arr[i].weight[:] = [float(x) for x in w]
        arr[i].lobe = m
    return arr


def composite_eval_batch(lobe_arr, wo, wi, beta=None):
    wo = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, np.float32), np.shape(wi)), np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    n = wi.shape[0]
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    b = None if beta is None else np.ascontiguousarray(beta, np.float32)
    lib().or_composite_eval_batch(ctypes.cast(lobe_arr, ctypes.c_void_p), None, len(lobe_arr),
                                  _ptr(b) if b is not None else None, _ptr(wo), _ptr(wi), n, _ptr(f), _ptr(pdf))
    return f, pdf


def composite_sample_batch(lobe_arr, wo, u3, beta=None):
    u3 = np.ascontiguousarray(u3, np.float32)
    n = u3.shape[0]
    wo = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, np.float32), (n, 3)), np.float32)
    wi = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    ok = np.zeros(n, np.int32)
    b = None if beta is None else np.ascontiguousarray(beta, np.float32)
    lib().or_composite_sample_batch(ctypes.cast(lobe_arr, ctypes.c_void_p), None, len(lobe_arr),
                                    _ptr(b) if b is not None else None, _ptr(wo), _ptr(u3), n, _ptr(wi),
                                    _ptr(f), _ptr(pdf), _ptr(ok))
    return ok.astype(bool), wi, f, pdf


def mf_eval_batch(dist, alpha, w, alpha_y=None):
    w = np.ascontiguousarray(w, np.float32)
    n = w.shape[0]
    D = np.zeros(n, np.float64)
    lam = np.zeros(n, np.float64)
    lib().or_mf_eval_batch(dist, alpha, alpha if alpha_y is None else alpha_y, _ptr(w), n, _ptr(D), _ptr(lam))
    return D, lam


def mf_D(dist, alpha, m, alpha_y=None):
    return lib().or_mf_D(dist, alpha, alpha if alpha_y is None else alpha_y,
                         _ptr(np.ascontiguousarray(m, np.float32)))


def mf_lambda(dist, alpha, w, alpha_y=None):
    return lib().or_mf_lambda(dist, alpha, alpha if alpha_y is None else alpha_y,
                              _ptr(np.ascontiguousarray(w, np.float32)))


def mf_sample_vndf(dist, alpha, wo, u2, alpha_y=None):
    m = np.zeros(3, np.float32)
    lib().or_mf_sample_vndf(dist, alpha, alpha if alpha_y is None else alpha_y,
                            _ptr(np.ascontiguousarray(wo, np.float32)),
                            _ptr(np.ascontiguousarray(u2, np.float32)), _ptr(m))
    return m


def fresnel_dielectric(cos_i, eta):
    return lib().or_fresnel_dielectric(cos_i, eta)


def shading_point(tri9, normals9, u, v):
    """(P, shading normal) of the integrator's post-hit step."""
    t = np.ascontiguousarray(tri9, np.float32)
    n = np.ascontiguousarray(normals9, np.float32)
    P = np.zeros(3, np.float32)
    ns = np.zeros(3, np.float32)
    lib().or_shading_point(_ptr(t), _ptr(n), u, v, _ptr(P), _ptr(ns))
    return P, ns


def erfinv(x):
    return lib().or_erfinv(x)


# ---- path-guiding network ----------------------------------------------------
def mlp_param_count():
    return int(lib().or_mlp_param_count())


def mlp_encode(in6):
    in6 = np.ascontiguousarray(in6, np.float32)
    out = np.zeros(48, np.float64)
    lib().or_mlp_encode(_ptr(in6), _ptr(out))
    return out


def mlp_forward(params, in6, emulate_bf16=True):
    params = np.ascontiguousarray(params, np.float32)
    in6 = np.ascontiguousarray(in6, np.float32).reshape(-1, 6)
    out = np.zeros((in6.shape[0], 32), np.float64)
    lib().or_mlp_forward(_ptr(params), in6.shape[0], _ptr(in6), int(emulate_bf16), _ptr(out))
    return out


def mlp_grad(params, in6, dirs, weights, emulate_bf16=True):
    params = np.ascontiguousarray(params, np.float32)
    in6 = np.ascontiguousarray(in6, np.float32).reshape(-1, 6)
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    weights = np.ascontiguousarray(weights, np.float32)
    g = np.zeros(params.size, np.float64)
    loss = lib().or_mlp_grad(_ptr(params), in6.shape[0], _ptr(in6), _ptr(dirs), _ptr(weights),
                             int(emulate_bf16), _ptr(g))
    return loss, g


def vmf_pdf(out32, d):
    return lib().or_vmf_pdf(_ptr(np.ascontiguousarray(out32, np.float64)),
                            _ptr(np.ascontiguousarray(d, np.float32)))


def vmf_sample(out32, u3):
    d = np.zeros(3, np.float32)
    pdf = ctypes.c_double()
    lib().or_vmf_sample(_ptr(np.ascontiguousarray(out32, np.float64)),
                        _ptr(np.ascontiguousarray(u3, np.float32)), _ptr(d), ctypes.byref(pdf))
    return d, pdf.value


def adam_step(params, m, v, grad, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    for a in (params, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    grad = np.ascontiguousarray(grad, np.float32)
    lib().or_adam_step(_ptr(params), _ptr(m), _ptr(v), _ptr(grad), params.size, t, lr, beta1,
                       beta2, eps)


class Scene:
    """Oracle scene: triangle soup, per-triangle material and light ids."""

    def __init__(self, tris, tri_material, tri_light, materials):
        self._tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self._tm = np.ascontiguousarray(tri_material, np.int32)
        self._tl = np.ascontiguousarray(tri_light, np.int32)
        arr = (Material * len(materials))(*materials)
        self._mats = arr
        self._h = lib().or_scene_create(_ptr(self._tris), self._tris.shape[0], _ptr(self._tm),
                                        _ptr(self._tl), ctypes.cast(arr, ctypes.c_void_p),
                                        len(materials))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_scene_free(self._h)
            self._h = None

    def set_composite(self, material, pairs, Le=None):
        """Material `material` becomes a composite of [(weight rgb, Material)]."""
        arr = lobes(pairs)
        le = None if Le is None else np.ascontiguousarray(Le, np.float32)
        rc = lib().or_scene_set_composite(self._h, material, ctypes.cast(arr, ctypes.c_void_p), len(pairs),
                                          _ptr(le) if le is not None else None)
        if rc != 0:
            raise RuntimeError("or_scene_set_composite failed")

    def set_normals(self, normals9, has):
        """Per-triangle vertex normals [n, 9] (world space) and a per-triangle flag."""
        self._norms = np.ascontiguousarray(normals9, np.float32).reshape(-1, 9)
        self._has = np.ascontiguousarray(has, np.uint8)
        assert self._norms.shape[0] == self._tris.shape[0] == self._has.shape[0]
        if lib().or_scene_set_normals(self._h, _ptr(self._norms), _ptr(self._has)) != 0:
            raise RuntimeError("or_scene_set_normals failed")

    def render(self, desc, threads=0, want_sample0=False, guide=None):
        out = np.zeros((desc.height, desc.width, 3), np.float64)
        s0 = np.zeros((desc.height, desc.width, 3), np.float32) if want_sample0 else None
        rc = lib().or_render_guided(self._h, ctypes.byref(desc),
                                    ctypes.byref(guide) if guide is not None else None, _ptr(out),
                                    _ptr(s0) if s0 is not None else None, threads)
        if rc != 0:
            raise RuntimeError("or_render failed (%d)" % rc)
        return (out, s0) if want_sample0 else out

    def render_pixels(self, desc, pixels, threads=0):
        """Mean radiance [n, 3] (FP64) of the listed row-major pixel indices."""
        px = np.ascontiguousarray(pixels, np.uint32)
        out = np.zeros((px.size, 3), np.float64)
        rc = lib().or_render_pixels(self._h, ctypes.byref(desc), _ptr(px), px.size, _ptr(out), threads)
        if rc != 0:
            raise RuntimeError("or_render_pixels failed (%d)" % rc)
        return out

    def integrate_rays(self, rays, desc, threads=0, guide=None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((rays.shape[0], 3), np.float32)
        rc = lib().or_integrate_rays_guided(self._h, _ptr(rays), rays.shape[0], ctypes.byref(desc),
                                            ctypes.byref(guide) if guide is not None else None,
                                            _ptr(out), threads)
        if rc != 0:
            raise RuntimeError("or_integrate_rays failed (%d)" % rc)
        return out


def _material_from(d):
    return material(d["type"], d["dist"], d["albedo"], d["F0"], d["alpha"], d["eta"], d["Le"],
                    alpha_y=d.get("alpha_y", 0.0))


def scene_from(s):
    """Oracle scene of a scene description (meshes, instances, materials as
    dicts of type / dist / albedo / F0 / alpha / alpha_y / eta / Le, or
    composites {"lobes": [(weight, dict)], "Le": ...}, optional shading
    normals from s.shading_normals())."""
    single = [_material_from(m) if not m.get("lobes") else material(0, Le=m["Le"]) for m in s.materials]
    sc = Scene(s.soup(), s.tri_material(), s.tri_light(), single)
    for i, m in enumerate(s.materials):
        if m.get("lobes"):
            sc.set_composite(i, [(w, _material_from(d)) for w, d in m["lobes"]], m["Le"])
    sn = s.shading_normals() if hasattr(s, "shading_normals") else None
    if sn is not None:
        sc.set_normals(*sn)
    return sc


def instanced_closest(meshes, instances, rays, threads=0):
    """Two-level brute-force closest hit.  meshes: list of [n, 9] object-space
    triangle arrays; instances: list of (mesh index, world-to-object 3x4
    matrix); global prim of instance k = sum of the earlier instances' mesh
    sizes + local index.  Returns hits [n, 4] (t, prim bits, u, v)."""
    tris = np.ascontiguousarray(np.concatenate([np.asarray(m, np.float32).reshape(-1, 9) for m in meshes]))
    counts = np.array([np.asarray(m).reshape(-1, 9).shape[0] for m in meshes], np.uint32)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32)
    w2o = np.ascontiguousarray(np.array([np.asarray(x, np.float32).reshape(12) for _, x in instances], np.float32))
    imesh = np.array([m for m, _ in instances], np.uint32)
    base = np.concatenate([[0], np.cumsum(counts[imesh])[:-1]]).astype(np.uint32)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    hits = np.zeros((rays.shape[0], 4), np.float32)
    lib().or_instanced_closest(_ptr(tris), _ptr(first), _ptr(counts), _ptr(w2o), _ptr(imesh), _ptr(base),
                               len(instances), _ptr(rays), rays.shape[0], _ptr(hits), threads)
    return hits
