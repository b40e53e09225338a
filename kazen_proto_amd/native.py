"""ctypes binding of the C ABI in include/kp_hip.h (libkp_hip.so).

The library is loaded from this package directory; if it is missing the
import of any entry point raises (there is no CPU fallback).  Arrays may be
numpy (host memory, KP_PTR_HOST) or DeviceBuffer objects (device memory
allocated through the library, KP_PTR_DEVICE).  The binding deliberately does
not use torch: torch wheels bundle their own HIP runtime, and two HIP
runtimes in one process do not share devices.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkp_hip.so")

KP_OK = 0
KP_PTR_HOST = 0
KP_PTR_DEVICE = 1
DEVICE_HOST = -1   # host-only context: scene building and host queries, no GPU

_p = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int

_lib = None


class KpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("kp error %d: %s" % (code, msg))
        self.code = code


INSTANCES_AUTO, INSTANCES_TWO_LEVEL, INSTANCES_BAKE = 0, 1, 2
NODE_WIDTH_AUTO = 1    # kp_build_opts.node_width: tune 4 / 6 / 8 children per node on the device
BAKE_BUDGET = 1 << 24


class BuildOpts(ctypes.Structure):
    _fields_ = [("leaf_size", _u32), ("sah_bins", _u32), ("instances", _u32), ("node_width", _u32),
                ("reserved", _u32 * 4)]


class Stats(ctypes.Structure):
    _fields_ = [("n_tris", _u64), ("n_nodes", _u64), ("n_bvh2_nodes", _u64),
                ("bvh_bytes", _u64), ("max_depth", _u32), ("spill_entries", _u32),
                ("grid_closest", _u32), ("grid_occluded", _u32), ("build_seconds", ctypes.c_double),
                ("device_cus", _u32), ("two_level", _u32), ("node_width", _u32),
                ("width_ms", ctypes.c_float * 3), ("reserved", _u32 * 2)]


class Material(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("dist", ctypes.c_int32),
                ("albedo", ctypes.c_float * 3), ("F0", ctypes.c_float * 3),
                ("alpha", ctypes.c_float), ("eta", ctypes.c_float),
                ("Le", ctypes.c_float * 3), ("alpha_y", ctypes.c_float)]


class Lobe(ctypes.Structure):
    _fields_ = [("weight", ctypes.c_float * 3), ("pad", ctypes.c_float), ("lobe", Material)]


MAX_LOBES = 8
MAT_LAMBERT, MAT_CONDUCTOR, MAT_DIELECTRIC, MAT_GLOSS = 0, 1, 2, 3


class RenderDesc(ctypes.Structure):
    _fields_ = [("width", _u32), ("height", _u32), ("spp", _u32), ("max_depth", _u32),
                ("rr_depth", _u32), ("seed", _u32), ("spp_offset", _u32), ("flags", _u32),
                ("eye", ctypes.c_float * 3), ("lookat", ctypes.c_float * 3),
                ("up", ctypes.c_float * 3), ("tan_half_fov", ctypes.c_float)]


class RenderStats(ctypes.Structure):
    _fields_ = [("paths", _u64), ("extension_rays", _u64), ("shadow_rays", _u64),
                ("seconds", ctypes.c_double), ("passes", _u32), ("paths_per_pass", _u32),
                ("train_steps", _u32), ("guided_passes", _u32), ("train_records", _u32),
                ("reserved", _u32 * 3)]


RENDER_GUIDED = 1


class BatchTuning(ctypes.Structure):
    _fields_ = [("pass_paths", _u32), ("shade_blocks_per_cu", _u32), ("trace_blocks_per_cu", _u32),
                ("refill_below", _u32), ("use_graph", _u32), ("reserved", _u32 * 3)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class GuidingDesc(ctypes.Structure):
    _fields_ = [("bsdf_fraction", ctypes.c_float), ("min_alpha", ctypes.c_float),
                ("lr", ctypes.c_float), ("train_steps", _u32), ("guide_after", _u32),
                ("max_records", _u32), ("pass_paths", _u32), ("train_passes", _u32),
                ("bmin", ctypes.c_float * 3), ("bmax", ctypes.c_float * 3), ("weight_clamp", ctypes.c_float),
                ("relative_weights", _u32), ("max_guided_depth", _u32), ("late_train_steps", _u32)]


def make_material(d):
    """Material structure from a scenes.material() dict."""
    m = Material()
    m.type = d["type"]
    m.dist = d["dist"]
    m.albedo[:] = d["albedo"]
    m.F0[:] = d["F0"]
    m.alpha = d["alpha"]
    m.eta = d["eta"]
    m.Le[:] = d["Le"]
    m.alpha_y = d.get("alpha_y", 0.0)
    return m


def make_render_desc(scene, spp=None, spp_offset=0, width=None, height=None, guided=False):
    d = RenderDesc()
    d.width = width or scene.width
    d.height = height or scene.height
    d.spp = scene.spp if spp is None else spp
    d.max_depth = scene.max_depth
    d.rr_depth = scene.rr_depth
    d.seed = scene.seed
    d.spp_offset = spp_offset
    d.flags = RENDER_GUIDED if guided else 0
    d.eye[:] = scene.eye
    d.lookat[:] = scene.lookat
    d.up[:] = scene.up
    d.tan_half_fov = scene.tan_half_fov
    return d


# name -> (restype, argtypes); mirrors include/kp_hip.h
SIGNATURES = {
    "kp_create": (_i32, [ctypes.POINTER(_p), _i32]),
    "kp_destroy": (None, [_p]),
    "kp_last_error": (ctypes.c_char_p, [_p]),
    "kp_set_stream": (_i32, [_p, _p]),
    "kp_synchronize": (_i32, [_p]),
    "kp_mesh_add": (_i32, [_p, _p, _u32, _p, _u32, _p, ctypes.POINTER(_u32)]),
    "kp_mesh_set_normals": (_i32, [_p, _u32, _p]),
    "kp_blas_set_normals": (_i32, [_p, _u32, _p]),
    "kp_material_add_composite": (_i32, [_p, _p, _u32, _p, ctypes.POINTER(_u32)]),
    "kp_bvh_build": (_i32, [_p, _p]),
    "kp_blas_add": (_i32, [_p, _p, _u32, _p, _u32, ctypes.POINTER(_u32)]),
    "kp_blas_set_material": (_i32, [_p, _u32, _u32]),
    "kp_instance_add": (_i32, [_p, _u32, _p, ctypes.POINTER(_u32)]),
    "kp_instance_get": (_i32, [_p, _u32, ctypes.POINTER(_u32), _p, _p]),
    "kp_geometry_clear": (_i32, [_p]),
    "kp_instance_set_material": (_i32, [_p, _u32, _u32]),
    "kp_prim_locate": (_i32, [_p, _u32, ctypes.POINTER(_u32), ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "kp_get_stats": (_i32, [_p, _p]),
    "kp_intersect_closest": (_i32, [_p, _p, _u32, _p, _i32]),
    "kp_occluded": (_i32, [_p, _p, _u32, _p, _i32]),
    "kp_intersect_closest_counted": (_i32, [_p, _p, _u32, _p, _i32, ctypes.POINTER(_u64),
                                            ctypes.POINTER(_u64)]),
    "kp_bvh8_build_host": (_i32, [_p, _u32, _p, ctypes.POINTER(_p)]),
    "kp_bvh8_blob_info": (_i32, [_p, ctypes.POINTER(_u64), ctypes.POINTER(_u64),
                                 ctypes.POINTER(_u32)]),
    "kp_bvh8_blob_nodes": (_p, [_p]),
    "kp_bvh8_blob_tris": (_p, [_p]),
    "kp_bvh8_blob_free": (None, [_p]),
    "kp_bvh8_blob_closest": (_i32, [_p, _p, _u32, _p]),
    "kp_bvh8_blob_occluded": (_i32, [_p, _p, _u32, _p]),
    "kp_intersect_closest_host": (_i32, [_p, _p, _u32, _p]),
    "kp_intersect_closest_noalloc": (_i32, [_p, _p, _u32, _p]),
    "kp_occluded_host": (_i32, [_p, _p, _u32, _p]),
    "kp_set_trace_tuning": (_i32, [_p, _u32]),
    "kp_set_pass_paths": (_i32, [_p, _u32]),
    "kp_release_buffers": (_i32, [_p]),
    "kp_material_add": (_i32, [_p, _p, ctypes.POINTER(_u32)]),
    "kp_mesh_set_material": (_i32, [_p, _u32, _u32]),
    "kp_film_reset": (_i32, [_p, _u32, _u32]),
    "kp_render_accumulate": (_i32, [_p, _p]),
    "kp_film_read": (_i32, [_p, _p, _i32]),
    "kp_render": (_i32, [_p, _p, _p, _i32]),
    "kp_render_get_stats": (_i32, [_p, _p]),
    "kp_integrate_rays": (_i32, [_p, _p, _u32, _p, _p, _i32]),
    "kp_guiding_init": (_i32, [_p, _u64]),
    "kp_guiding_set_params": (_i32, [_p, _p, _u32]),
    "kp_guiding_get_params": (_i32, [_p, _p, _u32]),
    "kp_guiding_eval": (_i32, [_p, _p, _u32, _p]),
    "kp_guiding_pdf": (_i32, [_p, _p, _p, _u32, _p]),
    "kp_guiding_sample": (_i32, [_p, _p, _p, _u32, _p, _p]),
    "kp_guiding_train_step": (_i32, [_p, _p, _p, _p, _u32, ctypes.c_float, _i32, _p]),
    "kp_guiding_configure": (_i32, [_p, _p]),
    "kp_guiding_save": (_i32, [_p, _p, _u64, ctypes.POINTER(_u64)]),
    "kp_guiding_load": (_i32, [_p, _p, _u64]),
    "kp_batch_create": (_i32, [_p, _p, _u32, ctypes.POINTER(_p)]),
    "kp_batch_destroy": (_i32, [_p]),
    "kp_batch_set_tuning": (_i32, [_p, _p]),
    "kp_batch_get_tuning": (_i32, [_p, _p]),
    "kp_batch_render": (_i32, [_p, _p]),
    "kp_batch_autotune": (_i32, [_p, _p, _u32, _p]),
    "kp_batch_film_read": (_i32, [_p, _u32, _p, _i32]),
    "kp_batch_get_stats": (_i32, [_p, _p]),
    "kp_device_alloc": (_i32, [_p, _u64, ctypes.POINTER(_p)]),
    "kp_device_free": (_i32, [_p, _p]),
    "kp_copy_to_device": (_i32, [_p, _p, _p, _u64]),
    "kp_copy_to_host": (_i32, [_p, _p, _p, _u64]),
    "kp_device_copy": (_i32, [_p, _p, _p, _u64]),
    "kp_timer_begin": (_i32, [_p]),
    "kp_timer_end": (_i32, [_p, ctypes.POINTER(ctypes.c_float)]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libkp_hip.so is not built: run __graft_entry__.build() "
                               "(the HIP path has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceBuffer:
    """Device memory owned by the caller, allocated through a Context."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._p = _p()
        ctx._check(lib().kp_device_alloc(ctx.handle, self.nbytes, ctypes.byref(self._p)))

    @classmethod
    def from_array(cls, ctx, a):
        a = np.ascontiguousarray(a)
        b = cls(ctx, a.shape, a.dtype)
        b.upload(a)
        return b

    @property
    def ptr(self):
        return self._p

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        self.ctx._check(lib().kp_copy_to_device(self.ctx.handle, self._p, _np_ptr(a), self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(lib().kp_copy_to_host(self.ctx.handle, _np_ptr(out), self._p, self.nbytes))
        return out

    def free(self):
        if self._p and self.ctx.handle:
            self.ctx._check(lib().kp_device_free(self.ctx.handle, self._p))
        self._p = _p()

    def __del__(self):
        try:
            if self._p and self.ctx.handle:
                lib().kp_device_free(self.ctx.handle, self._p)
        except Exception:
            pass


def bvh8_build_host(tris9, leaf_size=4, sah_bins=32, node_width=0):
    """Host-only BVH8 build -> (nodes uint8[n*128], tris float32[n*12], max_depth);
    node_width 2..8 children per node (0: 8)."""
    tris9 = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
    o = BuildOpts()
    o.leaf_size = leaf_size
    o.sah_bins = sah_bins
    o.node_width = node_width
    blob = _p()
    rc = lib().kp_bvh8_build_host(_np_ptr(tris9), tris9.shape[0], ctypes.byref(o), ctypes.byref(blob))
    if rc != KP_OK:
        raise KpError(rc, lib().kp_last_error(None).decode())
    try:
        nn, nt, md = _u64(), _u64(), _u32()
        lib().kp_bvh8_blob_info(blob, ctypes.byref(nn), ctypes.byref(nt), ctypes.byref(md))
        nodes = np.ctypeslib.as_array(ctypes.cast(lib().kp_bvh8_blob_nodes(blob),
                                                  ctypes.POINTER(ctypes.c_uint8)),
                                      shape=(nn.value * 128,)).copy()
        tris = np.ctypeslib.as_array(ctypes.cast(lib().kp_bvh8_blob_tris(blob),
                                                 ctypes.POINTER(ctypes.c_float)),
                                     shape=(nt.value * 12,)).copy()
        return nodes, tris, int(md.value)
    finally:
        lib().kp_bvh8_blob_free(blob)


class HostBvh8:
    """Host-only BVH8 (kp_bvh8_build_host) with host closest-hit / occlusion
    queries (no device, thread-safe)."""

    def __init__(self, tris9, leaf_size=4, sah_bins=32, node_width=0):
        tris9 = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
        o = BuildOpts()
        o.leaf_size = leaf_size
        o.sah_bins = sah_bins
        o.node_width = node_width
        self._h = _p()
        rc = lib().kp_bvh8_build_host(_np_ptr(tris9), tris9.shape[0], ctypes.byref(o), ctypes.byref(self._h))
        if rc != KP_OK:
            raise KpError(rc, lib().kp_last_error(None).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().kp_bvh8_blob_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def closest(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((rays.shape[0], 4), np.float32)
        rc = lib().kp_bvh8_blob_closest(self._h, _np_ptr(rays), rays.shape[0], _np_ptr(out))
        if rc != KP_OK:
            raise KpError(rc, lib().kp_last_error(None).decode())
        return out

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.uint32)
        rc = lib().kp_bvh8_blob_occluded(self._h, _np_ptr(rays), rays.shape[0], _np_ptr(out))
        if rc != KP_OK:
            raise KpError(rc, lib().kp_last_error(None).decode())
        return out


class Context:
    """One HIP device context: scene, BVH and queries (one stream).
    device=DEVICE_HOST makes a host-only context (no GPU needed)."""

    def __init__(self, device=0):
        self._h = _p()
        rc = lib().kp_create(ctypes.byref(self._h), device)
        if rc != KP_OK:
            raise KpError(rc, lib().kp_last_error(None).decode())
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            lib().kp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != KP_OK:
            raise KpError(rc, lib().kp_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr):
        self._check(lib().kp_set_stream(self._h, _p(stream_ptr) if stream_ptr else None))

    def synchronize(self):
        self._check(lib().kp_synchronize(self._h))

    def mesh_add(self, verts, idx, xform=None):
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
        xf = None
        if xform is not None:
            xf = np.ascontiguousarray(xform, np.float32).reshape(12)
        mid = _u32()
        self._check(lib().kp_mesh_add(self._h, _np_ptr(verts), verts.shape[0], _np_ptr(idx),
                                      idx.shape[0], _np_ptr(xf) if xf is not None else None,
                                      ctypes.byref(mid)))
        return mid.value

    def mesh_add_soup(self, tris9):
        tris9 = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
        n = tris9.shape[0]
        return self.mesh_add(tris9.reshape(-1, 3), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3))

    def blas_add(self, verts, idx):
        """Adds a prototype mesh (object space) for instancing; returns its id."""
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
        bid = _u32()
        self._check(lib().kp_blas_add(self._h, _np_ptr(verts), verts.shape[0], _np_ptr(idx), idx.shape[0],
                                      ctypes.byref(bid)))
        return bid.value

    def blas_add_soup(self, tris9):
        tris9 = np.ascontiguousarray(tris9, np.float32).reshape(-1, 9)
        n = tris9.shape[0]
        return self.blas_add(tris9.reshape(-1, 3), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3))

    def mesh_set_normals(self, mesh_id, normals9):
        """Vertex normals [n,9] of a kp_mesh_add mesh (its object space)."""
        n = np.ascontiguousarray(normals9, np.float32).reshape(-1, 9)
        self._check(lib().kp_mesh_set_normals(self._h, mesh_id, _np_ptr(n)))

    def blas_set_normals(self, blas_id, normals9):
        n = np.ascontiguousarray(normals9, np.float32).reshape(-1, 9)
        self._check(lib().kp_blas_set_normals(self._h, blas_id, _np_ptr(n)))

    def material_add(self, d):
        """Adds a scenes.material() or scenes.composite() dict; returns its id."""
        mid = _u32()
        if d.get("lobes"):
            arr = (Lobe * len(d["lobes"]))()
            for i, (w, m) in enumerate(d["lobes"]):
                arr[i].weight[:] = [float(x) for x in w]
                arr[i].lobe = make_material(m)
            le = np.ascontiguousarray(d["Le"], np.float32)
            self._check(lib().kp_material_add_composite(self._h, ctypes.cast(arr, _p), len(d["lobes"]),
                                                        _np_ptr(le), ctypes.byref(mid)))
        else:
            m = make_material(d)
            self._check(lib().kp_material_add(self._h, ctypes.byref(m), ctypes.byref(mid)))
        return mid.value

    def blas_set_material(self, blas_id, material_id):
        self._check(lib().kp_blas_set_material(self._h, blas_id, material_id))

    def instance_add(self, blas_id, o2w):
        """Places prototype blas_id with a row-major 3x4 object-to-world matrix."""
        xf = np.ascontiguousarray(o2w, np.float32).reshape(12)
        iid = _u32()
        self._check(lib().kp_instance_add(self._h, blas_id, _np_ptr(xf), ctypes.byref(iid)))
        return iid.value

    def instance_set_material(self, instance_id, material_id):
        self._check(lib().kp_instance_set_material(self._h, instance_id, material_id))

    def prim_locate(self, prim):
        """(kind, index, local): kind 0 = kp_mesh_add mesh, 1 = instance."""
        k, i, l = _u32(), _u32(), _u32()
        self._check(lib().kp_prim_locate(self._h, prim, ctypes.byref(k), ctypes.byref(i), ctypes.byref(l)))
        return k.value, i.value, l.value

    def instance_get(self, instance_id):
        """(prototype id, object-to-world 3x4, world-to-object 3x4 as traversed)."""
        b = _u32()
        o2w = np.zeros((3, 4), np.float32)
        w2o = np.zeros((3, 4), np.float32)
        self._check(lib().kp_instance_get(self._h, instance_id, ctypes.byref(b), _np_ptr(o2w), _np_ptr(w2o)))
        return b.value, o2w, w2o

    def geometry_clear(self):
        """Drops every mesh, prototype, instance and the BVH (materials stay)."""
        self._check(lib().kp_geometry_clear(self._h))

    def bvh_build(self, leaf_size=4, sah_bins=32, instances=INSTANCES_AUTO, node_width=0):
        """node_width: children per node 2..8 (0: 8) or NODE_WIDTH_AUTO."""
        o = BuildOpts()
        o.leaf_size = leaf_size
        o.sah_bins = sah_bins
        o.instances = instances
        o.node_width = node_width
        self._check(lib().kp_bvh_build(self._h, ctypes.byref(o)))

    def stats(self):
        s = Stats()
        self._check(lib().kp_get_stats(self._h, ctypes.byref(s)))
        out = {k: getattr(s, k) for k, _ in Stats._fields_ if k != "reserved"}
        out["width_ms"] = list(s.width_ms)
        return out

    def set_trace_tuning(self, refill_below_lanes):
        self._check(lib().kp_set_trace_tuning(self._h, refill_below_lanes))

    def set_pass_paths(self, max_paths):
        """Paths in flight per integrator pass (default 2^26)."""
        self._check(lib().kp_set_pass_paths(self._h, max_paths))

    def release_buffers(self):
        """Frees cached path state and staging buffers (reallocated on demand)."""
        self._check(lib().kp_release_buffers(self._h))

    def intersect_closest(self, rays, out=None, counted=False, n=None):
        """rays: [n,8] float32 numpy array or DeviceBuffer.  Returns [n,4]
        float32 hits {t, prim bits, u, v} of the same kind (plus the node and
        triangle counts when counted=True).  For DeviceBuffers, n traces only
        the first n rays (all by default)."""
        if isinstance(rays, DeviceBuffer):
            cap = rays.nbytes // 32
            n = cap if n is None else int(n)
            if not 0 <= n <= cap or (out is not None and n * 16 > out.nbytes):
                raise ValueError("ray count %d exceeds the buffers" % n)
            if out is None:
                out = DeviceBuffer(self, (n, 4), np.float32)
            rp, hp, kind = rays.ptr, out.ptr, KP_PTR_DEVICE
        else:
            rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
            n = rays.shape[0]
            if out is None:
                out = np.zeros((n, 4), np.float32)
            rp, hp, kind = _np_ptr(rays), _np_ptr(out), KP_PTR_HOST
        if counted:
            nv, nt = _u64(), _u64()
            self._check(lib().kp_intersect_closest_counted(self._h, rp, n, hp, kind,
                                                           ctypes.byref(nv), ctypes.byref(nt)))
            return out, int(nv.value), int(nt.value)
        self._check(lib().kp_intersect_closest(self._h, rp, n, hp, kind))
        return out

    def intersect_closest_noalloc(self, rays, out, n=None):
        """kp_intersect_closest_noalloc: rays and out are DeviceBuffers; traces
        the first n rays (all by default). Asynchronous on the context stream."""
        n = rays.nbytes // 32 if n is None else n
        assert n * 32 <= rays.nbytes and n * 16 <= out.nbytes
        self._check(lib().kp_intersect_closest_noalloc(self._h, rays.ptr, n, out.ptr))
        return out

    def intersect_closest_host(self, rays):
        """Closest hit on host cores (kp_intersect_closest_host): [n,4] hits,
        bit-identical to intersect_closest."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((rays.shape[0], 4), np.float32)
        rc = lib().kp_intersect_closest_host(self._h, _np_ptr(rays), rays.shape[0], _np_ptr(out))
        if rc != KP_OK:
            raise KpError(rc, lib().kp_last_error(None).decode())
        return out

    def occluded_host(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.uint32)
        rc = lib().kp_occluded_host(self._h, _np_ptr(rays), rays.shape[0], _np_ptr(out))
        if rc != KP_OK:
            raise KpError(rc, lib().kp_last_error(None).decode())
        return out

    def occluded(self, rays, out=None):
        if isinstance(rays, DeviceBuffer):
            n = rays.nbytes // 32
            if out is None:
                out = DeviceBuffer(self, (n,), np.uint32)
            self._check(lib().kp_occluded(self._h, rays.ptr, n, out.ptr, KP_PTR_DEVICE))
            return out
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        if out is None:
            out = np.zeros(n, np.uint32)
        self._check(lib().kp_occluded(self._h, _np_ptr(rays), n, _np_ptr(out), KP_PTR_HOST))
        return out

    # ------------------------------------------------------------ render --
    def load_scene(self, scene, leaf_size=4, instances=INSTANCES_AUTO):
        """Adds every mesh and material of a scenes.Scene and builds the BVH."""
        mids = [self.material_add(md) for md in scene.materials]
        mnorm = list(getattr(scene, "mesh_normals", []))
        for k, (tris, mi) in enumerate(zip(scene.meshes, scene.mesh_material)):
            mesh = self.mesh_add_soup(tris)
            self._check(lib().kp_mesh_set_material(self._h, mesh, mids[mi]))
            if k < len(mnorm) and mnorm[k] is not None:
                self.mesh_set_normals(mesh, mnorm[k])
        protos = []
        pnorm = list(getattr(scene, "prototype_normals", []))
        for k, (tris, mi) in enumerate(zip(getattr(scene, "prototypes", []),
                                           getattr(scene, "prototype_material", []))):
            b = self.blas_add_soup(tris)
            self.blas_set_material(b, mids[mi])
            if k < len(pnorm) and pnorm[k] is not None:
                self.blas_set_normals(b, pnorm[k])
            protos.append(b)
        inst_mat = getattr(scene, "instance_material", [])
        for k, (pi, o2w) in enumerate(getattr(scene, "instances", [])):
            iid = self.instance_add(protos[pi], o2w)
            if k < len(inst_mat) and inst_mat[k] is not None:
                self.instance_set_material(iid, mids[inst_mat[k]])
        self.bvh_build(leaf_size=leaf_size, instances=instances)

    def film_reset(self, width, height):
        self._check(lib().kp_film_reset(self._h, width, height))

    def render_accumulate(self, desc):
        self._check(lib().kp_render_accumulate(self._h, ctypes.byref(desc)))

    def film_read(self, width, height):
        out = np.zeros((height, width, 3), np.float32)
        self._check(lib().kp_film_read(self._h, _np_ptr(out), KP_PTR_HOST))
        return out

    def render(self, desc):
        """Mean radiance image [h, w, 3] float32."""
        out = np.zeros((desc.height, desc.width, 3), np.float32)
        self._check(lib().kp_render(self._h, ctypes.byref(desc), _np_ptr(out), KP_PTR_HOST))
        return out

    def integrate_rays(self, rays, desc):
        """Radiance [n,3] of one path per ray (batched Integrator::Li)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((rays.shape[0], 3), np.float32)
        self._check(lib().kp_integrate_rays(self._h, _np_ptr(rays), rays.shape[0],
                                            ctypes.byref(desc), _np_ptr(out), KP_PTR_HOST))
        return out

    def render_stats(self):
        s = RenderStats()
        self._check(lib().kp_render_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in RenderStats._fields_ if k != "reserved"}

    # ----------------------------------------------------------- guiding --
    GUIDE_PARAMS = 17696

    def guiding_init(self, seed=0):
        self._check(lib().kp_guiding_init(self._h, seed))

    def guiding_set_params(self, params):
        params = np.ascontiguousarray(params, np.float32)
        self._check(lib().kp_guiding_set_params(self._h, _np_ptr(params), params.size))

    def guiding_get_params(self):
        p = np.zeros(self.GUIDE_PARAMS, np.float32)
        self._check(lib().kp_guiding_get_params(self._h, _np_ptr(p), p.size))
        return p

    def guiding_save(self):
        """The network's checkpoint (parameters, Adam moments and step) as bytes."""
        n = _u64()
        self._check(lib().kp_guiding_save(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._check(lib().kp_guiding_save(self._h, buf, n.value, ctypes.byref(n)))
        return buf.raw

    def guiding_load(self, data):
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        self._check(lib().kp_guiding_load(self._h, buf, len(data)))

    def guiding_eval(self, in6):
        in6 = np.ascontiguousarray(in6, np.float32).reshape(-1, 6)
        out = np.zeros((in6.shape[0], 32), np.float32)
        self._check(lib().kp_guiding_eval(self._h, _np_ptr(in6), in6.shape[0], _np_ptr(out)))
        return out

    def guiding_pdf(self, in6, dirs):
        in6 = np.ascontiguousarray(in6, np.float32).reshape(-1, 6)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        pdf = np.zeros(in6.shape[0], np.float32)
        self._check(lib().kp_guiding_pdf(self._h, _np_ptr(in6), _np_ptr(dirs), in6.shape[0], _np_ptr(pdf)))
        return pdf

    def guiding_sample(self, in6, u3):
        in6 = np.ascontiguousarray(in6, np.float32).reshape(-1, 6)
        u3 = np.ascontiguousarray(u3, np.float32).reshape(-1, 3)
        d = np.zeros((in6.shape[0], 3), np.float32)
        pdf = np.zeros(in6.shape[0], np.float32)
        self._check(lib().kp_guiding_sample(self._h, _np_ptr(in6), _np_ptr(u3), in6.shape[0], _np_ptr(d),
                                            _np_ptr(pdf)))
        return d, pdf

    def guiding_train_step(self, in6, dirs, weights, lr=1e-3, apply=True):
        in6 = np.ascontiguousarray(in6, np.float32).reshape(-1, 6)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        weights = np.ascontiguousarray(weights, np.float32)
        g = np.zeros(self.GUIDE_PARAMS, np.float32)
        self._check(lib().kp_guiding_train_step(self._h, _np_ptr(in6), _np_ptr(dirs), _np_ptr(weights),
                                                in6.shape[0], lr, int(apply), _np_ptr(g)))
        return g

    def guiding_configure(self, bsdf_fraction=0.5, min_alpha=0.5, lr=1e-3, train_steps=1, guide_after=1,
                          max_records=1 << 18, pass_paths=1 << 18, box=None, train_passes=0, weight_clamp=0.0,
                          relative_weights=False, max_guided_depth=0, late_train_steps=0):
        """Guided-render settings (kp_guiding_configure); box = (lo, hi) or None for the scene's bounds;
        max_guided_depth 0 guides every vertex."""
        d = GuidingDesc()
        d.bsdf_fraction = bsdf_fraction
        d.min_alpha = min_alpha
        d.lr = lr
        d.train_steps = train_steps
        d.guide_after = guide_after
        d.max_records = max_records
        d.pass_paths = pass_paths
        d.train_passes = train_passes
        d.weight_clamp = weight_clamp
        d.relative_weights = 1 if relative_weights else 0
        d.max_guided_depth = max_guided_depth
        d.late_train_steps = late_train_steps
        if box is not None:
            d.bmin[:] = [float(x) for x in box[0]]
            d.bmax[:] = [float(x) for x in box[1]]
        self._check(lib().kp_guiding_configure(self._h, ctypes.byref(d)))

    def device_copy(self, dst, src):
        """Kernel copy between two DeviceBuffers of the same size."""
        assert dst.nbytes == src.nbytes
        self._check(lib().kp_device_copy(self._h, dst.ptr, src.ptr, src.nbytes))

    def timer_begin(self):
        self._check(lib().kp_timer_begin(self._h))

    def timer_end(self):
        ms = ctypes.c_float()
        self._check(lib().kp_timer_end(self._h, ctypes.byref(ms)))
        return float(ms.value)


def hit_fields(hits):
    """Split an [n,4] float32 hit array into (t, prim uint32, u, v) numpy views."""
    h = np.asarray(hits)
    return h[:, 0], h[:, 1].view(np.uint32), h[:, 2], h[:, 3]


class Batch:
    """Many scenes rendered as one workload (kp_batch_*).  `owner` is a
    Context; `scenes` are Contexts with loaded scenes (same device).  The
    batch copies them, so they may be closed afterwards."""

    def __init__(self, owner, scenes):
        self._owner = owner
        arr = (_p * len(scenes))(*[sc._h for sc in scenes])
        h = _p()
        owner._check(lib().kp_batch_create(owner._h, arr, len(scenes), ctypes.byref(h)))
        self._h = h
        self.n = len(scenes)

    def close(self):
        if self._h:
            lib().kp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _descs(descs):
        return (RenderDesc * len(descs))(*descs)

    def tuning(self):
        t = BatchTuning()
        self._owner._check(lib().kp_batch_get_tuning(self._h, ctypes.byref(t)))
        return t

    def set_tuning(self, t=None, **kw):
        if t is None:
            t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._owner._check(lib().kp_batch_set_tuning(self._h, ctypes.byref(t)))

    def render(self, descs):
        assert len(descs) == self.n
        self._owner._check(lib().kp_batch_render(self._h, self._descs(descs)))

    def autotune(self, descs, passes=2):
        t = BatchTuning()
        self._owner._check(lib().kp_batch_autotune(self._h, self._descs(descs), passes, ctypes.byref(t)))
        return t

    def film_read(self, scene, width, height):
        out = np.zeros((height, width, 3), np.float32)
        self._owner._check(lib().kp_batch_film_read(self._h, scene, _np_ptr(out), KP_PTR_HOST))
        return out

    def stats(self):
        s = RenderStats()
        self._owner._check(lib().kp_batch_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in RenderStats._fields_ if k != "reserved"}
