"""Procedural workloads: seeded scenes, cameras and ray sets.

A scene is a list of meshes (world-space triangle soups) with one material
each; emissive meshes are area lights (one light per mesh, as the reference
creates one GeometryLight per emissive mesh, src/core/scene.h:144-160).
Materials use the closure parameter table of the integrator:
  type 0 Lambert(albedo), 1 conductor(dist, alpha, Schlick F0),
  2 rough dielectric(dist, alpha, eta); dist 0 GGX, 1 Beckmann; Le = emission.
"""
import math
from dataclasses import dataclass, field

import numpy as np

LAMBERT, CONDUCTOR, DIELECTRIC, GLOSS = 0, 1, 2, 3
GGX, BECKMANN = 0, 1


def material(type_=LAMBERT, dist=GGX, albedo=(0.0, 0.0, 0.0), F0=(0.0, 0.0, 0.0), alpha=0.5,
             eta=1.5, Le=(0.0, 0.0, 0.0), alpha_y=0.0):
    """One closure; alpha along the shading frame's tangent, alpha_y along its
    bitangent (0: isotropic)."""
    return dict(type=int(type_), dist=int(dist), albedo=tuple(map(float, albedo)),
                F0=tuple(map(float, F0)), alpha=float(alpha), eta=float(eta),
                Le=tuple(map(float, Le)), alpha_y=float(alpha_y))


def composite(lobes, Le=(0.0, 0.0, 0.0)):
    """Material of 1..8 weighted lobes [(weight rgb, material dict)] emitting Le
    (a CompositeClosure of the reference, src/shading/bsdf.h:132-196)."""
    return dict(type=-1, lobes=[(tuple(map(float, w)), m) for w, m in lobes], Le=tuple(map(float, Le)))


@dataclass
class Scene:
    meshes: list = field(default_factory=list)        # float32 [n,9] each
    mesh_material: list = field(default_factory=list)  # material index per mesh
    materials: list = field(default_factory=list)
    eye: tuple = (0.0, 0.0, 3.0)
    lookat: tuple = (0.0, 0.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    fov_deg: float = 45.0
    width: int = 256
    height: int = 256
    spp: int = 16
    max_depth: int = 5
    rr_depth: int = 3
    seed: int = 0
    name: str = ""
    # instancing: prototype meshes in object space, placed by (prototype,
    # row-major 3x4 object-to-world matrix); primitive ids run over the meshes,
    # then over each instance's prototype in instance order
    prototypes: list = field(default_factory=list)
    prototype_material: list = field(default_factory=list)
    instances: list = field(default_factory=list)
    instance_material: list = field(default_factory=list)   # per instance: material or None (prototype's)
    # shading normals: per mesh / prototype None or [n,9] vertex normals (the
    # mesh's space; meshes are world space here)
    mesh_normals: list = field(default_factory=list)
    prototype_normals: list = field(default_factory=list)

    def add_mesh(self, tris9, mat_index, normals=None):
        self.meshes.append(np.ascontiguousarray(tris9, np.float32).reshape(-1, 9))
        self.mesh_material.append(int(mat_index))
        while len(self.mesh_normals) < len(self.meshes) - 1:
            self.mesh_normals.append(None)
        self.mesh_normals.append(None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 9))

    def add_prototype(self, tris9, mat_index, normals=None):
        self.prototypes.append(np.ascontiguousarray(tris9, np.float32).reshape(-1, 9))
        self.prototype_material.append(int(mat_index))
        while len(self.prototype_normals) < len(self.prototypes) - 1:
            self.prototype_normals.append(None)
        self.prototype_normals.append(None if normals is None else
                                      np.ascontiguousarray(normals, np.float32).reshape(-1, 9))
        return len(self.prototypes) - 1

    def shading_normals(self):
        """(normals [N,9] float32, has [N] uint8) in primitive-id order, world
        space and unit length as the library forms them (instances: the
        prototype's normals under the transpose of the world-to-object
        matrix), or None when no mesh has normals."""
        mn = list(self.mesh_normals) + [None] * (len(self.meshes) - len(self.mesh_normals))
        pn = list(self.prototype_normals) + [None] * (len(self.prototypes) - len(self.prototype_normals))
        if all(x is None for x in mn) and all(x is None for x in pn):
            return None
        outs, has = [], []

        def unit(v):
            v = v.reshape(-1, 3).astype(np.float64)
            return (v / np.sqrt(np.sum(v * v, axis=1, keepdims=True))).astype(np.float32).reshape(-1, 9)

        for m, nrm in zip(self.meshes, mn):
            outs.append(np.zeros((m.shape[0], 9), np.float32) if nrm is None else unit(nrm))
            has.append(np.full(m.shape[0], 0 if nrm is None else 1, np.uint8))
        for pi, o2w in self.instances:
            nrm = pn[pi]
            n = self.prototypes[pi].shape[0]
            if nrm is None:
                outs.append(np.zeros((n, 9), np.float32))
                has.append(np.zeros(n, np.uint8))
                continue
            m = np.asarray(o2w, np.float64)
            w2o = np.linalg.inv(np.vstack([m, [0, 0, 0, 1]]))[:3, :3].astype(np.float32).astype(np.float64)
            v = unit(nrm).reshape(-1, 3).astype(np.float64) @ w2o   # (W^T n) as rows
            outs.append(unit(v.reshape(-1, 9)))
            has.append(np.ones(n, np.uint8))
        return np.concatenate(outs), np.concatenate(has)

    def add_instance(self, proto, o2w, mat_index=None):
        self.instances.append((int(proto), np.asarray(o2w, np.float32).reshape(3, 4)))
        self.instance_material.append(None if mat_index is None else int(mat_index))
        return len(self.instances) - 1

    def placed_meshes(self):
        """(world-space [n,9] triangles, material) of every mesh and instance,
        in primitive-id order (instances transformed in float64)."""
        out = list(zip(self.meshes, self.mesh_material))
        for k, (pi, o2w) in enumerate(self.instances):
            m = np.asarray(o2w, np.float64)
            v = self.prototypes[pi].reshape(-1, 3).astype(np.float64) @ m[:, :3].T + m[:, 3]
            mi = self.instance_material[k] if k < len(self.instance_material) else None
            out.append((v.reshape(-1, 9).astype(np.float32), self.prototype_material[pi] if mi is None else mi))
        return out

    @property
    def n_tris(self):
        return int(sum(m.shape[0] for m, _ in self.placed_meshes()))

    def soup(self):
        return np.concatenate([m for m, _ in self.placed_meshes()], axis=0)

    def tri_material(self):
        return np.concatenate([np.full(m.shape[0], mi, np.int32) for m, mi in self.placed_meshes()])

    def tri_light(self):
        """Light id per triangle (-1: not emissive); one light per emissive mesh
        or instance."""
        out, lid = [], 0
        for m, mi in self.placed_meshes():
            if any(v != 0.0 for v in self.materials[mi]["Le"]):
                out.append(np.full(m.shape[0], lid, np.int32))
                lid += 1
            else:
                out.append(np.full(m.shape[0], -1, np.int32))
        return np.concatenate(out)

    @property
    def tan_half_fov(self):
        return float(np.float32(math.tan(math.radians(self.fov_deg) * 0.5)))


# ------------------------------------------------------------------ meshes --
def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); normal = (p1-p0) x (p2-p0)."""
    p = [np.asarray(x, np.float64) for x in (p0, p1, p2, p3)]
    return np.array([np.concatenate([p[0], p[1], p[2]]),
                     np.concatenate([p[0], p[2], p[3]])], np.float32)


def uv_sphere(n_lon, n_lat, center=(0, 0, 0), radius=1.0, drop_north_degenerate=True,
              displace=None):
    """UV sphere: n_lon x n_lat quads, two triangles each.  The north-pole band
    has one degenerate triangle per quad, dropped when drop_north_degenerate
    (so 64x32 gives 4,032 triangles); the south band keeps its zero-area ones.
    displace(p) -> radial scale factor per unit-sphere vertex."""
    j = np.arange(n_lat + 1)
    i = np.arange(n_lon)
    theta = np.pi * j / n_lat
    phi = 2.0 * np.pi * i / n_lon
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    sp, cp = np.sin(phi)[None, :], np.cos(phi)[None, :]
    x = st * cp
    y = np.broadcast_to(ct, (n_lat + 1, n_lon))
    z = st * sp
    p = np.stack([x, y, z], -1)                       # [lat+1, lon, 3]
    p[0, :] = (0.0, 1.0, 0.0)
    p[-1, :] = (0.0, -1.0, 0.0)
    if displace is not None:
        s = displace(p.reshape(-1, 3)).reshape(n_lat + 1, n_lon, 1)
        p = p * s
    p = p * radius + np.asarray(center, np.float64)
    i1 = (i + 1) % n_lon
    v00 = p[:-1][:, i]
    v01 = p[:-1][:, i1]
    v10 = p[1:][:, i]
    v11 = p[1:][:, i1]
    t_a = np.concatenate([v00, v10, v11], -1)         # [lat, lon, 9]
    t_b = np.concatenate([v00, v11, v01], -1)
    tris = np.stack([t_a, t_b], 2)                    # [lat, lon, 2, 9]
    if drop_north_degenerate:
        north = tris[0, :, 0:1]                       # keep (v00, v10, v11) at the north band
        rest = tris[1:].reshape(-1, 9)
        return np.concatenate([north.reshape(-1, 9), rest], 0).astype(np.float32)
    return tris.reshape(-1, 9).astype(np.float32)


# ------------------------------------------------------------ Perlin noise --
def _perlin_perm(rng):
    p = rng.permutation(256)
    return np.concatenate([p, p]).astype(np.int64)


def perlin3(pts, perm):
    """Improved Perlin noise (K. Perlin, "Improving Noise", SIGGRAPH 2002)."""
    pts = np.asarray(pts, np.float64)
    fl = np.floor(pts)
    ci = fl.astype(np.int64) & 255
    f = pts - fl
    u = f * f * f * (f * (f * 6 - 15) + 10)

    def grad(h, x, y, z):
        h = h & 15
        a = np.where(h < 8, x, y)
        b = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
        return np.where(h & 1, -a, a) + np.where(h & 2, -b, b)

    X, Y, Z = ci[:, 0], ci[:, 1], ci[:, 2]
    x, y, z = f[:, 0], f[:, 1], f[:, 2]
    A = perm[X] + Y
    AA = perm[A] + Z
    AB = perm[A + 1] + Z
    B = perm[X + 1] + Y
    BA = perm[B] + Z
    BB = perm[B + 1] + Z
    lerp = lambda t, a, b: a + t * (b - a)
    g000 = grad(perm[AA], x, y, z)
    g100 = grad(perm[BA], x - 1, y, z)
    g010 = grad(perm[AB], x, y - 1, z)
    g110 = grad(perm[BB], x - 1, y - 1, z)
    g001 = grad(perm[AA + 1], x, y, z - 1)
    g101 = grad(perm[BA + 1], x - 1, y, z - 1)
    g011 = grad(perm[AB + 1], x, y - 1, z - 1)
    g111 = grad(perm[BB + 1], x - 1, y - 1, z - 1)
    return lerp(u[:, 2],
                lerp(u[:, 1], lerp(u[:, 0], g000, g100), lerp(u[:, 0], g010, g110)),
                lerp(u[:, 1], lerp(u[:, 0], g001, g101), lerp(u[:, 0], g011, g111)))


# ------------------------------------------------------------ ray helpers --
def pinhole_rays(eye, lookat, up, fov_deg, width, height):
    """One ray per pixel centre, float32 [w*h, 8] {o, tmin=0, d, tmax=inf}."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(lookat, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    th = math.tan(math.radians(fov_deg) * 0.5)
    aspect = width / height
    px = (np.arange(width) + 0.5) / width
    py = (np.arange(height) + 0.5) / height
    nx = (2 * px - 1) * th * aspect
    ny = (1 - 2 * py) * th
    d = fwd[None, None, :] + nx[None, :, None] * right + ny[:, None, None] * upv
    d = d.reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = d.shape[0]
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = eye
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


def incoherent_rays(n, rng, box=1.2):
    """Origins uniform in [-box, box]^3, directions uniform on S^2."""
    o = rng.uniform(-box, box, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.empty((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


# ------------------------------------------------------ intersection bench --
def microbench_triangles(seed=1, n_lon=1024, n_lat=512, n_random=100_000, amplitude=0.05,
                         frequency=4.0):
    """Displaced unit sphere (n_lon x n_lat quads, all 2*n_lon*n_lat triangles
    kept) plus n_random small triangles (edge U[1e-3,1e-2], centres uniform in
    [-1,1]^3).  Defaults: 1,048,576 + 100,000 = 1,148,576 triangles."""
    rng = np.random.default_rng(seed)
    perm = _perlin_perm(rng)
    disp = lambda p: 1.0 + amplitude * perlin3(p * frequency, perm)
    sphere = uv_sphere(n_lon, n_lat, drop_north_degenerate=False, displace=disp)
    c = rng.uniform(-1.0, 1.0, size=(n_random, 3))
    edge = rng.uniform(1e-3, 1e-2, size=(n_random, 1))
    rad = edge / math.sqrt(3.0)
    a = rng.normal(size=(n_random, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    b = rng.normal(size=(n_random, 3))
    b -= (b * a).sum(1, keepdims=True) * a
    b /= np.linalg.norm(b, axis=1, keepdims=True)
    vs = []
    for k in range(3):
        ang = 2.0 * math.pi * k / 3.0
        vs.append(c + rad * (math.cos(ang) * a + math.sin(ang) * b))
    small = np.concatenate(vs, axis=1).astype(np.float32)
    return np.concatenate([sphere, small], 0)


def microbench_rays(seed=1, n_primary=4_194_304, n_incoherent=4_194_304):
    """(primary, incoherent) float32 [n,8] ray sets.  Primary rays: pinhole at
    (0,0,3) aimed at the origin, 45 degree vertical fov, square image of
    n_primary pixel centres."""
    side = int(round(math.sqrt(n_primary)))
    assert side * side == n_primary, "n_primary must be a square"
    prim = pinhole_rays((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, side, side)
    rng = np.random.default_rng(seed + 1000)
    inc = incoherent_rays(n_incoherent, rng)
    return prim, inc


# ------------------------------------------------------------ Cornell box --
def cornell_box(width=256, height=256, spp=16, max_depth=5, rr_depth=3, seed=7,
                sphere_lon=64, sphere_lat=32):
    """Closed-ish box [-1,1]^3 (5 walls, open toward the camera), a 0.25 x 0.25
    ceiling light (Le 15) and a Lambert 0.6 UV sphere (64x32 -> 4,032 tris)."""
    s = Scene(name="cornell", width=width, height=height, spp=spp, max_depth=max_depth,
              rr_depth=rr_depth, seed=seed, eye=(0.0, 0.0, 3.4), lookat=(0.0, 0.0, 0.0),
              up=(0.0, 1.0, 0.0), fov_deg=45.0)
    grey = len(s.materials); s.materials.append(material(LAMBERT, albedo=(0.75, 0.75, 0.75)))
    red = len(s.materials); s.materials.append(material(LAMBERT, albedo=(0.75, 0.1, 0.1)))
    green = len(s.materials); s.materials.append(material(LAMBERT, albedo=(0.1, 0.75, 0.1)))
    sph = len(s.materials); s.materials.append(material(LAMBERT, albedo=(0.6, 0.6, 0.6)))
    light = len(s.materials); s.materials.append(material(LAMBERT, albedo=(0, 0, 0), Le=(15, 15, 15)))
    # walls, normals pointing into the box
    s.add_mesh(quad((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1)), grey)     # floor
    s.add_mesh(quad((-1, 1, 1), (-1, 1, -1), (1, 1, -1), (1, 1, 1)), grey)         # ceiling
    s.add_mesh(quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)), grey)     # back
    s.add_mesh(quad((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1)), red)      # left
    s.add_mesh(quad((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1)), green)        # right
    h = 0.125
    y = 0.999
    s.add_mesh(quad((-h, y, -h), (h, y, -h), (h, y, h), (-h, y, h)), light)        # faces -y
    s.add_mesh(uv_sphere(sphere_lon, sphere_lat, center=(0.35, -0.6, -0.25), radius=0.4), sph)
    return s


def materials_scene(width=32, height=32, spp=16, max_depth=6, rr_depth=3, seed=11):
    """Every closure of the integrator in one small room: Lambert walls, a GGX
    conductor, a Beckmann conductor, a Beckmann dielectric (eta 1.5) and a GGX
    dielectric sphere under two area lights (one tilted)."""
    s = cornell_box(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=rr_depth,
                    seed=seed, sphere_lon=24, sphere_lat=12)
    s.name = "materials"
    s.meshes.pop()          # drop the Lambert sphere
    s.mesh_material.pop()
    mats = [material(CONDUCTOR, GGX, F0=(0.95, 0.7, 0.4), alpha=0.2),
            material(CONDUCTOR, BECKMANN, F0=(0.6, 0.9, 0.9), alpha=0.35),
            material(DIELECTRIC, BECKMANN, alpha=0.15, eta=1.5),
            material(DIELECTRIC, GGX, alpha=0.4, eta=1.5)]
    centers = [(-0.5, -0.65, -0.4), (0.5, -0.65, -0.5), (-0.35, -0.1, 0.35), (0.45, 0.0, 0.3)]
    for m, c in zip(mats, centers):
        s.materials.append(m)
        s.add_mesh(uv_sphere(24, 12, center=c, radius=0.3), len(s.materials) - 1)
    lt = len(s.materials)
    s.materials.append(material(LAMBERT, albedo=(0.2, 0.2, 0.2), Le=(4.0, 6.0, 8.0)))
    s.add_mesh(quad((-0.95, 0.2, -0.6), (-0.95, 0.2, -0.2), (-0.95, 0.6, -0.2), (-0.95, 0.6, -0.6)), lt)
    return s


def sphere_normals(tris9, center):
    """Smooth-shading vertex normals of a sphere mesh: (v - center) / |v - center|."""
    v = np.asarray(tris9, np.float64).reshape(-1, 3) - np.asarray(center, np.float64)
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).reshape(-1, 9).astype(np.float32)


def closures_scene(width=32, height=32, spp=16, max_depth=6, rr_depth=3, seed=13):
    """The closures the reference's shader groups build beyond single lobes:
    composites (plastic = Lambert + gloss; brushed metal = anisotropic GGX
    conductor + a little Lambert), an anisotropic Beckmann dielectric, a gloss
    sphere, and smooth-shaded spheres (vertex normals) in the Cornell room."""
    s = cornell_box(width=width, height=height, spp=spp, max_depth=max_depth, rr_depth=rr_depth,
                    seed=seed, sphere_lon=24, sphere_lat=12)
    s.name = "closures"
    s.meshes.pop()
    s.mesh_material.pop()
    s.mesh_normals = [None] * len(s.meshes)
    plastic = composite([((0.7, 0.2, 0.2), material(LAMBERT, albedo=(1, 1, 1))),
                         ((1, 1, 1), material(GLOSS, GGX, alpha=0.15, eta=1.5))])
    brushed = composite([((1, 1, 1), material(CONDUCTOR, GGX, F0=(0.9, 0.8, 0.5), alpha=0.08, alpha_y=0.5)),
                         ((0.1, 0.1, 0.1), material(LAMBERT, albedo=(1, 1, 1)))])
    mats = [(plastic, (-0.5, -0.65, -0.4), True),
            (brushed, (0.5, -0.65, -0.5), True),
            (material(DIELECTRIC, BECKMANN, alpha=0.1, alpha_y=0.35, eta=1.5), (-0.35, -0.1, 0.35), False),
            (material(GLOSS, BECKMANN, alpha=0.3, eta=1.6), (0.45, 0.0, 0.3), True)]
    for m, c, smooth in mats:
        s.materials.append(m)
        t = uv_sphere(24, 12, center=c, radius=0.3)
        s.add_mesh(t, len(s.materials) - 1, normals=sphere_normals(t, c) if smooth else None)
    lt = len(s.materials)
    s.materials.append(material(LAMBERT, albedo=(0.2, 0.2, 0.2), Le=(4.0, 6.0, 8.0)))
    s.add_mesh(quad((-0.95, 0.2, -0.6), (-0.95, 0.2, -0.2), (-0.95, 0.6, -0.2), (-0.95, 0.6, -0.6)), lt)
    return s


def batch_scenes(n=32, seed=100):
    """The batched-workload scenes: n terrain-style scenes with seeds seed ..
    seed + n - 1, triangle counts log-uniform in [1e4, 2e6], resolutions from
    {128^2, 256^2, 512^2, 1024^2}, spp from {4, 16, 64}, depth 8 (the scene
    sizes and settings are drawn from one generator seeded with `seed`)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        target = float(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
        grid = max(16, int(round(np.sqrt(max(target - 16 * 240, 1000) / 2.0))) + 1)
        res = int(rng.choice([128, 256, 512, 1024]))
        spp = int(rng.choice([4, 16, 64]))
        out.append(terrain_scene(grid=grid, n_spheres=16, sphere_lon=16, sphere_lat=8, width=res,
                                 height=res, spp=spp, max_depth=8, seed=seed + i))
    return out


def orient_toward(tris9, point):
    """Flips triangles so their winding normals face `point` (inward for a
    closed mesh around it)."""
    t = np.array(tris9, np.float64).reshape(-1, 3, 3)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    c = t.mean(1)
    away = ((np.asarray(point, np.float64) - c) * n).sum(1) < 0
    t[away] = t[away][:, [0, 2, 1]]
    return t.reshape(-1, 9).astype(np.float32)


def furnace_box(albedo=0.5, Le=1.0, width=16, height=16, spp=64, max_depth=4, rr_depth=64,
                seed=5):
    """Closed cube [-1,1]^3, every face Lambert(albedo) emitting Le toward the
    inside, camera at the centre.  Closed form of the estimator's mean:
    L = Le * sum_{k=0}^{max_depth} albedo^k (one emission per vertex, NEE and
    BSDF sampling combined by MIS)."""
    s = Scene(name="furnace", width=width, height=height, spp=spp, max_depth=max_depth,
              rr_depth=rr_depth, seed=seed, eye=(0.0, 0.0, 0.0), lookat=(0.3, 0.2, -1.0),
              up=(0.0, 1.0, 0.0), fov_deg=90.0)
    s.materials.append(material(LAMBERT, albedo=(albedo,) * 3, Le=(Le,) * 3))
    faces = [quad((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1)),
             quad((-1, 1, 1), (-1, 1, -1), (1, 1, -1), (1, 1, 1)),
             quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)),
             quad((-1, -1, 1), (-1, 1, 1), (1, 1, 1), (1, -1, 1)),
             quad((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1)),
             quad((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1))]
    s.add_mesh(orient_toward(np.concatenate(faces), (0, 0, 0)), 0)
    return s


def furnace_expected(albedo, Le, max_depth):
    return Le * sum(albedo ** k for k in range(max_depth + 1))


def small_test_scene(seed=3, n_random=400):
    """Tiny scene for brute-force parity: a 16x8 sphere, a floor quad and random
    triangles, including shared edges and vertices."""
    rng = np.random.default_rng(seed)
    parts = [uv_sphere(16, 8, center=(0.1, 0.0, -0.2), radius=0.6),
             quad((-2, -1, 2), (2, -1, 2), (2, -1, -2), (-2, -1, -2))]
    c = rng.uniform(-1.0, 1.0, size=(n_random, 3))
    v = c[:, None, :] + rng.normal(scale=0.08, size=(n_random, 3, 3))
    parts.append(v.reshape(-1, 9).astype(np.float32))
    return np.concatenate(parts, 0).astype(np.float32)


def guiding_scene(width=512, height=512, spp=16, max_depth=10, rr_depth=3, seed=5, n_spheres=64,
                  sphere_lon=16, sphere_lat=8):
    """Glossy room lit through a slit (the neural path-guiding workload).

    Box room [-1,1]^3 (12 triangles: Lambert 0.7 walls and ceiling, Lambert
    0.6 floor).  An occluding wall at x = 0.5 (4 triangles, Lambert 0.7) has a
    0.1-wide horizontal slit at y in [0.2, 0.3] across the whole room; behind
    it a 0.2 x 0.2 area light (radiance 500) at x = 0.95, y in [0.75, 0.95],
    faces the wall, so the only light reaching the camera's side comes through
    the slit (a stripe on the floor near x in [-0.8, -0.2], then indirect).
    64 GGX conductor spheres (radius U[0.05, 0.1], alpha U[0.02, 0.3], F0
    U[0.6, 0.95] per channel) stand on the floor of the camera's side,
    non-overlapping.  Camera in the corner (-0.95, 0.5, 0.95) looking at
    (0.3, -0.6, -0.5), 65 degree vertical field of view."""
    rng = np.random.default_rng(seed)
    s = Scene(name="guiding", width=width, height=height, spp=spp, max_depth=max_depth,
              rr_depth=rr_depth, seed=seed, eye=(-0.95, 0.5, 0.95), lookat=(0.3, -0.6, -0.5),
              up=(0.0, 1.0, 0.0), fov_deg=65.0)
    wall = len(s.materials); s.materials.append(material(LAMBERT, albedo=(0.7, 0.7, 0.7)))
    floor = len(s.materials); s.materials.append(material(LAMBERT, albedo=(0.6, 0.6, 0.6)))
    light = len(s.materials); s.materials.append(material(LAMBERT, albedo=(0, 0, 0), Le=(500, 500, 500)))
    s.add_mesh(quad((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1)), floor)
    s.add_mesh(np.concatenate([
        quad((-1, 1, 1), (-1, 1, -1), (1, 1, -1), (1, 1, 1)),        # ceiling
        quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)),    # back (z = -1)
        quad((-1, -1, 1), (-1, 1, 1), (1, 1, 1), (1, -1, 1)),        # front (z = +1)
        quad((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1)),    # left (x = -1)
        quad((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1)),        # right (x = +1)
    ]), wall)
    xw = 0.5
    s.add_mesh(np.concatenate([
        quad((xw, -1, -1), (xw, -1, 1), (xw, 0.2, 1), (xw, 0.2, -1)),
        quad((xw, 0.3, -1), (xw, 0.3, 1), (xw, 1, 1), (xw, 1, -1)),
    ]), wall)
    xl = 0.95
    s.add_mesh(quad((xl, 0.75, -0.1), (xl, 0.75, 0.1), (xl, 0.95, 0.1), (xl, 0.95, -0.1)), light)  # faces -x
    placed = []
    tries = 0
    while len(placed) < n_spheres:
        tries += 1
        if tries > 100000:
            raise RuntimeError("could not place the spheres")
        r = rng.uniform(0.05, 0.1)
        c = np.array([rng.uniform(-0.9 + r, 0.45 - r), -1.0 + r, rng.uniform(-0.9 + r, 0.9 - r)])
        if any(np.linalg.norm(c - pc) < r + pr + 0.01 for pc, pr in placed):
            continue
        placed.append((c, r))
        alpha = rng.uniform(0.02, 0.3)
        F0 = rng.uniform(0.6, 0.95, 3)
        m = len(s.materials)
        s.materials.append(material(CONDUCTOR, GGX, F0=F0, alpha=alpha))
        s.add_mesh(uv_sphere(sphere_lon, sphere_lat, center=tuple(c), radius=r), m)
    return s


def fbm_heights(xs, zs, perm, octaves=6, lacunarity=2.0, gain=0.5, base_freq=0.25):
    """fBm of Perlin noise on the grid xs x zs (sum over octaves o of
    gain^o noise(lacunarity^o base_freq (x, y_o, z)), y_o = 0.37 o decorrelates
    the octaves); returns [len(zs), len(xs)]."""
    X, Z = np.meshgrid(np.asarray(xs, np.float64), np.asarray(zs, np.float64))
    h = np.zeros(X.shape)
    amp, freq = 1.0, base_freq
    for o in range(octaves):
        pts = np.stack([X.ravel() * freq, np.full(X.size, 0.37 * o), Z.ravel() * freq], 1)
        h += amp * perlin3(pts, perm).reshape(X.shape)
        amp *= gain
        freq *= lacunarity
    return h


def heightfield_tris(xs, zs, h):
    """Two triangles per grid cell, upward winding; [2 (nz-1)(nx-1), 9] float32."""
    X, Z = np.meshgrid(np.asarray(xs, np.float32), np.asarray(zs, np.float32))
    P = np.stack([X, np.asarray(h, np.float32), Z], -1)
    v00, v10 = P[:-1, :-1], P[:-1, 1:]
    v01, v11 = P[1:, :-1], P[1:, 1:]
    a = np.concatenate([v00, v01, v10], -1)
    b = np.concatenate([v10, v01, v11], -1)
    return np.stack([a, b], 2).reshape(-1, 9)


def random_material(rng):
    """Uniform over Lambert (albedo U[0.2,0.8]), GGX conductor (alpha
    U[0.05,0.8], F0 U[0.5,1]) and Beckmann dielectric (eta 1.5, alpha
    U[0.05,0.5])."""
    kind = int(rng.integers(3))
    if kind == 0:
        return material(LAMBERT, albedo=rng.uniform(0.2, 0.8, 3))
    if kind == 1:
        return material(CONDUCTOR, GGX, F0=rng.uniform(0.5, 1.0, 3), alpha=rng.uniform(0.05, 0.8))
    return material(DIELECTRIC, BECKMANN, alpha=rng.uniform(0.05, 0.5), eta=1.5)


def terrain_scene(grid=2048, n_spheres=256, sphere_lon=64, sphere_lat=32, tiles=8, width=1024, height=1024,
                  spp=64, max_depth=8, rr_depth=3, seed=2, extent=10.0, instanced=False):
    """Full plain path-trace workload: a grid x grid heightfield over
    [-extent, extent]^2 (2 (grid-1)^2 triangles, 8,380,418 at 2048), heights
    1.5 fBm (6 octaves, lacunarity 2, gain 0.5), split into tiles x tiles
    meshes; n_spheres copies of one UV sphere (sphere_lon x sphere_lat, 4,032
    triangles at 64 x 32) at uniform random positions above the ground,
    radius U[0.15, 0.45], wound outward, flattened into world space (or, with
    instanced, one prototype placed n_spheres times with a material per
    instance: the same scene, traversed two-level); every mesh and sphere
    draws its material with random_material; 4 area-light quads (2 x 2, radiance 30)
    at height 7 facing down.  Camera (0, 6, 15) looking at (0, 0, 0), 50
    degree field of view."""
    rng = np.random.default_rng(seed)
    perm = _perlin_perm(rng)
    s = Scene(name="terrain", width=width, height=height, spp=spp, max_depth=max_depth,
              rr_depth=rr_depth, seed=seed, eye=(0.0, 6.0, 15.0), lookat=(0.0, 0.0, 0.0),
              up=(0.0, 1.0, 0.0), fov_deg=50.0)
    xs = np.linspace(-extent, extent, grid)
    h = 1.5 * fbm_heights(xs, xs, perm)
    step = (grid - 1 + tiles - 1) // tiles
    for tz in range(tiles):
        for tx in range(tiles):
            z0, x0 = tz * step, tx * step
            z1, x1 = min(z0 + step, grid - 1), min(x0 + step, grid - 1)
            if z1 <= z0 or x1 <= x0:
                continue
            m = len(s.materials)
            s.materials.append(random_material(rng))
            s.add_mesh(heightfield_tris(xs[x0:x1 + 1], xs[z0:z1 + 1], h[z0:z1 + 1, x0:x1 + 1]), m)
    # uv_sphere winds inward; swap two vertices so winding normals point out
    # (the dielectric's outside is the side its winding normal faces)
    unit = uv_sphere(sphere_lon, sphere_lat).reshape(-1, 3, 3)[:, [0, 2, 1]]
    proto = None
    if instanced and n_spheres:
        proto = s.add_prototype(unit.reshape(-1, 9), 0)
    for _ in range(n_spheres):
        r = rng.uniform(0.15, 0.45)
        cx, cz = rng.uniform(-0.9 * extent, 0.9 * extent, 2)
        ix = int(np.clip(np.searchsorted(xs, cx), 0, grid - 1))
        iz = int(np.clip(np.searchsorted(xs, cz), 0, grid - 1))
        cy = h[iz, ix] + r * rng.uniform(1.0, 2.5)
        m = len(s.materials)
        s.materials.append(random_material(rng))
        if proto is not None:
            s.add_instance(proto, [[r, 0, 0, cx], [0, r, 0, cy], [0, 0, r, cz]], m)
        else:
            s.add_mesh((unit * r + np.array([cx, cy, cz])).reshape(-1, 9), m)
    light = len(s.materials)
    s.materials.append(material(LAMBERT, albedo=(0, 0, 0), Le=(30, 30, 30)))
    for lx, lz in ((-5, -5), (5, -5), (-5, 5), (5, 5)):
        y = 7.0
        s.add_mesh(quad((lx - 1, y, lz - 1), (lx + 1, y, lz - 1), (lx + 1, y, lz + 1), (lx - 1, y, lz + 1)), light)
    return s
