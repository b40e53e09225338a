"""Two-level (instanced) closest hit of the oracle, pinned on the CPU.

oracle.instanced_closest moves the ray into each instance's object space and
runs the watertight test there.  Checks:
  - identity instances reproduce the one-level brute force bit for bit;
  - signed axis permutations with integer translations on dyadic geometry
    are exact in FP32, so instancing must reproduce the brute force over the
    flattened (transformed) triangles bit for bit (up to the sign of a zero
    barycentric); a transposed matrix or a wrong translation sign fails this;
  - general rotations and scales agree with the flattened brute force up to
    rounding (same primitive for all but grazing rays, t to 1e-5);
  - two coincident instances tie: the lower primitive id (earlier instance)
    wins.
"""
import numpy as np

import oracle


def dyadic_mesh(rng, n, scale=4.0):
    """Random triangles with coordinates on a 1/8 grid."""
    c = np.round(rng.uniform(-scale, scale, (n, 1, 3)) * 8) / 8
    v = c + np.round(rng.uniform(-1, 1, (n, 3, 3)) * 8) / 8
    return v.reshape(n, 9).astype(np.float32)


def rays_through(rng, n, box=6.0):
    """Rays from a sphere of radius 3*box toward points in the box; directions
    with distinct component magnitudes (no axis ties in the watertight setup)."""
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * 3 * box
    tgt = rng.uniform(-box, box, (n, 3))
    d = tgt - o
    o = np.round(o * 8) / 8
    d = np.round(d * 8) / 8
    a = np.abs(d)
    keep = (a[:, 0] != a[:, 1]) & (a[:, 1] != a[:, 2]) & (a[:, 0] != a[:, 2])
    o, d = o[keep], d[keep]
    n = o.shape[0]
    r = np.zeros((n, 8), np.float32)
    r[:, :3], r[:, 3], r[:, 4:7], r[:, 7] = o, 0.0, d, np.inf
    return r


def apply(o2w, tris):
    m = np.asarray(o2w, np.float64)
    v = tris.reshape(-1, 3).astype(np.float64) @ m[:, :3].T + m[:, 3]
    return v.reshape(-1, 9).astype(np.float32)


def signed_perm(rng):
    p = rng.permutation(3)
    s = rng.choice([-1.0, 1.0], 3)
    r = np.zeros((3, 3))
    r[np.arange(3), p] = s
    return r


def hits_equal(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def hits_equal_but_zero_sign(a, b):
    """t and primitive bit for bit, u and v as values (+0 == -0: where an edge
    function is exactly zero, x - x is +0 in either orientation, so a mirrored
    frame yields the other zero sign)."""
    return (np.array_equal(a[:, :2].view(np.uint32), b[:, :2].view(np.uint32)) and
            np.array_equal(a[:, 2:], b[:, 2:]))


def test_identity_instance_is_brute_force():
    rng = np.random.default_rng(1)
    tris = dyadic_mesh(rng, 300)
    rays = rays_through(rng, 2000)
    eye = np.hstack([np.eye(3), np.zeros((3, 1))]).astype(np.float32)
    a = oracle.instanced_closest([tris], [(0, eye)], rays)
    b = oracle.closest_bruteforce(tris, rays)
    assert hits_equal(a, b)
    assert np.count_nonzero(a[:, 1].view(np.uint32) != 0xFFFFFFFF) > 200


def test_signed_permutations_are_exact():
    rng = np.random.default_rng(2)
    meshes = [dyadic_mesh(rng, 60, 1.5), dyadic_mesh(rng, 40, 1.5)]
    inst, flat = [], []
    for k in range(12):
        mi = k % 2
        r = signed_perm(rng)
        t = rng.integers(-5, 6, 3).astype(np.float64)
        o2w = np.hstack([r, t[:, None]])
        w2o = np.hstack([r.T, (-r.T @ t)[:, None]])
        inst.append((mi, w2o.astype(np.float32)))
        flat.append(apply(o2w, meshes[mi]))
    rays = rays_through(rng, 3000)
    a = oracle.instanced_closest(meshes, inst, rays)
    b = oracle.closest_bruteforce(np.concatenate(flat), rays)
    assert hits_equal_but_zero_sign(a, b)
    assert np.count_nonzero(a[:, 1].view(np.uint32) != 0xFFFFFFFF) > 500


def test_planted_transpose_is_caught():
    rng = np.random.default_rng(3)
    mesh = dyadic_mesh(rng, 60, 1.5)
    r = np.array([[0, 1, 0], [0, 0, -1], [-1, 0, 0]], np.float64)   # not symmetric
    t = np.array([2.0, -1.0, 3.0])
    o2w = np.hstack([r, t[:, None]])
    wrong = np.hstack([r, (-r @ t)[:, None]])          # R instead of R^T
    rays = rays_through(rng, 2000)
    b = oracle.closest_bruteforce(apply(o2w, mesh), rays)
    a = oracle.instanced_closest([mesh], [(0, wrong.astype(np.float32))], rays)
    assert not hits_equal_but_zero_sign(a, b)
    right = np.hstack([r.T, (-r.T @ t)[:, None]])
    a = oracle.instanced_closest([mesh], [(0, right.astype(np.float32))], rays)
    assert hits_equal_but_zero_sign(a, b)


def test_general_transforms_close_to_flattened():
    rng = np.random.default_rng(4)
    meshes = [dyadic_mesh(rng, 80, 1.0)]
    inst, flat = [], []
    for k in range(8):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        rot = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        m = rot @ np.diag(rng.uniform(0.5, 2.0, 3))
        t = rng.uniform(-4, 4, 3)
        o2w = np.hstack([m, t[:, None]])
        inv = np.linalg.inv(m)
        inst.append((0, np.hstack([inv, (-inv @ t)[:, None]]).astype(np.float32)))
        flat.append(apply(o2w, meshes[0]))
    rays = rays_through(rng, 3000)
    a = oracle.instanced_closest(meshes, inst, rays)
    b = oracle.closest_bruteforce(np.concatenate(flat), rays)
    pa, pb = a[:, 1].view(np.uint32), b[:, 1].view(np.uint32)
    both = (pa != 0xFFFFFFFF) & (pb != 0xFFFFFFFF)
    assert both.sum() > 500
    assert np.mean(pa == pb) > 0.995
    same = both & (pa == pb)
    np.testing.assert_allclose(a[same, 0], b[same, 0], rtol=1e-5)


def test_coincident_instances_tie_to_lower_prim():
    rng = np.random.default_rng(5)
    mesh = dyadic_mesh(rng, 50, 1.5)
    eye = np.hstack([np.eye(3), np.zeros((3, 1))]).astype(np.float32)
    rays = rays_through(rng, 1000, box=2.0)
    a = oracle.instanced_closest([mesh], [(0, eye), (0, eye)], rays)
    b = oracle.closest_bruteforce(mesh, rays)
    assert hits_equal(a, b)   # the first instance's ids (0..49) win every tie
