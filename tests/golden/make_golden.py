"""Writes the intersection golden fixture (tests/golden/trace_small.npz).

Inputs: the seeded small test scene (kazen_proto_amd.scenes.small_test_scene)
and 4,096 seeded rays (incoherent, pinhole, plus rays aimed exactly at shared
edges/vertices).  Expected outputs: closest hits from the oracle's brute-force
watertight test (O(rays x triangles), no BVH involved).  The reference holds
no hit-list fixtures of its own; this file is regenerated only on purpose:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from kazen_proto_amd import scenes  # noqa: E402


def make_rays(tris):
    rng = np.random.default_rng(11)
    inc = scenes.incoherent_rays(2048, rng, box=1.5)
    pin = scenes.pinhole_rays((0.2, 0.4, 3.0), (0, 0, 0), (0, 1, 0), 60, 32, 32)
    # rays aimed at triangle vertices and edge midpoints (shared features)
    t = tris.reshape(-1, 3, 3)
    pick = rng.integers(0, t.shape[0], size=1024)
    which = rng.integers(0, 3, size=1024)
    p = np.where((np.arange(1024) % 2 == 0)[:, None],
                 t[pick, which], 0.5 * (t[pick, which] + t[pick, (which + 1) % 3]))
    o = p + rng.normal(size=(1024, 3)) * 1.5
    feat = np.empty((1024, 8), np.float32)
    feat[:, 0:3] = o
    feat[:, 3] = 0
    feat[:, 4:7] = p - o
    feat[:, 7] = np.inf
    return np.concatenate([inc, pin, feat]).astype(np.float32)


def main():
    tris = scenes.small_test_scene(seed=3)
    rays = make_rays(tris)
    hits = oracle.closest_bruteforce(tris, rays)
    np.savez_compressed(os.path.join(HERE, "trace_small.npz"), tris=tris, rays=rays, hits=hits)
    print("wrote", rays.shape[0], "rays,", tris.shape[0], "triangles")


if __name__ == "__main__":
    main()
