"""Writes the render golden fixture (tests/golden/cornell64_ref.npz).

Scene: the Cornell box workload (kazen_proto_amd.scenes.cornell_box) at 64x64,
max depth 5, roulette from depth 3.  Contents:
  ref       oracle render, 16,384 spp, seed 1000, FP64 accumulation (the image
            oracle the GPU render is compared against)
  noise256  relMSE against `ref` of an independent 256-spp oracle render
            (seed 2000): the error level an unbiased 256-spp render has.
The reference holds no image fixtures; regenerate only on purpose:
    python tests/golden/make_render_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from kazen_proto_amd import scenes  # noqa: E402


def rel_mse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


def oracle_scene(s):
    return oracle.scene_from(s)


def desc(s, spp, seed):
    d = oracle.RenderDesc()
    d.width, d.height, d.spp = s.width, s.height, spp
    d.max_depth, d.rr_depth, d.seed = s.max_depth, s.rr_depth, seed
    d.spp_offset = 0
    d.eye[:], d.lookat[:], d.up[:] = s.eye, s.lookat, s.up
    d.tan_half_fov = s.tan_half_fov
    return d


def main():
    s = scenes.cornell_box(width=64, height=64)
    sc = oracle_scene(s)
    ref = sc.render(desc(s, 16384, 1000))
    n256 = sc.render(desc(s, 256, 2000))
    np.savez_compressed(os.path.join(HERE, "cornell64_ref.npz"), ref=ref,
                        noise256=rel_mse(n256, ref), spp=16384, seed=1000)
    print("ref mean %.5f, 256-spp relMSE %.3e" % (ref.mean(), rel_mse(n256, ref)))


if __name__ == "__main__":
    main()
