"""Writes the FP64 oracle image of the guided-vs-plain equal-error workload
(tests/golden/guiding512_ref.npz): the slit-lit glossy room
(kazen_proto_amd.scenes.guiding_scene, 512x512, depth 10, 64 GGX spheres)
rendered by the CPU oracle integrator with 16,384 samples per pixel and FP64
accumulation, seed = scene seed + 7919 (independent of the renders it
judges), stored as float32.  bench.py measures both tracers' relMSE against
it.  Takes about 40 minutes on 8 threads:
    python tests/golden/make_guiding_reference.py [threads]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from kazen_proto_amd import scenes  # noqa: E402

SPP, CHUNK = 16384, 512


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 0
    s = scenes.guiding_scene(width=512, height=512, spp=1, max_depth=10)
    sc = oracle.scene_from(s)
    acc = np.zeros((s.height, s.width, 3), np.float64)
    t0 = time.time()
    for off in range(0, SPP, CHUNK):
        d = oracle.RenderDesc()
        d.width, d.height, d.spp = s.width, s.height, CHUNK
        d.max_depth, d.rr_depth, d.seed, d.spp_offset = s.max_depth, s.rr_depth, s.seed + 7919, off
        d.eye[:], d.lookat[:], d.up[:] = s.eye, s.lookat, s.up
        d.tan_half_fov = s.tan_half_fov
        acc += sc.render(d, threads=threads)
        print("%d/%d spp (%.0fs)" % (off + CHUNK, SPP, time.time() - t0), flush=True)
    ref = acc / (SPP // CHUNK)
    np.savez_compressed(os.path.join(HERE, "guiding512_ref.npz"), ref=ref.astype(np.float32), spp=SPP,
                        seed=s.seed + 7919)
    print("mean %.6f" % ref.mean())


if __name__ == "__main__":
    main()
