// Host traversal of the HIP intersector's BVH8 records (no HIP runtime needed).
//
// Answers single-ray closest-hit and occlusion queries on host cores for
// callers that intersect one ray at a time from many threads, as the
// reference's TBB tile workers call the const Hitable::intersect
// (src/core/renderer.cpp:24-26, src/core/hitable.h:12-33).  The arithmetic
// of the node test, the instance transform and the watertight triangle test
// is the one the GPU kernels execute (same operation order, contraction off,
// correctly rounded division), so every hit is bit-identical to k_closest's.
// A HostBvh is immutable once built; any number of threads may query it.
#pragma once
#include <stdint.h>

#include <vector>

#include "kp_layout.h"

namespace kp {

struct HostBvh {
    std::vector<KpNode8> nodes;      // one-level: the BVH8; two-level: [top | unit 0 | unit 1 ...]
    std::vector<float> tris;         // 12 floats per triangle record (kp_layout.h)
    bool two_level = false;
    uint32_t n_tlas = 0;             // nodes [0, n_tlas) form the top level
    uint32_t world_root = KP_INVALID;   // root of the kp_mesh_add meshes' BVH (two-level)
    std::vector<float> insts;        // 16 floats per instance: w2o rows, {root, prim base, flags, 0}
    std::vector<uint32_t> tlas_inst; // top-level leaf position -> instance
    bool empty() const { return nodes.empty(); }
};

// rays: 8 floats each {ox, oy, oz, tmin, dx, dy, dz, tmax}; hits: 4 words
// each {t, prim, u, v} (prim 0xFFFFFFFF and t = +inf on a miss).
void host_closest(const HostBvh& b, const float* rays, uint32_t n, float* hits);
// out[i] = 1 if any primitive is hit with t in [tmin, tmax].
void host_occluded(const HostBvh& b, const float* rays, uint32_t n, uint32_t* out);

}  // namespace kp
