// Host-side BVH8 build interface of the HIP intersector.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "kp_layout.h"

namespace kp {

struct BuildOptions {
    uint32_t leaf_size = 4;   // max triangles per leaf (1..4)
    uint32_t sah_bins = 32;   // SAH bins per axis (4..64)
    uint32_t max_children = 8;   // node width (2..8)
};

struct Bvh8Host {
    std::vector<KpNode8> nodes;      // node 0 is the root
    std::vector<float> tris;         // 12 floats per triangle (kp_layout.h)
    std::vector<uint32_t> tri_order; // slot -> caller primitive index
    uint32_t max_depth = 0;          // levels of 8-wide nodes
    uint32_t n_bvh2_nodes = 0;
    float pad = 0.0f;                // box widening applied
};

// Builds the BVH8 over a triangle soup (9 floats per triangle). Returns false
// with a message on invalid input.
bool build_bvh8(const float* tris, uint32_t n_tris, const BuildOptions& opt, Bvh8Host& out,
                std::string& err);

}  // namespace kp
