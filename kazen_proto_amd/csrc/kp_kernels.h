// Kernel-launch interface between the C-ABI layer and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Threads per block of the traversal kernels and LDS short-stack depth.
#define KP_TRACE_BLOCK 256
#define KP_TRACE_STACK 8

// Static device scratch of the allocation-free closest-hit entry point:
// 64 MiB per device (counter, error word, stack spill area).
#define KP_SCRATCH_BYTES (64u << 20)
#define KP_SCRATCH_ERR_OFFSET 256u
#define KP_SCRATCH_SPILL_OFFSET 512u
#define KP_SCRATCH_MAX_DEVICES 64

namespace kp {

struct TraceArgs {
    const uint4* nodes;     // BVH8 nodes, 8 x uint4 each
    const float4* tris;     // 3 x float4 per triangle
    const float4* rays;     // 2 x float4 per ray {o, tmin}, {d, tmax}
    float4* hits;           // closest: {t, prim bits, u, v}
    uint32_t* occl;         // occluded: 0/1 per ray
    uint32_t n;             // ray count
    const uint32_t* n_dev;  // if set, the ray count is read from device memory
    uint32_t* counter;      // work counter (reset by the launcher)
    uint2* spill;           // per-lane stack spill area
    uint32_t spill_cap;     // entries per lane
    uint32_t spill_blocks;  // blocks of KP_TRACE_BLOCK lanes the spill area covers (launch grid limit)
    uint32_t refill_below;  // refill a wave once fewer lanes than this are busy
    uint32_t* err;          // bit0 stack overflow, bit1 step limit
    uint32_t max_steps;     // per-ray step limit (> any legitimate traversal)
    unsigned long long* counters;   // [nodes visited, triangles tested]
    const uint32_t* roots;  // if set, root node of each ray (batched scenes); else node 0
    // two-level traversal (instances): nodes [0, n_tlas) are the top level,
    // whose leaf children name instances through tlas_inst; insts holds per
    // instance the world-to-object rows and {root node, primitive base, flags}
    // world_root is the root of the kp_mesh_add meshes' BVH (KP_INVALID if
    // none), traversed before the top level.
    const float4* insts;
    const uint32_t* tlas_inst;
    uint32_t n_tlas;
    uint32_t world_root;
    // one-level closest / occluded launches of the integrator: the kernels
    // compiled for 5 waves per SIMD (96 registers) instead of 4
    uint32_t five_waves;
};

hipError_t launch_closest(const TraceArgs& a, uint32_t grid, bool count, hipStream_t s);
hipError_t launch_occluded(const TraceArgs& a, uint32_t grid, hipStream_t s);
hipError_t trace_occupancy(int* blocks_per_cu_closest, int* blocks_per_cu_occluded);
hipError_t trace_occupancy_five(int* blocks_per_cu_closest, int* blocks_per_cu_occluded);
hipError_t shade_occupancy(int* blocks_per_cu_plain, int* blocks_per_cu_guided);
hipError_t launch_reset_u32(uint32_t* p, uint32_t n, hipStream_t s);
// Device address of the static scratch on `device` (made current by the
// caller), looked up once per device.
hipError_t noalloc_scratch(int device, void** ptr);
hipError_t launch_copy16(void* dst, const void* src, uint64_t n16, int cus, hipStream_t s);

}  // namespace kp
