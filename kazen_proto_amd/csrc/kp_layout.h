// Device data layout of the HIP intersector (host and device visible).
//
// BVH8 node: 128 bytes, 16-byte aligned, read by one lane as eight 16-byte
// loads that fall in one 128-byte line:
//   dword 0-2  p.xyz       float    origin of the quantisation grid
//   dword 3    e.x,e.y,e.z int8     per-axis power-of-two scale exponents
//              imask       uint8    bit i set = child slot i is an inner node
//   dword 4    child_base  uint32   index of the first inner child; inner
//                                   children are stored contiguously in slot
//                                   order (child = base + rank of slot i)
//   dword 5    prim_base   uint32   first triangle of the leaf children
//   dword 6-7  meta[8]     uint8    leaf slot: (count << 5) | offset,
//                                   count 1..7 triangles at prim_base+offset;
//                                   inner slot: 1; empty slot: 0
//   dword 8-31 q[6][8]     binary16 child boxes, planes lo.x lo.y lo.z hi.x
//                                   hi.y hi.z, each 8 slots, holding integers
//                                   0..2047 (exact in binary16, so the
//                                   mixed-precision fma reads them without a
//                                   conversion); decoded bound = p + q * 2^e
//                                   (conservative by construction); an empty
//                                   slot has lo = 2047, hi = 0 and meta 0.
// Inner children sit in octant slots: bit k of the slot set = the child lies
// on the high side of axis k (as nearly as the builder can arrange).
// Triangle record: 48 bytes = three float4 {v0.xyz, prim}, {v1.xyz, 0},
// {v2.xyz, 0}; `prim` is the caller-visible primitive index (bit pattern).
#pragma once
#include <stdint.h>

#define KP_NODE_BYTES 128u
#define KP_TRI_BYTES 48u
#define KP_LEAF_FLAG 0x80000000u
#define KP_INVALID 0xFFFFFFFFu
#define KP_MAX_LEAF 7u

struct alignas(16) KpNode8 {
    float px, py, pz;
    int8_t ex, ey, ez;
    uint8_t imask;
    uint32_t child_base;
    uint32_t prim_base;
    uint8_t meta[8];
    uint16_t q[6][8];
};
static_assert(sizeof(KpNode8) == KP_NODE_BYTES, "node must be 128 bytes");

// A leaf reference packs a triangle start (28 bits) and count-1 (3 bits).
static inline uint32_t kp_leaf_ref(uint32_t first, uint32_t count) {
    return KP_LEAF_FLAG | (first << 3) | (count - 1u);
}
