// BVH8 traversal kernels of the HIP intersector (gfx950, wave64).
//
// k_closest  : closest hit, replaces EmbreeAccel::intersect
//              (reference src/core/accel.cpp:348-407; interface
//              HitableInterface::intersect, src/core/hitable.h:12-13)
// k_occluded : any hit in [tmin, tmax], replaces HitableInterface::occluded
//              (src/core/hitable.h:15-33)
//
// Design:
//  - persistent lanes: every lane owns one ray at a time; finished lanes are
//    refilled from a global ray counter, one atomic per wave (ballot +
//    prefix count), so incoherent rays do not leave idle lanes behind;
//  - a per-lane short stack in LDS (kStack group entries); deeper entries
//    spill to a per-lane region in HBM, so any BVH depth works;
//  - 128-byte nodes: eight 16-byte loads per visit (each lane loads its own
//    near and far planes by the ray's octant), 16-bit child boxes decoded
//    with one fma per plane; hit leaves are tested at once, the nearest hit
//    inner child is visited next and the others go on the stack as one group,
//    visited in octant order (children sit in octant slots) and culled as a
//    group by their nearest entry distance;
//  - watertight ray/triangle test (Woop, Benthin, Wald 2013) with contraction
//    disabled, so its result is bit-identical to the host oracle's;
//  - closest hit is the smallest t, ties broken by the lowest primitive
//    index, which makes the answer independent of traversal order.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kp_kernels.h"
#include "kp_layout.h"

#pragma clang fp contract(off)

namespace kp {

// Scratch of the allocation-free closest-hit entry point
// (kp_intersect_closest_noalloc): work counter at byte 0, error word at byte
// 256, per-lane stack spill area from byte KP_SCRATCH_SPILL_OFFSET.  Each
// device holds its own copy for the life of the process; host code must use
// the address noalloc_scratch() returns for the device, never this name.
static __device__ __align__(256) unsigned char scratch[KP_SCRATCH_BYTES];

hipError_t noalloc_scratch(int device, void** ptr) {
    static std::mutex mu;
    static void* per_device[KP_SCRATCH_MAX_DEVICES] = {};
    if (device < 0 || device >= KP_SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!per_device[device]) {
        // the caller has made `device` current
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        per_device[device] = p;
    }
    *ptr = per_device[device];
    return hipSuccess;
}


struct RayPre {
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
    float ix, iy, iz;
    float Sx, Sy, Sz;
    int kx, ky, kz;
};

__device__ __forceinline__ float pick3(float a, float b, float c, int k) {
    return k == 0 ? a : (k == 1 ? b : c);
}

__device__ __forceinline__ float safe_rcp(float d) {
    float dd = fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d;
    return 1.0f / dd;
}

// Rays with a non-finite origin or direction, a zero direction or a NaN
// interval bound miss everything (same rule as the oracle).
__device__ __forceinline__ bool ray_valid(const float4 a, const float4 b) {
    const bool fin = isfinite(a.x) && isfinite(a.y) && isfinite(a.z) && isfinite(b.x) &&
                     isfinite(b.y) && isfinite(b.z);
    const bool nz = b.x != 0.0f || b.y != 0.0f || b.z != 0.0f;
    return fin && nz && !isnan(a.w) && !isnan(b.w);
}

__device__ __forceinline__ void setup_ray(const float4 a, const float4 b, RayPre& r) {
    r.ox = a.x; r.oy = a.y; r.oz = a.z; r.tmin = a.w;
    r.dx = b.x; r.dy = b.y; r.dz = b.z; r.tmax = b.w;
    const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = (kz == 2) ? 0 : kz + 1;
    int ky = (kx == 2) ? 0 : kx + 1;
    const float dkz = pick3(r.dx, r.dy, r.dz, kz);
    if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = pick3(r.dx, r.dy, r.dz, kx) / dkz;
    r.Sy = pick3(r.dx, r.dy, r.dz, ky) / dkz;
    r.Sz = 1.0f / dkz;
    r.ix = safe_rcp(r.dx);
    r.iy = safe_rcp(r.dy);
    r.iz = safe_rcp(r.dz);
}

// Watertight test; same operation order as the host oracle.
__device__ __forceinline__ bool tri_test(const RayPre& r, const float4 a, const float4 b,
                                         const float4 c, float tmax_incl, float& t_out,
                                         float& u_out, float& v_out) {
    const float A0 = a.x - r.ox, A1 = a.y - r.oy, A2 = a.z - r.oz;
    const float B0 = b.x - r.ox, B1 = b.y - r.oy, B2 = b.z - r.oz;
    const float C0 = c.x - r.ox, C1 = c.y - r.oy, C2 = c.z - r.oz;
    const float Akz = pick3(A0, A1, A2, r.kz), Bkz = pick3(B0, B1, B2, r.kz), Ckz = pick3(C0, C1, C2, r.kz);
    const float Ax = pick3(A0, A1, A2, r.kx) - r.Sx * Akz;
    const float Ay = pick3(A0, A1, A2, r.ky) - r.Sy * Akz;
    const float Bx = pick3(B0, B1, B2, r.kx) - r.Sx * Bkz;
    const float By = pick3(B0, B1, B2, r.ky) - r.Sy * Bkz;
    const float Cx = pick3(C0, C1, C2, r.kx) - r.Sx * Ckz;
    const float Cy = pick3(C0, C1, C2, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        const double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        const double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        const double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float t = T / det;
    if (!(t >= r.tmin && t <= tmax_incl)) return false;
    t_out = t;
    u_out = V / det;
    v_out = W / det;
    return true;
}


// Per-ray node-test constants: the ray direction's octant (bit k set when
// direction k is negative) and the byte offset in a node of each near / far
// plane word (planes lo.x lo.y lo.z hi.x hi.y hi.z are the 16-byte words
// 2..7), one byte each, so each lane loads its own near and far planes
// instead of selecting them per child.
struct NodeTest {
    uint32_t oct;
    uint32_t offa;   // bytes: near x, far x, near y, far y
    uint32_t offb;   // bytes: near z, far z
};

__device__ __forceinline__ NodeTest node_test_setup(const RayPre& r) {
    NodeTest t;
    const bool nx = r.ix < 0.0f, ny = r.iy < 0.0f, nz = r.iz < 0.0f;
    t.oct = (nx ? 1u : 0u) | (ny ? 2u : 0u) | (nz ? 4u : 0u);
    t.offa = (nx ? 0x2050u : 0x5020u) | ((ny ? 0x3060u : 0x6030u) << 16);
    t.offb = nz ? 0x4070u : 0x7040u;
    return t;
}

// The node array as a buffer resource: a node visit addresses its words by
// 32-bit byte offsets from one scalar base (no 64-bit address arithmetic per
// load); the range is left open, as every node index is valid.
typedef __amdgpu_buffer_rsrc_t NodeBuf;

__device__ __forceinline__ NodeBuf node_buffer(const uint4* nodes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)nodes, (short)0, (int)0xFFFFFFFF, 0x00020000);
}

__device__ __forceinline__ uint4 node_word(NodeBuf nb, uint32_t byte_off) {
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(nb, byte_off, 0, 0));
}

// Plane value of a slot: a binary16 integer read straight from its half of
// the word (the fma below becomes one mixed-precision fma, no conversion).
__device__ __forceinline__ float qplane(const uint4& plane, int slot) {
    const uint32_t d = (slot >> 1) == 0 ? plane.x : (slot >> 1) == 1 ? plane.y : (slot >> 1) == 2 ? plane.z : plane.w;
    const unsigned short h = (slot & 1) ? (unsigned short)(d >> 16) : (unsigned short)(d & 0xFFFFu);
    return (float)__builtin_bit_cast(_Float16, h);
}

// Relative widening of a far distance, 2*gamma(3) (Ize 2013): tf + |tf| * eps.
#define KP_FAR_EPS 3.5762793e-7f

// Result of one node visit.
struct NodeHits {
    uint4 w0, w1;        // header words (origin / exponents / inner mask, bases / meta)
    uint32_t inner;      // inner children hit (slot bits)
    uint32_t leaves;     // leaf children hit (slot bits)
    uint32_t near_key;   // (tnear key | slot) of the nearest inner child hit
    uint32_t next_key;   // the same for the second nearest (0xFFFFFFFF if none)
};

// Order key of a child entry distance: non-negative distances by their bits
// (monotone), negative ones (sign bit set: negative as integers) all first.
__device__ __forceinline__ uint32_t tkey(float tn) {
    return (uint32_t)max((int32_t)__float_as_uint(tn), 0) & 0xFFFFFFF8u;
}

// Tests the 8 children of node `idx` against [tmin, tmax].
__device__ __forceinline__ void visit_node(NodeBuf nb, uint32_t idx, const RayPre& r, const NodeTest& nt,
                                           float tmax, NodeHits& h) {
    const uint32_t p = idx << 7;   // 128-byte nodes
    h.w0 = node_word(nb, p);
    h.w1 = node_word(nb, p + 16u);
    const uint4 nx = node_word(nb, p + (nt.offa & 0xFFu)), fx = node_word(nb, p + ((nt.offa >> 8) & 0xFFu));
    const uint4 ny = node_word(nb, p + ((nt.offa >> 16) & 0xFFu)), fy = node_word(nb, p + (nt.offa >> 24));
    const uint4 nz = node_word(nb, p + (nt.offb & 0xFFu)), fz = node_word(nb, p + (nt.offb >> 8));
    const float px = __uint_as_float(h.w0.x), py = __uint_as_float(h.w0.y), pz = __uint_as_float(h.w0.z);
    const uint32_t e = h.w0.w;
    const int ex = (int)(int8_t)(e & 0xFF), ey = (int)(int8_t)((e >> 8) & 0xFF), ez = (int)(int8_t)((e >> 16) & 0xFF);
    // ix 2^ex: a product with a power of two, exact as the multiplication
    const float ax = ldexpf(r.ix, ex), ay = ldexpf(r.iy, ey), az = ldexpf(r.iz, ez);
    const float bx = (px - r.ox) * r.ix, by = (py - r.oy) * r.iy, bz = (pz - r.oz) * r.iz;
    const uint32_t imask = e >> 24;
    uint32_t hit = 0u, k1 = 0xFFFFFFFFu, k2 = 0xFFFFFFFFu;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const float tnx = fmaf(qplane(nx, s), ax, bx);
        const float tfx = fmaf(qplane(fx, s), ax, bx);
        const float tny = fmaf(qplane(ny, s), ay, by);
        const float tfy = fmaf(qplane(fy, s), ay, by);
        const float tnz = fmaf(qplane(nz, s), az, bz);
        const float tfz = fmaf(qplane(fz, s), az, bz);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tmin));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        const bool hs = tn <= fmaf(fabsf(tf), KP_FAR_EPS, tf);
        hit |= hs ? (1u << s) : 0u;
        // key of a hit inner child: entry distance bits (non-positive -> 0), slot
        const uint32_t key = (hs && ((imask >> s) & 1u)) ? (tkey(tn) | (uint32_t)s) : 0xFFFFFFFFu;
        k2 = min(k2, max(k1, key));
        k1 = min(k1, key);
    }
    // Empty slots hold inverted boxes (lower plane 2047, upper plane 0), which
    // miss unless the far-distance widening exceeds the node's extent; such a
    // slot then counts as a leaf child with no triangles and tests nothing.
    h.inner = hit & imask;
    h.leaves = hit & ~imask;
    h.near_key = k1;
    h.next_key = k2;
}

__device__ __forceinline__ uint32_t inner_child(uint32_t child_base, uint32_t imask, uint32_t slot) {
    return child_base + __popc(imask & ((1u << slot) - 1u));
}

// Slot bits reordered by rank = slot ^ oct (an involution).
__device__ __forceinline__ uint32_t oct_permute(uint32_t m, uint32_t oct) {
    if (oct & 1u) m = ((m & 0x55u) << 1) | ((m >> 1) & 0x55u);
    if (oct & 2u) m = ((m & 0x33u) << 2) | ((m >> 2) & 0x33u);
    if (oct & 4u) m = ((m & 0x0Fu) << 4) | ((m >> 4) & 0x0Fu);
    return m;
}

// Stack entry (a group of inner children of one node still to visit):
//   x = child_base (24 bits) | inner mask << 24
//   y = lower bound of the group's entry distances (float bits, low 8 bits
//       cleared: a truncation toward zero, or -inf) | remaining children as
//       octant-rank bits (8 bits).
__device__ __forceinline__ uint32_t group_bound_bits(uint32_t key2) {
    if (key2 == 0xFFFFFFFFu || key2 < 8u) return 0xFF800000u;   // none, or a negative distance: -inf
    return key2 & 0xFFFFFF00u;
}

// Entry distance of one child slot of node `idx` (the same arithmetic as
// visit_node), for a top-level leaf child that names an instance.
__device__ __forceinline__ float slot_tnear(const uint4* __restrict__ nodes, uint32_t idx, uint32_t slot,
                                            const RayPre& r, const NodeTest& nt) {
    const uint4* p = nodes + (size_t)idx * 8;
    const uint4 w0 = p[0];
    const uint4 nx = p[(nt.offa & 0xFFu) >> 4], ny = p[((nt.offa >> 16) & 0xFFu) >> 4], nz = p[(nt.offb & 0xFFu) >> 4];
    const uint32_t e = w0.w;
    const int ex = (int)(int8_t)(e & 0xFF), ey = (int)(int8_t)((e >> 8) & 0xFF), ez = (int)(int8_t)((e >> 16) & 0xFF);
    const float ax = ldexpf(r.ix, ex), ay = ldexpf(r.iy, ey), az = ldexpf(r.iz, ez);
    const float bx = (__uint_as_float(w0.x) - r.ox) * r.ix, by = (__uint_as_float(w0.y) - r.oy) * r.iy,
                bz = (__uint_as_float(w0.z) - r.oz) * r.iz;
    const int s = (int)slot;
    const float tnx = fmaf(qplane(nx, s), ax, bx);
    const float tny = fmaf(qplane(ny, s), ay, by);
    const float tnz = fmaf(qplane(nz, s), az, bz);
    return fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tmin));
}

// World ray to an instance's object space: o' = M o + m, d' = M d with the
// row-major 3x4 world-to-object matrix in inst[0..2] (FP32, no contraction,
// the same order as the oracle); the interval [tmin, tmax] is unchanged.
__device__ __forceinline__ void ray_to_object(const float4* __restrict__ inst, const float4 a, const float4 b,
                                              float4& oa, float4& ob) {
    const float4 m0 = inst[0], m1 = inst[1], m2 = inst[2];
    oa.x = ((m0.x * a.x + m0.y * a.y) + m0.z * a.z) + m0.w;
    oa.y = ((m1.x * a.x + m1.y * a.y) + m1.z * a.z) + m1.w;
    oa.z = ((m2.x * a.x + m2.y * a.y) + m2.z * a.z) + m2.w;
    oa.w = a.w;
    ob.x = (m0.x * b.x + m0.y * b.y) + m0.z * b.z;
    ob.y = (m1.x * b.x + m1.y * b.y) + m1.z * b.z;
    ob.z = (m2.x * b.x + m2.y * b.y) + m2.z * b.z;
    ob.w = b.w;
}

// Special stack entries of the two-level traversal (low byte of y zero; a
// group entry always has remaining-child bits there):
//   instance entry  x = instance record, y = bound of its box entry distance
//   return marker   x = kReturnMarker, y = -inf: back to the world-space ray
//   same-space end  x = kReturnSame, y = -inf: an identity placement ends
//   top level       x = kEnterTop, y = -inf: the meshes are done, go to the
//                   top level's root (same space and primitive base)
constexpr uint32_t kReturnMarker = 0xFFFFFFFFu;
constexpr uint32_t kReturnSame = 0xFFFFFFFEu;
constexpr uint32_t kEnterTop = 0xFFFFFFFDu;
// Instance record flag (bits of the third word of its fourth float4): the
// world-to-object matrix is exactly the identity, so the ray is used as is.
constexpr uint32_t kInstIdentity = 1u;

template <int S>
struct Stack {
    uint2 (*lds)[KP_TRACE_BLOCK];
    uint2* spill;
    uint32_t cap;
    int sp;
    int tid;

    __device__ __forceinline__ bool push(uint32_t a, uint32_t b) {
        if (sp >= S) {
            if ((uint32_t)(sp - S) >= cap) return false;
            spill[sp - S] = lds[sp & (S - 1)][tid];
        }
        lds[sp & (S - 1)][tid] = make_uint2(a, b);
        ++sp;
        return true;
    }
    __device__ __forceinline__ uint2 top() const { return lds[(sp - 1) & (S - 1)][tid]; }
    __device__ __forceinline__ void set_top_y(uint32_t y) { lds[(sp - 1) & (S - 1)][tid].y = y; }
    __device__ __forceinline__ void drop() {
        --sp;
        if (sp >= S) lds[sp & (S - 1)][tid] = spill[sp - S];
    }
};

constexpr int kStack = KP_TRACE_STACK;

// Work distribution of the persistent kernels.  A wave takes chunks of
// consecutive rays: a first chunk by its index (about half the rays split
// evenly over the waves), then chunks of kChunk from the global counter (one
// atomic per chunk; kTailChunk once fewer than one kChunk per wave remained
// after the wave's last chunk, so the waves run dry at about the same time),
// and hands them to its lanes from a wave-uniform cursor; every lane keeps
// its next ray prefetched in registers, so replacing a finished ray costs no
// memory round trip.  All members except the prefetch are wave-uniform.
constexpr uint32_t kChunk = 256;
constexpr uint32_t kTailChunk = 64;

struct RayFeeder {
    uint32_t next, end, n;
    uint32_t dyn_base;  // first ray handed out by the counter (after the static chunks)
    bool drained;
    uint32_t* counter;
    const float4* rays;
    const uint32_t* roots;
    uint32_t pid, proot;   // prefetched ray of this lane (pid KP_INVALID: none) and its root
    float4 p0, p1;

    __device__ __forceinline__ void init(uint32_t* ctr, const float4* rs, const uint32_t* rts, uint32_t count) {
        counter = ctr;
        rays = rs;
        roots = rts;
        proot = 0u;
        n = count;
        pid = KP_INVALID;
        // Wave w's first chunk is rays [w F, (w+1) F) without an atomic
        // (every wave of the grid would otherwise hit the one counter address
        // at start); the counter hands out the rays after all first chunks,
        // and a wave whose first chunk lies beyond the rays has none.
        const uint32_t waves = gridDim.x * (blockDim.x >> 6);
        const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
        // up to kChunk rays per wave: every ray in a first chunk (multiples of
        // 64, as many waves as that takes); more: the first chunks hold
        // about half the rays (256 .. 4096 per wave) and the counter's chunks
        // balance the rest
        uint32_t first;
        if (n <= waves * kChunk) {
            first = ((n + waves - 1u) / waves + 63u) & ~63u;
            if (first == 0u) first = 64u;
        } else {
            first = (n / (2u * waves)) & ~63u;
            first = first < kChunk ? kChunk : (first > 4096u ? 4096u : first);
        }
        dyn_base = waves * first;
        next = wave * first;
        if (next < n) {
            end = next + first < n ? next + first : n;
            drained = false;
        } else {
            next = end = 0;
            drained = true;
        }
        assign();
    }
    // Gives a prefetched ray to every lane without one, as far as rays last.
    __device__ __forceinline__ void assign() {
        const uint64_t m = __ballot(pid == KP_INVALID);
        if (m == 0 || drained) return;
        if (next >= end) {
            if (dyn_base >= n) {   // the first chunks covered every ray
                drained = true;
                return;
            }
            // the last chunk taken ended at `end`: small chunks once fewer
            // than kChunk rays per wave were left after it
            const uint64_t late = (uint64_t)(gridDim.x * (blockDim.x >> 6)) * kChunk;
            const uint32_t size = (uint64_t)end + late >= n ? kTailChunk : kChunk;
            uint32_t base = 0;
            if (__lane_id() == 0) base = atomicAdd(counter, size);
            base = dyn_base + __builtin_amdgcn_readfirstlane(base);
            if (base >= n) {
                drained = true;
                return;
            }
            next = base;
            end = base + size < n ? base + size : n;
        }
        const uint32_t avail = end - next;
        const uint32_t want = (uint32_t)__popcll(m);
        const uint32_t prefix = (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
        if (pid == KP_INVALID && prefix < avail) {
            pid = next + prefix;
            p0 = rays[2 * (size_t)pid];
            p1 = rays[2 * (size_t)pid + 1];
            if (roots) proot = roots[pid];
        }
        next += want < avail ? want : avail;
    }
};

// Lanes holding postponed leaf children before the wave tests them together.
constexpr int kLeafBatch = 16;
// Lanes with nothing but postponed leaves left (they wait) before the wave
// tests them; a wave test has a fixed cost of about 150 instructions and an
// idle lane costs a 64th of an iteration, so waiting for a fuller batch pays.
constexpr uint32_t kStuckBatch = 8u;

// LDS of the wave-wide leaf test (per block): the item map of each wave and,
// per lane, the merged result of the triangles tested for its ray.
struct LeafLds {
    uint32_t map[KP_TRACE_BLOCK];             // wave's 64 entries: owner lane << 8 | item of the owner
    unsigned long long key[KP_TRACE_BLOCK];   // closest hit: minimum hit key; any hit: nonzero once hit
    float4 hit[KP_TRACE_BLOCK];               // closest hit: {t, u, v} of the minimum key
};

__device__ __forceinline__ unsigned long long hit_key(float t, uint32_t prim);

// Tests the postponed leaf sets of all active lanes of the wave together.
// Each lane holds up to two sets (leaf slots `pend` of a node with primitive
// base `pbase` and meta words m0 / m1: per slot first offset and count; the
// second set is the one found in the current iteration); a lane's sets are
// one run of triangle tests (at most 2 x 8 slots x 4 = 64); the runs are laid
// out as one list (exclusive prefix sums over the lanes) and dealt out to the
// active lanes, as many as there are per round, so the triangle tests run with nearly every lane busy instead of
// one loop per lane at the pace of its longest set.  A tester reads the
// owner's ray and closest distance from the owner's registers (lane
// permutes) and merges a hit into the owner's LDS slot: the 64-bit minimum of
// the (distance, primitive) key, then the t, u, v of the lane whose key is
// the minimum.  Every triangle is tested against the owner's closest distance
// at the start of the batch, so the owner's new closest hit is the minimum
// key over its hits, the same as testing them one after another.
__device__ __forceinline__ uint32_t leaf_set_tris(uint32_t pend, uint32_t m0, uint32_t m1) {
    // slot bits of pend spread to byte masks (bit k -> byte k), then a byte sum
    const uint32_t b0 = (((pend & 0xFu) * 0x00204081u) & 0x01010101u) * 0xFFu;
    const uint32_t b1 = ((((pend >> 4) & 0xFu) * 0x00204081u) & 0x01010101u) * 0xFFu;
    return __builtin_amdgcn_sad_u8((m0 >> 5) & 0x07070707u & b0, 0u,
                                   __builtin_amdgcn_sad_u8((m1 >> 5) & 0x07070707u & b1, 0u, 0u));
}

template <bool kAny, bool kCount>
__device__ __forceinline__ void test_leaves_wave(const float4* __restrict__ tris, const RayPre& r, uint32_t pend,
                                                 uint32_t pbase, uint32_t m0, uint32_t m1, uint32_t pend2,
                                                 uint32_t pbase2, uint32_t m02, uint32_t m12, float& best_t,
                                                 uint32_t& best_p, float& bu, float& bv, bool& occl,
                                                 uint64_t& cnt_tris, uint32_t prim_off, LeafLds& L, int tid) {
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long active = __ballot(true);
    const uint32_t nact = (uint32_t)__popcll(active);
    const uint32_t rank = (uint32_t)__popcll(active & below);
    // triangles of this lane's sets: the counts (bits 5..7 of each meta byte) of their slots
    uint32_t T1 = 0u, T = 0u;
    if (!(kAny && occl)) {
        T1 = leaf_set_tris(pend, m0, m1);
        T = T1 + leaf_set_tris(pend2, m02, m12);
    }
    uint32_t E = 0u, P = 0u;   // exclusive prefix and total over the wave (T <= 64)
#pragma unroll
    for (int bit = 0; bit < 7; ++bit) {
        const unsigned long long m = __ballot((T >> bit) & 1u);
        E += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << bit;
        P += (uint32_t)__popcll(m) << bit;
    }
    if (P == 0u) return;
    const int wbase = tid & ~63;
    uint32_t* map = &L.map[wbase];
    if (T != 0u) L.key[tid] = kAny ? 0ull : ~0ull;
    for (uint32_t base = 0u; base < P; base += nact) {
        if (T != 0u) {
            const uint32_t lo = max(E, base), hi = min(E + T, base + nact);
            for (uint32_t i = lo; i < hi; ++i) map[i - base] = (lane << 8) | (i - E);
        }
        __builtin_amdgcn_wave_barrier();
        const bool has = base + rank < P;
        const uint32_t me = has ? map[rank] : (lane << 8);
        const int ow = (int)(me >> 8);
        const uint32_t o_t1 = __shfl(T1, ow);
        const uint32_t o_pend = __shfl(pend, ow), o_base = __shfl(pbase, ow);
        const uint32_t o_m0 = __shfl(m0, ow), o_m1 = __shfl(m1, ow);
        const uint32_t o_pend2 = __shfl(pend2, ow), o_base2 = __shfl(pbase2, ow);
        const uint32_t o_m02 = __shfl(m02, ow), o_m12 = __shfl(m12, ow);
        RayPre o;
        o.ox = __shfl(r.ox, ow);
        o.oy = __shfl(r.oy, ow);
        o.oz = __shfl(r.oz, ow);
        o.tmin = __shfl(r.tmin, ow);
        o.Sx = __shfl(r.Sx, ow);
        o.Sy = __shfl(r.Sy, ow);
        o.Sz = __shfl(r.Sz, ow);
        o.kx = __shfl(r.kx, ow);
        o.ky = __shfl(r.ky, ow);
        o.kz = __shfl(r.kz, ow);
        const float o_best = __shfl(best_t, ow);
        const uint32_t o_off = __shfl(prim_off, ow);
        if (has) {
            // the item's triangle: slots in increasing order, triangles in order
            uint32_t q = me & 0xFFu, tri = 0u;
            const bool second = q >= o_t1;
            uint32_t lm = second ? o_pend2 : o_pend;
            const uint32_t sb = second ? o_base2 : o_base, sm0 = second ? o_m02 : o_m0, sm1 = second ? o_m12 : o_m1;
            if (second) q -= o_t1;
            while (lm) {
                const uint32_t s = __ffs(lm) - 1u;
                lm &= lm - 1u;
                const uint32_t meta = ((s < 4u ? sm0 : sm1) >> ((s & 3u) * 8u)) & 0xFFu;
                const uint32_t c = meta >> 5;
                if (q < c) {
                    tri = sb + (meta & 31u) + q;
                    break;
                }
                q -= c;
            }
            if (kCount) cnt_tris += 1u;
            const float4* tp = tris + 3 * (size_t)tri;
            const float4 ta = tp[0], tb = tp[1], tc = tp[2];
            float t, u, v;
            if (tri_test(o, ta, tb, tc, o_best, t, u, v)) {
                const int ot = wbase + ow;
                if (kAny) {
                    L.key[ot] = 1ull;
                } else {
                    const unsigned long long hk = hit_key(t, __float_as_uint(ta.w) + o_off);
                    __hip_atomic_fetch_min(&L.key[ot], hk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    __builtin_amdgcn_wave_barrier();
                    if (__hip_atomic_load(&L.key[ot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) == hk)
                        L.hit[ot] = make_float4(t, u, v, 0.0f);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (T != 0u) {
        const unsigned long long k = __hip_atomic_load(&L.key[tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        if (kAny) {
            if (k != 0ull) occl = true;
        } else if (k != ~0ull && (best_p == KP_INVALID || k < hit_key(best_t, best_p))) {
            const float4 h = L.hit[tid];
            best_t = h.x;
            best_p = (uint32_t)k;
            bu = h.y;
            bv = h.z;
        }
    }
}

// Order key of a closest-hit candidate: distance first (float bits mapped to
// unsigned order, -0 taken as +0 as the comparisons do), then primitive, so
// the minimum key is the closest hit with the lowest primitive among equal
// distances (the tie rule).
__device__ __forceinline__ unsigned long long hit_key(float t, uint32_t prim) {
    uint32_t b = __float_as_uint(t);
    if (b == 0x80000000u) b = 0u;
    b ^= (b & 0x80000000u) ? 0xFFFFFFFFu : 0x80000000u;
    return ((unsigned long long)b << 32) | prim;
}

// Position of the set bit of rank r (0-based) in m; r < popcount(m).
__device__ __forceinline__ uint32_t nth_set_bit(unsigned long long m, uint32_t r) {
    uint32_t pos = 0u;
    uint32_t c = __popc((uint32_t)m);
    if (r >= c) { r -= c; m >>= 32; pos += 32u; }
    c = __popc((uint32_t)m & 0xFFFFu);
    if (r >= c) { r -= c; m >>= 16; pos += 16u; }
    c = __popc((uint32_t)m & 0xFFu);
    if (r >= c) { r -= c; m >>= 8; pos += 8u; }
    c = __popc((uint32_t)m & 0xFu);
    if (r >= c) { r -= c; m >>= 4; pos += 4u; }
    c = __popc((uint32_t)m & 0x3u);
    if (r >= c) { r -= c; m >>= 2; pos += 2u; }
    return pos + (r >= ((uint32_t)m & 1u) ? 1u : 0u);
}

// Lane roles of a split closest-hit ray (low 8 bits: block thread of the
// lane that owns the ray).
constexpr uint32_t kHelper = 1u << 8;    // traverses a group of another lane's ray
constexpr uint32_t kDonated = 1u << 9;   // owner whose ray has had groups handed out
constexpr uint32_t kWaiting = 1u << 10;  // owner done with its part, helpers still at work
constexpr uint32_t kStepsWaiting = 0xFFFFFFFFu;   // step count of a waiting owner (skips the traversal)

// Persistent traversal shared by closest hit (kAny = false) and occlusion
// (kAny = true).  A node visit tests the 8 children; the nearest hit inner
// child is visited next and the other hit inner children are pushed as one
// group entry, taken later in octant order (the builder puts children in
// octant slots) and skipped as a whole when their nearest entry distance lies
// beyond the closest hit so far.  Hit leaf children are postponed (one set
// per lane) and tested wave-wide (test_leaves_wave) when at least kLeafBatch
// lanes of the wave hold some, when kStuckBatch lanes (or every lane in the
// loop) have nothing else left, or when a lane finds a second set, so the
// triangle tests run with many lanes active instead of one or two.
//
// Two-level variant (kInst): a ray first traverses the BVH of the meshes
// (world_root, one-level, in world space) with a top-level marker below it on
// the stack; nodes below args.n_tlas form the top level, whose leaf children
// are instances.  Each lane keeps its world-space ray in LDS, read back on
// entering and leaving an instance.  A hit instance is pushed as an instance
// entry (culled later like a group by its entry distance); taking it pushes a
// return marker, moves the ray to the instance's object space and continues
// at the instance's root; reaching the marker restores the world-space ray
// (refetched from the ray buffer).  An identity placement keeps the ray and
// only switches the primitive base.  Postponed leaf children are tested
// before any of these switches, so a set is always tested in the space and
// with the primitive base it was found in.
//
// Split rays (closest hit, one level): once the rays are handed out, lanes
// whose rays are done would idle while the wave finishes its longest rays.
// Such a lane takes the top group entry of a lane still at work and
// traverses that group for the same ray (reloaded from the ray buffer) with
// the donor's closest distance so far; the groups of a ray are disjoint
// subtrees, so every triangle is still tested by exactly one lane.  A helper
// merges its closest hit into its owner's slot in LDS (64-bit minimum of the
// (distance, primitive) key, then the t, u, v of the winning key); the owner
// writes the ray's hit once no helper of its ray is left.  The result is the
// minimum key over all triangles the ray hits in [tmin, tmax], the same as
// without splitting.  A helper also culls with its owner's closest distance
// once that is nearer (read at every pass of the drain phase).  Occlusion
// does not split: split occlusion rays made the integrator 2 % slower.
template <bool kAny, bool kCount, bool kInst>
__device__ __forceinline__ void traverse(const TraceArgs& args) {
    __shared__ uint2 lds[kStack][KP_TRACE_BLOCK];
    __shared__ LeafLds leafl;
    __shared__ float4 wray[kInst ? 2 : 1][kInst ? KP_TRACE_BLOCK : 1];   // world ray per lane (two-level)
    constexpr bool kSplit = !kAny && !kInst;
    __shared__ unsigned long long skey[kSplit ? KP_TRACE_BLOCK : 1];      // owner: minimum helper hit key
    __shared__ float4 shit[kSplit ? KP_TRACE_BLOCK : 1];                  // owner: {t, u, v} of that key
    __shared__ uint32_t srole[kSplit ? KP_TRACE_BLOCK : 1];               // owner thread | kHelper | kDonated | kWaiting
    __shared__ unsigned long long sbusy[kSplit ? KP_TRACE_BLOCK / 64 : 1];  // wave: owners with helpers at work
    const int tid = threadIdx.x;
    const uint32_t gl = blockIdx.x * KP_TRACE_BLOCK + tid;
    Stack<kStack> st;
    st.lds = lds;
    st.spill = args.spill + (size_t)gl * args.spill_cap;
    st.cap = args.spill_cap;
    st.sp = 0;
    st.tid = tid;
    const uint4* __restrict__ nodes = args.nodes;
    const NodeBuf nbuf = node_buffer(nodes);
    const uint32_t n_rays = args.n_dev ? *args.n_dev : args.n;   // device-side count for queues
    const float4* __restrict__ tris = args.tris;

    uint32_t ray_id = KP_INVALID;
    RayPre r;
    NodeTest nt;
    float best_t = 0.0f, bu = 0.0f, bv = 0.0f;
    uint32_t best_p = KP_INVALID, cur = KP_INVALID, steps = 0;
    uint32_t pend = 0u, pbase = 0u, pm0 = 0u, pm1 = 0u;   // postponed leaf children
    uint32_t inst_pbase = 0u;                              // primitive base of the instance entered
    bool occl = false;
    bool split_on = false;   // rays are being split (wave-uniform; roles live in srole)
    uint64_t cnt_nodes = 0, cnt_tris = 0;
    RayFeeder feed;
    feed.init(args.counter, args.rays, args.roots, n_rays);

    for (;;) {
        if (ray_id == KP_INVALID && feed.pid != KP_INVALID) {
            ray_id = feed.pid;
            setup_ray(feed.p0, feed.p1, r);
            nt = node_test_setup(r);
            best_t = r.tmax;
            best_p = KP_INVALID;
            bu = bv = 0.0f;
            occl = false;
            pend = 0u;
            inst_pbase = 0u;
            cur = ray_valid(feed.p0, feed.p1) ? feed.proot : KP_INVALID;   // root
            st.sp = 0;
            steps = 0;
            if (kSplit && split_on) {
                srole[tid] = (uint32_t)tid;
                skey[tid] = ~0ull;
            }
            feed.pid = KP_INVALID;
            if (kInst) {
                wray[0][tid] = feed.p0;
                wray[kInst ? 1 : 0][tid] = feed.p1;
                if (cur != KP_INVALID && args.world_root != KP_INVALID) {
                    st.push(kEnterTop, 0xFF800000u);   // the top level, after the meshes
                    cur = args.world_root;
                }
            }
        }
        feed.assign();
        if (kSplit && feed.drained) {
            const uint32_t lane = __lane_id();
            if (!split_on) {   // every ray in flight is its own owner
                split_on = true;
                srole[tid] = (uint32_t)tid;
                skey[tid] = ~0ull;
            }
            uint32_t role = srole[tid];
            unsigned long long* busy_w = &sbusy[tid >> 6];
            if (__ballot(ray_id != KP_INVALID && (role & kWaiting)) != 0ull) {
                // owners whose helpers are all done write their ray's hit (the
                // lanes of a wave reach LDS in instruction order)
                if (lane == 0u) __hip_atomic_store(busy_w, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                __builtin_amdgcn_wave_barrier();
                if (ray_id != KP_INVALID && (role & kHelper))
                    __hip_atomic_fetch_or(busy_w, 1ull << (role & 63u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                __builtin_amdgcn_wave_barrier();
                const unsigned long long busy = __hip_atomic_load(busy_w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                if (ray_id != KP_INVALID && (role & kWaiting) && !((busy >> lane) & 1ull)) {
                    const unsigned long long k = skey[tid];
                    if (k != ~0ull && (best_p == KP_INVALID || k < hit_key(best_t, best_p))) {
                        const float4 h = shit[tid];
                        best_t = h.x;
                        best_p = (uint32_t)k;
                        bu = h.y;
                        bv = h.z;
                    }
                    float4 hh;
                    if (best_p == KP_INVALID) hh = make_float4(__int_as_float(0x7f800000), __uint_as_float(KP_INVALID), 0.0f, 0.0f);
                    else hh = make_float4(best_t, __uint_as_float(best_p), bu, bv);
                    args.hits[ray_id] = hh;
                    ray_id = KP_INVALID;
                }
            }
            {
                // a helper culls with its owner's closest distance once that
                // is nearer; its own hit, farther than the owner's, cannot win
                const float ob = __shfl(best_t, (int)(role & 63u));
                if (ray_id != KP_INVALID && (role & kHelper) && ob < best_t) {
                    best_t = ob;
                    best_p = KP_INVALID;
                }
            }
            // idle lanes take the top group entry of lanes still at work
            const unsigned long long idle = __ballot(ray_id == KP_INVALID && feed.pid == KP_INVALID);
            if (idle != 0ull) {
                const bool can = ray_id != KP_INVALID && !(role & kWaiting) && st.sp > 0;
                uint2 e = make_uint2(0u, 0u);
                if (can) e = st.top();
                const bool donor = can && !(__uint_as_float(e.y & 0xFFFFFF00u) > best_t);
                const unsigned long long donors = __ballot(donor);
                if (donors != 0ull) {
                    const uint32_t pairs = min((uint32_t)__popcll(idle), (uint32_t)__popcll(donors));
                    const unsigned long long below = (1ull << lane) - 1ull;
                    const bool give = donor && (uint32_t)__popcll(donors & below) < pairs;
                    const uint32_t rank = (uint32_t)__popcll(idle & below);
                    const bool take = ((idle >> lane) & 1ull) && rank < pairs;
                    const int src = take ? (int)nth_set_bit(donors, rank) : (int)lane;
                    const uint32_t gx = __shfl(e.x, src), gy = __shfl(e.y, src);
                    const uint32_t rid = __shfl(ray_id, src);
                    const float bt = __shfl(best_t, src);
                    const uint32_t own = __shfl(role & 0xFFu, src);
                    if (give) {
                        st.drop();
                        srole[tid] = role | kDonated;
                    }
                    if (take) {
                        ray_id = rid;
                        srole[tid] = own | kHelper;
                        setup_ray(args.rays[2 * (size_t)rid], args.rays[2 * (size_t)rid + 1], r);
                        nt = node_test_setup(r);
                        best_t = bt;
                        best_p = KP_INVALID;
                        bu = bv = 0.0f;
                        pend = 0u;
                        cur = KP_INVALID;
                        steps = 0;
                        st.sp = 0;
                        st.push(gx, gy);   // the stack is empty: cannot overflow
                    }
                }
            }
        }
        if (__ballot(ray_id != KP_INVALID || feed.pid != KP_INVALID) == 0) break;
        // an owner waiting for its helpers has steps == kStepsWaiting
        if (ray_id != KP_INVALID && steps != kStepsWaiting) {
            for (;;) {
                bool fault = false;
                uint32_t pend2 = 0u, pbase2 = 0u, pm02 = 0u, pm12 = 0u;   // a second leaf set found this iteration
                if (cur != KP_INVALID) {
                    NodeHits h;
                    visit_node(nbuf, cur, r, nt, best_t, h);
                    if (kCount) cnt_nodes++;
                    const uint32_t vis = cur;
                    cur = KP_INVALID;
                    if (kInst && h.leaves && vis < args.n_tlas) {
                        // top-level leaves: push every instance they name
                        uint32_t lm = h.leaves;
                        while (lm && !fault) {
                            const uint32_t s = __ffs(lm) - 1u;
                            lm &= lm - 1u;
                            const uint32_t mw = s < 4u ? h.w1.z : h.w1.w;
                            const uint32_t meta = (mw >> ((s & 3u) * 8u)) & 0xFFu;
                            const uint32_t first = h.w1.y + (meta & 31u), count = meta >> 5;
                            const uint32_t gy = group_bound_bits(tkey(slot_tnear(nodes, vis, s, r, nt)));
                            for (uint32_t i = 0; i < count; ++i) {
                                if (!st.push(args.tlas_inst[first + i], gy)) {
                                    atomicOr(args.err, 1u);
                                    fault = true;
                                    break;
                                }
                            }
                        }
                    } else if (h.leaves) {
                        if (pend == 0u) {
                            pend = h.leaves;
                            pbase = h.w1.y;
                            pm0 = h.w1.z;
                            pm1 = h.w1.w;
                        } else {   // this lane already holds a set: the wave tests it below
                            pend2 = h.leaves;
                            pbase2 = h.w1.y;
                            pm02 = h.w1.z;
                            pm12 = h.w1.w;
                        }
                    }
                    if (h.inner && !fault && !(kAny && occl)) {
                        const uint32_t imask = h.w0.w >> 24;
                        const uint32_t s1 = h.near_key & 7u;
                        const uint32_t rest = h.inner & ~(1u << s1);
                        if (rest) {
                            const uint32_t gx = (h.w1.x & 0x00FFFFFFu) | (imask << 24);
                            const uint32_t gy = group_bound_bits(h.next_key) | oct_permute(rest, nt.oct);
                            if (!st.push(gx, gy)) {
                                atomicOr(args.err, 1u);
                                fault = true;
                            }
                        }
                        cur = inner_child(h.w1.x, imask, s1);
                    }
                }
                if (!fault && cur == KP_INVALID && !(kAny && occl)) {
                    while (st.sp > 0) {
                        const uint2 e = st.top();
                        const float bound = __uint_as_float(e.y & 0xFFFFFF00u);
                        if (!kAny && bound > best_t) {   // every child of the group is beyond the hit
                            st.drop();
                            continue;
                        }
                        if (kInst && (e.y & 0xFFu) == 0u) {
                            if (e.x == kEnterTop) {   // same space and base: leaves may stay postponed
                                st.drop();
                                cur = 0u;
                                break;
                            }
                            // postponed leaves belong to the current space and
                            // base: the lane waits (no node) until the wave tests them
                            if (pend != 0u) break;
                            st.drop();
                            if (e.x == kReturnSame) {
                                inst_pbase = 0u;
                                continue;
                            }
                            if (e.x == kReturnMarker) {   // leave the instance: world-space ray again
                                const float4 wa = wray[0][tid], wb = wray[kInst ? 1 : 0][tid];
                                setup_ray(wa, wb, r);
                                nt = node_test_setup(r);
                                inst_pbase = 0u;
                                continue;
                            }
                            const float4* inst = args.insts + 4 * (size_t)e.x;
                            const float4 rec = inst[3];
                            if (__float_as_uint(rec.z) & kInstIdentity) {
                                if (!st.push(kReturnSame, 0xFF800000u)) {
                                    atomicOr(args.err, 1u);
                                    fault = true;
                                    break;
                                }
                            } else {
                                const float4 wa = wray[0][tid], wb = wray[kInst ? 1 : 0][tid];
                                float4 oa, ob;
                                ray_to_object(inst, wa, wb, oa, ob);
                                if (!ray_valid(oa, ob)) continue;   // degenerate in object space: misses
                                if (!st.push(kReturnMarker, 0xFF800000u)) {
                                    atomicOr(args.err, 1u);
                                    fault = true;
                                    break;
                                }
                                setup_ray(oa, ob, r);
                                nt = node_test_setup(r);
                            }
                            inst_pbase = __float_as_uint(rec.y);
                            cur = __float_as_uint(rec.x);
                            break;
                        }
                        uint32_t pm = e.y & 0xFFu;
                        const uint32_t s = (__ffs(pm) - 1u) ^ nt.oct;
                        pm &= pm - 1u;
                        if (pm) st.set_top_y((e.y & 0xFFFFFF00u) | pm);
                        else st.drop();
                        cur = inner_child(e.x & 0x00FFFFFFu, e.x >> 24, s);
                        break;
                    }
                }
                // postponed leaves: tested together once enough lanes hold some,
                // once kStuckBatch lanes (or every lane here) have nothing but
                // leaves left, or when a lane found a second set (held for
                // this iteration only; the wave tests both sets)
                const bool stuck = pend != 0u && (cur == KP_INVALID || fault);
                const uint32_t n_stuck = (uint32_t)__popcll(__ballot(stuck));
                const uint32_t n_here = (uint32_t)__popcll(__ballot(true));
                if (n_stuck >= min(kStuckBatch, n_here) || __ballot(pend2 != 0u) != 0ull ||
                    __popcll(__ballot(pend != 0u)) >= kLeafBatch) {
                    test_leaves_wave<kAny, kCount>(tris, r, pend, pbase, pm0, pm1, pend2, pbase2, pm02, pm12, best_t,
                                                   best_p, bu, bv, occl, cnt_tris, inst_pbase, leafl, tid);
                    pend = 0u;
                }
                bool done = fault || (kAny && occl) || (cur == KP_INVALID && pend == 0u && st.sp == 0);
                if (!done && ++steps > args.max_steps) {
                    atomicOr(args.err, 2u);
                    done = true;
                }
                if (done) {
                    const uint32_t role = kSplit && split_on ? srole[tid] : 0u;
                    if (kAny) {
                        args.occl[ray_id] = occl ? 1u : 0u;
                    } else if (kSplit && (role & kHelper)) {
                        // merge into the owner's slot: minimum key, then the
                        // t, u, v of the lane whose key is the minimum
                        const uint32_t o = role & 0xFFu;
                        if (best_p != KP_INVALID) {
                            const unsigned long long k = hit_key(best_t, best_p);
                            __hip_atomic_fetch_min(&skey[o], k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                            __builtin_amdgcn_wave_barrier();
                            if (__hip_atomic_load(&skey[o], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) == k)
                                shit[o] = make_float4(best_t, bu, bv, 0.0f);
                        }
                    } else if (kSplit && (role & kDonated)) {
                        srole[tid] = role | kWaiting;   // the hit is written once the helpers are done
                        steps = kStepsWaiting;
                        pend = 0u;
                        break;
                    } else {
                        float4 hh;
                        if (best_p == KP_INVALID) hh = make_float4(__int_as_float(0x7f800000), __uint_as_float(KP_INVALID), 0.0f, 0.0f);
                        else hh = make_float4(best_t, __uint_as_float(best_p), bu, bv);
                        args.hits[ray_id] = hh;
                    }
                    ray_id = KP_INVALID;
                    pend = 0u;
                    break;
                }
                if (__popcll(__ballot(true)) < args.refill_below) break;
            }
        }
    }
    if (kCount) {
        atomicAdd((unsigned long long*)&args.counters[0], (unsigned long long)cnt_nodes);
        atomicAdd((unsigned long long*)&args.counters[1], (unsigned long long)cnt_tris);
    }
}

template <bool kCount>
__global__ __launch_bounds__(KP_TRACE_BLOCK) void k_closest(TraceArgs args) {
    traverse<false, kCount, false>(args);
}

__global__ __launch_bounds__(KP_TRACE_BLOCK) void k_occluded(TraceArgs args) {
    traverse<true, false, false>(args);
}

// The same traversals compiled for 5 waves per SIMD (96 registers; closest
// hit spills one register).  Measured: the integrator's bounces on the 9.4 M
// triangle terrain run 5 % faster, the incoherent microbench 2 % slower, so
// only the integrator launches them.
__global__ __launch_bounds__(KP_TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_closest5(
    TraceArgs args) {
    traverse<false, false, false>(args);
}

__global__ __launch_bounds__(KP_TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_occluded5(
    TraceArgs args) {
    traverse<true, false, false>(args);
}

// The two-level traversals compiled for 5 waves per SIMD (96 registers, 44 B
// of spill per lane; 31 KB of LDS per block, so 5 blocks fit a CU): the
// integrator's two-level render of C3's terrain runs 2.3 % faster than with
// the 4-wave build.
__global__ __launch_bounds__(KP_TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_closest_inst5(
    TraceArgs args) {
    traverse<false, false, true>(args);
}

__global__ __launch_bounds__(KP_TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_occluded_inst5(
    TraceArgs args) {
    traverse<true, false, true>(args);
}

template <bool kCount>
__global__ __launch_bounds__(KP_TRACE_BLOCK) void k_closest_inst(TraceArgs args) {
    traverse<false, kCount, true>(args);
}

__global__ __launch_bounds__(KP_TRACE_BLOCK) void k_occluded_inst(TraceArgs args) {
    traverse<true, false, true>(args);
}

__global__ __launch_bounds__(256) void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy16(void* dst, const void* src, uint64_t n16, int cus, hipStream_t s) {
    const uint64_t want = (n16 + 255) / 256;
    const uint64_t cap = (uint64_t)(cus > 0 ? cus : 256) * 16;
    const uint32_t grid = (uint32_t)(want < cap ? (want ? want : 1) : cap);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, s, (uint4*)dst, (const uint4*)src, n16);
    return hipGetLastError();
}

__global__ void k_reset_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------- launch --
hipError_t launch_closest(const TraceArgs& a, uint32_t grid, bool count, hipStream_t s) {
    if (grid == 0 || grid > a.spill_blocks) return hipErrorInvalidValue;   // lanes without a spill slice
    hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(64), 0, s, a.counter, 1u);
    if (a.insts) {
        if (count) hipLaunchKernelGGL(k_closest_inst<true>, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
        else if (a.five_waves) hipLaunchKernelGGL(k_closest_inst5, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
        else hipLaunchKernelGGL(k_closest_inst<false>, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
        return hipGetLastError();
    }
    if (count) hipLaunchKernelGGL(k_closest<true>, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
    else if (a.five_waves) hipLaunchKernelGGL(k_closest5, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
    else hipLaunchKernelGGL(k_closest<false>, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_occluded(const TraceArgs& a, uint32_t grid, hipStream_t s) {
    if (grid == 0 || grid > a.spill_blocks) return hipErrorInvalidValue;   // lanes without a spill slice
    hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(64), 0, s, a.counter, 1u);
    if (a.insts && a.five_waves) hipLaunchKernelGGL(k_occluded_inst5, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
    else if (a.insts) hipLaunchKernelGGL(k_occluded_inst, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
    else if (a.five_waves) hipLaunchKernelGGL(k_occluded5, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
    else hipLaunchKernelGGL(k_occluded, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t trace_occupancy(int* blocks_per_cu_closest, int* blocks_per_cu_occluded) {
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu_closest,
                                                                k_closest<false>, KP_TRACE_BLOCK, 0);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu_occluded, k_occluded,
                                                        KP_TRACE_BLOCK, 0);
}

// Blocks per CU of the 5-wave builds: the smaller of the one-level and the
// two-level kernel's, so a grid of that size stays resident for either.
hipError_t trace_occupancy_five(int* blocks_per_cu_closest, int* blocks_per_cu_occluded) {
    int one = 0, two = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&one, k_closest5, KP_TRACE_BLOCK, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&two, k_closest_inst5, KP_TRACE_BLOCK, 0);
    if (e != hipSuccess) return e;
    *blocks_per_cu_closest = one < two ? one : two;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&one, k_occluded5, KP_TRACE_BLOCK, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&two, k_occluded_inst5, KP_TRACE_BLOCK, 0);
    if (e != hipSuccess) return e;
    *blocks_per_cu_occluded = one < two ? one : two;
    return hipSuccess;
}

hipError_t launch_reset_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace kp
