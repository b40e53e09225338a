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
//  - a per-lane short stack in LDS (kStack entries of {ref, tnear}); deeper
//    entries spill to a per-lane region in HBM, so any BVH depth works;
//  - 128-byte nodes: eight 16-byte loads per visit, 16-bit child boxes decoded
//    with one fma per plane; children sorted near-to-far by an 8-key sorting
//    network on (tnear bits | slot);
//  - watertight ray/triangle test (Woop, Benthin, Wald 2013) with contraction
//    disabled, so its result is bit-identical to the host oracle's;
//  - closest hit is the smallest t, ties broken by the lowest primitive
//    index, which makes the answer independent of traversal order.
#include <hip/hip_runtime.h>

#include "kp_kernels.h"
#include "kp_layout.h"

#pragma clang fp contract(off)

namespace kp {

// 1 + 2*gamma(3), gamma(n) = n*u/(1-n*u), u = 2^-24 (Ize 2013).
#define KP_FAR_SCALE 1.00000036f

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

__device__ __forceinline__ float exp2i(int e) { return __int_as_float((e + 127) << 23); }

__device__ __forceinline__ void cswap(uint32_t& a, uint32_t& b) {
    uint32_t lo = min(a, b), hi = max(a, b);
    a = lo;
    b = hi;
}

// Optimal 19-comparator sorting network for 8 keys.
__device__ __forceinline__ void sort8(uint32_t* k) {
    cswap(k[0], k[2]); cswap(k[1], k[3]); cswap(k[4], k[6]); cswap(k[5], k[7]);
    cswap(k[0], k[4]); cswap(k[1], k[5]); cswap(k[2], k[6]); cswap(k[3], k[7]);
    cswap(k[0], k[1]); cswap(k[2], k[3]); cswap(k[4], k[5]); cswap(k[6], k[7]);
    cswap(k[2], k[4]); cswap(k[3], k[5]);
    cswap(k[1], k[4]); cswap(k[3], k[6]);
    cswap(k[1], k[2]); cswap(k[3], k[4]); cswap(k[5], k[6]);
}

struct NodeRegs {
    uint4 w[8];
};

__device__ __forceinline__ void load_node(const uint4* __restrict__ nodes, uint32_t idx, NodeRegs& n) {
    const uint4* p = nodes + (size_t)idx * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) n.w[i] = p[i];
}

__device__ __forceinline__ uint32_t q16(const uint4& plane, int slot) {
    const uint32_t d = (slot >> 1) == 0 ? plane.x : (slot >> 1) == 1 ? plane.y : (slot >> 1) == 2 ? plane.z : plane.w;
    return (slot & 1) ? (d >> 16) : (d & 0xFFFFu);
}

// Tests the 8 children of a node; returns sort keys (0xFFFFFFFF = miss).
__device__ __forceinline__ void test_children(const RayPre& r, const NodeRegs& n, float tmax,
                                              uint32_t* key) {
    const float px = __uint_as_float(n.w[0].x), py = __uint_as_float(n.w[0].y), pz = __uint_as_float(n.w[0].z);
    const uint32_t e = n.w[0].w;
    const int ex = (int)(int8_t)(e & 0xFF), ey = (int)(int8_t)((e >> 8) & 0xFF), ez = (int)(int8_t)((e >> 16) & 0xFF);
    const float ax = exp2i(ex) * r.ix, ay = exp2i(ey) * r.iy, az = exp2i(ez) * r.iz;
    const float bx = (px - r.ox) * r.ix, by = (py - r.oy) * r.iy, bz = (pz - r.oz) * r.iz;
    // near/far planes per axis from the direction signs
    const uint4 nx = r.ix >= 0.0f ? n.w[2] : n.w[5], fx = r.ix >= 0.0f ? n.w[5] : n.w[2];
    const uint4 ny = r.iy >= 0.0f ? n.w[3] : n.w[6], fy = r.iy >= 0.0f ? n.w[6] : n.w[3];
    const uint4 nz = r.iz >= 0.0f ? n.w[4] : n.w[7], fz = r.iz >= 0.0f ? n.w[7] : n.w[4];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const float tnx = fmaf((float)q16(nx, s), ax, bx);
        const float tfx = fmaf((float)q16(fx, s), ax, bx);
        const float tny = fmaf((float)q16(ny, s), ay, by);
        const float tfy = fmaf((float)q16(fy, s), ay, by);
        const float tnz = fmaf((float)q16(nz, s), az, bz);
        const float tfz = fmaf((float)q16(fz, s), az, bz);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tmin));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        const uint32_t mw = s < 4 ? n.w[1].z : n.w[1].w;
        const bool used = ((mw >> ((s & 3) * 8)) & 0xFFu) != 0u;   // empty slots never hit
        const bool hit = used && tn <= tf * KP_FAR_SCALE;
        key[s] = hit ? ((__float_as_uint(tn) & 0x7FFFFFF8u) | (uint32_t)s) : 0xFFFFFFFFu;
    }
}

__device__ __forceinline__ uint32_t child_ref(const NodeRegs& n, uint32_t slot) {
    const uint32_t imask = n.w[0].w >> 24;
    const uint32_t mw = slot < 4 ? n.w[1].z : n.w[1].w;
    const uint32_t meta = (mw >> ((slot & 3u) * 8u)) & 0xFFu;
    if ((imask >> slot) & 1u) return n.w[1].x + __popc(imask & ((1u << slot) - 1u));
    const uint32_t first = n.w[1].y + (meta & 31u);
    return KP_LEAF_FLAG | (first << 3) | ((meta >> 5) - 1u);
}

template <int S>
struct Stack {
    uint2 (*lds)[KP_TRACE_BLOCK];
    uint2* spill;
    uint32_t cap;
    int sp;
    int tid;

    __device__ __forceinline__ bool push(uint32_t ref, uint32_t tbits) {
        if (sp >= S) {
            if ((uint32_t)(sp - S) >= cap) return false;
            spill[sp - S] = lds[sp & (S - 1)][tid];
        }
        lds[sp & (S - 1)][tid] = make_uint2(ref, tbits);
        ++sp;
        return true;
    }
    __device__ __forceinline__ uint2 pop() {
        --sp;
        const uint2 e = lds[sp & (S - 1)][tid];
        if (sp >= S) lds[sp & (S - 1)][tid] = spill[sp - S];
        return e;
    }
};

constexpr int kStack = KP_TRACE_STACK;

// Work distribution of the persistent kernels.  A wave takes chunks of
// kChunk consecutive rays from the global counter (one atomic per chunk) and
// hands them to its lanes from a wave-uniform cursor; every lane keeps its
// next ray prefetched in registers, so replacing a finished ray costs no
// memory round trip.  All members except the prefetch are wave-uniform.
constexpr uint32_t kChunk = 256;

struct RayFeeder {
    uint32_t next, end, n;
    bool drained;
    uint32_t* counter;
    const float4* rays;
    const uint32_t* roots;
    bool pending;       // this lane holds a prefetched ray
    uint32_t pid, proot;
    float4 p0, p1;

    __device__ __forceinline__ void init(uint32_t* ctr, const float4* rs, const uint32_t* rts, uint32_t count) {
        counter = ctr;
        rays = rs;
        roots = rts;
        proot = 0u;
        n = count;
        next = end = 0;
        drained = false;
        pending = false;
        pid = KP_INVALID;
        assign();
    }
    // Gives a prefetched ray to every lane without one, as far as rays last.
    __device__ __forceinline__ void assign() {
        const uint64_t m = __ballot(!pending);
        if (m == 0 || drained) return;
        if (next >= end) {
            uint32_t base = 0;
            if (__lane_id() == 0) base = atomicAdd(counter, kChunk);
            base = __builtin_amdgcn_readfirstlane(base);
            if (base >= n) {
                drained = true;
                return;
            }
            next = base;
            end = base + kChunk < n ? base + kChunk : n;
        }
        const uint32_t avail = end - next;
        const uint32_t want = (uint32_t)__popcll(m);
        const uint32_t prefix = (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
        if (!pending && prefix < avail) {
            pid = next + prefix;
            p0 = rays[2 * (size_t)pid];
            p1 = rays[2 * (size_t)pid + 1];
            if (roots) proot = roots[pid];
            pending = true;
        }
        next += want < avail ? want : avail;
    }
};

template <bool kCount>
__global__ __launch_bounds__(KP_TRACE_BLOCK) void k_closest(TraceArgs args) {
    __shared__ uint2 lds[kStack][KP_TRACE_BLOCK];
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const uint32_t gl = blockIdx.x * KP_TRACE_BLOCK + tid;
    Stack<kStack> st;
    st.lds = lds;
    st.spill = args.spill + (size_t)gl * args.spill_cap;
    st.cap = args.spill_cap;
    st.sp = 0;
    st.tid = tid;
    const uint4* __restrict__ nodes = args.nodes;
    const uint32_t n_rays = args.n_dev ? *args.n_dev : args.n;   // device-side count for queues
    const float4* __restrict__ tris = args.tris;

    uint32_t ray_id = KP_INVALID;
    RayPre r;
    float best_t = 0.0f, bu = 0.0f, bv = 0.0f;
    uint32_t best_p = KP_INVALID, cur = KP_INVALID, steps = 0;
    uint64_t cnt_nodes = 0, cnt_tris = 0;
    (void)lane;
    RayFeeder feed;
    feed.init(args.counter, args.rays, args.roots, n_rays);

    for (;;) {
        if (ray_id == KP_INVALID && feed.pending) {
            ray_id = feed.pid;
            setup_ray(feed.p0, feed.p1, r);
            best_t = r.tmax;
            best_p = KP_INVALID;
            bu = bv = 0.0f;
            cur = ray_valid(feed.p0, feed.p1) ? feed.proot : KP_INVALID;   // root
            st.sp = 0;
            steps = 0;
            feed.pending = false;
        }
        feed.assign();
        if (__ballot(ray_id != KP_INVALID || feed.pending) == 0) break;
        if (ray_id != KP_INVALID) {
            for (;;) {
                bool done = false;
                if (cur == KP_INVALID) {
                    done = true;   // invalid ray: reported as a miss
                } else if (cur & KP_LEAF_FLAG) {
                    const uint32_t first = (cur & ~KP_LEAF_FLAG) >> 3;
                    const uint32_t count = (cur & 7u) + 1u;
                    if (kCount) cnt_tris += count;
                    for (uint32_t i = 0; i < count; ++i) {
                        const float4* tp = tris + 3 * (size_t)(first + i);
                        const float4 a = tp[0], b = tp[1], c = tp[2];
                        float t, u, v;
                        if (tri_test(r, a, b, c, best_t, t, u, v)) {
                            const uint32_t p = __float_as_uint(a.w);
                            if (t < best_t || p < best_p) { best_t = t; best_p = p; bu = u; bv = v; }
                        }
                    }
                    cur = KP_INVALID;
                } else {
                    NodeRegs n;
                    load_node(nodes, cur, n);
                    if (kCount) cnt_nodes++;
                    uint32_t key[8];
                    test_children(r, n, best_t, key);
                    sort8(key);
                    cur = KP_INVALID;
                    if (key[0] != 0xFFFFFFFFu) {
                        bool ok = true;
#pragma unroll
                        for (int j = 7; j >= 1; --j)
                            if (key[j] != 0xFFFFFFFFu)
                                ok = ok && st.push(child_ref(n, key[j] & 7u), key[j] & ~7u);
                        if (!ok) {
                            atomicOr(args.err, 1u);
                            done = true;
                        }
                        cur = child_ref(n, key[0] & 7u);
                    }
                }
                if (!done && cur == KP_INVALID) {
                    while (st.sp > 0) {
                        const uint2 e = st.pop();
                        if (__uint_as_float(e.y) <= best_t) { cur = e.x; break; }
                    }
                    if (cur == KP_INVALID) done = true;
                }
                if (!done && ++steps > args.max_steps) {
                    atomicOr(args.err, 2u);
                    done = true;
                }
                if (done) {
                    float4 h;
                    if (best_p == KP_INVALID) h = make_float4(__int_as_float(0x7f800000), __uint_as_float(KP_INVALID), 0.0f, 0.0f);
                    else h = make_float4(best_t, __uint_as_float(best_p), bu, bv);
                    args.hits[ray_id] = h;
                    ray_id = KP_INVALID;
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

__global__ __launch_bounds__(KP_TRACE_BLOCK) void k_occluded(TraceArgs args) {
    __shared__ uint2 lds[kStack][KP_TRACE_BLOCK];
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const uint32_t gl = blockIdx.x * KP_TRACE_BLOCK + tid;
    Stack<kStack> st;
    st.lds = lds;
    st.spill = args.spill + (size_t)gl * args.spill_cap;
    st.cap = args.spill_cap;
    st.sp = 0;
    st.tid = tid;
    const uint4* __restrict__ nodes = args.nodes;
    const uint32_t n_rays = args.n_dev ? *args.n_dev : args.n;   // device-side count for queues
    const float4* __restrict__ tris = args.tris;

    uint32_t ray_id = KP_INVALID;
    RayPre r;
    uint32_t cur = KP_INVALID, steps = 0;
    (void)lane;
    RayFeeder feed;
    feed.init(args.counter, args.rays, args.roots, n_rays);
    for (;;) {
        if (ray_id == KP_INVALID && feed.pending) {
            ray_id = feed.pid;
            setup_ray(feed.p0, feed.p1, r);
            cur = ray_valid(feed.p0, feed.p1) ? feed.proot : KP_INVALID;
            st.sp = 0;
            steps = 0;
            feed.pending = false;
        }
        feed.assign();
        if (__ballot(ray_id != KP_INVALID || feed.pending) == 0) break;
        if (ray_id != KP_INVALID) {
            for (;;) {
                bool hit = false, done = false;
                if (cur == KP_INVALID) {
                    done = true;   // invalid ray: reported as not occluded
                } else if (cur & KP_LEAF_FLAG) {
                    const uint32_t first = (cur & ~KP_LEAF_FLAG) >> 3;
                    const uint32_t count = (cur & 7u) + 1u;
                    for (uint32_t i = 0; i < count && !hit; ++i) {
                        const float4* tp = tris + 3 * (size_t)(first + i);
                        float t, u, v;
                        hit = tri_test(r, tp[0], tp[1], tp[2], r.tmax, t, u, v);
                    }
                    cur = KP_INVALID;
                } else {
                    NodeRegs n;
                    load_node(nodes, cur, n);
                    uint32_t key[8];
                    test_children(r, n, r.tmax, key);
                    cur = KP_INVALID;
                    bool ok = true;
#pragma unroll
                    for (int s = 0; s < 8; ++s) {
                        if (key[s] == 0xFFFFFFFFu) continue;
                        const uint32_t ref = child_ref(n, (uint32_t)s);
                        if (cur == KP_INVALID) cur = ref;
                        else ok = ok && st.push(ref, 0u);
                    }
                    if (!ok) {
                        atomicOr(args.err, 1u);
                        hit = true;   // report occluded rather than a false "visible"
                    }
                }
                if (hit) done = true;
                if (!done && cur == KP_INVALID) {
                    if (st.sp > 0) cur = st.pop().x;
                    else done = true;
                }
                if (!done && ++steps > args.max_steps) {
                    atomicOr(args.err, 2u);
                    hit = true;
                    done = true;
                }
                if (done) {
                    args.occl[ray_id] = hit ? 1u : 0u;
                    ray_id = KP_INVALID;
                    break;
                }
                if (__popcll(__ballot(true)) < args.refill_below) break;
            }
        }
    }
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
    hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(64), 0, s, a.counter, 1u);
    if (count) hipLaunchKernelGGL(k_closest<true>, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
    else hipLaunchKernelGGL(k_closest<false>, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_occluded(const TraceArgs& a, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(64), 0, s, a.counter, 1u);
    hipLaunchKernelGGL(k_occluded, dim3(grid), dim3(KP_TRACE_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t trace_occupancy(int* blocks_per_cu_closest, int* blocks_per_cu_occluded) {
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu_closest,
                                                                k_closest<false>, KP_TRACE_BLOCK, 0);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu_occluded, k_occluded,
                                                        KP_TRACE_BLOCK, 0);
}

hipError_t launch_reset_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace kp
