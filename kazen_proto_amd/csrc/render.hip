// Wavefront path integrator (gfx950).
//
// Replaces PathIntegrator::Li (reference src/core/integrator.cpp:362-561)
// driven by Renderer::render's tile/pixel/sample loops (src/core/renderer.cpp:
// 18-90), and defines the GPU render overload the reference declares but
// never implements (src/core/renderer.h:30-31).
//
// One pass traces k samples of every pixel as n = width*height*k independent
// paths (slot = pixel*k + j).  Per bounce:
//   extend  : closest hit of the compacted ray queue (trace.hip)
//   shade   : emission with MIS, next-event estimation (writes a shadow ray),
//             Russian roulette, BSDF sampling (writes the next ray); both
//             queues are compacted with a ballot per wave and one atomic per
//             block and queue (same-address device atomics are the limit:
//             one per wave made shading 3.4x slower)
//   connect : occlusion of the shadow queue, then the unoccluded
//             contributions are added to their paths
// After the last bounce every pixel folds its k path radiances into the film
// (FP32 sum with Kahan compensation, one owner per pixel, fixed order), so
// the film does not depend on scheduling.
//
// Sampling protocol of a path (shared with the oracle): PCG32 stream with
// state ((sample << 32) ^ seed ^ 0x9E3779B97F4A7C15) and sequence = pixel
// index; 2 draws for the pixel position, then 8 draws at every shaded vertex:
// light pick, light triangle, two barycentric, roulette, lobe, two direction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include <cstdio>
#include <cstdlib>

#include "../../include/kp_hip.h"
#include "kp_closures.h"
#include "kp_ctx.h"
#include "kp_guide.h"
#include "kp_guide_host.h"

namespace kp {

struct LightRec {
    uint32_t first;      // first entry in light_tris
    uint32_t count;      // triangles
    uint32_t cdf_first;  // first of count+1 CDF entries
    uint32_t pad;
    float area;
    float Le[3];
};

struct PathState {
    float4* rays[2];      // ray queues {o, tmin} {d, tmax}, compacted
    uint32_t* qmap[2];    // queue entry -> path slot
    float4* hits;         // closest hits of the current queue
    float4* s0;           // per slot: beta.xyz, pdf of the last BSDF sample
    float4* s1;           // per slot: radiance.xyz
    uint64_t* rng;        // per slot: PCG32 state
    float4* srays;        // shadow queue rays
    float4* scontrib;     // shadow contribution.xyz, slot bits in w
    uint32_t* socc;       // occlusion result per shadow ray
    uint32_t* cnt;        // queue sizes qc(0), qc(1) rays, qc(2) shadow rays
    unsigned long long* totals;   // [0] extension rays, [1] shadow rays
    // batched scenes only
    uint32_t* slot_scene; // per slot: scene of the path
    uint32_t* slot_pixel; // per slot: pixel of the path in its scene
    uint32_t* qroot[2];   // per ray-queue entry: BVH root of the ray's scene
    uint32_t* sroot;      // per shadow-queue entry: BVH root
    // Each queue counter has a 256-byte line of its own: the two counters
    // one shade appends to are then served by different memory channels.
    __host__ __device__ uint32_t& qc(uint32_t i) const { return cnt[i * 64u]; }
};

// One scene of a batch (the camera, film and sampling settings of its
// kp_render_desc, its BVH root and light range in the merged arrays).  Work
// item w of the batch is sample (w - work_first) % spp of pixel
// (w - work_first) / spp of the scene with the last work_first <= w.
struct SceneRec {
    float3 eye, fwd, right, upv;
    float tan_half, aspect;
    uint32_t width, height, spp, spp_offset, seed, max_depth, rr_depth;
    uint32_t root, light_first, n_lights, film_first, pad;
    unsigned long long work_first;
};

struct RenderArgs {
    const float4* stris;     // per primitive: {v0, mat}, {v1, light}, {v2, normal record or KP_INVALID}
    const float4* snorm;     // per normal record: three float4 {n_k, 0} (unit, world space)
    const MatRec* mats;
    const Lobe* lobes;
    const LightRec* lights;
    const uint32_t* light_tris;
    const float* light_cdf;
    uint32_t n_lights;
    uint32_t width, height, npix, k, n_paths;
    uint32_t seq_base;       // PCG sequence of slot s = seq_base + s / k
    uint32_t seed, s0, max_depth, rr_depth;
    float3 eye, fwd, right, upv;
    float tan_half, aspect;
    PathState ps;
    // guided shading (k_shade<true>)
    GuideNet net;
    float bsdf_frac, min_alpha;
    uint32_t guide_on;                      // sample from the network
    uint32_t guide_depth;                   // guided vertices: depth <= guide_depth
    uint32_t rec_cap, rec_stride, rec_phase;   // training records (rec_cap 0: none)
    uint32_t rec_depth;                     // records per recorded path: vertices 1..rec_depth
    float* rec_in6;
    float* rec_dirs;
    float4* rec_aux;
    // batched scenes (scenes != nullptr): n_paths is the pass capacity and
    // pass[0] / pass[1] hold the pass's first work item / the total on the device
    const SceneRec* scenes;
    uint32_t n_scenes;
    unsigned long long* pass;
};

// ------------------------------------------------------------------ RNG --
struct Pcg {
    uint64_t state, inc;
    __device__ __forceinline__ uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    __device__ __forceinline__ float nextf() {
        return __uint_as_float((next() >> 9) | 0x3f800000u) - 1.0f;
    }
};

__device__ __forceinline__ Pcg pcg_seed(uint64_t initstate, uint64_t initseq) {
    Pcg p;
    p.state = 0u;
    p.inc = (initseq << 1u) | 1u;
    p.next();
    p.state += initstate;
    p.next();
    return p;
}

__device__ __forceinline__ f3 ld3(const float4& v) { return mk(v.x, v.y, v.z); }

// Waechter & Binder, Ray Tracing Gems ch. 6 (constants as in the reference's
// intersection.h:78-80).
__device__ __forceinline__ f3 offset_ray(f3 p, f3 n) {
    const float origin = 1.0f / 32.0f, fscale = 1.0f / 65536.0f, iscale = 256.0f;
    const int ox = (int)(iscale * n.x), oy = (int)(iscale * n.y), oz = (int)(iscale * n.z);
    const f3 pi = mk(__int_as_float(__float_as_int(p.x) + (p.x < 0.0f ? -ox : ox)),
                     __int_as_float(__float_as_int(p.y) + (p.y < 0.0f ? -oy : oy)),
                     __int_as_float(__float_as_int(p.z) + (p.z < 0.0f ? -oz : oz)));
    return mk(fabsf(p.x) < origin ? p.x + fscale * n.x : pi.x,
              fabsf(p.y) < origin ? p.y + fscale * n.y : pi.y,
              fabsf(p.z) < origin ? p.z + fscale * n.z : pi.z);
}

// DiscrectPDF::sample (src/base/dpdf.h:32-37): first entry >= u, minus one.
__device__ __forceinline__ uint32_t cdf_sample(const float* cdf, uint32_t count, float u) {
    uint32_t lo = 0, hi = count + 1;   // search cdf[0..count]
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    const uint32_t idx = lo > 0 ? lo - 1 : 0;
    return idx < count - 1 ? idx : count - 1;
}

// ------------------------------------------------------------- kernels --
__global__ __launch_bounds__(256) void k_generate(RenderArgs a) {
#pragma clang fp contract(off)
    const uint32_t slot = blockIdx.x * 256 + threadIdx.x;
    if (slot == 0) {
        a.ps.qc(0) = a.n_paths;
        a.ps.qc(1) = 0;
        a.ps.qc(2) = 0;
    }
    if (slot >= a.n_paths) return;
    const uint32_t pixel = slot / a.k;
    const uint32_t j = slot - pixel * a.k;
    const uint32_t sample = a.s0 + j;
    const uint32_t px = pixel % a.width, py = pixel / a.width;
    const uint64_t initstate = ((uint64_t)sample << 32) ^ (uint64_t)a.seed ^ 0x9E3779B97F4A7C15ull;
    Pcg rng = pcg_seed(initstate, pixel);
    const float ux = rng.nextf(), uy = rng.nextf();
    const float sx = ((float)px + ux) / (float)a.width, sy = ((float)py + uy) / (float)a.height;
    const float nx = (2.0f * sx - 1.0f) * a.tan_half * a.aspect;
    const float ny = (1.0f - 2.0f * sy) * a.tan_half;
    const f3 fwd = mk(a.fwd.x, a.fwd.y, a.fwd.z), right = mk(a.right.x, a.right.y, a.right.z);
    const f3 upv = mk(a.upv.x, a.upv.y, a.upv.z);
    const f3 d = normalize(fwd + right * nx + upv * ny);
    a.ps.rays[0][2 * (size_t)slot] = make_float4(a.eye.x, a.eye.y, a.eye.z, 0.0f);
    a.ps.rays[0][2 * (size_t)slot + 1] = make_float4(d.x, d.y, d.z, __int_as_float(0x7f800000));
    a.ps.qmap[0][slot] = slot;
    a.ps.s0[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    a.ps.s1[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    a.ps.rng[slot] = rng.state;
}

// Starts one path per caller ray (integrator entry for arbitrary primary
// rays): sequence = seq_base + slot, sample index = s0.
__global__ __launch_bounds__(256) void k_generate_rays(RenderArgs a, const float4* rays) {
    const uint32_t slot = blockIdx.x * 256 + threadIdx.x;
    if (slot == 0) {
        a.ps.qc(0) = a.n_paths;
        a.ps.qc(1) = 0;
        a.ps.qc(2) = 0;
    }
    if (slot >= a.n_paths) return;
    const uint64_t initstate = ((uint64_t)a.s0 << 32) ^ (uint64_t)a.seed ^ 0x9E3779B97F4A7C15ull;
    const Pcg rng = pcg_seed(initstate, a.seq_base + slot);
    a.ps.rays[0][2 * (size_t)slot] = rays[2 * (size_t)slot];
    a.ps.rays[0][2 * (size_t)slot + 1] = rays[2 * (size_t)slot + 1];
    a.ps.qmap[0][slot] = slot;
    a.ps.s0[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    a.ps.s1[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    a.ps.rng[slot] = rng.state;
}

// Paths of the pass that starts at work item pass[0] (batched scenes).
__device__ __forceinline__ uint32_t batch_pass_paths(const RenderArgs& a) {
    const unsigned long long w0 = a.pass[0], total = a.pass[1];
    const unsigned long long left = total > w0 ? total - w0 : 0ull;
    return left < a.n_paths ? (uint32_t)left : a.n_paths;
}

// Starts the paths of one pass over a batch of scenes: the same camera ray,
// PCG32 stream and sample index as k_generate gives the pixel's sample in a
// render of that scene alone.
__global__ __launch_bounds__(256) void k_generate_batch(RenderArgs a) {
#pragma clang fp contract(off)
    const uint32_t slot = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = batch_pass_paths(a);
    if (slot == 0) {
        a.ps.qc(0) = n;
        a.ps.qc(1) = 0;
        a.ps.qc(2) = 0;
    }
    if (slot >= n) return;
    const unsigned long long w = a.pass[0] + slot;
    uint32_t lo = 0, hi = a.n_scenes;   // last scene with work_first <= w
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.scenes[mid].work_first <= w) lo = mid;
        else hi = mid;
    }
    const SceneRec sc = a.scenes[lo];
    const unsigned long long l = w - sc.work_first;
    const uint32_t pixel = (uint32_t)(l / sc.spp);
    const uint32_t sample = sc.spp_offset + (uint32_t)(l - (unsigned long long)pixel * sc.spp);
    const uint32_t px = pixel % sc.width, py = pixel / sc.width;
    const uint64_t initstate = ((uint64_t)sample << 32) ^ (uint64_t)sc.seed ^ 0x9E3779B97F4A7C15ull;
    Pcg rng = pcg_seed(initstate, pixel);
    const float ux = rng.nextf(), uy = rng.nextf();
    const float sx = ((float)px + ux) / (float)sc.width, sy = ((float)py + uy) / (float)sc.height;
    const float nx = (2.0f * sx - 1.0f) * sc.tan_half * sc.aspect;
    const float ny = (1.0f - 2.0f * sy) * sc.tan_half;
    const f3 fwd = mk(sc.fwd.x, sc.fwd.y, sc.fwd.z), right = mk(sc.right.x, sc.right.y, sc.right.z);
    const f3 upv = mk(sc.upv.x, sc.upv.y, sc.upv.z);
    const f3 d = normalize(fwd + right * nx + upv * ny);
    a.ps.rays[0][2 * (size_t)slot] = make_float4(sc.eye.x, sc.eye.y, sc.eye.z, 0.0f);
    a.ps.rays[0][2 * (size_t)slot + 1] = make_float4(d.x, d.y, d.z, __int_as_float(0x7f800000));
    a.ps.qmap[0][slot] = slot;
    a.ps.qroot[0][slot] = sc.root;
    a.ps.slot_scene[slot] = lo;
    a.ps.slot_pixel[slot] = pixel;
    a.ps.s0[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    a.ps.s1[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    a.ps.rng[slot] = rng.state;
}

// Folds the pass's path radiances into the scenes' films: the first slot of
// every pixel adds that pixel's consecutive samples in sample order (Kahan),
// continuing the sum of earlier passes, so the film equals the one-scene
// render's bit for bit whatever the pass size.
__global__ __launch_bounds__(256) void k_film_batch(RenderArgs a, float4* sum, float4* comp) {
#pragma clang fp contract(off)
    // The block's 256 slots are staged in LDS with coalesced loads; the
    // first slot of every pixel run then folds the run from LDS, continuing
    // from global memory only where a run crosses the block's end.  Same
    // additions in the same order as a fold from global memory.
    __shared__ float4 xs[256];
    __shared__ uint32_t ss[257], sp[257];   // [0]: the slot before the block
    const uint32_t base = blockIdx.x * 256;
    const uint32_t t = threadIdx.x, slot = base + t;
    const uint32_t n = batch_pass_paths(a);
    if (slot < n) {
        xs[t] = a.ps.s1[slot];
        ss[t + 1] = a.ps.slot_scene[slot];
        sp[t + 1] = a.ps.slot_pixel[slot];
    }
    if (t == 0) {
        ss[0] = base > 0 ? a.ps.slot_scene[base - 1] : 0xFFFFFFFFu;
        sp[0] = base > 0 ? a.ps.slot_pixel[base - 1] : 0xFFFFFFFFu;
    }
    __syncthreads();
    if (slot >= n) return;
    const uint32_t sid = ss[t + 1], pix = sp[t + 1];
    if (ss[t] == sid && sp[t] == pix) return;   // not the first slot of its run
    const uint32_t f = a.scenes[sid].film_first + pix;
    float4 S = sum[f], C = comp[f];
    auto add = [&](const float4& x) {
        float y, tt;
        y = x.x - C.x; tt = S.x + y; C.x = (tt - S.x) - y; S.x = tt;
        y = x.y - C.y; tt = S.y + y; C.y = (tt - S.y) - y; S.y = tt;
        y = x.z - C.z; tt = S.z + y; C.z = (tt - S.z) - y; S.z = tt;
    };
    uint32_t j = t;
    for (; j < 256 && base + j < n && ss[j + 1] == sid && sp[j + 1] == pix; ++j) add(xs[j]);
    if (j == 256)
        for (uint32_t g = base + 256; g < n && a.ps.slot_scene[g] == sid && a.ps.slot_pixel[g] == pix; ++g)
            add(a.ps.s1[g]);
    sum[f] = S;
    comp[f] = C;
}

// Moves the batch to its next pass (part of the captured pass graph).
__global__ void k_pass_advance(unsigned long long* pass, uint32_t paths) {
    if (threadIdx.x == 0 && blockIdx.x == 0) pass[0] += paths;
}

// Adds offsets to the node and triangle records of one scene copied into the
// merged arrays: inner-child and leaf bases, and the primitive id.
__global__ void k_rebase_nodes(uint4* nodes, uint32_t n, uint32_t node_off, uint32_t tri_off) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint4 w = nodes[8 * (size_t)i + 1];   // dwords 4-7: child_base, prim_base, meta
    w.x += node_off;
    w.y += tri_off;
    nodes[8 * (size_t)i + 1] = w;
}
__global__ void k_rebase_tris(float4* tris, uint32_t n, uint32_t prim_off) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float4 v = tris[3 * (size_t)i];
    v.w = __uint_as_float(__float_as_uint(v.w) + prim_off);
    tris[3 * (size_t)i] = v;
}
// Shading records: material id += mat_off, light id (if >= 0) += light_off,
// normal record (if any) += norm_off.
__global__ void k_rebase_stris(float4* st, uint32_t n, uint32_t mat_off, uint32_t light_off, uint32_t norm_off) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float4 a = st[3 * (size_t)i], b = st[3 * (size_t)i + 1], c = st[3 * (size_t)i + 2];
    a.w = __uint_as_float(__float_as_uint(a.w) + mat_off);
    const int lid = __float_as_int(b.w);
    if (lid >= 0) b.w = __int_as_float(lid + (int)light_off);
    const uint32_t nid = __float_as_uint(c.w);
    if (nid != KP_INVALID) c.w = __uint_as_float(nid + norm_off);
    st[3 * (size_t)i] = a;
    st[3 * (size_t)i + 1] = b;
    st[3 * (size_t)i + 2] = c;
}
__global__ void k_rebase_lights(LightRec* l, uint32_t n, uint32_t lt_off, uint32_t cdf_off, uint32_t* lt,
                                uint32_t n_lt, uint32_t prim_off) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        l[i].first += lt_off;
        l[i].cdf_first += cdf_off;
    }
    if (i < n_lt) lt[i] += prim_off;
}

__global__ void k_write_radiance(RenderArgs a, float* out) {
    const uint32_t slot = blockIdx.x * 256 + threadIdx.x;
    if (slot >= a.n_paths) return;
    const float4 l = a.ps.s1[slot];
    out[3 * (size_t)slot] = l.x;
    out[3 * (size_t)slot + 1] = l.y;
    out[3 * (size_t)slot + 2] = l.z;
}

// Resets the output queue counters of the coming shade, accumulating totals.
__global__ void k_bookkeep(RenderArgs a, int cur) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        a.ps.totals[0] += a.ps.qc(cur);
        a.ps.qc(cur ^ 1) = 0;
        a.ps.qc(2) = 0;
    }
}

__global__ void k_count_shadow(RenderArgs a) {
    if (threadIdx.x == 0 && blockIdx.x == 0) a.ps.totals[1] += a.ps.qc(2);
}

// Appends the lanes with `want0` to queue c0 and those with `want1` to queue
// c1: one ballot per wave and queue, one atomic per block and queue; returns
// the lane's position in c0 (and in c1 through o1).  Every thread of the
// block must call it (it synchronises the block); lds holds 10 words.
__device__ __forceinline__ uint32_t block_append2(bool want0, bool want1, uint32_t* c0, uint32_t* c1,
                                                  uint32_t* lds, uint32_t& o1) {
    const uint64_t m0 = __ballot(want0), m1 = __ballot(want1);
    const int lane = __lane_id();
    const int w = threadIdx.x >> 6;
    if (lane == 0) {
        lds[w] = (uint32_t)__popcll(m0);
        lds[4 + w] = (uint32_t)__popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t0 = lds[0] + lds[1] + lds[2] + lds[3], t1 = lds[4] + lds[5] + lds[6] + lds[7];
        lds[8] = t0 ? atomicAdd(c0, t0) : 0u;
        lds[9] = t1 ? atomicAdd(c1, t1) : 0u;
    }
    __syncthreads();
    uint32_t b0 = lds[8], b1 = lds[9];
    for (int i = 0; i < w; ++i) {
        b0 += lds[i];
        b1 += lds[4 + i];
    }
    __syncthreads();   // lds is reused by the next call
    const uint64_t below = (1ull << lane) - 1ull;
    o1 = b1 + (uint32_t)__popcll(m1 & below);
    return b0 + (uint32_t)__popcll(m0 & below);
}

// Shades the current queue.  kGuided adds the guiding network: 12 draws per
// vertex (8 plain + strategy, lobe, two direction), one-sample MIS between
// the BSDF and the network at guided vertices, and training records.  The
// network query runs once per wave after every lane has prepared its vertex.
template <bool kGuided>
__global__ __launch_bounds__(256) void k_shade(RenderArgs a, int cur, uint32_t depth) {
    __shared__ __align__(16) uint32_t glds[kGuided ? 4 * KP_G_WAVE_ROWS * KP_G_ROW : 1];   // 16-byte row fragments
    __shared__ uint32_t app_lds[10];
    __shared__ uint32_t n_blk;
    // one read of the queue size per block: the loop below synchronises the
    // block every iteration, so all its threads must agree on the trip count
    if (threadIdx.x == 0) n_blk = a.ps.qc(cur);
    __syncthreads();
    const uint32_t n = n_blk;
    const uint32_t total = gridDim.x * 256;
    const uint32_t gid = blockIdx.x * 256 + threadIdx.x;
    const uint32_t iters = (n + total - 1) / total;
    const float4* __restrict__ rin = a.ps.rays[cur];
    const uint32_t* __restrict__ qin = a.ps.qmap[cur];
    float4* __restrict__ rout = a.ps.rays[cur ^ 1];
    uint32_t* __restrict__ qout = a.ps.qmap[cur ^ 1];
    const float inv_nl_all = a.n_lights ? 1.0f / (float)a.n_lights : 0.0f;
    constexpr int kDraws = kGuided ? 12 : 8;
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t q = it * total + gid;
        const bool active = q < n;
        bool cont = false, shad = false;
        float4 nr0 = make_float4(0, 0, 0, 0), nr1 = nr0, sr0 = nr0, sr1 = nr0, sc = nr0;
        uint32_t slot = 0;
        // ---- per-lane vertex set-up --------------------------------------
        bool alive = false, guided = false, recorded = false;
        uint32_t rec = 0;
        f3 beta = mk(0, 0, 0), beta_pre = beta, L = beta, P = beta, ng = beta, nrm = beta, wo = beta;
        Bsdf bs;
        bs.lobes = a.lobes;
        bs.n = 0;
        bs.beta = beta;
        bs.inv_sum = 1.0f;
        Pcg rng;
        rng.state = 0;
        rng.inc = 1;
        float u[kDraws];
        bool nee = false;
        f3 nee_org = beta, nee_dir = beta, nee_c = beta;
        float nee_dist = 0.0f, nee_cos = 0.0f, pdf_l = 0.0f, pdf_bl = 0.0f;
        bool bs_ok = false;
        f3 bs_wi = beta, bs_f = beta;
        float bs_pdf = 0.0f;
        bool use_guide = false;
        uint32_t max_depth = a.max_depth, rr_depth = a.rr_depth, n_lights = a.n_lights, light_first = 0,
                 root = 0;
        float inv_nl = inv_nl_all;
        if (active) {
            slot = qin[q];
            const float4 hit = a.ps.hits[q];
            const float4 r1 = rin[2 * (size_t)q + 1];
            const float4 st0 = a.ps.s0[slot];
            const float4 st1 = a.ps.s1[slot];
            beta = mk(st0.x, st0.y, st0.z);
            const float prev_pdf = st0.w;
            L = mk(st1.x, st1.y, st1.z);
            const uint32_t prim = __float_as_uint(hit.y);
            uint32_t pixel = a.seq_base + slot / a.k;
            if (a.scenes) {
                const SceneRec& sc = a.scenes[a.ps.slot_scene[slot]];
                pixel = a.ps.slot_pixel[slot];
                max_depth = sc.max_depth;
                rr_depth = sc.rr_depth;
                n_lights = sc.n_lights;
                light_first = sc.light_first;
                root = sc.root;
                inv_nl = n_lights ? 1.0f / (float)n_lights : 0.0f;
            }
            rng.state = a.ps.rng[slot];
            rng.inc = ((uint64_t)pixel << 1u) | 1u;
            if constexpr (kGuided) {
                if (a.rec_cap && slot % a.rec_stride == a.rec_phase) {
                    recorded = true;
                    rec = (slot / a.rec_stride) * a.rec_depth;
                    // radiance gathered before the previous vertex's direction
                    // (its emission and next-event estimate included)
                    if (depth >= 2 && depth - 1 <= a.rec_depth)
                        a.rec_aux[2 * (size_t)(rec + depth - 2)] = make_float4(L.x, L.y, L.z, 1.0f);
                }
            }
            if (prim != 0xFFFFFFFFu) {
                const float4 t0 = a.stris[3 * (size_t)prim], t1 = a.stris[3 * (size_t)prim + 1],
                             t2 = a.stris[3 * (size_t)prim + 2];
                const f3 v0 = ld3(t0), v1 = ld3(t1), v2 = ld3(t2);
                const MatRec m = a.mats[__float_as_uint(t0.w)];
                const int lid = __float_as_int(t1.w);
                ng = normalize(cross(v1 - v0, v2 - v0));
                const float b0 = 1.0f - hit.z - hit.w;
                P = v0 * b0 + v1 * hit.z + v2 * hit.w;
                f3 ns = ng;
                const uint32_t nid = __float_as_uint(t2.w);
                if (nid != KP_INVALID) {
                    // TriangleMesh::post_hit (shape.cpp:615-654): interpolated
                    // shading normal; the shading point moved off the flat
                    // triangle (Hanika, "Hacking the Shadow Terminator", 2021)
                    const float4* nr = a.snorm + 3 * (size_t)nid;
                    const f3 n0 = ld3(nr[0]), n1 = ld3(nr[1]), n2 = ld3(nr[2]);
                    ns = normalize(n0 * b0 + n1 * hit.z + n2 * hit.w);
                    f3 tu = P - v0, tw = P - v1, tx = P - v2;
                    tu = tu - n0 * fminf(0.0f, dot(tu, n0));
                    tw = tw - n1 * fminf(0.0f, dot(tw, n1));
                    tx = tx - n2 * fminf(0.0f, dot(tx, n2));
                    P = P + (tu * b0 + tw * hit.z + tx * hit.w);
                }
                const f3 dir = mk(r1.x, r1.y, r1.z);
                const f3 Le = mk(m.Le[0], m.Le[1], m.Le[2]);
                const float cos_hit = dot(dir, ng);
                if (!is_zero(Le) && cos_hit < 0.0f) {
                    float w = 1.0f;
                    if (depth > 1 && lid >= 0) {
                        const LightRec lr = a.lights[lid];
                        const float pl = inv_nl * (1.0f / lr.area) * (hit.x * hit.x) / (-cos_hit);
                        w = prev_pdf * prev_pdf / (prev_pdf * prev_pdf + pl * pl);
                    }
                    L = L + mul(beta, Le) * w;
                }
                if (depth <= max_depth) {
#pragma unroll
                    for (int i = 0; i < kDraws; ++i) u[i] = rng.nextf();
                    const f3 wo_w = -dir;
                    // two-sided closures see the side the path arrives on; a
                    // transmitting material keeps its normals' orientation
                    nrm = (!(m.flags & KP_MATF_TRANSMIT) && dot(wo_w, ng) < 0.0f) ? -ns : ns;
                    const Frame fr(nrm);
                    wo = fr.to_local(wo_w);
                    bs = bsdf_at(m, a.lobes, beta);
                    if constexpr (kGuided) {
                        const bool eligible = !(m.flags & KP_MATF_TRANSMIT) && m.guide_alpha >= a.min_alpha &&
                                              depth <= a.guide_depth;
                        guided = eligible && a.guide_on;
                        recorded = recorded && eligible;
                    }
                    // next-event estimation (GeometryLight::sample, light.cpp:52-77)
                    if (n_lights > 0) {
                        uint32_t li = (uint32_t)(u[0] * (float)n_lights);
                        li = li < n_lights - 1 ? li : n_lights - 1;
                        const LightRec lr = a.lights[light_first + li];
                        const uint32_t lt = a.light_tris[lr.first + cdf_sample(a.light_cdf + lr.cdf_first, lr.count, u[1])];
                        const f3 la = ld3(a.stris[3 * (size_t)lt]), lb = ld3(a.stris[3 * (size_t)lt + 1]),
                                 lc = ld3(a.stris[3 * (size_t)lt + 2]);
                        const float su0 = sqrtf(u[2]);
                        const float bu = 1.0f - su0, bv = u[3] * su0;
                        const f3 pl = la * bu + lb * bv + lc * (1.0f - bu - bv);
                        const f3 nl_w = normalize(cross(lb - la, lc - la));
                        const f3 org = offset_ray(P, dot(pl - P, ng) > 0.0f ? ng : -ng);
                        const f3 dl = pl - org;
                        const float dist = length(dl);
                        if (dist > 0.0f) {
                            const f3 wi_w = dl * (1.0f / dist);
                            const float cos_l = -dot(wi_w, nl_w);
                            if (cos_l > 0.0f) {
                                const f3 wi = fr.to_local(wi_w);
                                f3 f;
                                float pdf_b;
                                bsdf_eval(bs, wo, wi, f, pdf_b);
                                if (!is_zero(f)) {
                                    nee = true;
                                    pdf_l = inv_nl * (1.0f / lr.area) * (dist * dist) / cos_l;
                                    pdf_bl = pdf_b;
                                    nee_c = mul(mul(beta, f), mk(lr.Le[0], lr.Le[1], lr.Le[2]));
                                    nee_org = org;
                                    nee_dir = wi_w;
                                    nee_dist = dist;
                                    nee_cos = fabsf(wi.y);
                                }
                            }
                        }
                    }
                    beta_pre = beta;
                    alive = true;
                    // Russian roulette (integrator.cpp:491-497)
                    if (depth >= rr_depth) {
                        const float eta_rr = 0.95f;
                        const float qs = fminf(maxc(beta) * eta_rr * eta_rr, 0.99f);
                        if (!(u[4] < qs)) alive = false;
                        else beta = beta * (1.0f / qs);
                    }
                    if (alive) {
                        if constexpr (kGuided) use_guide = guided && !(u[8] < a.bsdf_frac);
                        if (!use_guide) bs_ok = bsdf_sample(bs, wo, u[5], u[6], u[7], bs_wi, bs_f, bs_pdf);
                    }
                }
            }
        }
        // ---- guiding network: one query per wave ---------------------------
        GuideAnswer ga;
        ga.pdf_a = ga.pdf_b = 0.0f;
        ga.dx = ga.dy = 0.0f;
        ga.dz = 1.0f;
        if constexpr (kGuided) {
            const bool want = guided && (nee || use_guide || bs_ok);
            if (__ballot(want)) {
                GuideQuery gq;
                const f3 p01 = mk((P.x - a.net.bmin.x) * a.net.binv.x, (P.y - a.net.bmin.y) * a.net.binv.y,
                                  (P.z - a.net.bmin.z) * a.net.binv.z);
                gq.px = want ? p01.x : 0.0f;
                gq.py = want ? p01.y : 0.0f;
                gq.pz = want ? p01.z : 0.0f;
                gq.nx = want ? nrm.x : 0.0f;
                gq.ny = want ? nrm.y : 1.0f;
                gq.nz = want ? nrm.z : 0.0f;
                gq.ax = nee_dir.x;
                gq.ay = nee_dir.y;
                gq.az = nee_dir.z;
                f3 bw = mk(0, 0, 0);
                if (bs_ok) bw = Frame(nrm).to_world(bs_wi);
                gq.bx = bw.x;
                gq.by = bw.y;
                gq.bz = bw.z;
                gq.sample = want && use_guide;
                gq.u0 = gq.sample ? u[9] : 0.5f;
                gq.u1 = gq.sample ? u[10] : 0.5f;
                gq.u2 = gq.sample ? u[11] : 0.5f;
                g_wave_guide(a.net, glds + (threadIdx.x >> 6) * KP_G_WAVE_ROWS * KP_G_ROW, gq, ga);
            }
        }
        // ---- per-lane finish ---------------------------------------------
        if (active) {
            if (nee) {
                float pdf_b = pdf_bl;
                if (guided) pdf_b = a.bsdf_frac * pdf_bl + (1.0f - a.bsdf_frac) * ga.pdf_a;
                const float w = pdf_l * pdf_l / (pdf_l * pdf_l + pdf_b * pdf_b);
                const f3 c = nee_c * (nee_cos * w / pdf_l);
                shad = true;
                sr0 = make_float4(nee_org.x, nee_org.y, nee_org.z, 0.0f);
                sr1 = make_float4(nee_dir.x, nee_dir.y, nee_dir.z, nee_dist * 0.9999f);
                sc = make_float4(c.x, c.y, c.z, __uint_as_float(slot));
            }
            if (alive) {
                const Frame fr(nrm);
                f3 wi = bs_wi, f = bs_f;
                float pdf = bs_pdf;
                bool ok = bs_ok;
                if (use_guide) {
                    wi = fr.to_local(mk(ga.dx, ga.dy, ga.dz));
                    float pdf_bsdf;
                    bsdf_eval(bs, wo, wi, f, pdf_bsdf);
                    pdf = a.bsdf_frac * pdf_bsdf + (1.0f - a.bsdf_frac) * ga.pdf_b;
                    ok = !is_zero(f) && pdf > 0.0f;
                } else if (guided && ok) {
                    pdf = a.bsdf_frac * bs_pdf + (1.0f - a.bsdf_frac) * ga.pdf_b;
                }
                if (ok) {
                    beta = mul(beta, f) * (fabsf(wi.y) / pdf);
                    if (!is_zero(beta)) {
                        const f3 wi_w = fr.to_world(wi);
                        const f3 o = offset_ray(P, dot(wi_w, ng) > 0.0f ? ng : -ng);
                        nr0 = make_float4(o.x, o.y, o.z, 0.0f);
                        nr1 = make_float4(wi_w.x, wi_w.y, wi_w.z, __int_as_float(0x7f800000));
                        a.ps.s0[slot] = make_float4(beta.x, beta.y, beta.z, pdf);
                        cont = true;
                        if constexpr (kGuided) {
                            if (recorded) {
                                const uint32_t r = rec + depth - 1;
                                const f3 p01 = mk((P.x - a.net.bmin.x) * a.net.binv.x,
                                                  (P.y - a.net.bmin.y) * a.net.binv.y,
                                                  (P.z - a.net.bmin.z) * a.net.binv.z);
                                float* in6 = a.rec_in6 + 6 * (size_t)r;
                                in6[0] = p01.x; in6[1] = p01.y; in6[2] = p01.z;
                                in6[3] = nrm.x; in6[4] = nrm.y; in6[5] = nrm.z;
                                float* dr = a.rec_dirs + 3 * (size_t)r;
                                dr[0] = wi_w.x; dr[1] = wi_w.y; dr[2] = wi_w.z;
                                a.rec_aux[2 * (size_t)r + 1] = make_float4(beta_pre.x, beta_pre.y, beta_pre.z, 1.0f);
                            }
                        }
                    }
                }
            }
            a.ps.s1[slot] = make_float4(L.x, L.y, L.z, 0.0f);
            a.ps.rng[slot] = rng.state;
        }
        uint32_t so;
        const uint32_t qo = block_append2(cont, shad, &a.ps.qc(cur ^ 1), &a.ps.qc(2), app_lds, so);
        if (cont) {
            rout[2 * (size_t)qo] = nr0;
            rout[2 * (size_t)qo + 1] = nr1;
            qout[qo] = slot;
            if (a.scenes) a.ps.qroot[cur ^ 1][qo] = root;
        }
        if (shad) {
            a.ps.srays[2 * (size_t)so] = sr0;
            a.ps.srays[2 * (size_t)so + 1] = sr1;
            a.ps.scontrib[so] = sc;
            if (a.scenes) a.ps.sroot[so] = root;
        }
    }
}

__global__ __launch_bounds__(256) void k_connect(RenderArgs a) {
    const uint32_t n = a.ps.qc(2);
    const uint32_t total = gridDim.x * 256;
    for (uint32_t s = blockIdx.x * 256 + threadIdx.x; s < n; s += total) {
        if (a.ps.socc[s]) continue;
        const float4 c = a.ps.scontrib[s];
        const uint32_t slot = __float_as_uint(c.w);
        float4 l = a.ps.s1[slot];
        l.x += c.x;
        l.y += c.y;
        l.z += c.z;
        a.ps.s1[slot] = l;
    }
}

// Turns a pass's records into training samples.  Record i belongs to slot
// (i / rec_depth) * rec_stride + rec_phase and vertex i % rec_depth + 1; its
// weight is the luminance of (radiance at the end of the path - radiance
// before the vertex's sampled direction) / throughput before the vertex, an
// estimate of f cos L_i / p for the sampled direction.  With film != nullptr
// the weight is instead the luminance of the radiance gathered after the
// vertex (throughput included) over sqrt(R^2 + 0.01), R the luminance of the
// pixel's current estimate (film sum x inv_count): the sample's share of the
// per-pixel relative error.  Records that were not completed get weight 0 and
// finite inputs.
__global__ __launch_bounds__(256) void k_guide_records(RenderArgs a, uint32_t n_rec, float* weights,
                                                       const float4* film, float inv_count) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n_rec) return;
    const uint32_t slot = (i / a.rec_depth) * a.rec_stride + a.rec_phase;
    const float4 x0 = a.rec_aux[2 * (size_t)i], x1 = a.rec_aux[2 * (size_t)i + 1];
    float w = 0.0f;
    const bool made = x0.w == 1.0f && x1.w == 1.0f && slot < a.n_paths;
    if (made) {
        const float4 le = a.ps.s1[slot];
        if (film) {
            const float4 f = film[slot / a.k];
            const float R = (0.2126f * f.x + 0.7152f * f.y + 0.0722f * f.z) * inv_count;
            w = (0.2126f * (le.x - x0.x) + 0.7152f * (le.y - x0.y) + 0.0722f * (le.z - x0.z)) /
                sqrtf(R * R + 0.01f);
        } else {
            if (x1.x > 0.0f) w += 0.2126f * (le.x - x0.x) / x1.x;
            if (x1.y > 0.0f) w += 0.7152f * (le.y - x0.y) / x1.y;
            if (x1.z > 0.0f) w += 0.0722f * (le.z - x0.z) / x1.z;
        }
        if (!(w > 0.0f) || !isfinite(w)) w = 0.0f;
    } else {
        float* in6 = a.rec_in6 + 6 * (size_t)i;
        in6[0] = in6[1] = in6[2] = in6[3] = in6[5] = 0.0f;
        in6[4] = 1.0f;
        float* dr = a.rec_dirs + 3 * (size_t)i;
        dr[0] = dr[2] = 0.0f;
        dr[1] = 1.0f;
    }
    weights[i] = w;
}

// Folds the k path radiances of every pixel into the film, in sample order,
// with Kahan compensation.
__global__ __launch_bounds__(256) void k_film(RenderArgs a, float4* sum, float4* comp) {
#pragma clang fp contract(off)
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.npix) return;
    float4 s = sum[p], c = comp[p];
    for (uint32_t j = 0; j < a.k; ++j) {
        const float4 x = a.ps.s1[(size_t)p * a.k + j];
        float y, t;
        y = x.x - c.x; t = s.x + y; c.x = (t - s.x) - y; s.x = t;
        y = x.y - c.y; t = s.y + y; c.y = (t - s.y) - y; s.y = t;
        y = x.z - c.z; t = s.z + y; c.z = (t - s.z) - y; s.z = t;
    }
    sum[p] = s;
    comp[p] = c;
}

__global__ void k_film_resolve(const float4* sum, const float4* comp, uint32_t npix, float inv_count,
                               float* out) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
    const float4 s = sum[p], c = comp[p];
    out[3 * (size_t)p] = (s.x - c.x) * inv_count;
    out[3 * (size_t)p + 1] = (s.y - c.y) * inv_count;
    out[3 * (size_t)p + 2] = (s.z - c.z) * inv_count;
}

__global__ void k_zero_f4(float4* p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

hipError_t shade_occupancy(int* blocks_per_cu_plain, int* blocks_per_cu_guided) {
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu_plain, k_shade<false>, 256, 0);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu_guided, k_shade<true>, 256, 0);
}

}  // namespace kp

// =================================================================== host ==
struct RenderState {
    bool scene_ready = false;
    std::vector<kp::MatRec> mats;    // materials (kp_material_add*), lobes in `lobes`
    std::vector<kp::Lobe> lobes;
    // device scene
    float4* stris = nullptr;
    float4* d_snorm = nullptr;       // shading-normal records (3 float4 each)
    uint32_t n_snorm = 0;
    kp::MatRec* d_mats = nullptr;
    kp::Lobe* d_lobes = nullptr;
    kp::LightRec* d_lights = nullptr;
    uint32_t* d_light_tris = nullptr;
    float* d_light_cdf = nullptr;
    uint32_t n_lights = 0, n_light_tris = 0, n_cdf = 0;
    float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {0, 0, 0};
    // paths
    uint32_t cap_paths = 0;
    void* path_mem = nullptr;
    kp::PathState ps{};
    // film
    uint32_t fw = 0, fh = 0;
    uint64_t film_count = 0;
    float4* film_sum = nullptr;
    float4* film_comp = nullptr;
    float* film_out = nullptr;
    kp_render_stats stats{};
};

namespace kp {

static void free_scene(RenderState* r) {
    if (r->stris) (void)hipFree(r->stris);
    if (r->d_snorm) (void)hipFree(r->d_snorm);
    if (r->d_lobes) (void)hipFree(r->d_lobes);
    r->d_snorm = nullptr;
    r->n_snorm = 0;
    r->d_lobes = nullptr;
    if (r->d_mats) (void)hipFree(r->d_mats);
    if (r->d_lights) (void)hipFree(r->d_lights);
    if (r->d_light_tris) (void)hipFree(r->d_light_tris);
    if (r->d_light_cdf) (void)hipFree(r->d_light_cdf);
    r->stris = nullptr;
    r->d_mats = nullptr;
    r->d_lights = nullptr;
    r->d_light_tris = nullptr;
    r->d_light_cdf = nullptr;
    r->scene_ready = false;
}

void render_state_free(kp_ctx* c) {
    RenderState* r = c->render;
    if (!r) return;
    free_scene(r);
    if (r->path_mem) (void)hipFree(r->path_mem);
    if (r->film_sum) (void)hipFree(r->film_sum);
    if (r->film_comp) (void)hipFree(r->film_comp);
    if (r->film_out) (void)hipFree(r->film_out);
    delete r;
    c->render = nullptr;
}

static RenderState* rstate(kp_ctx* c) {
    if (!c->render) c->render = new RenderState();
    return c->render;
}

void render_scene_dirty(kp_ctx* c) {
    if (c->render) c->render->scene_ready = false;
}

static float tri_area(const float* t) {
    const float e1[3] = {t[3] - t[0], t[4] - t[1], t[5] - t[2]};
    const float e2[3] = {t[6] - t[0], t[7] - t[1], t[8] - t[2]};
    const float cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                cz = e1[0] * e2[1] - e1[1] * e2[0];
    return 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
}

// World-space unit vertex normals of triangle t of placed mesh mi (the
// kp_mesh_add meshes, then one entry per instance), if that mesh has them.
// Instances carry their prototype's normals under the inverse transpose of
// the object-to-world matrix, i.e. the transpose of the world-to-object one.
static bool shading_normals(const kp_ctx* c, size_t mi, uint32_t t, float* out9) {
    const size_t n_mesh = c->mesh_first.size();
    if (mi < n_mesh) {
        const size_t p = (size_t)c->mesh_first[mi] + t;
        if (p >= c->soup_has_n.size() || !c->soup_has_n[p]) return false;
        std::memcpy(out9, &c->soup_n[9 * p], 9 * sizeof(float));
        return true;
    }
    const size_t k = mi - n_mesh;
    const uint32_t b = c->inst_blas[k];
    const size_t p = (size_t)c->blas_first[b] + t;
    if (p >= c->blas_has_n.size() || !c->blas_has_n[p]) return false;
    const float* src = &c->blas_n[9 * p];
    const float* w = &c->inst_w2o[12 * k];
    for (int v = 0; v < 3; ++v) {
        double n[3], len = 0.0;
        for (int a = 0; a < 3; ++a) {
            n[a] = (double)w[a] * src[3 * v] + (double)w[4 + a] * src[3 * v + 1] + (double)w[8 + a] * src[3 * v + 2];
            len += n[a] * n[a];
        }
        len = std::sqrt(len);
        for (int a = 0; a < 3; ++a) out9[3 * v + a] = len > 0.0 ? (float)(n[a] / len) : 0.0f;
    }
    return true;
}

// Uploads primitives, materials and the light table (one area light per
// emissive mesh, sampled by area: TriangleMesh::setup_dpdf, shape.cpp:700-709).
static int prepare_scene(kp_ctx* c) {
    RenderState* r = rstate(c);
    if (r->scene_ready) return KP_OK;
    if (!c->built) return fail(c, KP_ERR_STATE, "render before kp_bvh_build");
    if (r->mats.empty()) return fail(c, KP_ERR_STATE, "render without materials (kp_material_add)");
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    free_scene(r);
    const uint32_t n = c->n_tris_built;
    const bool has_inst = !c->inst_blas.empty();
    if (has_inst && n >= (1u << 28))
        return fail(c, KP_ERR_INVALID, "render: instanced scene exceeds 2^28 placed triangles");
    // World-space triangles of every primitive id (instances baked the same
    // way whether traversal is baked or two-level: shading records are per
    // placed triangle; only traversal can be shared).
    std::vector<float> placed;
    if (has_inst) bake_placed(c, placed);
    const float* world = has_inst ? placed.data() : c->soup.data();
    if (has_inst && placed.size() != 9 * (size_t)n) return fail(c, KP_ERR_STATE, "render: BVH out of date");
    // meshes as (first primitive, count, material): kp_mesh_add meshes, then instances
    std::vector<uint32_t> g_first(c->mesh_first), g_count(c->mesh_count), g_mat(c->mesh_material);
    if (has_inst) {
        uint32_t next = (uint32_t)(c->soup.size() / 9);
        for (size_t k = 0; k < c->inst_blas.size(); ++k) {
            const uint32_t b = c->inst_blas[k];
            g_first.push_back(next);
            g_count.push_back(c->blas_count[b]);
            g_mat.push_back(c->inst_material[k] != KP_INVALID ? c->inst_material[k] : c->blas_material[b]);
            next += c->blas_count[b];
        }
    }
    std::vector<float4> st(3 * (size_t)n);
    std::vector<float4> snorm;   // shading-normal records, in primitive order
    std::vector<uint32_t> light_tris;
    std::vector<float> cdf;
    std::vector<LightRec> lights;
    for (size_t mi = 0; mi < g_first.size(); ++mi) {
        const uint32_t mat = g_mat[mi];
        if (mat >= r->mats.size()) return fail(c, KP_ERR_INVALID, "mesh material id out of range");
        const MatRec& m = r->mats[mat];
        const bool emissive = m.Le[0] != 0.0f || m.Le[1] != 0.0f || m.Le[2] != 0.0f;
        int32_t lid = -1;
        if (emissive) {
            lid = (int32_t)lights.size();
            LightRec lr{};
            lr.first = (uint32_t)light_tris.size();
            lr.count = g_count[mi];
            lr.cdf_first = (uint32_t)cdf.size();
            cdf.push_back(0.0f);
            for (uint32_t t = 0; t < lr.count; ++t) {
                const uint32_t p = g_first[mi] + t;
                light_tris.push_back(p);
                cdf.push_back(cdf.back() + tri_area(&world[9 * (size_t)p]));
            }
            const float sum = cdf.back();
            lr.area = sum;
            if (!(sum > 0.0f)) return fail(c, KP_ERR_INVALID, "emissive mesh with zero area");
            const float inv = 1.0f / sum;
            for (uint32_t t = 1; t <= lr.count; ++t) cdf[lr.cdf_first + t] *= inv;
            cdf[lr.cdf_first + lr.count] = 1.0f;
            lr.Le[0] = m.Le[0];
            lr.Le[1] = m.Le[1];
            lr.Le[2] = m.Le[2];
            lights.push_back(lr);
        }
        for (uint32_t t = 0; t < g_count[mi]; ++t) {
            const uint32_t p = g_first[mi] + t;
            const float* v = &world[9 * (size_t)p];
            uint32_t matbits = mat, nid = KP_INVALID;
            float nrm[9];
            if (shading_normals(c, mi, t, nrm)) {
                nid = (uint32_t)(snorm.size() / 3);
                for (int k = 0; k < 3; ++k) snorm.push_back(make_float4(nrm[3 * k], nrm[3 * k + 1], nrm[3 * k + 2], 0.0f));
            }
            float fm, fl, fn;
            std::memcpy(&fm, &matbits, 4);
            std::memcpy(&fl, &lid, 4);
            std::memcpy(&fn, &nid, 4);
            st[3 * (size_t)p] = make_float4(v[0], v[1], v[2], fm);
            st[3 * (size_t)p + 1] = make_float4(v[3], v[4], v[5], fl);
            st[3 * (size_t)p + 2] = make_float4(v[6], v[7], v[8], fn);
        }
    }
    for (int a = 0; a < 3; ++a) {
        r->scene_lo[a] = INFINITY;
        r->scene_hi[a] = -INFINITY;
    }
    for (size_t i = 0; i < 3 * (size_t)n; ++i)
        for (int a = 0; a < 3; ++a) {
            r->scene_lo[a] = std::fmin(r->scene_lo[a], world[3 * i + a]);
            r->scene_hi[a] = std::fmax(r->scene_hi[a], world[3 * i + a]);
        }
    KP_HIP_TRY(c, hipMalloc((void**)&r->stris, st.size() * sizeof(float4)));
    KP_HIP_TRY(c, hipMemcpy(r->stris, st.data(), st.size() * sizeof(float4), hipMemcpyHostToDevice));
    KP_HIP_TRY(c, hipMalloc((void**)&r->d_mats, r->mats.size() * sizeof(MatRec)));
    KP_HIP_TRY(c, hipMemcpy(r->d_mats, r->mats.data(), r->mats.size() * sizeof(MatRec), hipMemcpyHostToDevice));
    KP_HIP_TRY(c, hipMalloc((void**)&r->d_lobes, r->lobes.size() * sizeof(Lobe)));
    KP_HIP_TRY(c, hipMemcpy(r->d_lobes, r->lobes.data(), r->lobes.size() * sizeof(Lobe), hipMemcpyHostToDevice));
    r->n_snorm = (uint32_t)(snorm.size() / 3);
    if (!snorm.empty()) {
        KP_HIP_TRY(c, hipMalloc((void**)&r->d_snorm, snorm.size() * sizeof(float4)));
        KP_HIP_TRY(c, hipMemcpy(r->d_snorm, snorm.data(), snorm.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
    r->n_lights = (uint32_t)lights.size();
    r->n_light_tris = (uint32_t)light_tris.size();
    r->n_cdf = (uint32_t)cdf.size();
    if (!lights.empty()) {
        KP_HIP_TRY(c, hipMalloc((void**)&r->d_lights, lights.size() * sizeof(LightRec)));
        KP_HIP_TRY(c, hipMemcpy(r->d_lights, lights.data(), lights.size() * sizeof(LightRec), hipMemcpyHostToDevice));
        KP_HIP_TRY(c, hipMalloc((void**)&r->d_light_tris, light_tris.size() * 4));
        KP_HIP_TRY(c, hipMemcpy(r->d_light_tris, light_tris.data(), light_tris.size() * 4, hipMemcpyHostToDevice));
        KP_HIP_TRY(c, hipMalloc((void**)&r->d_light_cdf, cdf.size() * 4));
        KP_HIP_TRY(c, hipMemcpy(r->d_light_cdf, cdf.data(), cdf.size() * 4, hipMemcpyHostToDevice));
    }
    r->scene_ready = true;
    return KP_OK;
}

// Grows a path-state allocation to at least n slots.
static int alloc_paths(kp_ctx* c, PathState& ps, void*& mem, uint32_t& cap, uint32_t n) {
    if (cap >= n && mem) return KP_OK;
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (mem) (void)hipFree(mem);
    mem = nullptr;
    cap = 0;
    // bytes per slot: 2x32 rays + 2x4 qmap + 16 hit + 16 s0 + 16 s1 + 8 rng
    //                 + 32 shadow ray + 16 contrib + 4 occl
    //                 + 5x4 batch (scene, pixel, roots of the three queues)
    const size_t N = n;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t o_r0 = take(N * 32), o_r1 = take(N * 32), o_q0 = take(N * 4), o_q1 = take(N * 4);
    const size_t o_h = take(N * 16), o_s0 = take(N * 16), o_s1 = take(N * 16), o_rng = take(N * 8);
    const size_t o_sr = take(N * 32), o_sc = take(N * 16), o_so = take(N * 4), o_cnt = take(3 * 256),
                 o_tot = take(64);
    const size_t o_ss = take(N * 4), o_sp = take(N * 4), o_qr0 = take(N * 4), o_qr1 = take(N * 4),
                 o_sr2 = take(N * 4);
    KP_HIP_TRY(c, hipMalloc(&mem, off));
    char* b = (char*)mem;
    ps.rays[0] = (float4*)(b + o_r0);
    ps.rays[1] = (float4*)(b + o_r1);
    ps.qmap[0] = (uint32_t*)(b + o_q0);
    ps.qmap[1] = (uint32_t*)(b + o_q1);
    ps.hits = (float4*)(b + o_h);
    ps.s0 = (float4*)(b + o_s0);
    ps.s1 = (float4*)(b + o_s1);
    ps.rng = (uint64_t*)(b + o_rng);
    ps.srays = (float4*)(b + o_sr);
    ps.scontrib = (float4*)(b + o_sc);
    ps.socc = (uint32_t*)(b + o_so);
    ps.cnt = (uint32_t*)(b + o_cnt);
    ps.totals = (unsigned long long*)(b + o_tot);
    ps.slot_scene = (uint32_t*)(b + o_ss);
    ps.slot_pixel = (uint32_t*)(b + o_sp);
    ps.qroot[0] = (uint32_t*)(b + o_qr0);
    ps.qroot[1] = (uint32_t*)(b + o_qr1);
    ps.sroot = (uint32_t*)(b + o_sr2);
    // on the context's stream (created non-blocking: null-stream work would
    // not be ordered with its kernels), never inside a captured graph
    KP_HIP_TRY(c, hipMemsetAsync(ps.cnt, 0, 3 * 256, c->stream));
    KP_HIP_TRY(c, hipMemsetAsync(ps.totals, 0, 64, c->stream));
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    cap = n;
    return KP_OK;
}

static int ensure_paths(kp_ctx* c, uint32_t n) {
    RenderState* r = rstate(c);
    return alloc_paths(c, r->ps, r->path_mem, r->cap_paths, n);
}

static int film_alloc(kp_ctx* c, uint32_t w, uint32_t h) {
    RenderState* r = rstate(c);
    const size_t npix = (size_t)w * h;
    if (r->fw != w || r->fh != h || !r->film_sum) {
        KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (r->film_sum) (void)hipFree(r->film_sum);
        if (r->film_comp) (void)hipFree(r->film_comp);
        if (r->film_out) (void)hipFree(r->film_out);
        r->film_sum = r->film_comp = nullptr;
        r->film_out = nullptr;
        KP_HIP_TRY(c, hipMalloc((void**)&r->film_sum, npix * sizeof(float4)));
        KP_HIP_TRY(c, hipMalloc((void**)&r->film_comp, npix * sizeof(float4)));
        KP_HIP_TRY(c, hipMalloc((void**)&r->film_out, npix * 3 * sizeof(float)));
        r->fw = w;
        r->fh = h;
    }
    const uint32_t blocks = (uint32_t)((npix + 255) / 256);
    hipLaunchKernelGGL(k_zero_f4, dim3(blocks), dim3(256), 0, c->stream, r->film_sum, (uint32_t)npix);
    hipLaunchKernelGGL(k_zero_f4, dim3(blocks), dim3(256), 0, c->stream, r->film_comp, (uint32_t)npix);
    KP_HIP_TRY(c, hipGetLastError());
    r->film_count = 0;
    return KP_OK;
}

static void camera_basis(const kp_render_desc& d, RenderArgs& a) {
#pragma clang fp contract(off)
    // Same float operations as the oracle camera: fwd, right = fwd x up, upv.
    auto nrm = [](float x, float y, float z, float* o) {
        const float l = std::sqrt(x * x + y * y + z * z);
        o[0] = x / l; o[1] = y / l; o[2] = z / l;
    };
    float f[3], rr[3], u[3];
    nrm(d.lookat[0] - d.eye[0], d.lookat[1] - d.eye[1], d.lookat[2] - d.eye[2], f);
    const float cx = f[1] * d.up[2] - f[2] * d.up[1], cy = f[2] * d.up[0] - f[0] * d.up[2],
                cz = f[0] * d.up[1] - f[1] * d.up[0];
    nrm(cx, cy, cz, rr);
    u[0] = rr[1] * f[2] - rr[2] * f[1];
    u[1] = rr[2] * f[0] - rr[0] * f[2];
    u[2] = rr[0] * f[1] - rr[1] * f[0];
    a.eye = make_float3(d.eye[0], d.eye[1], d.eye[2]);
    a.fwd = make_float3(f[0], f[1], f[2]);
    a.right = make_float3(rr[0], rr[1], rr[2]);
    a.upv = make_float3(u[0], u[1], u[2]);
    a.tan_half = d.tan_half_fov;
    a.aspect = (float)d.width / (float)d.height;
}

static uint32_t samples_per_pass(uint32_t npix, uint32_t spp, uint64_t target) {
    uint32_t best = 1;
    for (uint32_t k = 1; k <= spp; ++k)
        if (spp % k == 0 && (uint64_t)npix * k <= target) best = k;
    return best;
}

// Traversal and launch-shape settings of the integrator's launches.
struct TraceCfg {
    TraceArgs base;            // BVH, work counter, error word, spill, tuning
    uint32_t grid_closest, grid_occluded, shade_blocks, shade_blocks_guided;
    bool roots;                // per-ray roots from the path state (batched scenes)
};

static TraceCfg ctx_trace_cfg(kp_ctx* c) {
    TraceCfg t;
    std::memset(&t, 0, sizeof(t));
    t.base = trace_base(c);
    // the 5-wave traversal builds, one-level or two-level
    t.base.five_waves = 1;
    t.grid_closest = c->grid_closest5;
    t.grid_occluded = c->grid_occluded5;
    // as many blocks as stay resident: every wave strides over the queue in
    // one round, no second partial round of blocks
    t.shade_blocks = c->shade_blocks;
    t.shade_blocks_guided = c->shade_blocks_guided;
    t.roots = false;
    return t;
}

// Extend / shade / connect for max_depth + 1 bounces of the paths that the
// generate kernel has queued (queue sizes live on the device).  Only kernel
// launches: the sequence can be captured into a graph.
static int run_bounces(kp_ctx* c, const RenderArgs& a, bool guided, const TraceCfg& tc, uint32_t max_depth) {
    int cur = 0;
    for (uint32_t depth = 1; depth <= max_depth + 1; ++depth) {
        TraceArgs ta = tc.base;
        ta.rays = a.ps.rays[cur];
        ta.hits = a.ps.hits;
        ta.n = a.n_paths;
        ta.n_dev = &a.ps.qc(cur);
        ta.roots = tc.roots ? a.ps.qroot[cur] : nullptr;
        KP_HIP_TRY(c, launch_closest(ta, tc.grid_closest, false, c->stream));
        hipLaunchKernelGGL(k_bookkeep, dim3(1), dim3(64), 0, c->stream, a, cur);
        if (guided)
            hipLaunchKernelGGL(k_shade<true>, dim3(tc.shade_blocks_guided), dim3(256), 0, c->stream, a, cur, depth);
        else
            hipLaunchKernelGGL(k_shade<false>, dim3(tc.shade_blocks), dim3(256), 0, c->stream, a, cur, depth);
        KP_HIP_TRY(c, hipGetLastError());
        if (depth <= max_depth) {
            ta.rays = a.ps.srays;
            ta.hits = nullptr;
            ta.occl = a.ps.socc;
            ta.n_dev = &a.ps.qc(2);
            ta.roots = tc.roots ? a.ps.sroot : nullptr;
            KP_HIP_TRY(c, launch_occluded(ta, tc.grid_occluded, c->stream));
            hipLaunchKernelGGL(k_count_shadow, dim3(1), dim3(64), 0, c->stream, a);
            hipLaunchKernelGGL(k_connect, dim3(tc.shade_blocks), dim3(256), 0, c->stream, a);
            KP_HIP_TRY(c, hipGetLastError());
        }
        cur ^= 1;
    }
    return KP_OK;
}

static RenderArgs base_render_args(kp_ctx* c) {
    RenderState* r = rstate(c);
    RenderArgs a{};
    a.stris = r->stris;
    a.snorm = r->d_snorm;
    a.mats = r->d_mats;
    a.lobes = r->d_lobes;
    a.lights = r->d_lights;
    a.light_tris = r->d_light_tris;
    a.light_cdf = r->d_light_cdf;
    a.n_lights = r->n_lights;
    a.ps = r->ps;
    return a;
}

// Guided-shading fields of the render arguments: network, position box,
// strategy fractions; sampling on, no records.
static int guide_args(kp_ctx* c, RenderArgs& a) {
    GuideState* g = c->guide;
    if (!g) return fail(c, KP_ERR_STATE, "guided render before kp_guiding_init");
    RenderState* r = rstate(c);
    if (g->auto_box) guide_set_box(c, r->scene_lo, r->scene_hi);
    a.net = guide_net(c);
    a.bsdf_frac = g->bsdf_fraction;
    a.min_alpha = g->min_alpha;
    a.guide_on = 1;
    a.guide_depth = g->max_guided_depth ? g->max_guided_depth : 0xFFFFFFFFu;
    a.rec_cap = 0;
    a.rec_stride = 1;
    a.rec_phase = 0;
    a.rec_depth = 1;
    a.rec_in6 = g->in6;
    a.rec_dirs = g->dirs;
    a.rec_aux = g->rec_aux;
    return KP_OK;
}

// Traces one path per caller ray; writes the path radiance (RGB) per ray.
static int integrate_rays(kp_ctx* c, const kp_ray* rays, uint32_t n, const kp_render_desc* d,
                          float* out, int kind) {
    if (d->max_depth == 0 || d->max_depth > 64) return fail(c, KP_ERR_INVALID, "max_depth must be 1..64");
    const bool guided = (d->flags & KP_RENDER_GUIDED) != 0;
    if (guided && !c->guide) return fail(c, KP_ERR_STATE, "guided render before kp_guiding_init");
    int rc = prepare_scene(c);
    if (rc) return rc;
    RenderState* r = rstate(c);
    const uint32_t chunk_cap = c->pass_paths;
    const uint32_t cap = n < chunk_cap ? n : chunk_cap;
    if ((rc = ensure_paths(c, cap))) return rc;
    float4* d_in = nullptr;
    float* d_out = nullptr;
    if (kind == KP_PTR_HOST) {
        if ((rc = ensure_buffer(c, &c->d_stage_in, &c->stage_in_bytes, (size_t)cap * 32))) return rc;
        if ((rc = ensure_buffer(c, &c->d_stage_out, &c->stage_out_bytes, (size_t)cap * 12))) return rc;
        d_in = (float4*)c->d_stage_in;
        d_out = (float*)c->d_stage_out;
    }
    unsigned long long zero2[2] = {0, 0};
    KP_HIP_TRY(c, hipMemcpyAsync(r->ps.totals, zero2, 16, hipMemcpyHostToDevice, c->stream));
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t off = 0; off < n; off += cap) {
        const uint32_t cnt = (n - off) < cap ? (n - off) : cap;
        RenderArgs a = base_render_args(c);
        a.k = 1;
        a.n_paths = cnt;
        a.seq_base = off;
        a.s0 = d->spp_offset;
        a.seed = d->seed;
        a.max_depth = d->max_depth;
        a.rr_depth = d->rr_depth;
        if (guided && (rc = guide_args(c, a))) return rc;
        const float4* src = (const float4*)(rays) + 2 * (size_t)off;
        float* dst = out + 3 * (size_t)off;
        if (kind == KP_PTR_HOST) {
            KP_HIP_TRY(c, hipMemcpyAsync(d_in, src, (size_t)cnt * 32, hipMemcpyHostToDevice, c->stream));
            src = d_in;
            dst = d_out;
        }
        hipLaunchKernelGGL(k_generate_rays, dim3((cnt + 255) / 256), dim3(256), 0, c->stream, a, src);
        KP_HIP_TRY(c, hipGetLastError());
        if ((rc = run_bounces(c, a, guided, ctx_trace_cfg(c), a.max_depth))) return rc;
        hipLaunchKernelGGL(k_write_radiance, dim3((cnt + 255) / 256), dim3(256), 0, c->stream, a, dst);
        KP_HIP_TRY(c, hipGetLastError());
        if (kind == KP_PTR_HOST)
            KP_HIP_TRY(c, hipMemcpyAsync(out + 3 * (size_t)off, d_out, (size_t)cnt * 12,
                                         hipMemcpyDeviceToHost, c->stream));
    }
    if ((rc = kp_synchronize(c))) return rc;
    auto t1 = std::chrono::steady_clock::now();
    unsigned long long tot[2];
    KP_HIP_TRY(c, hipMemcpy(tot, r->ps.totals, 16, hipMemcpyDeviceToHost));
    r->stats.paths = n;
    r->stats.extension_rays = tot[0];
    r->stats.shadow_rays = tot[1];
    r->stats.seconds = std::chrono::duration<double>(t1 - t0).count();
    r->stats.passes = (n + cap - 1) / cap;
    r->stats.paths_per_pass = cap;
    r->stats.train_steps = 0;
    r->stats.guided_passes = guided ? r->stats.passes : 0;
    r->stats.train_records = 0;
    return KP_OK;
}

static int render_accumulate(kp_ctx* c, const kp_render_desc* d) {
    RenderState* r = rstate(c);
    if (d->width != r->fw || d->height != r->fh)
        return fail(c, KP_ERR_STATE, "film size differs from the render size (kp_film_reset first)");
    if (d->max_depth == 0 || d->max_depth > 64) return fail(c, KP_ERR_INVALID, "max_depth must be 1..64");
    const bool guided = (d->flags & KP_RENDER_GUIDED) != 0;
    GuideState* g = c->guide;
    if (guided && !g) return fail(c, KP_ERR_STATE, "guided render before kp_guiding_init");
    // records per recorded path: the guided vertices 1 .. min(max_depth, max_guided_depth)
    const uint32_t rec_depth = guided && g->max_guided_depth ? std::min(d->max_depth, g->max_guided_depth)
                                                             : d->max_depth;
    if (guided && (g->train_steps || g->late_train_steps) && rec_depth > g->max_records)
        return fail(c, KP_ERR_INVALID, "guiding max_records must be >= the recorded depth");
    int rc = prepare_scene(c);
    if (rc) return rc;
    const uint32_t npix = d->width * d->height;
    const uint32_t k = samples_per_pass(npix, d->spp, guided ? g->pass_paths : c->pass_paths);
    const uint32_t n_paths = npix * k;
    if ((rc = ensure_paths(c, n_paths))) return rc;
    RenderArgs a{};
    a.stris = r->stris;
    a.snorm = r->d_snorm;
    a.mats = r->d_mats;
    a.lobes = r->d_lobes;
    a.lights = r->d_lights;
    a.light_tris = r->d_light_tris;
    a.light_cdf = r->d_light_cdf;
    a.n_lights = r->n_lights;
    a.width = d->width;
    a.height = d->height;
    a.npix = npix;
    a.k = k;
    a.n_paths = n_paths;
    a.seed = d->seed;
    a.max_depth = d->max_depth;
    a.rr_depth = d->rr_depth;
    camera_basis(*d, a);
    a.ps = r->ps;
    if (guided && (rc = guide_args(c, a))) return rc;
    // training records: every rec_stride-th path (phase rotating per pass),
    // rec_depth records each, at most max_records per pass
    uint32_t stride = 1, rec_slots_max = 0;
    const bool train_any = guided && (g->train_steps > 0 || g->late_train_steps > 0);
    a.rec_depth = rec_depth;
    if (train_any) {
        const uint64_t D = rec_depth;
        stride = (uint32_t)std::max<uint64_t>(1, ((uint64_t)n_paths * D + g->max_records - 1) / g->max_records);
        while (((uint64_t)n_paths + stride - 1) / stride * D > g->max_records) ++stride;
        rec_slots_max = (uint32_t)(((uint64_t)n_paths + stride - 1) / stride);
    }
    const uint32_t gen_blocks = (n_paths + 255) / 256;
    uint32_t steps = 0, guided_passes = 0;
    bool train = false;
    uint32_t pass_steps = 0;
    auto t0 = std::chrono::steady_clock::now();
    unsigned long long zero2[2] = {0, 0};
    KP_HIP_TRY(c, hipMemcpyAsync(r->ps.totals, zero2, 16, hipMemcpyHostToDevice, c->stream));
    for (uint32_t pass = 0; pass < d->spp / k; ++pass) {
        a.s0 = d->spp_offset + pass * k;
        uint32_t n_rec = 0;
        if (guided) {
            a.guide_on = g->train_passes >= g->guide_after ? 1u : 0u;
            guided_passes += a.guide_on;
            a.rec_cap = 0;
            // train_steps per pass for the first train_passes passes (all
            // passes if 0), late_train_steps per pass after them
            const bool early = g->max_train_passes == 0 || g->train_passes < g->max_train_passes;
            pass_steps = early ? g->train_steps : g->late_train_steps;
            train = train_any && pass_steps > 0;
            if (train) {
                a.rec_stride = stride;
                a.rec_phase = g->train_passes % stride;
                const uint32_t slots = (n_paths - a.rec_phase + stride - 1) / stride;
                n_rec = std::min(slots, rec_slots_max) * rec_depth;
                a.rec_cap = n_rec;
                hipLaunchKernelGGL(k_zero_f4, dim3((2 * n_rec + 255) / 256), dim3(256), 0, c->stream, a.rec_aux,
                                   2 * n_rec);
            }
        }
        hipLaunchKernelGGL(k_generate, dim3(gen_blocks), dim3(256), 0, c->stream, a);
        KP_HIP_TRY(c, hipGetLastError());
        if ((rc = run_bounces(c, a, guided, ctx_trace_cfg(c), a.max_depth))) return rc;
        hipLaunchKernelGGL(k_film, dim3((npix + 255) / 256), dim3(256), 0, c->stream, a, r->film_sum,
                           r->film_comp);
        KP_HIP_TRY(c, hipGetLastError());
        if (train) {
            const float inv_count = 1.0f / (float)(r->film_count + (uint64_t)(pass + 1) * k);
            hipLaunchKernelGGL(k_guide_records, dim3((n_rec + 255) / 256), dim3(256), 0, c->stream, a, n_rec,
                               g->weights, g->relative_weights ? (const float4*)r->film_sum : nullptr, inv_count);
            KP_HIP_TRY(c, hipGetLastError());
            if ((rc = guide_normalise_records(c, n_rec))) return rc;
            for (uint32_t sidx = 0; sidx < pass_steps; ++sidx)
                if ((rc = guide_train_device(c, g->lr, true))) return rc;
            steps += pass_steps;
            ++g->train_passes;
        }
    }
    r->film_count += d->spp;
    if ((rc = kp_synchronize(c))) return rc;
    auto t1 = std::chrono::steady_clock::now();
    unsigned long long tot[2];
    KP_HIP_TRY(c, hipMemcpy(tot, r->ps.totals, 16, hipMemcpyDeviceToHost));
    r->stats.paths = (uint64_t)npix * d->spp;
    r->stats.extension_rays = tot[0];
    r->stats.shadow_rays = tot[1];
    r->stats.seconds = std::chrono::duration<double>(t1 - t0).count();
    r->stats.passes = d->spp / k;
    r->stats.paths_per_pass = n_paths;
    r->stats.train_steps = steps;
    r->stats.guided_passes = guided_passes;
    r->stats.train_records = 0;
    if (steps) KP_HIP_TRY(c, hipMemcpy(&r->stats.train_records, g->rec_valid, 4, hipMemcpyDeviceToHost));
    return KP_OK;
}

// ============================================================== batches ==
// Many scenes rendered as one workload: their BVHs, shading records,
// materials and lights are copied into merged arrays (indices rebased), every
// ray carries its scene's BVH root, and each pass runs the paths of a window
// of the batch's work items whatever scene they belong to.
struct BatchState {
    kp_ctx* c = nullptr;
    uint4* nodes = nullptr;
    float4* tris = nullptr;
    float4* stris = nullptr;
    float4* snorm = nullptr;
    MatRec* mats = nullptr;
    Lobe* lobes = nullptr;
    LightRec* lights = nullptr;
    uint32_t* light_tris = nullptr;
    float* light_cdf = nullptr;
    uint32_t n_nodes = 0, n_tris = 0, bvh_depth = 0;
    std::vector<uint32_t> root, light_first, n_lights;
    uint2* spill = nullptr;
    uint32_t spill_cap = 0, max_steps = 0;
    // films of the last render
    std::vector<uint32_t> fw, fh, film_first, film_spp;
    float4* film_sum = nullptr;
    float4* film_comp = nullptr;
    float* film_out = nullptr;
    uint32_t film_cap = 0;
    // paths and the pass table
    PathState ps{};
    void* path_mem = nullptr;
    uint32_t cap_paths = 0;
    SceneRec* d_scenes = nullptr;
    uint32_t scenes_cap = 0;
    unsigned long long* d_pass = nullptr;
    std::vector<SceneRec> h_scenes;
    kp_batch_tuning tune{};
    // captured pass graph and what it was captured for
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> graph_key;
    kp_render_stats stats{};
};

static void batch_drop_graph(BatchState* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_key.clear();
}

static void batch_free(BatchState* b) {
    batch_drop_graph(b);
    for (void* p : {(void*)b->nodes, (void*)b->tris, (void*)b->stris, (void*)b->snorm, (void*)b->mats,
                    (void*)b->lobes, (void*)b->lights,
                    (void*)b->light_tris, (void*)b->light_cdf, (void*)b->spill, (void*)b->film_sum,
                    (void*)b->film_comp, (void*)b->film_out, b->path_mem, (void*)b->d_scenes, (void*)b->d_pass})
        if (p) (void)hipFree(p);
    delete b;
}

// Progress lines on stderr when the environment sets KP_VERBOSE (batch set-up
// and autotuning can take seconds with nothing else to show).
static bool verbose() {
    static const bool v = std::getenv("KP_VERBOSE") != nullptr;
    return v;
}

static int batch_create(kp_ctx* c, kp_ctx* const* scenes, uint32_t n, BatchState** out) {
    uint64_t nn = 0, nt = 0, nm = 0, nl = 0, nlt = 0, ncdf = 0, nlb = 0, nsn = 0;
    uint32_t depth = 0;
    for (uint32_t i = 0; i < n; ++i) {
        kp_ctx* s = scenes[i];
        if (!s) return fail(c, KP_ERR_INVALID, "kp_batch_create: NULL scene context");
        if (s->device != c->device) return fail(c, KP_ERR_INVALID, "kp_batch_create: scenes on another device");
        if (!s->built) return fail(c, KP_ERR_STATE, "kp_batch_create: scene without kp_bvh_build");
        if (s->two_level) return fail(c, KP_ERR_INVALID, "kp_batch_create: instanced scenes cannot be batched");
        KP_HIP_TRY(c, hipSetDevice(s->device));
        int rc = prepare_scene(s);
        if (rc) return fail(c, rc, std::string("kp_batch_create: scene ") + std::to_string(i) + ": " + s->err);
        KP_HIP_TRY(c, hipStreamSynchronize(s->stream));
        const RenderState* r = s->render;
        nn += s->n_nodes;
        nt += s->n_tris_built;
        nm += r->mats.size();
        nlb += r->lobes.size();
        nsn += r->n_snorm;
        nl += r->n_lights;
        nlt += r->n_light_tris;
        ncdf += r->n_cdf;
        depth = std::max(depth, s->max_depth);
    }
    KP_USE_DEVICE(c);
    if (nt >= (1ull << 28) || nn >= (1ull << 24))
        return fail(c, KP_ERR_INVALID, "kp_batch_create: more than 2^28 triangles or 2^24 BVH8 nodes in the batch");
    BatchState* b = new BatchState();
    b->c = c;
    auto bail = [&](hipError_t e, const char* what) {
        batch_free(b);
        return hip_fail(c, e, what);
    };
    hipError_t e = hipSuccess;
    auto alloc = [&](void** p, uint64_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes ? bytes : 16);
    };
    alloc((void**)&b->nodes, nn * KP_NODE_BYTES);
    alloc((void**)&b->tris, nt * KP_TRI_BYTES);
    alloc((void**)&b->stris, nt * 3 * sizeof(float4));
    alloc((void**)&b->mats, nm * sizeof(MatRec));
    alloc((void**)&b->lobes, nlb * sizeof(Lobe));
    alloc((void**)&b->snorm, nsn * 3 * sizeof(float4));
    alloc((void**)&b->lights, nl * sizeof(LightRec));
    alloc((void**)&b->light_tris, nlt * 4);
    alloc((void**)&b->light_cdf, ncdf * 4);
    alloc((void**)&b->d_pass, 64);
    if (e != hipSuccess) return bail(e, "kp_batch_create: allocation");
    uint32_t on = 0, ot = 0, om = 0, ol = 0, olt = 0, ocdf = 0, osn = 0;
    uint64_t steps = 0;   // per-ray step limit: a ray stays in its own scene's BVH
    // materials and lobes merged on the host (lobe ranges rebased), uploaded once
    std::vector<MatRec> all_mats;
    std::vector<Lobe> all_lobes;
    for (uint32_t i = 0; i < n; ++i) {
        const RenderState* r = scenes[i]->render;
        for (MatRec m : r->mats) {
            m.first += (uint32_t)all_lobes.size();
            all_mats.push_back(m);
        }
        all_lobes.insert(all_lobes.end(), r->lobes.begin(), r->lobes.end());
    }
    e = hipMemcpyAsync(b->mats, all_mats.data(), all_mats.size() * sizeof(MatRec), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->lobes, all_lobes.data(), all_lobes.size() * sizeof(Lobe), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);   // the host vectors are the sources
    if (e != hipSuccess) return bail(e, "kp_batch_create: materials");
    for (uint32_t i = 0; i < n; ++i) {
        kp_ctx* s = scenes[i];
        const RenderState* r = s->render;
        const uint32_t sn = s->n_nodes, st = s->n_tris_built, sm = (uint32_t)r->mats.size();
        // The copies go on the batch's stream, in order with the rebase kernels
        // below; the scene's own stream is drained first (its buffers are the
        // sources).  The final synchronisation of the batch's stream then also
        // covers the copies, so the caller may release the scene contexts.
        e = hipStreamSynchronize(s->stream);
        auto copy = [&](void* dst, const void* src, size_t bytes) {
            if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream);
        };
        copy(b->nodes + 8 * (size_t)on, s->d_nodes, (size_t)sn * KP_NODE_BYTES);
        copy(b->tris + 3 * (size_t)ot, s->d_tris, (size_t)st * KP_TRI_BYTES);
        copy(b->stris + 3 * (size_t)ot, r->stris, (size_t)st * 3 * sizeof(float4));
        if (r->n_snorm) copy(b->snorm + 3 * (size_t)osn, r->d_snorm, (size_t)r->n_snorm * 3 * sizeof(float4));
        if (r->n_lights) {
            copy(b->lights + ol, r->d_lights, (size_t)r->n_lights * sizeof(LightRec));
            copy(b->light_tris + olt, r->d_light_tris, (size_t)r->n_light_tris * 4);
            copy(b->light_cdf + ocdf, r->d_light_cdf, (size_t)r->n_cdf * 4);
        }
        if (e != hipSuccess) return bail(e, "kp_batch_create: copy");
        hipLaunchKernelGGL(k_rebase_nodes, dim3((sn + 255) / 256), dim3(256), 0, c->stream, b->nodes + 8 * (size_t)on,
                           sn, on, ot);
        hipLaunchKernelGGL(k_rebase_tris, dim3((st + 255) / 256), dim3(256), 0, c->stream, b->tris + 3 * (size_t)ot,
                           st, ot);
        hipLaunchKernelGGL(k_rebase_stris, dim3((st + 255) / 256), dim3(256), 0, c->stream,
                           b->stris + 3 * (size_t)ot, st, om, ol, osn);
        const uint32_t nmax = std::max(r->n_lights, r->n_light_tris);
        if (nmax)
            hipLaunchKernelGGL(k_rebase_lights, dim3((nmax + 255) / 256), dim3(256), 0, c->stream, b->lights + ol,
                               r->n_lights, olt, ocdf, b->light_tris + olt, r->n_light_tris, ot);
        if ((e = hipGetLastError()) != hipSuccess) return bail(e, "kp_batch_create: rebase");
        steps = std::max<uint64_t>(steps, 2ull * sn + st + 64u);
        b->root.push_back(on);
        b->light_first.push_back(ol);
        b->n_lights.push_back(r->n_lights);
        on += sn;
        ot += st;
        om += sm;
        osn += r->n_snorm;
        ol += r->n_lights;
        olt += r->n_light_tris;
        ocdf += r->n_cdf;
    }
    b->n_nodes = on;
    b->n_tris = ot;
    b->bvh_depth = depth;
    b->spill_cap = 7u * depth + 8u;
    b->max_steps = (uint32_t)std::min<uint64_t>(steps, 0xFFFFFFF0ull);
    const uint32_t grid = max_trace_grid(c);
    e = hipMalloc((void**)&b->spill, (size_t)grid * KP_TRACE_BLOCK * b->spill_cap * sizeof(uint2));
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return bail(e, "kp_batch_create: spill / rebase");
    b->tune.pass_paths = 1u << 24;
    b->tune.shade_blocks_per_cu = c->shade_blocks / (uint32_t)c->cus;
    b->tune.trace_blocks_per_cu = c->grid_closest5 / (uint32_t)c->cus;
    b->tune.refill_below = c->refill_below;
    b->tune.use_graph = 1;
    if (verbose()) std::fprintf(stderr, "[kp] batch: %u scenes merged\n", n);
    *out = b;
    return KP_OK;
}

// The launches of one pass: start paths, bounce, fold into the films, advance.
static int batch_enqueue_pass(BatchState* b, const RenderArgs& a, const TraceCfg& tc, uint32_t max_depth) {
    kp_ctx* c = b->c;
    const uint32_t blocks = (a.n_paths + 255) / 256;
    hipLaunchKernelGGL(k_generate_batch, dim3(blocks), dim3(256), 0, c->stream, a);
    KP_HIP_TRY(c, hipGetLastError());
    int rc = run_bounces(c, a, false, tc, max_depth);
    if (rc) return rc;
    hipLaunchKernelGGL(k_film_batch, dim3(blocks), dim3(256), 0, c->stream, a, b->film_sum, b->film_comp);
    hipLaunchKernelGGL(k_pass_advance, dim3(1), dim3(64), 0, c->stream, b->d_pass, a.n_paths);
    KP_HIP_TRY(c, hipGetLastError());
    return KP_OK;
}

// Renders desc[i].spp samples of every scene into freshly zeroed films; with
// max_passes > 0 stops after that many passes (autotuning).  Returns the
// number of passes run in *passes_run.
static int batch_render(BatchState* b, const kp_render_desc* descs, uint32_t max_passes, uint32_t* passes_run) {
    kp_ctx* c = b->c;
    const uint32_t n = (uint32_t)b->root.size();
    uint32_t max_depth = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const kp_render_desc& d = descs[i];
        if (d.width == 0 || d.height == 0 || (uint64_t)d.width * d.height > (1u << 26) || d.spp == 0)
            return fail(c, KP_ERR_INVALID, "kp_batch_render: bad film size or spp");
        if (d.max_depth == 0 || d.max_depth > 64) return fail(c, KP_ERR_INVALID, "kp_batch_render: max_depth must be 1..64");
        if (d.flags & KP_RENDER_GUIDED) return fail(c, KP_ERR_INVALID, "kp_batch_render: guided batches are not supported");
        max_depth = std::max(max_depth, d.max_depth);
    }
    // films
    b->fw.assign(n, 0);
    b->fh.assign(n, 0);
    b->film_first.assign(n, 0);
    b->film_spp.assign(n, 0);
    uint64_t npix_all = 0;
    for (uint32_t i = 0; i < n; ++i) {
        b->fw[i] = descs[i].width;
        b->fh[i] = descs[i].height;
        b->film_spp[i] = descs[i].spp;
        b->film_first[i] = (uint32_t)npix_all;
        npix_all += (uint64_t)descs[i].width * descs[i].height;
    }
    if (npix_all > (1ull << 30)) return fail(c, KP_ERR_INVALID, "kp_batch_render: films too large");
    if (npix_all > b->film_cap) {
        KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
        for (void* p : {(void*)b->film_sum, (void*)b->film_comp, (void*)b->film_out})
            if (p) (void)hipFree(p);
        b->film_sum = b->film_comp = nullptr;
        b->film_out = nullptr;
        b->film_cap = 0;
        KP_HIP_TRY(c, hipMalloc((void**)&b->film_sum, npix_all * sizeof(float4)));
        KP_HIP_TRY(c, hipMalloc((void**)&b->film_comp, npix_all * sizeof(float4)));
        KP_HIP_TRY(c, hipMalloc((void**)&b->film_out, npix_all * 3 * sizeof(float)));
        b->film_cap = (uint32_t)npix_all;
        batch_drop_graph(b);
    }
    const uint32_t zb = (uint32_t)((npix_all + 255) / 256);
    hipLaunchKernelGGL(k_zero_f4, dim3(zb), dim3(256), 0, c->stream, b->film_sum, (uint32_t)npix_all);
    hipLaunchKernelGGL(k_zero_f4, dim3(zb), dim3(256), 0, c->stream, b->film_comp, (uint32_t)npix_all);
    KP_HIP_TRY(c, hipGetLastError());
    // scene table and work items
    b->h_scenes.assign(n, SceneRec{});
    unsigned long long work = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const kp_render_desc& d = descs[i];
        RenderArgs cam{};
        camera_basis(d, cam);
        SceneRec& r = b->h_scenes[i];
        r.eye = cam.eye;
        r.fwd = cam.fwd;
        r.right = cam.right;
        r.upv = cam.upv;
        r.tan_half = cam.tan_half;
        r.aspect = cam.aspect;
        r.width = d.width;
        r.height = d.height;
        r.spp = d.spp;
        r.spp_offset = d.spp_offset;
        r.seed = d.seed;
        r.max_depth = d.max_depth;
        r.rr_depth = d.rr_depth;
        r.root = b->root[i];
        r.light_first = b->light_first[i];
        r.n_lights = b->n_lights[i];
        r.film_first = b->film_first[i];
        r.work_first = work;
        work += (unsigned long long)d.width * d.height * d.spp;
    }
    if (n > b->scenes_cap) {
        KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (b->d_scenes) (void)hipFree(b->d_scenes);
        b->d_scenes = nullptr;
        KP_HIP_TRY(c, hipMalloc((void**)&b->d_scenes, n * sizeof(SceneRec)));
        b->scenes_cap = n;
        batch_drop_graph(b);
    }
    // the table and pass words are read by kernels already queued from an
    // earlier render: wait for them before overwriting
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    KP_HIP_TRY(c, hipMemcpy(b->d_scenes, b->h_scenes.data(), n * sizeof(SceneRec), hipMemcpyHostToDevice));
    const unsigned long long pass_words[2] = {0ull, work};
    KP_HIP_TRY(c, hipMemcpy(b->d_pass, pass_words, sizeof(pass_words), hipMemcpyHostToDevice));
    // a pass holds at most the batch's whole work (small batches need no
    // path state of the full pass size), in multiples of 256 paths
    const uint32_t chunk = (uint32_t)std::min<unsigned long long>(b->tune.pass_paths, (work + 255ull) & ~255ull);
    const uint32_t old_cap = b->cap_paths;
    void* old_mem = b->path_mem;
    int rc = alloc_paths(c, b->ps, b->path_mem, b->cap_paths, chunk);
    if (rc) return rc;
    if (b->path_mem != old_mem || b->cap_paths != old_cap) batch_drop_graph(b);
    unsigned long long zero2[2] = {0, 0};
    KP_HIP_TRY(c, hipMemcpy(b->ps.totals, zero2, 16, hipMemcpyHostToDevice));
    RenderArgs a{};
    a.stris = b->stris;
    a.snorm = b->snorm;
    a.mats = b->mats;
    a.lobes = b->lobes;
    a.lights = b->lights;
    a.light_tris = b->light_tris;
    a.light_cdf = b->light_cdf;
    a.n_lights = 0;
    a.k = 1;
    a.n_paths = chunk;
    a.max_depth = max_depth;
    a.rr_depth = 0;
    a.ps = b->ps;
    a.scenes = b->d_scenes;
    a.n_scenes = n;
    a.pass = b->d_pass;
    TraceCfg tc;
    std::memset(&tc, 0, sizeof(tc));
    tc.base.nodes = b->nodes;
    tc.base.tris = b->tris;
    tc.base.counter = c->d_words;
    tc.base.err = c->d_words + 1;
    tc.base.spill = b->spill;
    tc.base.spill_blocks = max_trace_grid(c);
    tc.base.five_waves = 1;   // the integrator's 5-wave traversal kernels
    tc.base.spill_cap = b->spill_cap;
    tc.base.refill_below = b->tune.refill_below;
    tc.base.max_steps = b->max_steps;
    tc.base.counters = c->d_counters;
    const uint32_t tb_max = std::max(c->grid_closest5, c->grid_occluded5) / (uint32_t)c->cus;
    const uint32_t tb = std::min(std::max(b->tune.trace_blocks_per_cu, 1u), tb_max);
    tc.grid_closest = std::min((uint32_t)c->cus * tb, c->grid_closest5);
    tc.grid_occluded = std::min((uint32_t)c->cus * tb, c->grid_occluded5);
    tc.shade_blocks = (uint32_t)c->cus * std::max(b->tune.shade_blocks_per_cu, 1u);
    tc.shade_blocks_guided = c->shade_blocks_guided;
    tc.roots = true;
    uint64_t passes = (work + chunk - 1) / chunk;
    if (max_passes && passes > max_passes) passes = max_passes;
    if (verbose())
        std::fprintf(stderr, "[kp] batch render: %llu passes of %u paths, graph %u\n", (unsigned long long)passes,
                     chunk, b->tune.use_graph);
    auto t0 = std::chrono::steady_clock::now();
    if (b->tune.use_graph) {
        const std::vector<uint64_t> key = {chunk, n, max_depth, tc.grid_closest, tc.grid_occluded, tc.shade_blocks,
                                           tc.base.refill_below, (uint64_t)(uintptr_t)b->path_mem,
                                           (uint64_t)(uintptr_t)b->film_sum, (uint64_t)(uintptr_t)b->d_scenes};
        if (!b->exec || b->graph_key != key) {
            batch_drop_graph(b);
            KP_HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
            rc = batch_enqueue_pass(b, a, tc, max_depth);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(c->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ec != hipSuccess) return hip_fail(c, ec, "kp_batch_render: pass graph capture");
            b->graph = g;
            const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                batch_drop_graph(b);
                return hip_fail(c, ei, "kp_batch_render: pass graph instantiation");
            }
            b->graph_key = key;
        }
        for (uint64_t p = 0; p < passes; ++p) KP_HIP_TRY(c, hipGraphLaunch(b->exec, c->stream));
    } else {
        for (uint64_t p = 0; p < passes; ++p)
            if ((rc = batch_enqueue_pass(b, a, tc, max_depth))) return rc;
    }
    if ((rc = kp_synchronize(c))) return rc;
    auto t1 = std::chrono::steady_clock::now();
    unsigned long long tot[2];
    KP_HIP_TRY(c, hipMemcpy(tot, b->ps.totals, 16, hipMemcpyDeviceToHost));
    b->stats = kp_render_stats{};
    b->stats.paths = std::min<unsigned long long>(work, (unsigned long long)passes * chunk);
    b->stats.extension_rays = tot[0];
    b->stats.shadow_rays = tot[1];
    b->stats.seconds = std::chrono::duration<double>(t1 - t0).count();
    b->stats.passes = (uint32_t)passes;
    b->stats.paths_per_pass = chunk;
    if (passes_run) *passes_run = (uint32_t)passes;
    return KP_OK;
}

// Coordinate descent over the launch settings, one setting at a time, each
// candidate timed over `passes` passes of the batch's own work (seconds per
// path, the first pass of each candidate is a warm-up).
static int batch_autotune(BatchState* b, const kp_render_desc* descs, uint32_t passes, kp_batch_tuning* out) {
    kp_ctx* c = b->c;
    const uint32_t tb_max = std::max(c->grid_closest5, c->grid_occluded5) / (uint32_t)c->cus;
    auto cost = [&](const kp_batch_tuning& t, double* sec_per_path) -> int {
        b->tune = t;
        uint32_t run = 0;
        if (verbose())
            std::fprintf(stderr, "[kp] autotune: pass_paths %u trace %u shade %u refill %u\n", t.pass_paths,
                         t.trace_blocks_per_cu, t.shade_blocks_per_cu, t.refill_below);
        int rc = batch_render(b, descs, 1, &run);   // warm-up: allocation, graph capture
        if (rc) return rc;
        if ((rc = batch_render(b, descs, passes, &run))) return rc;
        *sec_per_path = b->stats.seconds / (double)std::max<uint64_t>(1, b->stats.paths);
        if (verbose()) std::fprintf(stderr, "[kp] autotune:   %.3f ns per path\n", *sec_per_path * 1e9);
        return KP_OK;
    };
    kp_batch_tuning best = b->tune;
    double best_cost = 0.0;
    int rc = cost(best, &best_cost);
    if (rc) return rc;
    auto try_values = [&](uint32_t kp_batch_tuning::*field, std::initializer_list<uint32_t> values) -> int {
        for (uint32_t v : values) {
            if (v == best.*field) continue;
            kp_batch_tuning t = best;
            t.*field = v;
            double cst = 0.0;
            int r2 = cost(t, &cst);
            if (r2) return r2;
            if (cst < best_cost) {
                best_cost = cst;
                best = t;
            }
        }
        return KP_OK;
    };
    if ((rc = try_values(&kp_batch_tuning::pass_paths, {1u << 21, 1u << 22, 1u << 23, 1u << 24, 1u << 25}))) return rc;
    if ((rc = try_values(&kp_batch_tuning::trace_blocks_per_cu, {tb_max, std::max(1u, tb_max / 2), std::max(1u, tb_max - 1)})))
        return rc;
    if ((rc = try_values(&kp_batch_tuning::refill_below, {16u, 32u, 40u, 48u, 56u}))) return rc;
    if ((rc = try_values(&kp_batch_tuning::shade_blocks_per_cu, {4u, 8u, 16u}))) return rc;
    b->tune = best;
    if (out) *out = best;
    return KP_OK;
}

// Validates the lobes and appends one material (its lobes to the lobe table).
static int add_material(kp_ctx* c, const kp_lobe* lobes, uint32_t n, const float* Le, const char* who,
                        uint32_t* id) {
    static_assert(sizeof(Lobe) == 48 && sizeof(MatRec) == 32, "closure record layout");
    const std::string w(who);
    std::vector<Lobe> ls;
    MatRec mr{};
    mr.guide_alpha = INFINITY;
    for (uint32_t i = 0; i < n; ++i) {
        const kp_material& m = lobes[i].lobe;
        if (m.type < 0 || m.type > 3 || m.dist < 0 || m.dist > 1)
            return fail(c, KP_ERR_INVALID, w + ": bad closure type or distribution");
        if (m.type != 0 && !(m.alpha > 0.0f && std::isfinite(m.alpha)))
            return fail(c, KP_ERR_INVALID, w + ": microfacet alpha must be > 0");
        if (m.type != 0 && !(m.alpha_y >= 0.0f && std::isfinite(m.alpha_y)))
            return fail(c, KP_ERR_INVALID, w + ": alpha_y must be >= 0 (0 = alpha)");
        if ((m.type == 2 || m.type == 3) && !(m.eta > 0.0f && std::isfinite(m.eta)))
            return fail(c, KP_ERR_INVALID, w + ": eta must be > 0");
        for (int k = 0; k < 3; ++k)
            if (!(lobes[i].weight[k] >= 0.0f && std::isfinite(lobes[i].weight[k])))
                return fail(c, KP_ERR_INVALID, w + ": lobe weights must be finite and >= 0");
        Lobe l{};
        l.type = m.type;
        l.dist = m.dist;
        for (int k = 0; k < 3; ++k) {
            l.w[k] = lobes[i].weight[k];
            l.c[k] = m.type == 1 ? m.F0[k] : m.albedo[k];
        }
        l.ax = m.alpha;
        l.ay = m.alpha_y > 0.0f ? m.alpha_y : m.alpha;
        l.eta = m.eta;
        if (m.type == 2) mr.flags |= KP_MATF_TRANSMIT;
        if (m.type != 0) mr.guide_alpha = std::fmin(mr.guide_alpha, std::fmin(l.ax, l.ay));
        ls.push_back(l);
    }
    for (int k = 0; k < 3; ++k) {
        if (!(Le[k] >= 0.0f && std::isfinite(Le[k]))) return fail(c, KP_ERR_INVALID, w + ": Le must be finite and >= 0");
        mr.Le[k] = Le[k];
    }
    RenderState* r = rstate(c);
    mr.first = (uint32_t)r->lobes.size();
    mr.n = n;
    r->lobes.insert(r->lobes.end(), ls.begin(), ls.end());
    r->mats.push_back(mr);
    r->scene_ready = false;
    if (id) *id = (uint32_t)(r->mats.size() - 1);
    return KP_OK;
}

}  // namespace kp

extern "C" {

int kp_material_add(kp_ctx* c, const kp_material* m, uint32_t* id) {
    if (!c || !m) return KP_ERR_INVALID;
    kp_lobe one{};
    one.weight[0] = one.weight[1] = one.weight[2] = 1.0f;
    one.lobe = *m;
    return kp::add_material(c, &one, 1, m->Le, "kp_material_add", id);
}

int kp_material_add_composite(kp_ctx* c, const kp_lobe* lobes, uint32_t n_lobes, const float* Le,
                              uint32_t* id) {
    if (!c) return KP_ERR_INVALID;
    if (!lobes || n_lobes == 0 || n_lobes > KP_MAX_LOBES)
        return kp::fail(c, KP_ERR_INVALID, "kp_material_add_composite: 1..8 lobes");
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    return kp::add_material(c, lobes, n_lobes, Le ? Le : zero, "kp_material_add_composite", id);
}

int kp_mesh_set_material(kp_ctx* c, uint32_t mesh_id, uint32_t material_id) {
    if (!c) return KP_ERR_INVALID;
    if (mesh_id >= c->mesh_material.size()) return kp::fail(c, KP_ERR_INVALID, "bad mesh id");
    c->mesh_material[mesh_id] = material_id;
    kp::render_scene_dirty(c);
    return KP_OK;
}

int kp_instance_set_material(kp_ctx* c, uint32_t instance_id, uint32_t material_id) {
    if (!c) return KP_ERR_INVALID;
    if (instance_id >= c->inst_material.size()) return kp::fail(c, KP_ERR_INVALID, "bad instance id");
    c->inst_material[instance_id] = material_id;
    kp::render_scene_dirty(c);
    return KP_OK;
}

int kp_release_buffers(kp_ctx* c) {
    if (!c) return KP_ERR_INVALID;
    KP_USE_DEVICE(c);
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->render && c->render->path_mem) {
        KP_HIP_TRY(c, hipFree(c->render->path_mem));
        c->render->path_mem = nullptr;
        c->render->cap_paths = 0;
        c->render->ps = kp::PathState{};
    }
    if (c->d_stage_in) KP_HIP_TRY(c, hipFree(c->d_stage_in));
    if (c->d_stage_out) KP_HIP_TRY(c, hipFree(c->d_stage_out));
    c->d_stage_in = c->d_stage_out = nullptr;
    c->stage_in_bytes = c->stage_out_bytes = 0;
    return KP_OK;
}

int kp_blas_set_material(kp_ctx* c, uint32_t blas_id, uint32_t material_id) {
    if (!c) return KP_ERR_INVALID;
    if (blas_id >= c->blas_material.size()) return kp::fail(c, KP_ERR_INVALID, "bad prototype id");
    c->blas_material[blas_id] = material_id;
    kp::render_scene_dirty(c);
    return KP_OK;
}

int kp_film_reset(kp_ctx* c, uint32_t w, uint32_t h) {
    if (!c) return KP_ERR_INVALID;
    if (w == 0 || h == 0 || (uint64_t)w * h > (1u << 26)) return kp::fail(c, KP_ERR_INVALID, "bad film size");
    KP_USE_DEVICE(c);
    return kp::film_alloc(c, w, h);
}

int kp_render_accumulate(kp_ctx* c, const kp_render_desc* d) {
    if (!c || !d) return KP_ERR_INVALID;
    if (d->spp == 0) return KP_OK;
    KP_USE_DEVICE(c);
    return kp::render_accumulate(c, d);
}

int kp_film_read(kp_ctx* c, float* out, int kind) {
    if (!c || !out) return KP_ERR_INVALID;
    RenderState* r = kp::rstate(c);
    if (!r->film_sum) return kp::fail(c, KP_ERR_STATE, "no film (kp_film_reset first)");
    if (kind != KP_PTR_HOST && kind != KP_PTR_DEVICE) return kp::fail(c, KP_ERR_INVALID, "bad pointer kind");
    KP_USE_DEVICE(c);
    const uint32_t npix = r->fw * r->fh;
    const float inv = r->film_count ? 1.0f / (float)r->film_count : 0.0f;
    float* dst = kind == KP_PTR_DEVICE ? out : r->film_out;
    hipLaunchKernelGGL(kp::k_film_resolve, dim3((npix + 255) / 256), dim3(256), 0, c->stream,
                       r->film_sum, r->film_comp, npix, inv, dst);
    KP_HIP_TRY(c, hipGetLastError());
    if (kind == KP_PTR_HOST)
        KP_HIP_TRY(c, hipMemcpyAsync(out, r->film_out, (size_t)npix * 12, hipMemcpyDeviceToHost, c->stream));
    return kp_synchronize(c);
}

int kp_render(kp_ctx* c, const kp_render_desc* d, float* out, int kind) {
    if (!c || !d || !out) return KP_ERR_INVALID;
    int rc = kp_film_reset(c, d->width, d->height);
    if (rc) return rc;
    if ((rc = kp_render_accumulate(c, d))) return rc;
    return kp_film_read(c, out, kind);
}

int kp_integrate_rays(kp_ctx* c, const kp_ray* rays, uint32_t n, const kp_render_desc* d,
                      float* rgb, int kind) {
    if (!c || !d) return KP_ERR_INVALID;
    if (kind != KP_PTR_HOST && kind != KP_PTR_DEVICE) return kp::fail(c, KP_ERR_INVALID, "bad pointer kind");
    if (n == 0) return KP_OK;
    if (!rays || !rgb) return kp::fail(c, KP_ERR_INVALID, "NULL rays or output");
    KP_USE_DEVICE(c);
    return kp::integrate_rays(c, rays, n, d, rgb, kind);
}

int kp_render_get_stats(kp_ctx* c, kp_render_stats* s) {
    if (!c || !s) return KP_ERR_INVALID;
    *s = kp::rstate(c)->stats;
    return KP_OK;
}


int kp_batch_create(kp_ctx* c, kp_ctx* const* scenes, uint32_t n, kp_batch** out) {
    if (!c || !out || (n && !scenes)) return KP_ERR_INVALID;
    *out = nullptr;
    if (n == 0 || n > 4096) return kp::fail(c, KP_ERR_INVALID, "kp_batch_create: 1..4096 scenes");
    KP_USE_DEVICE(c);
    kp::BatchState* b = nullptr;
    int rc = kp::batch_create(c, scenes, n, &b);
    if (rc) return rc;
    *out = reinterpret_cast<kp_batch*>(b);
    return KP_OK;
}

int kp_batch_destroy(kp_batch* h) {
    if (!h) return KP_ERR_INVALID;
    kp::BatchState* b = reinterpret_cast<kp::BatchState*>(h);
    (void)hipSetDevice(b->c->device);
    (void)hipStreamSynchronize(b->c->stream);
    kp::batch_free(b);
    return KP_OK;
}

int kp_batch_set_tuning(kp_batch* h, const kp_batch_tuning* t) {
    if (!h || !t) return KP_ERR_INVALID;
    kp::BatchState* b = reinterpret_cast<kp::BatchState*>(h);
    if (t->pass_paths < 256 || t->pass_paths > (1u << 25) || t->shade_blocks_per_cu == 0 ||
        t->shade_blocks_per_cu > 64 || t->trace_blocks_per_cu == 0 || t->refill_below == 0 || t->refill_below > 64)
        return kp::fail(b->c, KP_ERR_INVALID, "kp_batch_set_tuning: value out of range");
    b->tune = *t;
    return KP_OK;
}

int kp_batch_get_tuning(kp_batch* h, kp_batch_tuning* t) {
    if (!h || !t) return KP_ERR_INVALID;
    *t = reinterpret_cast<kp::BatchState*>(h)->tune;
    return KP_OK;
}

int kp_batch_render(kp_batch* h, const kp_render_desc* descs) {
    if (!h || !descs) return KP_ERR_INVALID;
    kp::BatchState* b = reinterpret_cast<kp::BatchState*>(h);
    KP_HIP_TRY(b->c, hipSetDevice(b->c->device));
    return kp::batch_render(b, descs, 0, nullptr);
}

int kp_batch_autotune(kp_batch* h, const kp_render_desc* descs, uint32_t passes, kp_batch_tuning* out) {
    if (!h || !descs) return KP_ERR_INVALID;
    kp::BatchState* b = reinterpret_cast<kp::BatchState*>(h);
    KP_HIP_TRY(b->c, hipSetDevice(b->c->device));
    return kp::batch_autotune(b, descs, passes ? passes : 2, out);
}

int kp_batch_film_read(kp_batch* h, uint32_t scene, float* out, int kind) {
    if (!h || !out) return KP_ERR_INVALID;
    kp::BatchState* b = reinterpret_cast<kp::BatchState*>(h);
    kp_ctx* c = b->c;
    if (scene >= b->fw.size()) return kp::fail(c, KP_ERR_INVALID, "kp_batch_film_read: no such scene in the last render");
    if (kind != KP_PTR_HOST && kind != KP_PTR_DEVICE) return kp::fail(c, KP_ERR_INVALID, "bad pointer kind");
    KP_USE_DEVICE(c);
    const uint32_t npix = b->fw[scene] * b->fh[scene], f0 = b->film_first[scene];
    const float inv = 1.0f / (float)b->film_spp[scene];
    float* dst = kind == KP_PTR_DEVICE ? out : b->film_out + 3 * (size_t)f0;
    hipLaunchKernelGGL(kp::k_film_resolve, dim3((npix + 255) / 256), dim3(256), 0, c->stream, b->film_sum + f0,
                       b->film_comp + f0, npix, inv, dst);
    KP_HIP_TRY(c, hipGetLastError());
    if (kind == KP_PTR_HOST)
        KP_HIP_TRY(c, hipMemcpyAsync(out, dst, (size_t)npix * 12, hipMemcpyDeviceToHost, c->stream));
    return kp_synchronize(c);
}

int kp_batch_get_stats(kp_batch* h, kp_render_stats* s) {
    if (!h || !s) return KP_ERR_INVALID;
    *s = reinterpret_cast<kp::BatchState*>(h)->stats;
    return KP_OK;
}

}  // extern "C"
