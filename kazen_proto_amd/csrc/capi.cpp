// C ABI of the HIP intersector: context, scene upload, BVH build, queries.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "../../include/kp_hip.h"
#include "kp_ctx.h"

namespace {
bool valid_kind(int k) { return k == KP_PTR_HOST || k == KP_PTR_DEVICE; }
}  // namespace

namespace kp {

int fail(kp_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else set_thread_error(msg);
    return code;
}

int hip_fail(kp_ctx* c, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();   // clear the sticky launch error
    return fail(c, e == hipErrorOutOfMemory ? KP_ERR_ALLOC : KP_ERR_HIP, m);
}

int ensure_buffer(kp_ctx* c, void** buf, size_t* cap, size_t bytes) {
    if (*cap >= bytes && *buf) return KP_OK;
    if (*buf) {
        KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
        KP_HIP_TRY(c, hipFree(*buf));
        *buf = nullptr;
        *cap = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    KP_HIP_TRY(c, hipMalloc(buf, want));
    *cap = want;
    return KP_OK;
}

// Largest persistent grid of any traversal launch: the spill area has one
// slice per lane of it.
uint32_t max_trace_grid(const kp_ctx* c) {
    return std::max(std::max(c->grid_closest, c->grid_occluded), std::max(c->grid_closest5, c->grid_occluded5));
}

TraceArgs trace_base(kp_ctx* c) {
    TraceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nodes = c->d_nodes;
    a.tris = c->d_tris;
    a.counter = c->d_words;
    a.err = c->d_words + 1;
    a.spill = c->d_spill;
    a.spill_cap = c->spill_cap;
    a.spill_blocks = (uint32_t)(c->spill_lanes / KP_TRACE_BLOCK);
    a.refill_below = c->refill_below;
    a.counters = c->d_counters;
    a.max_steps = c->max_steps;
    if (c->two_level) {
        a.insts = c->d_insts;
        a.tlas_inst = c->d_tlas_inst;
        a.n_tlas = c->n_tlas_nodes;
        a.world_root = c->world_root;
    }
    return a;
}

int trace_closest_dev(kp_ctx* c, const float4* rays, uint32_t n, float4* hits, bool count) {
    if (n == 0) return KP_OK;
    TraceArgs a = trace_base(c);
    a.rays = rays;
    a.hits = hits;
    a.n = n;
    // the 5-wave builds (96 registers, one-level or two-level); the counted
    // variant has a 4-wave kernel of its own
    const bool five = !count;
    a.five_waves = five ? 1u : 0u;
    hipError_t e = launch_closest(a, five ? c->grid_closest5 : c->grid_closest, count, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "closest-hit launch");
    return KP_OK;
}

int trace_occluded_dev(kp_ctx* c, const float4* rays, uint32_t n, uint32_t* out) {
    if (n == 0) return KP_OK;
    TraceArgs a = trace_base(c);
    a.rays = rays;
    a.occl = out;
    a.n = n;
    a.five_waves = 1u;   // the 5-wave builds, as the integrator's shadow rays
    hipError_t e = launch_occluded(a, c->grid_occluded5, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "occlusion launch");
    return KP_OK;
}

static void free_bvh(kp_ctx* c) {
    c->host = HostBvh();
    if (c->d_nodes) (void)hipFree(c->d_nodes);
    if (c->d_tris) (void)hipFree(c->d_tris);
    if (c->d_spill) (void)hipFree(c->d_spill);
    if (c->d_insts) (void)hipFree(c->d_insts);
    if (c->d_tlas_inst) (void)hipFree(c->d_tlas_inst);
    c->d_nodes = nullptr;
    c->d_tris = nullptr;
    c->d_spill = nullptr;
    c->d_insts = nullptr;
    c->d_tlas_inst = nullptr;
    c->two_level = false;
    c->n_tlas_nodes = 0;
    c->world_root = KP_INVALID;
    c->built = false;
}

// Row-major 3x4 affine inverse in double; false if singular or non-finite.
bool affine_inverse(const float* m, double* inv) {
    double a[9];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) a[3 * r + k] = m[4 * r + k];
    for (int i = 0; i < 12; ++i)
        if (!std::isfinite(m[i])) return false;
    const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double scale = 0.0;
    for (double v : a) scale = std::max(scale, std::fabs(v));
    if (!(std::fabs(det) > 1e-12 * scale * scale * scale)) return false;
    const double id = 1.0 / det;
    double b[9];
    b[0] = c00 * id;
    b[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    b[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    b[3] = c01 * id;
    b[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    b[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    b[6] = c02 * id;
    b[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    b[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    for (int r = 0; r < 3; ++r) {
        inv[4 * r + 0] = b[3 * r + 0];
        inv[4 * r + 1] = b[3 * r + 1];
        inv[4 * r + 2] = b[3 * r + 2];
        inv[4 * r + 3] = -(b[3 * r + 0] * m[3] + b[3 * r + 1] * m[7] + b[3 * r + 2] * m[11]);
    }
    return true;
}

void bake_placed(const kp_ctx* c, std::vector<float>& out) {
    size_t n = c->soup.size();
    for (uint32_t b : c->inst_blas) n += 9 * (size_t)c->blas_count[b];
    out.resize(n);
    std::memcpy(out.data(), c->soup.data(), c->soup.size() * sizeof(float));
    size_t o = c->soup.size();
    for (size_t k = 0; k < c->inst_blas.size(); ++k) {
        const uint32_t b = c->inst_blas[k];
        const float* m = &c->inst_o2w[12 * k];
        const float* src = &c->blas_soup[9 * (size_t)c->blas_first[b]];
        for (size_t v = 0; v < 3 * (size_t)c->blas_count[b]; ++v, o += 3)
            for (int a = 0; a < 3; ++a) {
                double w = (double)m[4 * a] * (double)src[3 * v];
                w = w + (double)m[4 * a + 1] * (double)src[3 * v + 1];
                w = w + (double)m[4 * a + 2] * (double)src[3 * v + 2];
                w = w + (double)m[4 * a + 3];
                out[o + a] = (float)w;
            }
    }
}

static void box_of(const float* tris9, uint32_t n, double lo[3], double hi[3]) {
    for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    for (size_t i = 0; i < 3 * (size_t)n; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], (double)tris9[3 * i + k]);
            hi[k] = std::max(hi[k], (double)tris9[3 * i + k]);
        }
}

// Two-level build: one BVH8 per prototype (unit 0 = the kp_mesh_add meshes,
// if any, traversed first in world space), a top-level BVH8 over the
// instances' world boxes (leaf size 1), all nodes in one array [top | unit 0 |
// unit 1 ...] with child and primitive bases rebased, and per instance a
// record {world-to-object rows, root node, primitive base, flags}.
static int build_two_level(kp_ctx* c, const BuildOptions& bo, double* seconds) {
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t n_world = (uint32_t)(c->soup.size() / 9);
    const uint32_t n_blas = (uint32_t)c->blas_first.size();
    const bool has_world = n_world > 0;
    const uint32_t ub = has_world ? 1u : 0u;   // unit of prototype 0
    // units: [world] + prototypes
    std::vector<Bvh8Host> units(n_blas + ub);
    std::vector<const float*> src(units.size());
    std::vector<uint32_t> ucount(units.size());
    if (has_world) { src[0] = c->soup.data(); ucount[0] = n_world; }
    for (uint32_t b = 0; b < n_blas; ++b) {
        src[b + ub] = &c->blas_soup[9 * (size_t)c->blas_first[b]];
        ucount[b + ub] = c->blas_count[b];
    }
    std::string err;
    for (size_t u = 0; u < units.size(); ++u)
        if (!build_bvh8(src[u], ucount[u], bo, units[u], err))
            return fail(c, KP_ERR_INVALID, "kp_bvh_build (prototype " + std::to_string(u) + "): " + err);
    // one record per instance
    const uint32_t n_inst = (uint32_t)c->inst_blas.size();
    std::vector<float> tlas_tris(9 * (size_t)n_inst);
    std::vector<float> rec(16 * (size_t)n_inst, 0.0f);
    std::vector<uint32_t> place_unit(n_inst);
    std::vector<uint64_t> place_base(n_inst);
    uint64_t prim_next = n_world;
    for (uint32_t k = 0; k < n_inst; ++k) {
        double lo[3], hi[3];
        float* r = &rec[16 * (size_t)k];
        const uint32_t b = c->inst_blas[k];
        place_unit[k] = b + ub;
        place_base[k] = prim_next;
        prim_next += c->blas_count[b];
        double olo[3], ohi[3];
        box_of(&c->blas_soup[9 * (size_t)c->blas_first[b]], c->blas_count[b], olo, ohi);
        const float* m = &c->inst_o2w[12 * (size_t)k];
        for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
        for (int corner = 0; corner < 8; ++corner) {
            const double p[3] = {(corner & 1) ? ohi[0] : olo[0], (corner & 2) ? ohi[1] : olo[1],
                                 (corner & 4) ? ohi[2] : olo[2]};
            for (int a = 0; a < 3; ++a) {
                const double w = (double)m[4 * a] * p[0] + (double)m[4 * a + 1] * p[1] +
                                 (double)m[4 * a + 2] * p[2] + (double)m[4 * a + 3];
                lo[a] = std::min(lo[a], w);
                hi[a] = std::max(hi[a], w);
            }
        }
        std::memcpy(r, &c->inst_w2o[12 * (size_t)k], 12 * sizeof(float));
        // the box as a degenerate triangle (lo, hi, lo), widened outward by
        // one float step so the rounding to float stays conservative
        float* t = &tlas_tris[9 * (size_t)k];
        for (int a = 0; a < 3; ++a) {
            const float fl = std::nextafter((float)lo[a], -INFINITY), fh = std::nextafter((float)hi[a], INFINITY);
            t[a] = fl; t[3 + a] = fh; t[6 + a] = fl;
        }
    }
    if (prim_next >= 0xFFFFFFFFull)
        return fail(c, KP_ERR_INVALID, "kp_bvh_build: instanced scene exceeds 2^32-1 primitives");
    BuildOptions tb = bo;
    tb.leaf_size = 1;
    Bvh8Host top;
    if (!build_bvh8(tlas_tris.data(), n_inst, tb, top, err))
        return fail(c, KP_ERR_INVALID, "kp_bvh_build (top level): " + err);
    // merge
    uint64_t total_nodes = top.nodes.size(), total_tris = 0;
    std::vector<uint32_t> node_off(units.size()), tri_off(units.size());
    for (size_t u = 0; u < units.size(); ++u) {
        node_off[u] = (uint32_t)total_nodes;
        tri_off[u] = (uint32_t)total_tris;
        total_nodes += units[u].nodes.size();
        total_tris += ucount[u];
        if (total_nodes >= (1ull << 24))
            return fail(c, KP_ERR_INVALID, "kp_bvh_build: instanced BVH exceeds 2^24 nodes");
        if (total_tris >= (1ull << 28))
            return fail(c, KP_ERR_INVALID, "kp_bvh_build: prototypes exceed 2^28 triangles");
    }
    std::vector<KpNode8> nodes(total_nodes);
    std::memcpy(nodes.data(), top.nodes.data(), top.nodes.size() * sizeof(KpNode8));
    std::vector<float> tris(12 * (size_t)total_tris);
    uint64_t steps = 2ull * top.nodes.size() + n_inst + 64u + (has_world ? 2ull * units[0].nodes.size() : 0u);
    uint32_t max_unit_depth = 0;
    for (size_t u = 0; u < units.size(); ++u) {
        for (size_t i = 0; i < units[u].nodes.size(); ++i) {
            KpNode8 nd = units[u].nodes[i];
            nd.child_base += node_off[u];
            nd.prim_base += tri_off[u];
            nodes[node_off[u] + i] = nd;
        }
        std::memcpy(&tris[12 * (size_t)tri_off[u]], units[u].tris.data(), units[u].tris.size() * sizeof(float));
        max_unit_depth = std::max(max_unit_depth, units[u].max_depth);
    }
    static const float eye[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    for (uint32_t k = 0; k < n_inst; ++k) {
        float* r = &rec[16 * (size_t)k];
        const uint32_t u = place_unit[k];
        steps += 2ull * units[u].nodes.size() + 2u;
        const uint32_t root = node_off[u], base = (uint32_t)place_base[k];
        // an exact identity world-to-object matrix: the ray is used as is
        const uint32_t flags = std::memcmp(r, eye, sizeof(eye)) == 0 ? 1u : 0u;
        std::memcpy(&r[12], &root, 4);
        std::memcpy(&r[13], &base, 4);
        std::memcpy(&r[14], &flags, 4);
    }
    auto t1 = std::chrono::steady_clock::now();
    *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (c->device >= 0) KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    free_bvh(c);
    // Stack bound: the top-level marker, at most 8 entries
    // per top-level node on the path (instance entries and one group), the
    // return marker, 7 per prototype level.
    // Postponed leaves need no stack entries.
    c->spill_cap = 8u * top.max_depth + 7u * max_unit_depth + 16u;
    if (c->device >= 0) {
        const size_t nb = nodes.size() * sizeof(KpNode8), tbytes = tris.size() * sizeof(float);
        KP_HIP_TRY(c, hipMalloc((void**)&c->d_nodes, nb));
        KP_HIP_TRY(c, hipMalloc((void**)&c->d_tris, tbytes));
        KP_HIP_TRY(c, hipMalloc((void**)&c->d_insts, rec.size() * sizeof(float)));
        KP_HIP_TRY(c, hipMalloc((void**)&c->d_tlas_inst, top.tri_order.size() * 4));
        KP_HIP_TRY(c, hipMemcpy(c->d_nodes, nodes.data(), nb, hipMemcpyHostToDevice));
        KP_HIP_TRY(c, hipMemcpy(c->d_tris, tris.data(), tbytes, hipMemcpyHostToDevice));
        KP_HIP_TRY(c, hipMemcpy(c->d_insts, rec.data(), rec.size() * sizeof(float), hipMemcpyHostToDevice));
        KP_HIP_TRY(c, hipMemcpy(c->d_tlas_inst, top.tri_order.data(), top.tri_order.size() * 4,
                                hipMemcpyHostToDevice));
        c->spill_lanes = (size_t)kp::max_trace_grid(c) * KP_TRACE_BLOCK;
        KP_HIP_TRY(c, hipMalloc((void**)&c->d_spill, c->spill_lanes * c->spill_cap * sizeof(uint2)));
    }
    c->two_level = true;
    c->world_root = has_world ? node_off[0] : KP_INVALID;
    c->n_tlas_nodes = (uint32_t)top.nodes.size();
    c->n_world_prims = n_world;
    c->n_tris_built = (uint32_t)prim_next;
    c->n_tris_stored = (uint32_t)total_tris;
    c->n_nodes = (uint32_t)nodes.size();
    c->n_bvh2 = top.n_bvh2_nodes;
    for (const auto& u : units) c->n_bvh2 += u.n_bvh2_nodes;
    c->max_depth = top.max_depth + max_unit_depth;
    // A traversal visits every node of the meshes' BVH and of the top level
    // once and, per instance entered, each node of its prototype once.
    c->max_steps = (uint32_t)std::min<uint64_t>(steps, 0xFFFFFFF0ull);
    // the host copy answers the thread-safe per-ray queries
    c->host.nodes = std::move(nodes);
    c->host.tris = std::move(tris);
    c->host.two_level = true;
    c->host.n_tlas = c->n_tlas_nodes;
    c->host.world_root = c->world_root;
    c->host.insts = std::move(rec);
    c->host.tlas_inst = std::move(top.tri_order);
    c->built = true;
    render_scene_dirty(c);
    return KP_OK;
}

}  // namespace kp

extern "C" {

int kp_create(kp_ctx** out, int device) {
    if (!out) return kp::fail(nullptr, KP_ERR_INVALID, "kp_create: out is NULL");
    *out = nullptr;
    if (device == KP_DEVICE_HOST) {   // scene building and host queries only
        kp_ctx* h = new kp_ctx();
        h->device = KP_DEVICE_HOST;
        *out = h;
        return KP_OK;
    }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return kp::fail(nullptr, KP_ERR_HIP, "kp_create: no HIP device available");
    if (device < 0 || device >= ndev) return kp::fail(nullptr, KP_ERR_INVALID, "kp_create: bad device index");
    kp_ctx* c = new kp_ctx();
    c->device = device;
    auto bail = [&](hipError_t he, const char* what) {
        std::string m = std::string("kp_create: ") + what + ": " + hipGetErrorString(he);
        kp_destroy(c);
        return kp::fail(nullptr, KP_ERR_HIP, m);
    };
    if ((e = hipSetDevice(device)) != hipSuccess) return bail(e, "hipSetDevice");
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return bail(e, "properties");
    c->cus = prop.multiProcessorCount;
    if ((e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "stream");
    c->stream = c->own_stream;
    if ((e = hipMalloc((void**)&c->d_words, 64)) != hipSuccess) return bail(e, "alloc");
    if ((e = hipMemsetAsync(c->d_words, 0, 64, c->stream)) != hipSuccess) return bail(e, "memset");
    if ((e = hipMalloc((void**)&c->d_counters, 64)) != hipSuccess) return bail(e, "alloc");
    if ((e = hipMemsetAsync(c->d_counters, 0, 64, c->stream)) != hipSuccess) return bail(e, "memset");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail(e, "memset");
    int bc = 0, bo = 0;
    if ((e = kp::trace_occupancy(&bc, &bo)) != hipSuccess) return bail(e, "occupancy");
    c->grid_closest = (uint32_t)(c->cus * (bc > 0 ? bc : 1));
    c->grid_occluded = (uint32_t)(c->cus * (bo > 0 ? bo : 1));
    int b5c = 0, b5o = 0;
    if ((e = kp::trace_occupancy_five(&b5c, &b5o)) != hipSuccess) return bail(e, "occupancy");
    c->grid_closest5 = (uint32_t)(c->cus * (b5c > 0 ? b5c : 1));
    c->grid_occluded5 = (uint32_t)(c->cus * (b5o > 0 ? b5o : 1));
    void* scr = nullptr;
    if ((e = kp::noalloc_scratch(device, &scr)) != hipSuccess) return bail(e, "scratch symbol");
    c->d_scratch = (unsigned char*)scr;
    int sp = 0, sg = 0;
    if ((e = kp::shade_occupancy(&sp, &sg)) != hipSuccess) return bail(e, "occupancy");
    c->shade_blocks = (uint32_t)(c->cus * (sp > 0 ? sp : 1));
    c->shade_blocks_guided = (uint32_t)(c->cus * (sg > 0 ? sg : 1));
    *out = c;
    return KP_OK;
}

void kp_destroy(kp_ctx* c) {
    if (!c) return;
    if (c->device < 0) {
        kp::free_bvh(c);
        kp::render_state_free(c);
        kp::guide_state_free(c);
        delete c;
        return;
    }
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    kp::render_state_free(c);
    kp::guide_state_free(c);
    kp::free_bvh(c);
    if (c->d_words) (void)hipFree(c->d_words);
    if (c->d_counters) (void)hipFree(c->d_counters);
    if (c->d_stage_in) (void)hipFree(c->d_stage_in);
    if (c->d_stage_out) (void)hipFree(c->d_stage_out);
    if (c->ev_begin) (void)hipEventDestroy(c->ev_begin);
    if (c->ev_end) (void)hipEventDestroy(c->ev_end);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

const char* kp_last_error(const kp_ctx* c) {
    return c ? c->err.c_str() : kp::thread_error();
}

int kp_set_stream(kp_ctx* c, void* s) {
    if (!c) return KP_ERR_INVALID;
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return KP_OK;
}

int kp_synchronize(kp_ctx* c) {
    if (!c) return KP_ERR_INVALID;
    KP_USE_DEVICE(c);
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    uint32_t flags = 0;
    KP_HIP_TRY(c, hipMemcpy(&flags, c->d_words + 1, 4, hipMemcpyDeviceToHost));
    if (c->noalloc_used) {   // errors of kp_intersect_closest_noalloc launches
        uint32_t f2 = 0;
        uint32_t* ew = (uint32_t*)(c->d_scratch + KP_SCRATCH_ERR_OFFSET);
        KP_HIP_TRY(c, hipMemcpy(&f2, ew, 4, hipMemcpyDeviceToHost));
        if (f2) KP_HIP_TRY(c, kp::launch_reset_u32(ew, 1u, c->stream));
        flags |= f2;
    }
    if (flags) {
        KP_HIP_TRY(c, hipMemsetAsync(c->d_words + 1, 0, 4, c->stream));
        KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
        return kp::fail(c, KP_ERR_KERNEL,
                        (flags & 1u) ? "traversal stack exceeded its spill capacity"
                                     : "traversal exceeded its step limit");
    }
    return KP_OK;
}

int kp_mesh_add(kp_ctx* c, const float* xyz, uint32_t n_verts, const uint32_t* idx,
                uint32_t n_tris, const float* xf, uint32_t* mesh_id) {
    if (!c) return KP_ERR_INVALID;
    if (!xyz || !idx || n_verts == 0 || n_tris == 0)
        return kp::fail(c, KP_ERR_INVALID, "kp_mesh_add: empty mesh or NULL arrays");
    if ((uint64_t)c->soup.size() / 9 + n_tris >= (1ull << 28))
        return kp::fail(c, KP_ERR_INVALID, "kp_mesh_add: scene exceeds 2^28 triangles");
    for (uint64_t i = 0; i < 3ull * n_tris; ++i)
        if (idx[i] >= n_verts) return kp::fail(c, KP_ERR_INVALID, "kp_mesh_add: vertex index out of range");
    const size_t base = c->soup.size();
    c->soup.resize(base + 9 * (size_t)n_tris);
    for (uint32_t t = 0; t < n_tris; ++t) {
        for (int v = 0; v < 3; ++v) {
            const float* p = xyz + 3 * (size_t)idx[3 * (size_t)t + v];
            float* d = &c->soup[base + 9 * (size_t)t + 3 * v];
            if (xf) {
                for (int r = 0; r < 3; ++r)
                    d[r] = xf[4 * r + 0] * p[0] + xf[4 * r + 1] * p[1] + xf[4 * r + 2] * p[2] + xf[4 * r + 3];
            } else {
                d[0] = p[0]; d[1] = p[1]; d[2] = p[2];
            }
        }
    }
    const uint32_t id = (uint32_t)c->mesh_first.size();
    c->mesh_first.push_back((uint32_t)(base / 9));
    c->mesh_count.push_back(n_tris);
    c->mesh_material.push_back(0);
    static const float eye[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    c->mesh_xf.insert(c->mesh_xf.end(), xf ? xf : eye, (xf ? xf : eye) + 12);
    if (!c->soup_has_n.empty()) {
        c->soup_has_n.resize(c->soup.size() / 9, 0);
        c->soup_n.resize(c->soup.size(), 0.0f);
    }
    kp::render_scene_dirty(c);
    if (mesh_id) *mesh_id = id;
    return KP_OK;
}

// Unit vertex normals of n triangles (9 floats each) under the inverse
// transpose of the row-major 3x4 matrix xf (NULL: as given).
static int normals_to(kp_ctx* c, const float* n9, uint32_t n, const float* xf, float* out, const char* who) {
    double inv[12];
    if (xf && !kp::affine_inverse(xf, inv))
        return kp::fail(c, KP_ERR_INVALID, std::string(who) + ": mesh transform not invertible");
    for (size_t i = 0; i < 3 * (size_t)n; ++i) {
        const float* v = n9 + 3 * i;
        double w[3], len = 0.0;
        for (int a = 0; a < 3; ++a) {
            // (M^-1)^T n: column a of the inverse's linear part
            w[a] = xf ? inv[a] * v[0] + inv[4 + a] * v[1] + inv[8 + a] * v[2] : (double)v[a];
            len += w[a] * w[a];
        }
        len = std::sqrt(len);
        if (!(len > 0.0) || !std::isfinite(len))
            return kp::fail(c, KP_ERR_INVALID, std::string(who) + ": zero or non-finite normal");
        for (int a = 0; a < 3; ++a) out[3 * i + a] = (float)(w[a] / len);
    }
    return KP_OK;
}

int kp_mesh_set_normals(kp_ctx* c, uint32_t mesh_id, const float* normals9) {
    if (!c) return KP_ERR_INVALID;
    if (mesh_id >= c->mesh_first.size()) return kp::fail(c, KP_ERR_INVALID, "kp_mesh_set_normals: bad mesh id");
    if (!normals9) return kp::fail(c, KP_ERR_INVALID, "kp_mesh_set_normals: NULL normals");
    const uint32_t first = c->mesh_first[mesh_id], n = c->mesh_count[mesh_id];
    std::vector<float> w(9 * (size_t)n);
    int rc = normals_to(c, normals9, n, &c->mesh_xf[12 * (size_t)mesh_id], w.data(), "kp_mesh_set_normals");
    if (rc) return rc;
    c->soup_has_n.resize(c->soup.size() / 9, 0);
    c->soup_n.resize(c->soup.size(), 0.0f);
    std::memcpy(&c->soup_n[9 * (size_t)first], w.data(), w.size() * sizeof(float));
    std::memset(&c->soup_has_n[first], 1, n);
    kp::render_scene_dirty(c);
    return KP_OK;
}

int kp_blas_set_normals(kp_ctx* c, uint32_t blas_id, const float* normals9) {
    if (!c) return KP_ERR_INVALID;
    if (blas_id >= c->blas_first.size()) return kp::fail(c, KP_ERR_INVALID, "kp_blas_set_normals: bad prototype id");
    if (!normals9) return kp::fail(c, KP_ERR_INVALID, "kp_blas_set_normals: NULL normals");
    const uint32_t first = c->blas_first[blas_id], n = c->blas_count[blas_id];
    std::vector<float> w(9 * (size_t)n);
    int rc = normals_to(c, normals9, n, nullptr, w.data(), "kp_blas_set_normals");
    if (rc) return rc;
    c->blas_has_n.resize(c->blas_soup.size() / 9, 0);
    c->blas_n.resize(c->blas_soup.size(), 0.0f);
    std::memcpy(&c->blas_n[9 * (size_t)first], w.data(), w.size() * sizeof(float));
    std::memset(&c->blas_has_n[first], 1, n);
    kp::render_scene_dirty(c);
    return KP_OK;
}

int kp_blas_add(kp_ctx* c, const float* xyz, uint32_t n_verts, const uint32_t* idx, uint32_t n_tris,
                uint32_t* blas_id) {
    if (!c) return KP_ERR_INVALID;
    if (!xyz || !idx || n_verts == 0 || n_tris == 0)
        return kp::fail(c, KP_ERR_INVALID, "kp_blas_add: empty mesh or NULL arrays");
    if ((uint64_t)c->blas_soup.size() / 9 + n_tris >= (1ull << 28))
        return kp::fail(c, KP_ERR_INVALID, "kp_blas_add: prototypes exceed 2^28 triangles");
    for (uint64_t i = 0; i < 3ull * n_tris; ++i)
        if (idx[i] >= n_verts) return kp::fail(c, KP_ERR_INVALID, "kp_blas_add: vertex index out of range");
    const size_t base = c->blas_soup.size();
    c->blas_soup.resize(base + 9 * (size_t)n_tris);
    for (uint32_t t = 0; t < n_tris; ++t)
        for (int v = 0; v < 3; ++v)
            std::memcpy(&c->blas_soup[base + 9 * (size_t)t + 3 * v], xyz + 3 * (size_t)idx[3 * (size_t)t + v],
                        3 * sizeof(float));
    const uint32_t id = (uint32_t)c->blas_first.size();
    c->blas_first.push_back((uint32_t)(base / 9));
    c->blas_count.push_back(n_tris);
    c->blas_material.push_back(0);
    if (!c->blas_has_n.empty()) {
        c->blas_has_n.resize(c->blas_soup.size() / 9, 0);
        c->blas_n.resize(c->blas_soup.size(), 0.0f);
    }
    if (blas_id) *blas_id = id;
    return KP_OK;
}

int kp_instance_add(kp_ctx* c, uint32_t blas_id, const float* xform12, uint32_t* instance_id) {
    if (!c) return KP_ERR_INVALID;
    if (blas_id >= c->blas_first.size()) return kp::fail(c, KP_ERR_INVALID, "kp_instance_add: unknown prototype");
    if (!xform12) return kp::fail(c, KP_ERR_INVALID, "kp_instance_add: NULL transform");
    double inv[12];
    if (!kp::affine_inverse(xform12, inv))
        return kp::fail(c, KP_ERR_INVALID, "kp_instance_add: transform not finite or not invertible");
    const uint32_t id = (uint32_t)c->inst_blas.size();
    c->inst_blas.push_back(blas_id);
    c->inst_material.push_back(KP_INVALID);
    c->inst_o2w.insert(c->inst_o2w.end(), xform12, xform12 + 12);
    for (int i = 0; i < 12; ++i) c->inst_w2o.push_back((float)inv[i]);
    if (instance_id) *instance_id = id;
    return KP_OK;
}

int kp_prim_locate(kp_ctx* c, uint32_t prim, uint32_t* kind, uint32_t* index, uint32_t* local) {
    if (!c) return KP_ERR_INVALID;
    if (!c->built) return kp::fail(c, KP_ERR_STATE, "kp_prim_locate before kp_bvh_build");
    if (prim >= c->n_tris_built) return kp::fail(c, KP_ERR_INVALID, "kp_prim_locate: primitive id out of range");
    uint32_t k, i, l;
    if (prim < c->n_world_prims) {
        // meshes are contiguous in insertion order
        const auto it = std::upper_bound(c->mesh_first.begin(), c->mesh_first.end(), prim);
        i = (uint32_t)(it - c->mesh_first.begin()) - 1u;
        k = 0;
        l = prim - c->mesh_first[i];
    } else {
        const auto it = std::upper_bound(c->inst_first.begin(), c->inst_first.end(), prim);
        i = (uint32_t)(it - c->inst_first.begin()) - 1u;
        k = 1;
        l = prim - c->inst_first[i];
    }
    if (kind) *kind = k;
    if (index) *index = i;
    if (local) *local = l;
    return KP_OK;
}

int kp_geometry_clear(kp_ctx* c) {
    if (!c) return KP_ERR_INVALID;
    if (c->device >= 0) {
        KP_USE_DEVICE(c);
        KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    kp::free_bvh(c);
    c->soup.clear();
    c->mesh_first.clear();
    c->mesh_count.clear();
    c->mesh_material.clear();
    c->mesh_xf.clear();
    c->soup_n.clear();
    c->soup_has_n.clear();
    c->blas_n.clear();
    c->blas_has_n.clear();
    c->blas_soup.clear();
    c->blas_first.clear();
    c->blas_count.clear();
    c->blas_material.clear();
    c->inst_o2w.clear();
    c->inst_w2o.clear();
    c->inst_blas.clear();
    c->inst_material.clear();
    c->inst_first.clear();
    c->n_tris_built = c->n_tris_stored = c->n_world_prims = 0;
    c->n_nodes = c->n_bvh2 = c->max_depth = 0;
    kp::render_scene_dirty(c);
    return KP_OK;
}

int kp_instance_get(kp_ctx* c, uint32_t instance_id, uint32_t* blas_id, float* o2w12, float* w2o12) {
    if (!c) return KP_ERR_INVALID;
    if (instance_id >= c->inst_blas.size()) return kp::fail(c, KP_ERR_INVALID, "kp_instance_get: unknown instance");
    if (blas_id) *blas_id = c->inst_blas[instance_id];
    if (o2w12) std::memcpy(o2w12, &c->inst_o2w[12 * (size_t)instance_id], 12 * sizeof(float));
    if (w2o12) std::memcpy(w2o12, &c->inst_w2o[12 * (size_t)instance_id], 12 * sizeof(float));
    return KP_OK;
}

// Makes h the context's one-level BVH (host copy for the per-ray queries,
// device copy and stack spill area on a device context).
static int install_one_level(kp_ctx* c, kp::Bvh8Host& h, uint32_t n_all, uint32_t n_world, uint32_t width) {
    if (c->device >= 0) KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    kp::free_bvh(c);
    // Worst-case stack: 7 pushes per level below the root, plus slack; the
    // top KP_TRACE_STACK entries live in LDS.
    c->spill_cap = 7u * h.max_depth + 8u;
    if (c->device >= 0) {
        const size_t nb = h.nodes.size() * sizeof(KpNode8), tb = h.tris.size() * sizeof(float);
        KP_HIP_TRY(c, hipMalloc((void**)&c->d_nodes, nb));
        KP_HIP_TRY(c, hipMalloc((void**)&c->d_tris, tb));
        KP_HIP_TRY(c, hipMemcpy(c->d_nodes, h.nodes.data(), nb, hipMemcpyHostToDevice));
        KP_HIP_TRY(c, hipMemcpy(c->d_tris, h.tris.data(), tb, hipMemcpyHostToDevice));
        c->spill_lanes = (size_t)kp::max_trace_grid(c) * KP_TRACE_BLOCK;
        KP_HIP_TRY(c, hipMalloc((void**)&c->d_spill, c->spill_lanes * c->spill_cap * sizeof(uint2)));
    }
    c->n_tris_built = n_all;
    c->n_tris_stored = n_all;
    c->n_world_prims = n_world;
    c->n_nodes = (uint32_t)h.nodes.size();
    c->n_bvh2 = h.n_bvh2_nodes;
    c->max_depth = h.max_depth;
    c->node_width = width;
    // A traversal visits every node and leaf at most once.
    c->max_steps = 2u * (uint32_t)h.nodes.size() + n_all + 64u;
    // the host copy answers the thread-safe per-ray queries
    c->host.nodes = std::move(h.nodes);
    c->host.tris = std::move(h.tris);
    c->built = true;
    return KP_OK;
}

// Build-time node-width tuner: builds widths 4, 6 and 8, times closest hit
// of 2^20 probe rays (origins uniform in the scene's bounds, directions
// uniform on the sphere; a fixed generator) on each, and leaves the fastest
// installed.  ms[i] receives the mean launch time of width 4 + 2i.
static int tune_node_width(kp_ctx* c, const float* soup, uint32_t n_all, uint32_t n_world, kp::BuildOptions bo,
                           float* ms, double* build_secs) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < (size_t)n_all * 3; ++i) {
        const int a = (int)(i % 3);
        lo[a] = std::min(lo[a], soup[3 * (i / 3) + a]);
        hi[a] = std::max(hi[a], soup[3 * (i / 3) + a]);
    }
    const uint32_t n_probe = 1u << 20;
    std::vector<float> rays((size_t)n_probe * 8);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    auto next = [&]() {   // splitmix64 -> [0, 1)
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (float)((z ^ (z >> 31)) >> 40) * (1.0f / 16777216.0f);
    };
    for (uint32_t i = 0; i < n_probe; ++i) {
        float* r = &rays[8 * (size_t)i];
        for (int a = 0; a < 3; ++a) r[a] = lo[a] + (hi[a] - lo[a]) * next();
        const float z = 1.0f - 2.0f * next(), phi = 6.2831853f * next();
        const float rr = std::sqrt(std::max(0.0f, 1.0f - z * z));
        r[3] = 0.0f;
        r[4] = rr * std::cos(phi);
        r[5] = rr * std::sin(phi);
        r[6] = z;
        r[7] = INFINITY;
    }
    float4 *d_rays = nullptr, *d_hits = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = KP_OK;
    float best = INFINITY;
    uint32_t best_w = 8;
    kp::Bvh8Host keep;
    double secs = 0.0;
    auto fail_hip = [&](hipError_t e, const char* what) { rc = kp::hip_fail(c, e, what); };
    hipError_t e;
    if ((e = hipMalloc((void**)&d_rays, (size_t)n_probe * 32)) != hipSuccess) { fail_hip(e, "width tuner"); goto done; }
    if ((e = hipMalloc((void**)&d_hits, (size_t)n_probe * 16)) != hipSuccess) { fail_hip(e, "width tuner"); goto done; }
    if ((e = hipMemcpy(d_rays, rays.data(), (size_t)n_probe * 32, hipMemcpyHostToDevice)) != hipSuccess) {
        fail_hip(e, "width tuner");
        goto done;
    }
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
        fail_hip(e, "width tuner");
        goto done;
    }
    for (int i = 0; i < 3; ++i) {
        const uint32_t w = 4u + 2u * (uint32_t)i;
        bo.max_children = w;
        kp::Bvh8Host h;
        std::string err;
        auto t0 = std::chrono::steady_clock::now();
        if (!kp::build_bvh8(soup, n_all, bo, h, err)) { rc = kp::fail(c, KP_ERR_INVALID, "kp_bvh_build: " + err); goto done; }
        secs += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        kp::Bvh8Host copy = h;   // the installed BVH moves its host arrays
        if ((rc = install_one_level(c, h, n_all, n_world, w))) goto done;
        if ((rc = kp::trace_closest_dev(c, d_rays, n_probe, d_hits, false))) goto done;   // warm-up
        if ((e = hipEventRecord(e0, c->stream)) != hipSuccess) { fail_hip(e, "width tuner"); goto done; }
        for (int k = 0; k < 3; ++k)
            if ((rc = kp::trace_closest_dev(c, d_rays, n_probe, d_hits, false))) goto done;
        if ((e = hipEventRecord(e1, c->stream)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess) {
            fail_hip(e, "width tuner");
            goto done;
        }
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, e0, e1);
        ms[i] = t / 3.0f;
        if (ms[i] < best) {
            best = ms[i];
            best_w = w;
            keep = std::move(copy);
        }
    }
    if (best_w != 8u && (rc = install_one_level(c, keep, n_all, n_world, best_w))) goto done;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (d_rays) (void)hipFree(d_rays);
    if (d_hits) (void)hipFree(d_hits);
    *build_secs = secs;
    return rc;
}

int kp_bvh_build(kp_ctx* c, const kp_build_opts* o) {
    if (!c) return KP_ERR_INVALID;
    if (c->device >= 0) KP_USE_DEVICE(c);
    const uint32_t n = (uint32_t)(c->soup.size() / 9);
    kp::BuildOptions bo;
    bool auto_width = false;
    if (o) {
        if (o->leaf_size) bo.leaf_size = o->leaf_size;
        if (o->sah_bins) bo.sah_bins = o->sah_bins;
        if (o->node_width == KP_NODE_WIDTH_AUTO) auto_width = true;
        else if (o->node_width > 8 || (o->node_width != 0 && o->node_width < 2))
            return kp::fail(c, KP_ERR_INVALID, "kp_bvh_build: node_width must be 2..8, 0 or KP_NODE_WIDTH_AUTO");
        else if (o->node_width) bo.max_children = o->node_width;
    }
    for (float& m : c->width_ms) m = 0.0f;
    // first primitive id of every instance (the same in both instance modes)
    c->inst_first.resize(c->inst_blas.size());
    {
        uint64_t next = n;
        for (size_t k = 0; k < c->inst_blas.size(); ++k) {
            c->inst_first[k] = (uint32_t)std::min<uint64_t>(next, 0xFFFFFFFFu);
            next += c->blas_count[c->inst_blas[k]];
        }
    }
    std::vector<float> baked;
    const float* soup = c->soup.data();
    uint32_t n_all = n;
    if (!c->inst_blas.empty()) {
        const uint32_t mode = o ? o->instances : KP_INSTANCES_AUTO;
        if (mode > KP_INSTANCES_BAKE) return kp::fail(c, KP_ERR_INVALID, "kp_bvh_build: bad instances mode");
        uint64_t placed = 0;
        for (uint32_t b : c->inst_blas) placed += c->blas_count[b];
        const bool two = mode == KP_INSTANCES_TWO_LEVEL || (mode == KP_INSTANCES_AUTO && placed > KP_BAKE_BUDGET);
        if (two) {
            double secs = 0.0;
            const int rc = kp::build_two_level(c, bo, &secs);
            if (rc == KP_OK) {
                c->build_seconds = secs;
                c->node_width = bo.max_children;
            }
            return rc;
        }
        if ((uint64_t)n + placed >= (1ull << 28))
            return kp::fail(c, KP_ERR_INVALID, "kp_bvh_build: baked instances exceed 2^28 triangles");
        kp::bake_placed(c, baked);
        soup = baked.data();
        n_all = (uint32_t)(baked.size() / 9);
    }
    if (auto_width && c->device >= 0 && n_all > 0) {
        double secs = 0.0;
        const int rc = tune_node_width(c, soup, n_all, n, bo, c->width_ms, &secs);
        if (rc) return rc;
        c->build_seconds = secs;
        kp::render_scene_dirty(c);
        return KP_OK;
    }
    kp::Bvh8Host h;
    std::string err;
    auto t0 = std::chrono::steady_clock::now();
    if (!kp::build_bvh8(soup, n_all, bo, h, err)) return kp::fail(c, KP_ERR_INVALID, "kp_bvh_build: " + err);
    auto t1 = std::chrono::steady_clock::now();
    const int rc = install_one_level(c, h, n_all, n, bo.max_children);
    if (rc) return rc;
    c->build_seconds = std::chrono::duration<double>(t1 - t0).count();
    kp::render_scene_dirty(c);
    return KP_OK;
}

int kp_get_stats(kp_ctx* c, kp_stats* s) {
    if (!c || !s) return KP_ERR_INVALID;
    std::memset(s, 0, sizeof(*s));
    s->n_tris = c->n_tris_built;
    s->n_nodes = c->n_nodes;
    s->n_bvh2_nodes = c->n_bvh2;
    s->bvh_bytes = (uint64_t)c->n_nodes * KP_NODE_BYTES + (uint64_t)c->n_tris_stored * KP_TRI_BYTES +
                   (c->two_level ? (uint64_t)c->inst_blas.size() * 68u : 0u);
    s->max_depth = c->max_depth;
    s->spill_entries = c->spill_cap;
    s->grid_closest = c->grid_closest;
    s->grid_occluded = c->grid_occluded;
    s->build_seconds = c->build_seconds;
    s->device_cus = (uint32_t)c->cus;
    s->two_level = c->two_level ? 1u : 0u;
    s->node_width = c->node_width;
    for (int i = 0; i < 3; ++i) s->width_ms[i] = c->width_ms[i];
    return KP_OK;
}

static int closest_common(kp_ctx* c, const kp_ray* rays, uint32_t n, kp_hit* hits, int kind,
                          bool count, uint64_t* nodes, uint64_t* tris) {
    if (!c) return KP_ERR_INVALID;
    if (!c->built) return kp::fail(c, KP_ERR_STATE, "intersect before kp_bvh_build");
    if (!valid_kind(kind)) return kp::fail(c, KP_ERR_INVALID, "bad pointer kind");
    if (n == 0) return KP_OK;
    if (!rays || !hits) return kp::fail(c, KP_ERR_INVALID, "NULL rays or hits");
    KP_USE_DEVICE(c);
    if (count) KP_HIP_TRY(c, hipMemsetAsync(c->d_counters, 0, 16, c->stream));
    int rc;
    if (kind == KP_PTR_DEVICE) {
        rc = kp::trace_closest_dev(c, (const float4*)rays, n, (float4*)hits, count);
        if (rc) return rc;
    } else {
        const size_t rb = (size_t)n * sizeof(kp_ray), hb = (size_t)n * sizeof(kp_hit);
        if ((rc = kp::ensure_buffer(c, &c->d_stage_in, &c->stage_in_bytes, rb))) return rc;
        if ((rc = kp::ensure_buffer(c, &c->d_stage_out, &c->stage_out_bytes, hb))) return rc;
        KP_HIP_TRY(c, hipMemcpyAsync(c->d_stage_in, rays, rb, hipMemcpyHostToDevice, c->stream));
        rc = kp::trace_closest_dev(c, (const float4*)c->d_stage_in, n, (float4*)c->d_stage_out, count);
        if (rc) return rc;
        KP_HIP_TRY(c, hipMemcpyAsync(hits, c->d_stage_out, hb, hipMemcpyDeviceToHost, c->stream));
    }
    if (count || kind == KP_PTR_HOST) {
        if ((rc = kp_synchronize(c))) return rc;
    }
    if (count) {
        unsigned long long v[2];
        KP_HIP_TRY(c, hipMemcpy(v, c->d_counters, 16, hipMemcpyDeviceToHost));
        if (nodes) *nodes = v[0];
        if (tris) *tris = v[1];
    }
    return KP_OK;
}

int kp_intersect_closest(kp_ctx* c, const kp_ray* rays, uint32_t n, kp_hit* hits, int kind) {
    return closest_common(c, rays, n, hits, kind, false, nullptr, nullptr);
}

int kp_intersect_closest_noalloc(kp_ctx* c, const kp_ray* rays, uint32_t n, kp_hit* hits) {
    if (!c) return KP_ERR_INVALID;
    if (!c->built) return kp::fail(c, KP_ERR_STATE, "kp_intersect_closest_noalloc before kp_bvh_build");
    if (n == 0) return KP_OK;
    if (!rays || !hits) return kp::fail(c, KP_ERR_INVALID, "kp_intersect_closest_noalloc: NULL rays or hits");
    KP_USE_DEVICE(c);
    if (!c->d_scratch) return kp::fail(c, KP_ERR_STATE, "kp_intersect_closest_noalloc: no scratch on this device");
    // one stack spill slice of spill_cap entries per lane of the grid; fewer
    // blocks than the regular launch when the BVH is deep
    const uint64_t block_bytes = (uint64_t)KP_TRACE_BLOCK * c->spill_cap * sizeof(uint2);
    const uint64_t fit = (KP_SCRATCH_BYTES - KP_SCRATCH_SPILL_OFFSET) / block_bytes;
    if (fit == 0)
        return kp::fail(c, KP_ERR_INVALID,
                        "kp_intersect_closest_noalloc: the BVH needs " + std::to_string(block_bytes) +
                            " bytes of traversal stack per block; the scratch holds " +
                            std::to_string(KP_SCRATCH_BYTES - KP_SCRATCH_SPILL_OFFSET));
    const uint32_t grid = (uint32_t)std::min<uint64_t>(c->grid_closest, fit);
    kp::TraceArgs a = kp::trace_base(c);
    a.rays = (const float4*)rays;
    a.hits = (float4*)hits;
    a.n = n;
    a.counter = (uint32_t*)c->d_scratch;
    a.err = (uint32_t*)(c->d_scratch + KP_SCRATCH_ERR_OFFSET);
    a.spill = (uint2*)(c->d_scratch + KP_SCRATCH_SPILL_OFFSET);
    a.spill_blocks = grid;
    c->noalloc_used = true;
    KP_HIP_TRY(c, kp::launch_reset_u32(a.err, 1u, c->stream));   // launch_closest clears the counter
    hipError_t e = kp::launch_closest(a, grid, false, c->stream);
    if (e != hipSuccess) return kp::hip_fail(c, e, "kp_intersect_closest_noalloc launch");
    return KP_OK;
}

int kp_intersect_closest_counted(kp_ctx* c, const kp_ray* rays, uint32_t n, kp_hit* hits, int kind,
                                 uint64_t* nodes, uint64_t* tris) {
    return closest_common(c, rays, n, hits, kind, true, nodes, tris);
}

int kp_occluded(kp_ctx* c, const kp_ray* rays, uint32_t n, uint32_t* out, int kind) {
    if (!c) return KP_ERR_INVALID;
    if (!c->built) return kp::fail(c, KP_ERR_STATE, "occluded before kp_bvh_build");
    if (!valid_kind(kind)) return kp::fail(c, KP_ERR_INVALID, "bad pointer kind");
    if (n == 0) return KP_OK;
    if (!rays || !out) return kp::fail(c, KP_ERR_INVALID, "NULL rays or out");
    KP_USE_DEVICE(c);
    int rc;
    if (kind == KP_PTR_DEVICE) return kp::trace_occluded_dev(c, (const float4*)rays, n, out);
    const size_t rb = (size_t)n * sizeof(kp_ray), ob = (size_t)n * 4;
    if ((rc = kp::ensure_buffer(c, &c->d_stage_in, &c->stage_in_bytes, rb))) return rc;
    if ((rc = kp::ensure_buffer(c, &c->d_stage_out, &c->stage_out_bytes, ob))) return rc;
    KP_HIP_TRY(c, hipMemcpyAsync(c->d_stage_in, rays, rb, hipMemcpyHostToDevice, c->stream));
    rc = kp::trace_occluded_dev(c, (const float4*)c->d_stage_in, n, (uint32_t*)c->d_stage_out);
    if (rc) return rc;
    KP_HIP_TRY(c, hipMemcpyAsync(out, c->d_stage_out, ob, hipMemcpyDeviceToHost, c->stream));
    return kp_synchronize(c);
}

int kp_intersect_closest_host(const kp_ctx* c, const kp_ray* rays, uint32_t n, kp_hit* hits) {
    if (!c) return kp::fail(nullptr, KP_ERR_INVALID, "kp_intersect_closest_host: NULL context");
    if (!c->built) return kp::fail(nullptr, KP_ERR_STATE, "kp_intersect_closest_host before kp_bvh_build");
    if (n && (!rays || !hits)) return kp::fail(nullptr, KP_ERR_INVALID, "kp_intersect_closest_host: NULL rays or hits");
    kp::host_closest(c->host, (const float*)rays, n, (float*)hits);
    return KP_OK;
}

int kp_occluded_host(const kp_ctx* c, const kp_ray* rays, uint32_t n, uint32_t* out) {
    if (!c) return kp::fail(nullptr, KP_ERR_INVALID, "kp_occluded_host: NULL context");
    if (!c->built) return kp::fail(nullptr, KP_ERR_STATE, "kp_occluded_host before kp_bvh_build");
    if (n && (!rays || !out)) return kp::fail(nullptr, KP_ERR_INVALID, "kp_occluded_host: NULL rays or out");
    kp::host_occluded(c->host, (const float*)rays, n, out);
    return KP_OK;
}

int kp_device_alloc(kp_ctx* c, uint64_t bytes, void** ptr) {
    if (!c || !ptr) return KP_ERR_INVALID;
    *ptr = nullptr;
    KP_USE_DEVICE(c);
    KP_HIP_TRY(c, hipMalloc(ptr, bytes ? bytes : 1));
    return KP_OK;
}

int kp_device_free(kp_ctx* c, void* ptr) {
    if (!c) return KP_ERR_INVALID;
    if (!ptr) return KP_OK;
    KP_USE_DEVICE(c);
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    KP_HIP_TRY(c, hipFree(ptr));
    return KP_OK;
}

int kp_copy_to_device(kp_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (!c || (!dst && bytes) || (!src && bytes)) return KP_ERR_INVALID;
    if (!bytes) return KP_OK;
    KP_USE_DEVICE(c);
    KP_HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    return KP_OK;
}

int kp_copy_to_host(kp_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (!c || (!dst && bytes) || (!src && bytes)) return KP_ERR_INVALID;
    if (!bytes) return KP_OK;
    KP_USE_DEVICE(c);
    KP_HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    KP_HIP_TRY(c, hipStreamSynchronize(c->stream));
    return KP_OK;
}

int kp_device_copy(kp_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (!c || (bytes && (!dst || !src))) return KP_ERR_INVALID;
    if (bytes % 16) return kp::fail(c, KP_ERR_INVALID, "kp_device_copy: bytes must be a multiple of 16");
    if (!bytes) return KP_OK;
    KP_USE_DEVICE(c);
    KP_HIP_TRY(c, kp::launch_copy16(dst, src, bytes / 16, c->cus, c->stream));
    return KP_OK;
}

int kp_timer_begin(kp_ctx* c) {
    if (!c) return KP_ERR_INVALID;
    KP_USE_DEVICE(c);
    if (!c->ev_begin) KP_HIP_TRY(c, hipEventCreate(&c->ev_begin));
    if (!c->ev_end) KP_HIP_TRY(c, hipEventCreate(&c->ev_end));
    KP_HIP_TRY(c, hipEventRecord(c->ev_begin, c->stream));
    return KP_OK;
}

int kp_timer_end(kp_ctx* c, float* ms) {
    if (!c || !ms) return KP_ERR_INVALID;
    if (!c->ev_begin) return kp::fail(c, KP_ERR_STATE, "kp_timer_end without kp_timer_begin");
    KP_HIP_TRY(c, hipEventRecord(c->ev_end, c->stream));
    KP_HIP_TRY(c, hipEventSynchronize(c->ev_end));
    KP_HIP_TRY(c, hipEventElapsedTime(ms, c->ev_begin, c->ev_end));
    return KP_OK;
}

int kp_set_trace_tuning(kp_ctx* c, uint32_t lanes) {
    if (!c) return KP_ERR_INVALID;
    if (lanes < 1 || lanes > 64) return kp::fail(c, KP_ERR_INVALID, "refill threshold must be 1..64");
    c->refill_below = lanes;
    return KP_OK;
}

int kp_set_pass_paths(kp_ctx* c, uint32_t max_paths) {
    if (!c) return KP_ERR_INVALID;
    if (max_paths < (1u << 16) || max_paths > (1u << 28))
        return kp::fail(c, KP_ERR_INVALID, "kp_set_pass_paths: max_paths must be 2^16 .. 2^28");
    c->pass_paths = max_paths;
    return KP_OK;
}

}  // extern "C"
