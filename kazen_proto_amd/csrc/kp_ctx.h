// Context object behind the C ABI (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "host_trace.h"
#include "kp_bvh.h"
#include "kp_kernels.h"

struct kp_ctx {
    int device = 0;
    int cus = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    std::string err;

    // scene (host copy, world space)
    std::vector<float> soup;          // 9 floats per triangle
    std::vector<uint32_t> mesh_first; // first global primitive of each mesh
    std::vector<uint32_t> mesh_count;
    std::vector<uint32_t> mesh_material;
    std::vector<float> mesh_xf;       // 12 per mesh: the object-to-world rows kp_mesh_add applied
    // shading normals (kp_mesh_set_normals / kp_blas_set_normals): 9 floats
    // per triangle (world space for meshes, object space for prototypes) and
    // a per-triangle flag; empty while no mesh has normals
    std::vector<float> soup_n, blas_n;
    std::vector<uint8_t> soup_has_n, blas_has_n;

    // instancing: prototype meshes (object space) and their placements
    std::vector<float> blas_soup;
    std::vector<uint32_t> blas_first, blas_count, blas_material;
    std::vector<float> inst_o2w, inst_w2o;   // 12 per instance, row-major 3x4
    std::vector<uint32_t> inst_blas;
    std::vector<uint32_t> inst_material;     // KP_INVALID: the prototype's material
    std::vector<uint32_t> inst_first;        // first primitive id of each instance (set by kp_bvh_build)

    // BVH on the device
    bool built = false;
    bool two_level = false;          // built with instances (top-level BVH + prototypes)
    uint32_t n_tlas_nodes = 0;       // nodes [0, n_tlas_nodes) form the top level
    uint32_t world_root = 0xFFFFFFFFu; // root of the kp_mesh_add meshes' BVH in two-level mode
    uint32_t n_world_prims = 0;      // primitives of the kp_mesh_add meshes
    float4* d_insts = nullptr;       // instance records (4 float4 each)
    uint32_t* d_tlas_inst = nullptr; // top-level leaf position -> instance
    uint32_t n_tris_built = 0;       // primitive ids (instanced: every placement counted)
    uint32_t n_tris_stored = 0;      // triangles stored on the device
    uint32_t n_nodes = 0, n_bvh2 = 0, max_depth = 0;
    uint32_t node_width = 8;   // children per node of the built BVH
    float width_ms[3] = {0.0f, 0.0f, 0.0f};   // node-width tuner: probe ms of widths 4, 6, 8
    double build_seconds = 0.0;
    kp::HostBvh host;                // host copy of the device BVH (per-ray queries on host cores)
    uint4* d_nodes = nullptr;
    float4* d_tris = nullptr;
    uint2* d_spill = nullptr;
    uint32_t spill_cap = 0;
    size_t spill_lanes = 0;
    uint32_t grid_closest = 0, grid_occluded = 0;
    uint32_t shade_blocks = 0, shade_blocks_guided = 0;   // resident blocks of k_shade, whole device
    uint32_t grid_closest5 = 0, grid_occluded5 = 0;       // the integrator's 5-wave traversal kernels
    uint32_t refill_below = 40;   // measured best for incoherent rays (leaf 4; 36..60 swept)
    uint32_t pass_paths = 1u << 26;   // paths per pass of the integrator (measured: larger is faster)
    uint32_t max_steps = 64;

    // small device words: [0] work counter, [1] error flags
    uint32_t* d_words = nullptr;
    // this device's static scratch (kp_intersect_closest_noalloc), looked up at kp_create
    unsigned char* d_scratch = nullptr;
    bool noalloc_used = false;       // kp_synchronize also reads the scratch's error word
    unsigned long long* d_counters = nullptr;

    // staging for host-pointer calls
    void* d_stage_in = nullptr;
    size_t stage_in_bytes = 0;
    void* d_stage_out = nullptr;
    size_t stage_out_bytes = 0;

    // integrator state (render.hip) and guiding network state (guiding.hip)
    struct RenderState* render = nullptr;
    struct GuideState* guide = nullptr;
};

namespace kp {
// Row-major 3x4 affine inverse in double; false if singular or non-finite.
bool affine_inverse(const float* m, double* inv);
// Per-thread error text (host_api.cpp): kp_last_error(NULL).
void set_thread_error(const std::string& msg);
const char* thread_error();
// Records a HIP failure on the context and returns KP_ERR_HIP.
int hip_fail(kp_ctx* c, hipError_t e, const char* what);
int fail(kp_ctx* c, int code, const std::string& msg);
// Grows a device buffer to at least `bytes`.
int ensure_buffer(kp_ctx* c, void** buf, size_t* cap, size_t bytes);
// Trace helpers shared with the integrator (device pointers, async).
int trace_closest_dev(kp_ctx* c, const float4* rays, uint32_t n, float4* hits, bool count);
int trace_occluded_dev(kp_ctx* c, const float4* rays, uint32_t n, uint32_t* out);
// Traversal arguments of the context's BVH (one- or two-level).
TraceArgs trace_base(kp_ctx* c);
uint32_t max_trace_grid(const kp_ctx* c);
// World-space triangles (9 floats each) of every primitive id: the
// kp_mesh_add meshes, then each instance's prototype under its
// object-to-world matrix, each coordinate ((m0 x + m1 y) + m2 z) + m3 in FP64
// rounded to FP32.
void bake_placed(const kp_ctx* c, std::vector<float>& out);
void render_state_free(kp_ctx* c);
void guide_state_free(kp_ctx* c);
// Marks the integrator's device copy of the scene stale (meshes changed).
void render_scene_dirty(kp_ctx* c);
}  // namespace kp

// Makes the context's device current; host-only contexts (KP_DEVICE_HOST)
// fail every call that needs the GPU.
#define KP_USE_DEVICE(ctx)                                                                           \
    do {                                                                                             \
        if ((ctx)->device < 0)                                                                       \
            return kp::fail((ctx), KP_ERR_STATE, "this call needs a HIP device; the context is host-only"); \
        KP_HIP_TRY((ctx), hipSetDevice((ctx)->device));                                              \
    } while (0)

#define KP_HIP_TRY(ctx, expr)                                  \
    do {                                                       \
        hipError_t kp_e_ = (expr);                             \
        if (kp_e_ != hipSuccess) return kp::hip_fail((ctx), kp_e_, #expr); \
    } while (0)
