/*
 * kp_hip.h - C ABI of the MI355X-native kazen_proto hot path.
 *
 * Plain C types only (no torch, no HIP types in the signatures). Every call
 * returns a status (KP_OK = 0); on failure kp_last_error() describes it. The
 * library never exits the process and never throws across this boundary
 * (unlike the reference's CUDA_CHECK/OPTIX_CHECK -> exit(1),
 * src/core/optix_utils.h:16-58).
 *
 * Ownership: caller memory stays the caller's; host arrays are copied during
 * the call. Device memory created by the library belongs to the kp_ctx.
 * Threading: one kp_ctx issues work on one HIP stream; calls on one context
 * must be serialised by the caller (the reference serialises scene building
 * and calls intersect concurrently only through const CPU code paths).
 *
 * Pointer kinds: KP_PTR_HOST arrays are staged through the context;
 * KP_PTR_DEVICE arrays must be device memory of the context's device and the
 * call is asynchronous on the context stream (errors detected by the kernels
 * are reported by the next kp_synchronize).
 */
#ifndef KP_HIP_H
#define KP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KP_OK 0
#define KP_ERR_INVALID 1     /* bad argument */
#define KP_ERR_STATE 2       /* call out of order (e.g. trace before build) */
#define KP_ERR_HIP 3         /* HIP runtime error */
#define KP_ERR_ALLOC 4       /* device allocation failed */
#define KP_ERR_KERNEL 5      /* a kernel reported an internal limit */

#define KP_PTR_HOST 0
#define KP_PTR_DEVICE 1

typedef struct kp_ctx kp_ctx;

/* Ray: origin, tmin, direction, tmax (32 bytes). Directions need not be
 * normalised; t is measured in units of |d|. Valid hits have t in
 * [tmin, tmax]. Mirrors Ray (reference src/core/ray.h:10-45) without the
 * differentials and time. */
typedef struct {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} kp_ray;

/* Closest hit: prim = caller primitive index (global across meshes, in the
 * order meshes were added) or 0xFFFFFFFF on a miss (t = +inf). u, v are the
 * barycentric weights of vertices 1 and 2 (Embree's convention, used by the
 * reference at src/core/accel.cpp:393-394). 16 bytes. */
typedef struct {
    float t;
    uint32_t prim;
    float u, v;
} kp_hit;

/* How kp_bvh_build treats instances (kp_instance_add). Baking transforms
 * every placement's triangles into the one-level BVH (on the host, FP64 to
 * FP32), which traverses fastest; two-level traversal shares each prototype's
 * BVH and triangles among its placements, which saves memory. Primitive ids
 * are the same either way. AUTO bakes while the placed instance triangles
 * number at most KP_BAKE_BUDGET. */
#define KP_INSTANCES_AUTO 0u
#define KP_INSTANCES_TWO_LEVEL 1u
#define KP_INSTANCES_BAKE 2u
#define KP_BAKE_BUDGET (1u << 24)

/* BVH node width: children per node, 2..8 in the same 128-byte node record
 * (0 means 8). KP_NODE_WIDTH_AUTO builds widths 4, 6 and 8, times each on
 * 2^20 probe rays (origins uniform in the scene's bounds, directions uniform
 * on the sphere) and keeps the fastest; one-level BVHs on a device only
 * (otherwise 8). Hits do not depend on the width. */
#define KP_NODE_WIDTH_AUTO 1u

typedef struct {
    uint32_t leaf_size;   /* max triangles per leaf, 1..4 (default 4) */
    uint32_t sah_bins;    /* SAH bins, 4..64 (default 32) */
    uint32_t instances;   /* KP_INSTANCES_* (default AUTO) */
    uint32_t node_width;  /* 2..8 children per node, 0 = 8, or KP_NODE_WIDTH_AUTO */
    uint32_t reserved[4];
} kp_build_opts;

typedef struct {
    uint64_t n_tris;
    uint64_t n_nodes;          /* 128-byte BVH8 nodes */
    uint64_t n_bvh2_nodes;
    uint64_t bvh_bytes;        /* nodes + triangle records on the device */
    uint32_t max_depth;        /* BVH8 levels */
    uint32_t spill_entries;    /* per-lane stack spill capacity */
    uint32_t grid_closest;     /* persistent blocks per launch */
    uint32_t grid_occluded;
    double build_seconds;      /* host BVH build wall time */
    uint32_t device_cus;
    uint32_t two_level;        /* 1: instances traversed two-level; 0: one-level (baked) */
    uint32_t node_width;       /* children per node of the built BVH */
    float width_ms[3];         /* KP_NODE_WIDTH_AUTO: probe time (ms) of widths 4, 6, 8; else 0 */
    uint32_t reserved[2];
} kp_stats;

/* Context ---------------------------------------------------------------- */
/* Creates a context on HIP device `device`. Replaces the accelerator
 * construction of Scene::Scene / set_accelerator (src/core/scene.cpp:281,
 * src/core/scene.h:60-77). device = KP_DEVICE_HOST makes a host-only
 * context: scene building, kp_bvh_build and the *_host queries work without
 * any GPU (no HIP call is made); every call that needs the GPU returns
 * KP_ERR_STATE. */
#define KP_DEVICE_HOST (-1)
int kp_create(kp_ctx** out, int device);
void kp_destroy(kp_ctx* ctx);
/* Last error text of the context (or of the last failed kp_create when ctx is
 * NULL). Valid until the next call on the context. */
const char* kp_last_error(const kp_ctx* ctx);
/* Issue all further work on `hip_stream` (a hipStream_t, NULL = the context's
 * own stream). */
int kp_set_stream(kp_ctx* ctx, void* hip_stream);
/* Waits for the context's stream and reports kernel-detected errors. */
int kp_synchronize(kp_ctx* ctx);

/* Scene ------------------------------------------------------------------ */
/* Adds a triangle mesh (copied). xform: optional row-major 3x4 object->world
 * matrix (NULL = identity), applied on the host. The mesh's primitives get
 * global ids [first, first + n_tris) where first = triangles added before.
 * Replaces Accelerator::add_trianglemesh (src/core/accel.h:26,
 * EmbreeAccel::add_trianglemesh src/core/accel.cpp:258-315) together with
 * Scene::add_mesh (src/core/scene.h:144-160). */
int kp_mesh_add(kp_ctx* ctx, const float* xyz, uint32_t n_verts, const uint32_t* idx,
                uint32_t n_tris, const float* xform12, uint32_t* mesh_id);
/* Vertex normals of a mesh for shading: 9 floats per triangle (the normals
 * at its three vertices, in the mesh's triangle order and object space; the
 * mesh's transform is applied as its inverse transpose, then they are
 * normalised). The integrator then shades with the barycentric interpolation
 * of the normals and moves the shading point off the flat triangle (Hanika
 * 2021), as TriangleMesh::post_hit does (src/core/shape.cpp:615-654); the
 * geometric normal still decides which side a ray leaves from. Meshes without
 * normals shade with their geometric normal. */
int kp_mesh_set_normals(kp_ctx* ctx, uint32_t mesh_id, const float* normals9);
/* Builds the BVH over every mesh added so far and uploads it. Replaces
 * Accelerator::build (src/core/accel.h:32; EmbreeAccel::build accel.h:75-77,
 * add_instances accel.cpp:317-346) and Scene::build_bvh (scene.h:162-164). */
int kp_bvh_build(kp_ctx* ctx, const kp_build_opts* opts);
int kp_get_stats(kp_ctx* ctx, kp_stats* out);

/* Instancing. A prototype mesh is added once in object space with
 * kp_blas_add and placed any number of times with kp_instance_add (row-major
 * 3x4 object-to-world matrix, invertible). kp_bvh_build either bakes the
 * placements (kp_build_opts.instances) or builds one BVH8 per prototype (and
 * one over the kp_mesh_add meshes, traversed first) and a top-level BVH8 over
 * the instances' world boxes; a ray enters each instance it reaches in the
 * instance's object space (o' = M o + m, d' = M d in FP32, so t keeps its
 * meaning; an exactly-identity placement uses the ray as given). Primitive ids: the kp_mesh_add meshes first, then every
 * instance's primitives in instance order. The integrator shades instances
 * like any mesh (material per prototype, an emissive prototype makes each
 * instance a light). Replaces Accelerator::add_instances
 * (src/core/accel.h:28-31; EmbreeAccel::add_instances
 * src/core/accel.cpp:317-346). */
int kp_blas_add(kp_ctx* ctx, const float* xyz, uint32_t n_verts, const uint32_t* idx, uint32_t n_tris,
                uint32_t* blas_id);
int kp_blas_set_material(kp_ctx* ctx, uint32_t blas_id, uint32_t material_id);
/* Prototype vertex normals (object space, 9 floats per triangle; each
 * instance applies the inverse transpose of its placement). */
int kp_blas_set_normals(kp_ctx* ctx, uint32_t blas_id, const float* normals9);
int kp_instance_add(kp_ctx* ctx, uint32_t blas_id, const float* xform12, uint32_t* instance_id);
/* Material of one instance, overriding its prototype's
 * (kp_blas_set_material) for the integrator. */
int kp_instance_set_material(kp_ctx* ctx, uint32_t instance_id, uint32_t material_id);
/* Where a hit's primitive id comes from: *kind = 0 for a kp_mesh_add mesh
 * (*index = mesh id), 1 for an instance (*index = instance id); *local = the
 * triangle's index within that mesh or prototype. Valid after kp_bvh_build.
 * Replaces the (geomID, instID, primID) of an Embree hit
 * (src/core/accel.cpp:374-399). Any output may be NULL. */
int kp_prim_locate(kp_ctx* ctx, uint32_t prim, uint32_t* kind, uint32_t* index, uint32_t* local);
/* Removes every mesh, prototype and instance and the BVH (materials, film
 * and guiding state stay), so a scene can be rebuilt from scratch.
 * (The reference rebuilds by constructing a new Embree scene; no counterpart
 * call.) */
int kp_geometry_clear(kp_ctx* ctx);
/* The prototype and both matrices of an instance (the world-to-object one is
 * the FP32 matrix the traversal applies). Any output may be NULL. */
int kp_instance_get(kp_ctx* ctx, uint32_t instance_id, uint32_t* blas_id, float* o2w12, float* w2o12);

/* Queries ---------------------------------------------------------------- */
/* Closest hit for n rays. Replaces HitableInterface::intersect(const Ray&,
 * Intersection&) / intersect(const Ray&, float&) (src/core/hitable.h:12-13)
 * as implemented by EmbreeAccel::intersect (src/core/accel.cpp:348-432),
 * batched. */
int kp_intersect_closest(kp_ctx* ctx, const kp_ray* rays, uint32_t n, kp_hit* hits,
                         int ptr_kind);
/* Occlusion for n rays: out[i] = 1 if any primitive is hit with t in
 * [tmin, tmax]. Replaces HitableInterface::occluded (src/core/hitable.h:15-33),
 * batched; 4 bytes per result (uint32). */
int kp_occluded(kp_ctx* ctx, const kp_ray* rays, uint32_t n, uint32_t* out, int ptr_kind);
/* Host closest hit / occlusion over the same BVH, on the calling thread's
 * host core: bit-identical to kp_intersect_closest / kp_occluded (the same
 * node test, instance transform and watertight triangle arithmetic). Safe to
 * call from any number of threads at once on one context (the context is only
 * read; kp_bvh_build, kp_geometry_clear and kp_destroy must not run
 * concurrently). Errors are described by kp_last_error(NULL) of the calling
 * thread. Replaces the const per-ray HitableInterface::intersect(const Ray&,
 * Intersection&) / intersect(const Ray&, float&) (src/core/hitable.h:12-13;
 * EmbreeAccel::intersect, src/core/accel.cpp:348-432), which the reference
 * calls concurrently from its TBB tile workers (src/core/renderer.cpp:24-26). */
int kp_intersect_closest_host(const kp_ctx* ctx, const kp_ray* rays, uint32_t n, kp_hit* hits);
int kp_occluded_host(const kp_ctx* ctx, const kp_ray* rays, uint32_t n, uint32_t* out);
/* kp_intersect_closest for device pointers that allocates nothing: no
 * hipMalloc, hipFree or allocating library call during the call. No
 * counterpart in the reference; for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it). The
 * traversal's work counter, error word and stack spill area live in a static
 * 64 MiB device array of the library (one per device, held until the process
 * ends) instead of the context's buffers; results are bit-identical to
 * kp_intersect_closest. rays / hits must be device memory; the call is
 * asynchronous on the context stream and kernel-detected errors are reported
 * by the next kp_synchronize. A BVH whose traversal stack does not fit the
 * scratch for one block of 256 lanes is rejected with KP_ERR_INVALID (deeper
 * BVHs run on fewer blocks). Every call on one device shares that device's
 * scratch: calls that overlap in time on one device (any context, any
 * stream) race and must be serialised by the caller. */
int kp_intersect_closest_noalloc(kp_ctx* ctx, const kp_ray* rays, uint32_t n, kp_hit* hits);
/* Instrumented closest hit: same results, plus the totals of BVH8 nodes
 * visited and triangles tested (for the bytes-per-ray roofline figure). */
int kp_intersect_closest_counted(kp_ctx* ctx, const kp_ray* rays, uint32_t n, kp_hit* hits,
                                 int ptr_kind, uint64_t* nodes_visited,
                                 uint64_t* tris_tested);
/* Host-only BVH8 build (no device needed): builds the same 8-wide BVH that
 * kp_bvh_build uploads and hands back its node and triangle records, for
 * inspection, offline builds and host-side checks. tris9: 9 floats per
 * triangle (world space). */
typedef struct kp_bvh8_blob kp_bvh8_blob;
int kp_bvh8_build_host(const float* tris9, uint32_t n_tris, const kp_build_opts* opts,
                       kp_bvh8_blob** out);
/* n_nodes 128-byte nodes, n_tris 48-byte triangle records. */
int kp_bvh8_blob_info(const kp_bvh8_blob* b, uint64_t* n_nodes, uint64_t* n_tris,
                      uint32_t* max_depth);
const void* kp_bvh8_blob_nodes(const kp_bvh8_blob* b);
const void* kp_bvh8_blob_tris(const kp_bvh8_blob* b);
void kp_bvh8_blob_free(kp_bvh8_blob* b);
/* Host closest hit / occlusion over a blob (same arithmetic as
 * kp_intersect_closest_host; thread-safe, no device). Errors: kp_last_error(NULL). */
int kp_bvh8_blob_closest(const kp_bvh8_blob* b, const kp_ray* rays, uint32_t n, kp_hit* hits);
int kp_bvh8_blob_occluded(const kp_bvh8_blob* b, const kp_ray* rays, uint32_t n, uint32_t* out);

/* Tuning: refill a wave once fewer than `lanes` (1..64) of its lanes are
 * busy (default 40). */
int kp_set_trace_tuning(kp_ctx* ctx, uint32_t refill_below_lanes);
/* Tuning: at most `max_paths` paths (2^16 .. 2^28) in flight per pass of
 * kp_render / kp_render_accumulate and per chunk of kp_integrate_rays
 * (default 2^26: 13.4 GB of path state, 200 B per path). Larger passes keep
 * every bounce's launches full (a pass's later bounces hold only the paths
 * still alive); results do not depend on it. Guided renders use
 * kp_guiding_desc.pass_paths instead. */
int kp_set_pass_paths(kp_ctx* ctx, uint32_t max_paths);
/* Frees the context's cached work buffers (integrator path state, host-call
 * staging); the next call that needs them allocates them again. Scene, BVH,
 * film and guiding state stay. For callers that keep many contexts. */
int kp_release_buffers(kp_ctx* ctx);

/* Integrator -------------------------------------------------------------- */
#define KP_MAT_LAMBERT 0      /* Diffuse (src/shading/bsdfs.cpp:27-41) */
#define KP_MAT_CONDUCTOR 1    /* microfacet reflection, Schlick F0
                                 (KpPrincipleSpecularReflection, bsdfs.cpp:688-730) */
#define KP_MAT_DIELECTRIC 2   /* rough dielectric (KpGlass, bsdfs.cpp:441-554) */
#define KP_MAT_GLOSS 3        /* microfacet reflection with dielectric Fresnel F(eta, m.wo)
                                 (KpGloss, bsdfs.cpp:359-439; OSL microfacet(..., refract = 0)) */
#define KP_DIST_GGX 0
#define KP_DIST_BECKMANN 1

/* Closure parameters of one material (64 bytes). Stands in for the OSL
 * shader group of a mesh (the reference maps OSL closures to these lobes in
 * process_closure, src/shading/bsdf.cpp:184-328). Le != 0 makes every mesh
 * using the material a one-sided area light (GeometryLight,
 * src/core/light.cpp:52-85), emitting on the side its winding normal
 * (v1-v0)x(v2-v0) points to. Microfacet roughness is anisotropic as
 * MicrofacetParams' xalpha / yalpha (src/shading/bsdfs.h:41-46,
 * microfacet.h:60-93): alpha along the shading frame's tangent, alpha_y
 * along its bitangent (the frame of src/base/utils.h:183-210). */
typedef struct {
    int32_t type;        /* KP_MAT_* */
    int32_t dist;        /* KP_DIST_* (microfacet closures) */
    float albedo[3];     /* Lambert reflectance */
    float F0[3];         /* conductor normal-incidence reflectance */
    float alpha;         /* microfacet roughness along the tangent (> 0) */
    float eta;           /* dielectric and gloss: n_inside / n_outside */
    float Le[3];         /* emitted radiance */
    float alpha_y;       /* roughness along the bitangent; 0 = alpha (isotropic) */
} kp_material;

/* One lobe of a composite material: the closure weight (the OSL closure
 * colour) and the lobe's parameters (its Le is ignored). */
#define KP_MAX_LOBES 8
typedef struct {
    float weight[3];
    float pad;
    kp_material lobe;
} kp_lobe;

/* One render: width x height pixels, spp samples per pixel starting at sample
 * index spp_offset (so renders resume), paths of up to max_depth scattering
 * vertices, Russian roulette from vertex rr_depth, pinhole camera with
 * tan(vertical fov / 2). Replaces the Film/Camera/Renderer setup of the
 * reference (src/core/renderer.cpp:18-90, camera.h:50-96). */
typedef struct {
    uint32_t width, height, spp, max_depth, rr_depth;
    uint32_t seed;
    uint32_t spp_offset;
    uint32_t flags;      /* KP_RENDER_* bits */
    float eye[3], lookat[3], up[3];
    float tan_half_fov;
} kp_render_desc;

/* flags: sample with the path-guiding network (see kp_guiding_configure). */
#define KP_RENDER_GUIDED 1u

typedef struct {
    uint64_t paths;           /* camera paths traced by the last accumulate */
    uint64_t extension_rays;  /* closest-hit rays */
    uint64_t shadow_rays;     /* occlusion rays */
    double seconds;           /* wall time of the last accumulate */
    uint32_t passes;
    uint32_t paths_per_pass;
    uint32_t train_steps;     /* guiding: Adam steps taken by the last accumulate */
    uint32_t guided_passes;   /* guiding: passes that sampled from the network */
    uint32_t train_records;   /* guiding: valid training samples of the last pass */
    uint32_t reserved[3];
} kp_render_stats;

int kp_material_add(kp_ctx* ctx, const kp_material* m, uint32_t* material_id);
/* A material of 1..8 weighted lobes emitting Le (may be NULL): replaces the
 * CompositeClosure an OSL shader group builds (src/shading/bsdf.h:132-196,
 * bsdf.cpp:17-74). f = sum_i weight_i f_i; at a vertex with path throughput
 * beta, lobe i is sampled with probability p_i = dot(weight_i, beta) /
 * sum_j dot(weight_j, beta) (the reference's compute_pdfs, bsdf.h:157-176;
 * its sampler picks uniformly instead) and the density is sum_i p_i pdf_i.
 * kp_material_add is the one-lobe case with weight 1. */
int kp_material_add_composite(kp_ctx* ctx, const kp_lobe* lobes, uint32_t n_lobes, const float* Le,
                              uint32_t* material_id);
/* Assigns a material to a mesh (default 0). */
int kp_mesh_set_material(kp_ctx* ctx, uint32_t mesh_id, uint32_t material_id);
/* Film: FP32 RGB sum + Kahan compensation per pixel, zeroed here. */
int kp_film_reset(kp_ctx* ctx, uint32_t width, uint32_t height);
/* Traces desc->spp samples per pixel into the film (wavefront path
 * integrator, replacing PathIntegrator::Li, src/core/integrator.cpp:362-561,
 * and the never-defined Renderer::render(SceneGPU&) of renderer.h:30-31). */
int kp_render_accumulate(kp_ctx* ctx, const kp_render_desc* desc);
/* Mean radiance per pixel (row-major, top row first, RGB float). */
int kp_film_read(kp_ctx* ctx, float* rgb, int ptr_kind);
/* kp_film_reset + kp_render_accumulate + kp_film_read. */
int kp_render(kp_ctx* ctx, const kp_render_desc* desc, float* rgb, int ptr_kind);
int kp_render_get_stats(kp_ctx* ctx, kp_render_stats* out);
/* Batched Integrator::Li (src/core/integrator.h:29, PathIntegrator::Li
 * integrator.cpp:362-561): traces one path from each caller ray and writes its
 * radiance (rgb[3*i]). Uses desc->max_depth, rr_depth, seed and spp_offset
 * (the sample index); path i draws from PCG32 sequence i. */
int kp_integrate_rays(kp_ctx* ctx, const kp_ray* rays, uint32_t n, const kp_render_desc* desc,
                      float* rgb, int ptr_kind);

/* Path guiding -------------------------------------------------------------
 * No counterpart in the reference (kazen_proto has no path guiding). A small
 * MLP maps a frequency-encoded position (normalised to [0,1]^3) and normal to
 * an 8-lobe von Mises-Fisher mixture over directions; forward, backward and
 * Adam run on the matrix cores in bf16 with FP32 accumulation. in6 per sample:
 * {px, py, pz, nx, ny, nz}. Parameters: 17,696 FP32 values in the order
 * W0[48x64] b0 W1[64x64] b1 W2 b2 W3 b3 W4[64x32] b4 (row-major, input-major). */
#define KP_GUIDE_PARAMS 17696
int kp_guiding_init(kp_ctx* ctx, uint64_t seed);
int kp_guiding_set_params(kp_ctx* ctx, const float* params, uint32_t n);
int kp_guiding_get_params(kp_ctx* ctx, float* params, uint32_t n);
/* Guiding checkpoint (save / load of the network state): the
 * FP32 parameters, both Adam moments, the Adam step, the count of training
 * passes and the position box, KP_GUIDE_CHECKPOINT_BYTES bytes in all, in a
 * caller buffer (host memory). kp_guiding_save with buf NULL only reports
 * the size through *written. Loading then training continues exactly as the
 * saved context would have (training is deterministic for a given batch).
 * The render settings of kp_guiding_configure are not part of it. No
 * counterpart in the reference. */
#define KP_GUIDE_CHECKPOINT_BYTES (64u + 3u * 4u * KP_GUIDE_PARAMS)
int kp_guiding_save(kp_ctx* ctx, void* buf, uint64_t bytes, uint64_t* written);
int kp_guiding_load(kp_ctx* ctx, const void* buf, uint64_t bytes);
/* Network outputs (32 per sample); host pointers. */
int kp_guiding_eval(kp_ctx* ctx, const float* in6, uint32_t n, float* out32);
/* Mixture density of unit directions (host pointers). */
int kp_guiding_pdf(kp_ctx* ctx, const float* in6, const float* dirs, uint32_t n, float* pdf);
/* Mixture samples from uniforms u3 (lobe, two direction), with their density. */
int kp_guiding_sample(kp_ctx* ctx, const float* in6, const float* u3, uint32_t n, float* dirs,
                      float* pdf);
/* Gradient of mean_i(-w_i log q_i(dir_i)) over n <= 262144 samples; with
 * apply != 0 also one Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8).
 * grad (optional) receives the gradient in the parameter order above. */
int kp_guiding_train_step(kp_ctx* ctx, const float* in6, const float* dirs, const float* weights,
                          uint32_t n, float lr, int apply, float* grad);

/* Guided rendering (desc->flags & KP_RENDER_GUIDED). At a guided vertex (any
 * Lambert material; conductors with alpha >= min_alpha) the next direction is
 * drawn from the BSDF with probability bsdf_fraction and from the network
 * otherwise, and weighted by the one-sample MIS density
 * bsdf_fraction * p_bsdf + (1 - bsdf_fraction) * q; next-event estimation and
 * emission use the same density in their power heuristic, so the estimator
 * stays unbiased for any network. Guided paths draw 12 PCG32 numbers per
 * vertex: the 8 of the plain protocol, then strategy, lobe and two direction
 * numbers.
 * kp_render_accumulate then also trains the network online: each pass records
 * up to max_records guided-vertex samples (position, normal, sampled
 * direction, weight = luminance of the radiance gathered beyond the vertex
 * divided by the path throughput before it) and takes train_steps Adam steps
 * of the loss of kp_guiding_train_step on them, weights normalised to sum 1.
 * Sampling from the network starts once guide_after passes have trained it;
 * after train_passes training passes (if nonzero) each pass takes
 * late_train_steps steps instead (0: the network stays fixed).
 * Guided renders run at most pass_paths paths per pass. The position is
 * normalised with the box [bmin, bmax] (bmin == bmax: the scene's bounds).
 * kp_integrate_rays with the flag samples from the network without training. */
typedef struct {
    float bsdf_fraction;   /* (0, 1]; 0.5 */
    float min_alpha;       /* 0.5 */
    float lr;              /* Adam learning rate; 1e-3 */
    uint32_t train_steps;  /* per pass; 0 = frozen network */
    uint32_t guide_after;  /* training passes before sampling from the network */
    uint32_t max_records;  /* 1 .. 262144 */
    uint32_t pass_paths;   /* >= 1; 262144 */
    uint32_t train_passes; /* passes that train (then the network is frozen); 0 = every pass */
    float bmin[3], bmax[3];
    float weight_clamp;    /* training weights capped at this multiple of the batch's mean
                              (fewer rare-path outliers steering a step); 0 = no cap */
    uint32_t relative_weights; /* 1: a training sample counts by its path's contribution
                                  relative to its pixel's current estimate R, divided by
                                  sqrt(R^2 + 0.01) (the per-pixel relative error the image
                                  is judged by); 0: by its incident radiance */
    uint32_t max_guided_depth; /* only vertices at depth <= this (the camera's hit is
                                  depth 1) are guided and recorded; deeper vertices use
                                  the BSDF alone and skip the network query; 0 = all */
    uint32_t late_train_steps; /* Adam steps per pass after train_passes passes; 0 = frozen */
} kp_guiding_desc;
/* Needs kp_guiding_init first; a later kp_guiding_init keeps the settings. */
int kp_guiding_configure(kp_ctx* ctx, const kp_guiding_desc* desc);

/* Batched scenes -------------------------------------------------------------
 * No counterpart in the reference (its Renderer::render, src/core/renderer.cpp:
 * 18-90, renders one scene per call). kp_batch_create copies the built scenes
 * of n contexts (same device as owner; kp_bvh_build and materials done) into
 * one merged workload owned by owner; the scene contexts may change or go
 * away afterwards. kp_batch_render traces descs[i].spp samples of every scene
 * i into fresh per-scene films, in passes of pass_paths paths drawn from all
 * scenes at once: each bounce is one persistent traversal launch over every
 * scene's rays. With use_graph one pass (about six launches per bounce) is
 * captured into a HIP graph once and replayed. Scene i's film equals, bit for
 * bit, kp_render of that scene alone with descs[i] (not guided).
 * kp_batch_autotune times candidate settings on `passes` passes of the
 * batch's own work, one setting at a time, and keeps the fastest. The BVH
 * node width is 8 in every configuration. */
typedef struct kp_batch kp_batch;
typedef struct {
    uint32_t pass_paths;           /* paths in flight per pass (queue size), 256 .. 2^25 */
    uint32_t shade_blocks_per_cu;  /* 256-thread shade / connect blocks per CU */
    uint32_t trace_blocks_per_cu;  /* persistent traversal blocks per CU (occupancy) */
    uint32_t refill_below;         /* traversal wave refill threshold, lanes 1..64 */
    uint32_t use_graph;            /* replay a captured pass graph */
    uint32_t reserved[3];
} kp_batch_tuning;
int kp_batch_create(kp_ctx* owner, kp_ctx* const* scenes, uint32_t n, kp_batch** out);
int kp_batch_destroy(kp_batch* batch);
int kp_batch_set_tuning(kp_batch* batch, const kp_batch_tuning* tuning);
int kp_batch_get_tuning(kp_batch* batch, kp_batch_tuning* tuning);
int kp_batch_render(kp_batch* batch, const kp_render_desc* descs);
int kp_batch_autotune(kp_batch* batch, const kp_render_desc* descs, uint32_t passes, kp_batch_tuning* out);
/* Mean radiance of scene `scene` of the last kp_batch_render. */
int kp_batch_film_read(kp_batch* batch, uint32_t scene, float* rgb, int ptr_kind);
int kp_batch_get_stats(kp_batch* batch, kp_render_stats* out);

/* Device memory and timing helpers for callers without their own HIP
 * runtime handle (the reference has no counterpart: its callers never see
 * device memory). Buffers belong to the caller and must be freed with
 * kp_device_free before kp_destroy. Copies are synchronous. */
int kp_device_alloc(kp_ctx* ctx, uint64_t bytes, void** ptr);
int kp_device_free(kp_ctx* ctx, void* ptr);
int kp_copy_to_device(kp_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int kp_copy_to_host(kp_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte vector loads and
 * stores), asynchronous on the context stream. Also serves as the known-size
 * reference kernel when calibrating HBM traffic counters. bytes % 16 == 0. */
int kp_device_copy(kp_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* HIP events on the context stream: kp_timer_end returns the milliseconds
 * between the two events (waits for the end event). */
int kp_timer_begin(kp_ctx* ctx);
int kp_timer_end(kp_ctx* ctx, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* KP_HIP_H */
