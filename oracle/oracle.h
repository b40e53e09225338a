/*
 * CPU oracle for the kazen_proto path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP library
 * is compared against. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product (libkp_hip.so) never
 * links, loads or calls it.
 *
 * It restates, in plain C++ that runs on host cores, the reference's
 * intended algorithm for:
 *   - rays / closest-hit / occlusion   (src/core/ray.h:10-45, hitable.h:12-33,
 *                                        accel.cpp:348-432 - Embree robust mode)
 *   - the PCG32 sampler                 (src/core/sampler.cpp:14-60 via pcg32.h)
 *   - closures                          (src/shading/bsdfs.cpp, microfacet.h, fresnel.h)
 *   - the path integrator               (src/core/integrator.cpp:362-561, light.cpp:52-85)
 * with the reference's defects replaced by the intended math (each deviation
 * is cited next to the code that implements it).
 *
 * All entry points are extern "C" with plain pointers so that Python (ctypes)
 * can drive them.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG ---------------------------------------------------------------- */
/* Fill out[n] with successive pcg32 outputs of the stream seeded with
 * (initstate, initseq). */
void or_pcg32_u32(uint64_t initstate, uint64_t initseq, uint32_t* out, int n);
/* Same stream, mapped to [0,1) floats with 23 random mantissa bits. */
void or_pcg32_f32(uint64_t initstate, uint64_t initseq, float* out, int n);

/* ---- Geometry ----------------------------------------------------------- */
/* Ray record: {ox,oy,oz,tmin,dx,dy,dz,tmax}, 32 bytes. Hit record:
 * {t, prim, u, v}, 16 bytes; prim = 0xFFFFFFFF on a miss; u,v = barycentric
 * weights of vertices 1 and 2 (Embree convention). */

/* Single watertight ray-triangle test. Returns 1 on hit with t in
 * [tmin, tmax]; writes t,u,v. */
int or_intersect_triangle(const float* ray8, const float* v0, const float* v1,
                          const float* v2, float* t, float* u, float* v);

/* Brute-force closest hit over all triangles (O(R*N)), threads = 0 -> all
 * cores. tris: n_tris*9 floats (v0,v1,v2). */
void or_closest_bruteforce(const float* tris, uint32_t n_tris, const float* rays,
                           uint32_t n_rays, float* hits, int threads);

/* Count rays whose closest hit has a second candidate (another primitive)
 * with t within 1 ulp of the closest t.  Brute force. */
uint32_t or_count_near_ties(const float* tris, uint32_t n_tris, const float* rays,
                            uint32_t n_rays, int threads);

/* Two-level closest hit by brute force (instancing). Instance k places mesh
 * inst_mesh[k] (triangles mesh_first..+mesh_count of mesh_tris, object space)
 * with the row-major 3x4 world-to-object matrix inst_w2o[12k..]; the ray goes
 * to object space as o'_a = ((m_a0 ox + m_a1 oy) + m_a2 oz) + m_a3 and
 * d'_a = (m_a0 dx + m_a1 dy) + m_a2 dz in FP32 (t keeps its meaning); the
 * hit's prim = inst_base[k] + index in the mesh; ties by the lower prim.
 * An exactly-identity matrix uses the ray as given (no arithmetic). */
void or_instanced_closest(const float* mesh_tris, const uint32_t* mesh_first,
                          const uint32_t* mesh_count, const float* inst_w2o,
                          const uint32_t* inst_mesh, const uint32_t* inst_base,
                          uint32_t n_inst, const float* rays, uint32_t n_rays,
                          float* hits, int threads);

/* Oracle BVH (binary, binned SAH, independent of the product's BVH). */
typedef struct or_bvh or_bvh;
or_bvh* or_bvh_build(const float* tris, uint32_t n_tris);
void or_bvh_free(or_bvh* b);
uint32_t or_bvh_node_count(const or_bvh* b);
/* Closest hit through the oracle BVH, threaded over rays. */
void or_bvh_closest(const or_bvh* b, const float* rays, uint32_t n_rays,
                    float* hits, int threads);
/* Rays whose closest hit (hits from or_bvh_closest) has another primitive
 * within 1 ulp of its t: the near ties of the closest-hit parity report. */
uint32_t or_bvh_count_near_ties(const or_bvh* b, const float* rays, uint32_t n_rays,
                                const float* hits, int threads);
/* Occlusion (any hit in [tmin,tmax]) -> out[i] in {0,1}. */
void or_bvh_occluded(const or_bvh* b, const float* rays, uint32_t n_rays,
                     uint8_t* out, int threads);
/* Wall-clock seconds for or_bvh_closest on the given rays (timed here so the
 * measurement excludes Python overhead). */
double or_bvh_closest_timed(const or_bvh* b, const float* rays, uint32_t n_rays,
                            float* hits, int threads);
int or_hardware_threads(void);

/* Closest hit through the HIP library's BVH8 records (128-byte nodes, 48-byte
 * triangles, layout documented in kazen_proto_amd/csrc/kp_layout.h), decoded
 * in exact double arithmetic and traversed on the host. Used to check the
 * builder on the CPU and to count the nodes visited / triangles tested per
 * ray (counts may be NULL). */
void or_bvh8_closest(const void* nodes, uint64_t n_nodes, const void* tris,
                     uint64_t n_tris, const float* rays, uint32_t n_rays,
                     float* hits, uint64_t* nodes_visited, uint64_t* tris_tested,
                     int threads);

/* ---- Shading ------------------------------------------------------------ */
/* Material record, shared layout with the scene description:
 *   type: 0 Lambert, 1 conductor (Schlick F0), 2 rough dielectric,
 *         3 gloss (microfacet reflection with dielectric Fresnel, KpGloss)
 *   dist: 0 GGX, 1 Beckmann
 *   params: albedo[3] | F0[3] | alpha (along the frame's x) | eta | Le[3] |
 *           alpha_y (along z; 0 = alpha)                                  */
typedef struct {
    int32_t type;
    int32_t dist;
    float albedo[3];
    float F0[3];
    float alpha;
    float eta;
    float Le[3];
    float alpha_y;
} or_material;   /* 64 bytes */

/* One lobe of a composite material: closure weight and parameters (the
 * lobe's Le is ignored). */
#define or_max_lobes 8
typedef struct {
    float weight[3];
    float pad;
    or_material lobe;
} or_lobe;   /* 80 bytes */

/* BSDF evaluation in the y-up local frame (cos theta = w.y). wo points
 * toward the previous vertex, wi toward the next one.  Writes f[3] (without
 * the cosine) and the solid-angle pdf of sampling wi given wo. */
void or_bsdf_eval(const or_material* m, const float* wo, const float* wi,
                  float* f, float* pdf);
/* Sample wi given wo and 3 uniforms; writes wi[3], f[3], pdf. Returns 0 when
 * the sample is invalid. */
int or_bsdf_sample(const or_material* m, const float* wo, const float* u3,
                   float* wi, float* f, float* pdf);
/* Microfacet pieces, for the closed-form tests (roughness ax along x, ay
 * along z). */
double or_mf_D(int dist, float ax, float ay, const float* m);
double or_mf_lambda(int dist, float ax, float ay, const float* w);
void or_mf_sample_vndf(int dist, float ax, float ay, const float* wo, const float* u2,
                       float* m);
double or_fresnel_dielectric(double cos_i, double eta);
/* Shading point and normal of a hit at barycentrics (u, v) of a triangle
 * with vertex normals (the integrator's post-hit step). */
void or_shading_point(const float* tri9, const float* normals9, float u, float v, float* P, float* ns);
double or_erfinv(double x);
/* Batched forms (n directions, 3 floats each) for quadrature tests. */
void or_bsdf_eval_batch(const or_material* m, const float* wo, const float* wi, int n,
                        float* f, float* pdf);
void or_bsdf_sample_batch(const or_material* m, const float* wo, const float* u3, int n,
                          float* wi, float* f, float* pdf, int32_t* ok);
void or_mf_eval_batch(int dist, float ax, float ay, const float* w, int n, double* D,
                      double* lambda);
/* Composite materials (n_lobes weighted lobes; with lobes == NULL the single
 * material `single`): f = sum_i w_i f_i and pdf = sum_i p_i pdf_i with the
 * lobe-pick probabilities p_i = dot(w_i, beta) / sum_j dot(w_j, beta)
 * (beta NULL = (1,1,1)); the sampler picks lobe i by u[0] and re-uses u[0]
 * rescaled within the lobe. */
void or_composite_eval_batch(const or_lobe* lobes, const or_material* single, uint32_t n_lobes,
                             const float* beta, const float* wo, const float* wi, int n,
                             float* f, float* pdf);
void or_composite_sample_batch(const or_lobe* lobes, const or_material* single, uint32_t n_lobes,
                               const float* beta, const float* wo, const float* u3, int n,
                               float* wi, float* f, float* pdf, int32_t* ok);

/* ---- Path integrator ---------------------------------------------------- */
typedef struct {
    uint32_t width, height, spp, max_depth, rr_depth;
    uint32_t seed;
    uint32_t spp_offset;   /* first sample index (for resumable renders) */
    uint32_t flags;        /* bit0: write per-sample radiance of sample 0 */
    float eye[3], lookat[3], up[3];
    float tan_half_fov;    /* tan(vfov/2), computed once by the caller */
} or_render_desc;

/* Scene for rendering: world-space triangle soup + per-triangle material id.
 * Lights are the triangles whose material has Le != 0, grouped by light_id
 * (-1 = not a light); each light group is one area light sampled uniformly
 * by area (reference GeometryLight / TriangleMesh::sample). */
typedef struct or_scene or_scene;
or_scene* or_scene_create(const float* tris, uint32_t n_tris,
                          const int32_t* tri_material, const int32_t* tri_light,
                          const or_material* materials, uint32_t n_materials);
void or_scene_free(or_scene* s);
/* Replaces material `material` of the scene by a composite of n_lobes lobes
 * (1..8) emitting Le (may be NULL). */
int or_scene_set_composite(or_scene* s, uint32_t material, const or_lobe* lobes, uint32_t n_lobes,
                           const float* Le);
/* Per-triangle vertex normals (9 floats per triangle, world space, unit);
 * has[i] = 0: triangle i shades with its geometric normal. */
int or_scene_set_normals(or_scene* s, const float* normals9, const uint8_t* has);
/* FP64 render: out[h*w*3] = mean radiance per pixel. sample0 (optional,
 * may be NULL): radiance of the first sample of every pixel, FP32. */
int or_render(const or_scene* s, const or_render_desc* d, double* out,
              float* sample0, int threads);
/* The pixels listed (row-major indices) of or_render's image: out[3*i] is
 * the mean radiance of pixel pixels[i] over d->spp samples from d->spp_offset,
 * the same paths or_render traces for it. */
int or_render_pixels(const or_scene* s, const or_render_desc* d, const uint32_t* pixels, uint32_t n,
                     double* out, int threads);
/* One path per caller ray (rays: 8 floats each, tmin/tmax ignored): path i
 * uses PCG32 sequence i and sample index d->spp_offset; out[3*i] radiance. */
int or_integrate_rays(const or_scene* s, const float* rays, uint32_t n,
                      const or_render_desc* d, float* out, int threads);

/* Guided variants (path-guiding network; no reference counterpart). Every
 * vertex draws 12 numbers; with sample != 0, guided vertices (Lambert, or
 * conductor with alpha >= min_alpha, at depth <= max_depth unless max_depth
 * is 0) sample the network with probability 1 - bsdf_fraction and weight by
 * the one-sample MIS density. The network position input is (P - bmin) * binv
 * in FP32. */
typedef struct {
    const float* params;
    float bmin[3], binv[3];
    float bsdf_fraction, min_alpha;
    int32_t sample;
    uint32_t max_depth;
} or_guide_cfg;
int or_render_guided(const or_scene* s, const or_render_desc* d, const or_guide_cfg* g,
                     double* out, float* sample0, int threads);
int or_integrate_rays_guided(const or_scene* s, const float* rays, uint32_t n,
                             const or_render_desc* d, const or_guide_cfg* g, float* out,
                             int threads);

/* ---- Path-guiding network (oracle_guide.cpp) ---------------------------- */
/* in6 per sample: position normalised to [0,1]^3, unit normal. Canonical
 * parameter layout: W0[48x64] b0 W1[64x64] b1 W2 b2 W3 b3 W4[64x32] b4. With
 * emulate_bf16 the weights, features, hidden activations and back-propagated
 * gradients are rounded to bfloat16 the way the GPU kernels feed them to the
 * matrix cores (accumulation in double). */
int or_mlp_param_count(void);
void or_mlp_encode(const float* in6, double* feat48);
void or_mlp_forward(const float* params, int n, const float* in6, int emulate_bf16,
                    double* out32);
double or_vmf_pdf(const double* out32, const float* dir);
void or_vmf_sample(const double* out32, const float* u3, float* dir, double* pdf);
double or_mlp_grad(const float* params, int n, const float* in6, const float* dirs,
                   const float* weights, int emulate_bf16, double* grad);
void or_adam_step(float* params, float* m, float* v, const float* grad, int n, int t,
                  float lr, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif
