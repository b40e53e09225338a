// HipAccel: the reference's Accelerator seam over the HIP intersector.
//
// Mirrors src/core/accel.h:13-41 (Accelerator: add_trianglemesh, add_instances,
// build, intersect x2) and replaces EmbreeAccel (accel.h:64-88,
// accel.cpp:185-432).  Per-ray calls from existing callers go to the GPU as a
// batch of one; new callers should use the batched overloads, which are the
// fast path (one kernel launch for the whole batch).
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kazen/core.h"
#include "kp_hip.h"

namespace kazen {

// Subset of TriangleMesh (src/core/shape.h:188-236) the intersector needs.
struct TriangleMesh {
    Mat4f local_to_world;
    std::vector<Vec3f> verts;
    std::vector<Vec3f> norms;
    std::vector<Vec2f> uvs;
    std::vector<Vec3i> indice;
    std::string name;
    std::string shader_name;
    bool is_light = false;
    uint32_t light_id = 0;
};

class Accelerator : public HitableInterface {
public:
    virtual void add_trianglemesh(std::shared_ptr<TriangleMesh>& t) = 0;
    virtual void add_instances(const std::string&, const std::vector<std::string>&, const Mat4f& = Mat4f(),
                               bool = false) {}
    virtual void build(const std::vector<std::string>&) = 0;
};

class KpError : public std::runtime_error {
public:
    KpError(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
    int code;
};

class HipAccel : public Accelerator {
public:
    explicit HipAccel(int device = 0) {
        if (kp_create(&ctx_, device) != KP_OK) throw KpError(KP_ERR_HIP, kp_last_error(nullptr));
    }
    ~HipAccel() override { kp_destroy(ctx_); }
    HipAccel(const HipAccel&) = delete;
    HipAccel& operator=(const HipAccel&) = delete;

    kp_ctx* context() const { return ctx_; }
    const std::vector<std::shared_ptr<TriangleMesh>>& meshes() const { return meshes_; }

    void add_trianglemesh(std::shared_ptr<TriangleMesh>& t) override {
        std::vector<float> xyz(3 * t->verts.size());
        for (size_t i = 0; i < t->verts.size(); ++i)
            for (int k = 0; k < 3; ++k) xyz[3 * i + k] = t->verts[i][k];
        std::vector<uint32_t> idx(3 * t->indice.size());
        for (size_t i = 0; i < t->indice.size(); ++i)
            for (int k = 0; k < 3; ++k) idx[3 * i + k] = (uint32_t)t->indice[i][k];
        float xf[12];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) xf[4 * r + c] = t->local_to_world.m[r][c];
        uint32_t id = 0;
        check(kp_mesh_add(ctx_, xyz.data(), (uint32_t)t->verts.size(), idx.data(),
                          (uint32_t)t->indice.size(), xf, &id));
        first_prim_.push_back(total_tris_);
        total_tris_ += (uint32_t)t->indice.size();
        meshes_.push_back(t);
    }

    void build(const std::vector<std::string>&) override { check(kp_bvh_build(ctx_, nullptr)); }

    bool intersect(const Ray& r, Intersection& isect) const override {
        std::vector<Intersection> out(1);
        std::vector<uint8_t> hit;
        intersect({r}, out, hit);
        if (hit[0]) isect = out[0];
        return hit[0] != 0;
    }

    bool intersect(const Ray& r, float& t) const override {
        kp_ray kr = to_kp(r);
        kp_hit h;
        check(kp_intersect_closest(ctx_, &kr, 1, &h, KP_PTR_HOST));
        if (h.prim == 0xFFFFFFFFu) return false;
        t = h.t;
        return true;
    }

    // Batched closest hit: hit[i] = 1 when rays[i] hits; isects[i] filled.
    void intersect(const std::vector<Ray>& rays, std::vector<Intersection>& isects,
                   std::vector<uint8_t>& hit) const {
        std::vector<kp_ray> kr(rays.size());
        for (size_t i = 0; i < rays.size(); ++i) kr[i] = to_kp(rays[i]);
        std::vector<kp_hit> kh(rays.size());
        check(kp_intersect_closest(ctx_, kr.data(), (uint32_t)kr.size(), kh.data(), KP_PTR_HOST));
        isects.resize(rays.size());
        hit.assign(rays.size(), 0);
        for (size_t i = 0; i < rays.size(); ++i) {
            if (kh[i].prim == 0xFFFFFFFFu) continue;
            hit[i] = 1;
            fill(rays[i], kh[i], isects[i]);
        }
    }

    // Batched occlusion: blocked[i] = 1 if anything lies in [tmin, tmax].
    void occluded(const std::vector<Ray>& rays, std::vector<uint8_t>& blocked) const {
        std::vector<kp_ray> kr(rays.size());
        for (size_t i = 0; i < rays.size(); ++i) kr[i] = to_kp(rays[i]);
        std::vector<uint32_t> o(rays.size());
        check(kp_occluded(ctx_, kr.data(), (uint32_t)kr.size(), o.data(), KP_PTR_HOST));
        blocked.assign(o.begin(), o.end());
    }
    using HitableInterface::occluded;

    static kp_ray to_kp(const Ray& r) {
        return kp_ray{r.origin[0], r.origin[1], r.origin[2], r.tmin,
                      r.direction[0], r.direction[1], r.direction[2], r.tmax};
    }

private:
    void check(int rc) const {
        if (rc != KP_OK) throw KpError(rc, kp_last_error(ctx_));
    }

    // Global primitive id -> (mesh, local id), then the fields post_hit fills
    // (TriangleMesh::post_hit, src/core/shape.cpp:615-683, flat normals).
    void fill(const Ray& r, const kp_hit& h, Intersection& is) const {
        size_t lo = 0, hi = first_prim_.size();
        while (hi - lo > 1) {
            size_t mid = (lo + hi) / 2;
            if (first_prim_[mid] <= h.prim) lo = mid;
            else hi = mid;
        }
        const TriangleMesh& m = *meshes_[lo];
        const uint32_t local = h.prim - first_prim_[lo];
        Vec3f v[3];
        for (int k = 0; k < 3; ++k) {
            const Vec3f& p = m.verts[m.indice[local][k]];
            for (int rr = 0; rr < 3; ++rr)
                v[k][rr] = m.local_to_world.m[rr][0] * p[0] + m.local_to_world.m[rr][1] * p[1] +
                           m.local_to_world.m[rr][2] * p[2] + m.local_to_world.m[rr][3];
        }
        const float b0 = 1.0f - h.u - h.v;
        is.P = v[0] * b0 + v[1] * h.u + v[2] * h.v;
        is.N = normalize(cross(v[1] - v[0], v[2] - v[0]));
        is.shading_normal = is.N;
        is.bary = Vec3f(b0, h.u, h.v);
        is.uv = Vec2f(h.u, h.v);
        is.ray_t = h.t;
        is.backface = dot(r.direction, is.N) > 0.0f;
        is.is_light = m.is_light;
        is.light_id = m.light_id;
        is.geom_id = lo;
        is.prim_id = local;
        is.shader_name = m.shader_name;
    }

    kp_ctx* ctx_ = nullptr;
    std::vector<std::shared_ptr<TriangleMesh>> meshes_;
    std::vector<uint32_t> first_prim_;
    uint32_t total_tris_ = 0;
};

}  // namespace kazen
