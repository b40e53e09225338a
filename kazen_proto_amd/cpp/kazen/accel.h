// HipAccel: the reference's Accelerator seam over the HIP intersector.
//
// Mirrors src/core/accel.h:13-41 (Accelerator: add_trianglemesh, add_instances,
// build, intersect x2) and replaces EmbreeAccel (accel.h:64-88,
// accel.cpp:185-432).  The per-ray const intersect overloads, which the
// reference calls concurrently from its TBB tile workers
// (src/core/renderer.cpp:24-26), traverse the host copy of the same BVH on
// the calling thread (kp_intersect_closest_host: thread-safe, hits
// bit-identical to the GPU's); the batched overloads are the fast path (one
// kernel launch for the whole batch) and must not be called concurrently on
// one accelerator.
#pragma once
#include <functional>
#include <map>
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
    uint32_t material_id = 0;   // closure-table entry in the context (set by SceneGPU::add_mesh)
};

inline Mat4f operator*(const Mat4f& a, const Mat4f& b) {
    Mat4f c;
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            float acc = 0.0f;
            for (int j = 0; j < 4; ++j) acc += a.m[r][j] * b.m[j][k];
            c.m[r][k] = acc;
        }
    return c;
}

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

    // Meshes are kept until build(), which decides how each is placed.
    void add_trianglemesh(std::shared_ptr<TriangleMesh>& t) override { meshes_.push_back(t); }

    // Names a group of meshes and groups placed with `trans` (EmbreeAccel::
    // add_instances, src/core/accel.cpp:317-346: a subscene wrapped in an
    // instance); with root = true the names are the scene root, as build().
    void add_instances(const std::string& name, const std::vector<std::string>& names,
                       const Mat4f& trans = Mat4f(), bool root = false) override {
        if (root) {
            build(names);
            return;
        }
        groups_[name] = Group{names, trans};
    }

    // Places the named meshes and groups (all meshes when `names` is empty)
    // and builds the BVH. A mesh placed once goes into the one-level soup
    // under its full transform; a mesh placed several times becomes one
    // prototype with an instance per placement (two-level traversal).
    void build(const std::vector<std::string>& names) override {
        std::vector<std::pair<uint32_t, Mat4f>> placed;
        if (names.empty()) {
            for (uint32_t i = 0; i < meshes_.size(); ++i) placed.push_back({i, meshes_[i]->local_to_world});
        } else {
            std::map<std::string, uint32_t> by_name;
            for (uint32_t i = 0; i < meshes_.size(); ++i) by_name[meshes_[i]->name] = i;
            std::function<void(const std::string&, const Mat4f&, int)> expand = [&](const std::string& n,
                                                                                  const Mat4f& parent, int depth) {
                if (depth > 64) throw KpError(KP_ERR_INVALID, "instance groups nest deeper than 64 (a cycle?)");
                auto m = by_name.find(n);
                if (m != by_name.end()) {
                    placed.push_back({m->second, parent * meshes_[m->second]->local_to_world});
                    return;
                }
                auto g = groups_.find(n);
                if (g == groups_.end()) throw KpError(KP_ERR_INVALID, "build: unknown mesh or group " + n);
                for (const auto& child : g->second.names) expand(child, parent * g->second.trans, depth + 1);
            };
            for (const auto& n : names) expand(n, Mat4f(), 0);
        }
        std::vector<uint32_t> uses(meshes_.size(), 0);
        for (const auto& p : placed) uses[p.first]++;
        check(kp_geometry_clear(ctx_));
        places_.clear();
        uint32_t next = 0;
        for (const auto& p : placed) {
            if (uses[p.first] != 1) continue;
            const TriangleMesh& m = *meshes_[p.first];
            std::vector<float> xyz;
            std::vector<uint32_t> idx;
            flatten(m, xyz, idx);
            float xf[12];
            rows(p.second, xf);
            uint32_t id = 0;
            check(kp_mesh_add(ctx_, xyz.data(), (uint32_t)m.verts.size(), idx.data(), (uint32_t)m.indice.size(), xf,
                              &id));
            check(kp_mesh_set_material(ctx_, id, m.material_id));
            std::vector<float> n9;
            if (normals9(m, n9)) check(kp_mesh_set_normals(ctx_, id, n9.data()));
            places_.push_back({next, p.first, p.second});
            next += (uint32_t)m.indice.size();
        }
        std::vector<uint32_t> proto(meshes_.size(), 0xFFFFFFFFu);
        for (const auto& p : placed) {
            if (uses[p.first] < 2) continue;
            const TriangleMesh& m = *meshes_[p.first];
            if (proto[p.first] == 0xFFFFFFFFu) {
                std::vector<float> xyz;
                std::vector<uint32_t> idx;
                flatten(m, xyz, idx);
                check(kp_blas_add(ctx_, xyz.data(), (uint32_t)m.verts.size(), idx.data(), (uint32_t)m.indice.size(),
                                  &proto[p.first]));
                check(kp_blas_set_material(ctx_, proto[p.first], m.material_id));
                std::vector<float> n9;
                if (normals9(m, n9)) check(kp_blas_set_normals(ctx_, proto[p.first], n9.data()));
            }
            float xf[12];
            rows(p.second, xf);
            check(kp_instance_add(ctx_, proto[p.first], xf, nullptr));
            places_.push_back({next, p.first, p.second});
            next += (uint32_t)m.indice.size();
        }
        if (places_.empty()) throw KpError(KP_ERR_INVALID, "build: nothing to place");
        check(kp_bvh_build(ctx_, nullptr));
    }

    // Per-ray closest hit on the calling thread (safe from many threads).
    bool intersect(const Ray& r, Intersection& isect) const override {
        const kp_ray kr = to_kp(r);
        kp_hit h;
        check_host(kp_intersect_closest_host(ctx_, &kr, 1, &h));
        if (h.prim == 0xFFFFFFFFu) return false;
        fill(r, h, isect);
        return true;
    }

    bool intersect(const Ray& r, float& t) const override {
        const kp_ray kr = to_kp(r);
        kp_hit h;
        check_host(kp_intersect_closest_host(ctx_, &kr, 1, &h));
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
    // host queries report through the calling thread's error text
    static void check_host(int rc) {
        if (rc != KP_OK) throw KpError(rc, kp_last_error(nullptr));
    }

    static void flatten(const TriangleMesh& m, std::vector<float>& xyz, std::vector<uint32_t>& idx) {
        xyz.resize(3 * m.verts.size());
        for (size_t i = 0; i < m.verts.size(); ++i)
            for (int k = 0; k < 3; ++k) xyz[3 * i + k] = m.verts[i][k];
        idx.resize(3 * m.indice.size());
        for (size_t i = 0; i < m.indice.size(); ++i)
            for (int k = 0; k < 3; ++k) idx[3 * i + k] = (uint32_t)m.indice[i][k];
    }
    // Per-triangle vertex normals (9 floats each) when the mesh has one
    // normal per vertex (TriangleMesh::norms, src/core/shape.h:188-236).
    static bool normals9(const TriangleMesh& m, std::vector<float>& n9) {
        if (m.norms.size() != m.verts.size() || m.norms.empty()) return false;
        n9.resize(9 * m.indice.size());
        for (size_t i = 0; i < m.indice.size(); ++i)
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) n9[9 * i + 3 * k + a] = m.norms[m.indice[i][k]][a];
        return true;
    }
    static void rows(const Mat4f& x, float* xf) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) xf[4 * r + c] = x.m[r][c];
    }

    // Global primitive id -> (placement, mesh, local id), then the fields
    // post_hit fills (TriangleMesh::post_hit, src/core/shape.cpp:615-683) in
    // world space: interpolated shading normal and the shading point moved
    // off the flat triangle (Hanika 2021) when the mesh has vertex normals,
    // tangent / bitangent, uv and dpdu / dpdv from the mesh uvs.
    void fill(const Ray& r, const kp_hit& h, Intersection& is) const {
        size_t lo = 0, hi = places_.size();
        while (hi - lo > 1) {
            size_t mid = (lo + hi) / 2;
            if (places_[mid].first_prim <= h.prim) lo = mid;
            else hi = mid;
        }
        const Place& pl = places_[lo];
        const TriangleMesh& m = *meshes_[pl.mesh];
        const uint32_t local = h.prim - pl.first_prim;
        Vec3f v[3];
        for (int k = 0; k < 3; ++k) {
            const Vec3f& p = m.verts[m.indice[local][k]];
            for (int rr = 0; rr < 3; ++rr)
                v[k][rr] = pl.xf.m[rr][0] * p[0] + pl.xf.m[rr][1] * p[1] + pl.xf.m[rr][2] * p[2] + pl.xf.m[rr][3];
        }
        const float b0 = 1.0f - h.u - h.v;
        is.P = v[0] * b0 + v[1] * h.u + v[2] * h.v;
        is.N = normalize(cross(v[1] - v[0], v[2] - v[0]));
        is.shading_normal = is.N;
        if (m.norms.size() == m.verts.size() && !m.norms.empty()) {
            // normals transform with the inverse transpose of the placement
            Vec3f n[3];
            for (int k = 0; k < 3; ++k) n[k] = normalize(inverse_transpose(pl.xf, m.norms[m.indice[local][k]]));
            is.shading_normal = normalize(n[0] * b0 + n[1] * h.u + n[2] * h.v);
            Vec3f tu = is.P - v[0], tw = is.P - v[1], tx = is.P - v[2];
            tu = tu - n[0] * std::fmin(0.0f, dot(tu, n[0]));
            tw = tw - n[1] * std::fmin(0.0f, dot(tw, n[1]));
            tx = tx - n[2] * std::fmin(0.0f, dot(tx, n[2]));
            is.P = is.P + (tu * b0 + tw * h.u + tx * h.v);
        }
        is.tangent = normalize(v[1] - v[0]);
        is.bitangent = cross(is.tangent, is.N);
        is.bary = Vec3f(b0, h.u, h.v);
        is.uv = Vec2f(h.u, h.v);
        is.dpdu = is.tangent;
        is.dpdv = is.bitangent;
        if (m.uvs.size() == m.verts.size() && !m.uvs.empty()) {
            const Vec2f& t0 = m.uvs[m.indice[local][0]];
            const Vec2f& t1 = m.uvs[m.indice[local][1]];
            const Vec2f& t2 = m.uvs[m.indice[local][2]];
            is.uv = Vec2f(b0 * t0[0] + h.u * t1[0] + h.v * t2[0], b0 * t0[1] + h.u * t1[1] + h.v * t2[1]);
            const float d13u = t0[0] - t2[0], d13v = t0[1] - t2[1], d23u = t1[0] - t2[0], d23v = t1[1] - t2[1];
            const float det = d13u * d23v - d13v * d23u;
            if (std::fabs(det) >= 1e-8f) {
                const Vec3f dp13 = v[0] - v[2], dp23 = v[1] - v[2];
                const float inv = 1.0f / det;
                const Vec3f du = (dp13 * d23v - dp23 * d13v) * inv, dv = (dp23 * d13u - dp13 * d23u) * inv;
                if (length_squared(cross(du, dv)) > 0.0f) {
                    is.dpdu = normalize(du);
                    is.dpdv = normalize(dv);
                }
            }
        }
        is.ray_t = h.t;
        is.backface = dot(r.direction, is.N) > 0.0f;
        is.is_light = m.is_light;
        is.light_id = m.light_id;
        is.geom_id = pl.mesh;
        is.prim_id = local;
        is.shader_name = m.shader_name;
    }

    // (M^-1)^T n for the upper 3x3 of a row-major object-to-world matrix
    static Vec3f inverse_transpose(const Mat4f& x, const Vec3f& n) {
        const float(*a)[4] = x.m;
        const float c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1], c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2],
                    c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
        const float c10 = a[0][2] * a[2][1] - a[0][1] * a[2][2], c11 = a[0][0] * a[2][2] - a[0][2] * a[2][0],
                    c12 = a[0][1] * a[2][0] - a[0][0] * a[2][1];
        const float c20 = a[0][1] * a[1][2] - a[0][2] * a[1][1], c21 = a[0][2] * a[1][0] - a[0][0] * a[1][2],
                    c22 = a[0][0] * a[1][1] - a[0][1] * a[1][0];
        // the cofactor matrix is det * (M^-1)^T: the scale drops out in
        // normalize, the sign of det does not
        const float sg = (a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02) < 0.0f ? -1.0f : 1.0f;
        return Vec3f(c00 * n[0] + c01 * n[1] + c02 * n[2], c10 * n[0] + c11 * n[1] + c12 * n[2],
                     c20 * n[0] + c21 * n[1] + c22 * n[2]) * sg;
    }

    struct Group {
        std::vector<std::string> names;
        Mat4f trans;
    };
    struct Place {
        uint32_t first_prim;   // first global primitive id of this placement
        uint32_t mesh;
        Mat4f xf;              // object to world
    };

    kp_ctx* ctx_ = nullptr;
    std::vector<std::shared_ptr<TriangleMesh>> meshes_;
    std::map<std::string, Group> groups_;
    std::vector<Place> places_;   // in primitive-id order
};

}  // namespace kazen
