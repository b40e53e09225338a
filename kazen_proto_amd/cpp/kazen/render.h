// SceneGPU / Integrator / Renderer adapters over the wavefront integrator.
//
// Mirrors of kazen_proto:
//   SceneGPU            src/core/scene.h:237-318 (set_film, set_camera,
//                       add_mesh, build_bvh); OSL shader groups are replaced by
//                       named closure tables (add_material)
//   Integrator          src/core/integrator.h:22-41 (setup, Li)
//   PathIntegrator      src/core/integrator.h:121-132 -> WavefrontPathIntegrator
//   IntegratorFactory   src/core/integrator.h:134-141
//   RenderCallback      src/core/renderer.h:10-15
//   Renderer            src/core/renderer.h:17-37; render(SceneGPU&, ...) is
//                       declared by the reference (renderer.h:30-31) and never
//                       defined there; it is defined here.
#pragma once
#include <cmath>
#include <cstdio>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "kazen/accel.h"

namespace kazen {

struct Film {
    uint32_t width = 0, height = 0;
    std::string filename;
    std::vector<float> rgb;   // mean radiance, row-major, top row first
};

struct RecordContext {
    uint32_t pixel_x = 0, pixel_y = 0;
};

class SceneGPU {
public:
    explicit SceneGPU(int device = 0) : accelerator(std::make_unique<HipAccel>(device)) {}

    void set_film(uint32_t w, uint32_t h, const std::string& out) {
        film.width = w;
        film.height = h;
        film.filename = out;
    }
    // fov in degrees (vertical); ratio is taken from the film.
    void set_camera(const Vec3f& p, const Vec3f& l, const Vec3f& u, const float ratio = 1.0f,
                    const float near_plane = 1, const float far_plane = 1000, const float fov = 60.f) {
        (void)ratio;
        (void)near_plane;
        (void)far_plane;
        eye = p;
        lookat = l;
        up = u;
        fov_deg = fov;
    }
    // Named closure table standing in for an OSL shader group.
    void add_material(const std::string& shader_name, const kp_material& m) {
        uint32_t id = 0;
        check(kp_material_add(ctx(), &m, &id));
        materials[shader_name] = id;
    }
    void add_mesh(const Mat4f& world, const std::vector<Vec3f>& vs, const std::vector<Vec3f>& ns,
                  const std::vector<Vec2f>& ts, const std::vector<Vec3i>& idx, const std::string& name,
                  const std::string& shader_name, bool is_light = false) {
        auto m = std::make_shared<TriangleMesh>();
        m->local_to_world = world;
        m->verts = vs;
        m->norms = ns;
        m->uvs = ts;
        m->indice = idx;
        m->name = name;
        m->shader_name = shader_name;
        m->is_light = is_light;
        auto it = materials.find(shader_name);
        if (it == materials.end()) throw KpError(KP_ERR_INVALID, "unknown shader/material: " + shader_name);
        m->material_id = it->second;
        accelerator->add_trianglemesh(m);
    }
    void build_bvh(const std::vector<std::string>& names = {}) { accelerator->build(names); }

    kp_ctx* ctx() const { return accelerator->context(); }

    std::unique_ptr<HipAccel> accelerator;
    std::map<std::string, uint32_t> materials;
    Film film;
    Vec3f eye{0, 0, 0}, lookat{0, 0, -1}, up{0, 1, 0};
    float fov_deg = 60.0f;

private:
    void check(int rc) const {
        if (rc != KP_OK) throw KpError(rc, kp_last_error(ctx()));
    }
};

class Integrator {
public:
    virtual ~Integrator() = default;
    virtual void setup(SceneGPU* scene) { scene_ = scene; }
    virtual RGBSpectrum Li(const Ray& r, const RecordContext* rctx) const = 0;

protected:
    SceneGPU* scene_ = nullptr;
};

// PathIntegrator::Li (integrator.cpp:362-561) on the GPU.  Li() of a single
// ray is a batch of one; Li(rays, out) is the batched form.
class WavefrontPathIntegrator : public Integrator {
public:
    uint32_t max_depth = 8;   // PathIntegrator's max_depth (integrator.cpp:404)
    uint32_t rr_depth = 3;    // min_depth for roulette (integrator.cpp:405)
    uint32_t seed = 0;
    uint32_t sample_index = 0;

    static std::unique_ptr<Integrator> create() { return std::make_unique<WavefrontPathIntegrator>(); }

    RGBSpectrum Li(const Ray& r, const RecordContext*) const override {
        std::vector<RGBSpectrum> out;
        Li(std::vector<Ray>{r}, out);
        return out[0];
    }
    void Li(const std::vector<Ray>& rays, std::vector<RGBSpectrum>& out) const {
        std::vector<kp_ray> kr(rays.size());
        for (size_t i = 0; i < rays.size(); ++i) kr[i] = HipAccel::to_kp(rays[i]);
        std::vector<float> rgb(3 * rays.size());
        kp_render_desc d{};
        d.max_depth = max_depth;
        d.rr_depth = rr_depth;
        d.seed = seed;
        d.spp_offset = sample_index;
        const int rc = kp_integrate_rays(scene_->ctx(), kr.data(), (uint32_t)kr.size(), &d, rgb.data(),
                                         KP_PTR_HOST);
        if (rc != KP_OK) throw KpError(rc, kp_last_error(scene_->ctx()));
        out.resize(rays.size());
        for (size_t i = 0; i < rays.size(); ++i) out[i] = RGBSpectrum(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    }
};

class IntegratorFactory {
public:
    std::unique_ptr<Integrator> create() { return create_functor(); }
    std::function<std::unique_ptr<Integrator>()> create_functor = WavefrontPathIntegrator::create;
};

class RenderCallback {
public:
    virtual ~RenderCallback() = default;
    virtual void on_tile_end(Film* film, uint32_t tile_id) { (void)film; (void)tile_id; }
};

class Renderer {
public:
    Renderer() = default;
    Renderer(const uint32_t spl_cnt, const uint32_t nth, RenderCallback* const cbk = nullptr)
        : sample_count(spl_cnt), nthreads(nth), render_cbk(cbk) {}

    // Whole film on the GPU; writes a PFM image when output is non-empty.
    bool render(SceneGPU& scene, const std::string& output = "./test.pfm") const {
        kp_render_desc d = desc(scene);
        scene.film.rgb.assign(3 * (size_t)d.width * d.height, 0.0f);
        if (kp_render(scene.ctx(), &d, scene.film.rgb.data(), KP_PTR_HOST) != KP_OK) return false;
        if (render_cbk) render_cbk->on_tile_end(&scene.film, 0);
        return output.empty() || write_pfm(scene.film, output);
    }
    // Debug render of one pixel (the reference's --pixel mode): returns its mean.
    bool render(SceneGPU& scene, const uint32_t x, const uint32_t y) const {
        if (!render(scene, std::string())) return false;
        const float* p = &scene.film.rgb[3 * ((size_t)y * scene.film.width + x)];
        std::printf("pixel (%u, %u): %g %g %g\n", x, y, p[0], p[1], p[2]);
        return true;
    }

    kp_render_desc desc(const SceneGPU& s) const {
        kp_render_desc d{};
        d.width = s.film.width;
        d.height = s.film.height;
        d.spp = sample_count;
        d.max_depth = max_depth;
        d.rr_depth = rr_depth;
        d.seed = seed;
        for (int k = 0; k < 3; ++k) {
            d.eye[k] = s.eye[k];
            d.lookat[k] = s.lookat[k];
            d.up[k] = s.up[k];
        }
        d.tan_half_fov = std::tan(s.fov_deg * 0.5f * 3.14159265358979f / 180.0f);
        return d;
    }

    static bool write_pfm(const Film& f, const std::string& path) {
        FILE* fp = std::fopen(path.c_str(), "wb");
        if (!fp) return false;
        std::fprintf(fp, "PF\n%u %u\n-1.0\n", f.width, f.height);
        // PFM stores rows bottom to top
        for (uint32_t y = f.height; y-- > 0;)
            std::fwrite(&f.rgb[3 * (size_t)y * f.width], sizeof(float), 3 * (size_t)f.width, fp);
        std::fclose(fp);
        return true;
    }

    uint32_t sample_count = 10;
    uint32_t nthreads = 0;   // unused on the GPU path
    RenderCallback* render_cbk = nullptr;
    uint32_t max_depth = 8, rr_depth = 3, seed = 0;
};

}  // namespace kazen
