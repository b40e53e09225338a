// Drives the C++ adapters the way kazen_proto callers use its interfaces
// (Scene::add_mesh + build_bvh, Accelerator::intersect/occluded,
// Integrator::Li, Renderer::render) and prints results for
// tests/test_adapter_gpu.py to compare with the oracle.
#include <cstdio>
#include <vector>

#include "kazen/render.h"

using namespace kazen;

static kp_material lambert(float a, float le = 0.0f) {
    kp_material m{};
    m.type = KP_MAT_LAMBERT;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = a;
    m.Le[0] = m.Le[1] = m.Le[2] = le;
    return m;
}

int main() {
    try {
        // 1. tests/libusage_tests/embree_tests.cpp: one triangle, ray along -y.
        SceneGPU kat;
        kat.add_material("diffuse", lambert(0.5f));
        kat.add_mesh(Mat4f(), {Vec3f(-1, 0, 1), Vec3f(2, 0, 1), Vec3f(-1, 0, -2)}, {}, {}, {Vec3i(0, 1, 2)},
                     "tri", "diffuse");
        kat.build_bvh();
        Intersection is;
        const bool hit = kat.accelerator->intersect(Ray(Vec3f(0, 1, 0), Vec3f(0, -1, 0), 0.0f), is);
        std::printf("kat_hit %d %zu %zu %.9g %.9g %.9g\n", hit ? 1 : 0, is.geom_id, is.prim_id, is.ray_t,
                    is.uv[0], is.uv[1]);
        float t = 0;
        const bool hit_t = kat.accelerator->intersect(Ray(Vec3f(0, 1, 0), Vec3f(0, -1, 0), 0.0f), t);
        std::printf("kat_t %d %.9g\n", hit_t ? 1 : 0, t);
        std::printf("kat_occluded %d %d\n", kat.accelerator->occluded(Vec3f(0, 1, 0), Vec3f(0, -1, 0)) ? 1 : 0,
                    kat.accelerator->occluded(Vec3f(0, 1, 0), Vec3f(0, 0.5f, 0)) ? 1 : 0);

        // 2. floor + downward-facing area light, Li of fixed rays.
        SceneGPU sc;
        sc.add_material("floor", lambert(0.5f));
        sc.add_material("light", lambert(0.0f, 5.0f));
        sc.add_mesh(Mat4f(), {Vec3f(-2, 0, -2), Vec3f(-2, 0, 2), Vec3f(2, 0, 2), Vec3f(2, 0, -2)}, {}, {},
                    {Vec3i(0, 1, 2), Vec3i(0, 2, 3)}, "floor", "floor");
        sc.add_mesh(Mat4f(), {Vec3f(-0.5f, 2, -0.5f), Vec3f(0.5f, 2, -0.5f), Vec3f(0.5f, 2, 0.5f), Vec3f(-0.5f, 2, 0.5f)},
                    {}, {}, {Vec3i(0, 1, 2), Vec3i(0, 2, 3)}, "light", "light", true);
        sc.build_bvh();
        WavefrontPathIntegrator integ;
        integ.max_depth = 3;
        integ.seed = 1;
        integ.setup(&sc);
        std::vector<Ray> rays;
        for (int i = 0; i < 16; ++i) {
            const float a = 0.4f * (float)i;
            rays.emplace_back(Vec3f(0.3f * std::cos(a), 1.0f, 0.3f * std::sin(a)),
                              Vec3f(std::sin(a) * 0.5f, -1.0f, std::cos(a) * 0.4f), 0.0f);
        }
        std::vector<RGBSpectrum> li;
        integ.Li(rays, li);
        for (size_t i = 0; i < rays.size(); ++i) {
            const Ray& q = rays[i];
            std::printf("ray %zu %.9g %.9g %.9g %.9g %.9g %.9g\n", i, q.origin[0], q.origin[1], q.origin[2],
                        q.direction[0], q.direction[1], q.direction[2]);
            std::printf("li %zu %.9g %.9g %.9g\n", i, li[i][0], li[i][1], li[i][2]);
        }
        const RGBSpectrum one = integ.Li(rays[0], nullptr);
        std::printf("li_single %.9g %.9g %.9g\n", one[0], one[1], one[2]);

        // 3. Renderer::render(SceneGPU&) writes the film.
        sc.set_film(16, 16, "adapter_check.pfm");
        sc.set_camera(Vec3f(0, 1.5f, 3.0f), Vec3f(0, 0, 0), Vec3f(0, 1, 0), 1.0f, 1, 1000, 50.0f);
        struct Cb : RenderCallback {
            int calls = 0;
            void on_tile_end(Film*, uint32_t) override { ++calls; }
        } cb;
        Renderer r(4, 0, &cb);
        r.max_depth = 3;
        const char* out = std::getenv("KP_ADAPTER_PFM");
        const bool ok = r.render(sc, out ? out : "");
        double mean = 0;
        for (float v : sc.film.rgb) mean += v;
        mean /= (double)sc.film.rgb.size();
        std::printf("render %d %d %.9g\n", ok ? 1 : 0, cb.calls, mean);

        // 4. Accelerator::add_instances: a tile mesh placed directly, through
        //    a translated group and through a rotated group that nests it;
        //    the ground placed once. Vertical rays from y = 5 above each.
        SceneGPU ins;
        ins.add_material("diffuse", lambert(0.5f));
        ins.add_mesh(Mat4f(), {Vec3f(-0.25f, 0, -0.25f), Vec3f(0.25f, 0, -0.25f), Vec3f(0.25f, 0, 0.25f),
                               Vec3f(-0.25f, 0, 0.25f)},
                     {}, {}, {Vec3i(0, 1, 2), Vec3i(0, 2, 3)}, "tile", "diffuse");
        ins.add_mesh(Mat4f(), {Vec3f(-5, -1, -5), Vec3f(5, -1, -5), Vec3f(5, -1, 5), Vec3f(-5, -1, 5)}, {}, {},
                     {Vec3i(0, 1, 2), Vec3i(0, 2, 3)}, "ground", "diffuse");
        Mat4f a;
        a.m[0][3] = -1.0f;
        a.m[1][3] = 0.5f;
        Mat4f b;   // 90 degrees about y, then (0, 1, 1.5)
        b.m[0][0] = 0.0f; b.m[0][2] = 1.0f;
        b.m[2][0] = -1.0f; b.m[2][2] = 0.0f;
        b.m[1][3] = 1.0f;
        b.m[2][3] = 1.5f;
        ins.accelerator->add_instances("pair", {"tile"}, a);
        ins.accelerator->add_instances("pair2", {"tile", "pair"}, b);
        for (int pass = 0; pass < 2; ++pass) {   // a second build starts from scratch
            ins.build_bvh({"ground", "tile", "pair", "pair2"});
            const float xz[5][2] = {{0.05f, 0.03f}, {-0.95f, 0.03f}, {0.05f, 1.53f}, {0.05f, 2.53f}, {3.0f, 3.0f}};
            for (int i = 0; i < 5; ++i) {
                Intersection h;
                const bool got = ins.accelerator->intersect(Ray(Vec3f(xz[i][0], 5.0f, xz[i][1]), Vec3f(0, -1, 0), 0.0f), h);
                std::printf("inst_hit %d %d %d %zu %zu %.9g %.9g %.9g %.9g\n", pass, i, got ? 1 : 0, h.geom_id,
                            h.prim_id, h.ray_t, h.P[0], h.P[1], h.P[2]);
            }
        }
        return 0;
    } catch (const KpError& e) {
        std::printf("error %d %s\n", e.code, e.what());
        return 1;
    }
}
