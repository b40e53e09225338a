// Drives the C++ adapters the way kazen_proto callers use its interfaces
// (Scene::add_mesh + build_bvh, Accelerator::intersect/occluded,
// Integrator::Li, Renderer::render) and prints results for
// tests/test_adapter_gpu.py to compare with the oracle.
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "kazen/render.h"

using namespace kazen;

// A scene for the concurrency check: a 32x16 UV sphere placed once, a tile
// of small random triangles placed twice (two instances), a ground quad.
static void threaded_scene(SceneGPU& sc) {
    sc.add_material("diffuse", kp_material{});
    std::vector<Vec3f> v;
    std::vector<Vec3i> f;
    const int nu = 32, nv = 16;
    for (int j = 0; j <= nv; ++j)
        for (int i = 0; i <= nu; ++i) {
            const float th = 3.14159265f * (float)j / nv, ph = 6.2831853f * (float)i / nu;
            v.emplace_back(std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph));
        }
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < nu; ++i) {
            const int a = j * (nu + 1) + i, b = a + 1, c = a + nu + 1, d = c + 1;
            f.emplace_back(a, c, b);
            f.emplace_back(b, c, d);
        }
    sc.add_mesh(Mat4f(), v, {}, {}, f, "sphere", "diffuse");
    std::vector<Vec3f> tv;
    std::vector<Vec3i> tf;
    uint32_t s = 12345u;
    auto rnd = [&]() { s = s * 1664525u + 1013904223u; return (float)(s >> 8) / 16777216.0f; };
    for (int k = 0; k < 600; ++k) {
        const Vec3f c(rnd() * 2 - 1, rnd() * 2 - 1, rnd() * 2 - 1);
        for (int q = 0; q < 3; ++q) tv.push_back(c + Vec3f(rnd() - 0.5f, rnd() - 0.5f, rnd() - 0.5f) * 0.3f);
        tf.emplace_back(3 * k, 3 * k + 1, 3 * k + 2);
    }
    sc.add_mesh(Mat4f(), tv, {}, {}, tf, "tile", "diffuse");
    sc.add_mesh(Mat4f(), {Vec3f(-6, -2, -6), Vec3f(6, -2, -6), Vec3f(6, -2, 6), Vec3f(-6, -2, 6)}, {}, {},
                {Vec3i(0, 1, 2), Vec3i(0, 2, 3)}, "ground", "diffuse");
    Mat4f a, b;
    a.m[0][3] = 2.5f;
    b.m[0][0] = 0.0f; b.m[0][2] = 1.0f; b.m[2][0] = -1.0f; b.m[2][2] = 0.0f;
    b.m[2][3] = -2.5f;
    sc.accelerator->add_instances("tile_a", {"tile"}, a);
    sc.accelerator->add_instances("tile_b", {"tile"}, b);
    sc.build_bvh({"sphere", "ground", "tile_a", "tile_b"});
}

static std::vector<Ray> threaded_rays(size_t n) {
    std::vector<Ray> rays;
    uint32_t s = 777u;
    auto rnd = [&]() { s = s * 1664525u + 1013904223u; return (float)(s >> 8) / 16777216.0f; };
    for (size_t i = 0; i < n; ++i) {
        const Vec3f o(rnd() * 8 - 4, rnd() * 8 - 4, rnd() * 8 - 4);
        const Vec3f t(rnd() * 5 - 2.5f, rnd() * 3 - 1.5f, rnd() * 5 - 2.5f);
        rays.emplace_back(o, t - o, 0.0f);
    }
    return rays;
}

static bool same_hit(bool ha, const Intersection& a, bool hb, const Intersection& b) {
    if (ha != hb) return false;
    if (!ha) return true;
    return a.geom_id == b.geom_id && a.prim_id == b.prim_id && std::memcmp(&a.ray_t, &b.ray_t, 4) == 0 &&
           std::memcmp(&a.uv, &b.uv, sizeof(a.uv)) == 0 && std::memcmp(&a.P, &b.P, sizeof(a.P)) == 0;
}

// 16 threads call the per-ray const intersect at once (as the reference's
// TBB tile workers do, renderer.cpp:24-26), each over every ray in its own
// order; returns the number of results that differ from `want`.
static size_t threaded_mismatches(const HipAccel& acc, const std::vector<Ray>& rays,
                                  const std::vector<uint8_t>& want_hit, const std::vector<Intersection>& want) {
    const int T = 16;
    std::vector<size_t> bad(T, 0);
    std::vector<std::thread> pool;
    for (int k = 0; k < T; ++k)
        pool.emplace_back([&, k]() {
            const size_t n = rays.size();
            for (size_t j = 0; j < n; ++j) {
                const size_t i = (j * 7919u + (size_t)k * 104729u) % n;
                Intersection is;
                const bool h = acc.intersect(rays[i], is);
                float t = 0.0f;
                const bool ht = acc.intersect(rays[i], t);
                if (!same_hit(h, is, want_hit[i] != 0, want[i]) || ht != h || (h && std::memcmp(&t, &is.ray_t, 4)))
                    ++bad[k];
            }
        });
    for (auto& t : pool) t.join();
    size_t total = 0;
    for (size_t b : bad) total += b;
    return total;
}

// Host-only mode (no GPU): the concurrency check against a single-threaded
// pass over the same rays; this is what the sanitizer builds run.
static int host_mode() {
    SceneGPU sc(KP_DEVICE_HOST);
    threaded_scene(sc);
    const std::vector<Ray> rays = threaded_rays(20000);
    std::vector<Intersection> want(rays.size());
    std::vector<uint8_t> hit(rays.size());
    size_t n_hit = 0;
    for (size_t i = 0; i < rays.size(); ++i) {
        hit[i] = sc.accelerator->intersect(rays[i], want[i]) ? 1 : 0;
        n_hit += hit[i];
    }
    const size_t bad = threaded_mismatches(*sc.accelerator, rays, hit, want);
    std::printf("threads_host %zu %zu %zu\n", rays.size(), n_hit, bad);
    // a 300k-triangle soup: the BVH builder builds large subtrees on threads
    SceneGPU big(KP_DEVICE_HOST);
    big.add_material("diffuse", kp_material{});
    std::vector<Vec3f> bv;
    std::vector<Vec3i> bf;
    uint32_t s = 99u;
    auto rnd = [&]() { s = s * 1664525u + 1013904223u; return (float)(s >> 8) / 16777216.0f; };
    for (int k = 0; k < 300000; ++k) {
        const Vec3f c(rnd() * 2 - 1, rnd() * 2 - 1, rnd() * 2 - 1);
        for (int q = 0; q < 3; ++q) bv.push_back(c + Vec3f(rnd() - 0.5f, rnd() - 0.5f, rnd() - 0.5f) * 0.02f);
        bf.emplace_back(3 * k, 3 * k + 1, 3 * k + 2);
    }
    big.add_mesh(Mat4f(), bv, {}, {}, bf, "soup", "diffuse");
    big.build_bvh();
    std::vector<Ray> brays = threaded_rays(4000);
    std::vector<Intersection> bwant(brays.size());
    std::vector<uint8_t> bhit(brays.size());
    for (size_t i = 0; i < brays.size(); ++i) bhit[i] = big.accelerator->intersect(brays[i], bwant[i]) ? 1 : 0;
    const size_t bbad = threaded_mismatches(*big.accelerator, brays, bhit, bwant);
    std::printf("threads_host_big %zu %zu\n", brays.size(), bbad);
    return bad == 0 && bbad == 0 ? 0 : 1;
}

static kp_material lambert(float a, float le = 0.0f) {
    kp_material m{};
    m.type = KP_MAT_LAMBERT;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = a;
    m.Le[0] = m.Le[1] = m.Le[2] = le;
    return m;
}

int main(int argc, char** argv) {
    try {
        if (argc > 1 && std::strcmp(argv[1], "host") == 0) return host_mode();
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

        // 5. Concurrent per-ray intersect (host traversal) against the batched
        //    GPU intersect of the same accelerator, bit for bit.
        SceneGPU th;
        threaded_scene(th);
        const std::vector<Ray> trays = threaded_rays(20000);
        std::vector<Intersection> gpu;
        std::vector<uint8_t> ghit;
        th.accelerator->intersect(trays, gpu, ghit);
        size_t n_hit = 0;
        for (uint8_t h : ghit) n_hit += h;
        std::printf("threads %zu %zu %zu\n", trays.size(), n_hit, threaded_mismatches(*th.accelerator, trays, ghit, gpu));
        return 0;
    } catch (const KpError& e) {
        std::printf("error %d %s\n", e.code, e.what());
        return 1;
    }
}
