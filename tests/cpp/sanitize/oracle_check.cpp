// Runs the oracle's threaded entry points (BVH build and traversal, brute
// force, instanced closest hit, path integrator) under a sanitizer build:
// every threaded result must equal the single-threaded one.  Built by
// tests/cpp/sanitize/Makefile; run by tests/test_sanitizers.py.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "oracle.h"

static uint32_t g_s = 2024u;
static float rnd() {
    g_s = g_s * 1664525u + 1013904223u;
    return (float)(g_s >> 8) / 16777216.0f;
}

int main() {
    int bad = 0;
    // random soup and rays
    const uint32_t nt = 60000, nr = 6000;
    std::vector<float> tris(9 * (size_t)nt), rays(8 * (size_t)nr);
    for (uint32_t k = 0; k < nt; ++k) {
        const float c[3] = {rnd() * 2 - 1, rnd() * 2 - 1, rnd() * 2 - 1};
        for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a) tris[9 * (size_t)k + 3 * v + a] = c[a] + (rnd() - 0.5f) * 0.05f;
    }
    for (uint32_t i = 0; i < nr; ++i) {
        float* r = &rays[8 * (size_t)i];
        for (int a = 0; a < 3; ++a) r[a] = rnd() * 2.4f - 1.2f;
        r[3] = 0.0f;
        for (int a = 0; a < 3; ++a) r[4 + a] = rnd() * 2 - 1;
        r[7] = INFINITY;
    }
    or_bvh* b = or_bvh_build(tris.data(), nt);
    std::vector<float> h1(4 * (size_t)nr), h8(4 * (size_t)nr), hb(4 * (size_t)nr);
    or_bvh_closest(b, rays.data(), nr, h1.data(), 1);
    or_bvh_closest(b, rays.data(), nr, h8.data(), 8);
    bad += std::memcmp(h1.data(), h8.data(), h1.size() * 4) != 0;
    std::vector<uint8_t> o1(nr), o8(nr);
    or_bvh_occluded(b, rays.data(), nr, o1.data(), 1);
    or_bvh_occluded(b, rays.data(), nr, o8.data(), 8);
    bad += std::memcmp(o1.data(), o8.data(), nr) != 0;
    or_closest_bruteforce(tris.data(), 4000, rays.data(), 2000, hb.data(), 8);
    std::vector<float> hb1(4 * 2000);
    or_closest_bruteforce(tris.data(), 4000, rays.data(), 2000, hb1.data(), 1);
    bad += std::memcmp(hb.data(), hb1.data(), hb1.size() * 4) != 0;
    bad += or_count_near_ties(tris.data(), 4000, rays.data(), 2000, 8) !=
           or_count_near_ties(tris.data(), 4000, rays.data(), 2000, 1);
    // instanced: two placements of the first 3000 triangles
    const float w2o[24] = {1, 0, 0, 0.5f, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 1, 0, 0, -1, 0, 0, 0.25f};
    const uint32_t first[1] = {0}, count[1] = {3000}, imesh[2] = {0, 0}, ibase[2] = {0, 3000};
    std::vector<float> hi1(4 * (size_t)nr), hi8(4 * (size_t)nr);
    or_instanced_closest(tris.data(), first, count, w2o, imesh, ibase, 2, rays.data(), nr, hi1.data(), 1);
    or_instanced_closest(tris.data(), first, count, w2o, imesh, ibase, 2, rays.data(), nr, hi8.data(), 8);
    bad += std::memcmp(hi1.data(), hi8.data(), hi1.size() * 4) != 0;
    or_bvh_free(b);
    // path integrator: a floor and a ceiling light
    const float scene[36] = {-2, 0, -2, -2, 0, 2, 2, 0, 2, -2, 0, -2, 2, 0, 2, 2, 0, -2,
                             -0.5f, 2, -0.5f, 0.5f, 2, -0.5f, 0.5f, 2, 0.5f, -0.5f, 2, -0.5f, 0.5f, 2, 0.5f,
                             -0.5f, 2, 0.5f};
    const int32_t tm[4] = {0, 0, 1, 1}, tl[4] = {-1, -1, 0, 0};
    or_material mats[2];
    std::memset(mats, 0, sizeof(mats));
    mats[0].albedo[0] = mats[0].albedo[1] = mats[0].albedo[2] = 0.5f;
    mats[1].Le[0] = mats[1].Le[1] = mats[1].Le[2] = 5.0f;
    or_scene* s = or_scene_create(scene, 4, tm, tl, mats, 2);
    or_render_desc d;
    std::memset(&d, 0, sizeof(d));
    d.width = d.height = 24;
    d.spp = 4;
    d.max_depth = 4;
    d.rr_depth = 3;
    d.seed = 7;
    d.eye[1] = 1.0f;
    d.eye[2] = 3.0f;
    d.up[1] = 1.0f;
    d.tan_half_fov = 0.5f;
    std::vector<double> img1(3 * 24 * 24), img8(3 * 24 * 24);
    or_render(s, &d, img1.data(), nullptr, 1);
    or_render(s, &d, img8.data(), nullptr, 8);
    bad += std::memcmp(img1.data(), img8.data(), img1.size() * 8) != 0;
    or_scene_free(s);
    std::printf("oracle_sanitize mismatches %d\n", bad);
    return bad == 0 ? 0 : 1;
}
