// CPU oracle: PCG32, watertight ray-triangle test, brute-force closest hit and
// an oracle-private binned-SAH BVH2 with robust traversal.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never the product.
//
// Reference behaviour restated here:
//   - Ray semantics: src/core/ray.h:10-45 (normalised dir, tmin/tmax interval).
//   - Closest hit: EmbreeAccel::intersect, src/core/accel.cpp:348-407, with
//     Embree in RTC_SCENE_FLAG_ROBUST mode (accel.cpp:194) = watertight test.
//     Restated with the published watertight algorithm (Woop, Benthin, Wald,
//     "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013), two-sided, FP32
//     with the paper's FP64 re-evaluation of zero edge functions.
//   - Barycentrics u,v are the weights of vertices 1 and 2 (Embree's u,v,
//     copied to Intersection::uv at accel.cpp:393-394).
//   - Occlusion: hitable.h:15-33, restated as "any hit with t in [tmin,tmax]"
//     (the geom-id comparison at hitable.h:28 never works with Embree's
//     subscene-local ids, accel.cpp:295,314,374).
//   - Deterministic closest-hit: the smallest t wins, equal t -> lowest primitive
//     index, so the answer does not depend on the traversal order.
//   - The in-tree BVHAccel (accel.cpp:76-183) never culls (aabb.h:49) and the
//     in-tree TriangleMesh::intersect returns the first, not the closest, hit
//     (shape.cpp:533-557); neither is reproduced.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>


namespace {

// ---------------------------------------------------------------- PCG32 ----
// PCG-XSH-RR 64/32 (M. E. O'Neill, "PCG: A Family of Simple Fast
// Space-Efficient Statistically Good Algorithms for Random Number
// Generation", 2014), seeded the way W. Jakob's pcg32.h seeds it (the header
// the reference's Sampler wraps, src/core/sampler.cpp:14-20).
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    // [0,1) with 23 random mantissa bits (pcg32.h nextFloat).
    float nextf() {
        uint32_t bits = (next() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &bits, 4);
        return f - 1.0f;
    }
};

// ----------------------------------------------------- watertight test ----
struct WRay {
    float o[3], d[3];
    float tmin, tmax;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

inline WRay make_wray(const float* r) {
    WRay w;
    w.o[0] = r[0]; w.o[1] = r[1]; w.o[2] = r[2]; w.tmin = r[3];
    w.d[0] = r[4]; w.d[1] = r[5]; w.d[2] = r[6]; w.tmax = r[7];
    float ax = std::fabs(w.d[0]), ay = std::fabs(w.d[1]), az = std::fabs(w.d[2]);
    // Dominant axis; ties resolve toward the lower axis index.
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = (kz == 2) ? 0 : kz + 1;
    int ky = (kx == 2) ? 0 : kx + 1;
    if (w.d[kz] < 0.0f) std::swap(kx, ky);   // keep the winding
    w.kx = kx; w.ky = ky; w.kz = kz;
    w.Sx = w.d[kx] / w.d[kz];
    w.Sy = w.d[ky] / w.d[kz];
    w.Sz = 1.0f / w.d[kz];
    return w;
}

// Rays with a non-finite origin or direction, a zero direction or a NaN
// interval bound miss everything.
inline bool ray_valid(const float* r) {
    for (int k : {0, 1, 2, 4, 5, 6})
        if (!std::isfinite(r[k])) return false;
    if (r[4] == 0.0f && r[5] == 0.0f && r[6] == 0.0f) return false;
    return !std::isnan(r[3]) && !std::isnan(r[7]);
}

// Returns true for a hit with t in [tmin, tmax_incl].
inline bool wt_hit(const WRay& r, const float* v0, const float* v1, const float* v2,
                   float tmax_incl, float& t_out, float& u_out, float& v_out) {
    const float A[3] = {v0[0] - r.o[0], v0[1] - r.o[1], v0[2] - r.o[2]};
    const float B[3] = {v1[0] - r.o[0], v1[1] - r.o[1], v1[2] - r.o[2]};
    const float C[3] = {v2[0] - r.o[0], v2[1] - r.o[1], v2[2] - r.o[2]};
    const float Ax = A[r.kx] - r.Sx * A[r.kz];
    const float Ay = A[r.ky] - r.Sy * A[r.kz];
    const float Bx = B[r.kx] - r.Sx * B[r.kz];
    const float By = B[r.ky] - r.Sy * B[r.kz];
    const float Cx = C[r.kx] - r.Sx * C[r.kz];
    const float Cy = C[r.ky] - r.Sy * C[r.kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        // FP64 re-evaluation (products of two floats are exact in double).
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * A[r.kz];
    const float Bz = r.Sz * B[r.kz];
    const float Cz = r.Sz * C[r.kz];
    const float T = U * Az + V * Bz + W * Cz;
    const float t = T / det;
    if (!(t >= r.tmin && t <= tmax_incl)) return false;   // also rejects NaN
    t_out = t;
    u_out = V / det;
    v_out = W / det;
    return true;
}

inline void write_miss(float* h) {
    const uint32_t miss = 0xFFFFFFFFu;
    h[0] = std::numeric_limits<float>::infinity();
    std::memcpy(&h[1], &miss, 4);
    h[2] = 0.0f;
    h[3] = 0.0f;
}

inline void write_hit(float* h, float t, uint32_t prim, float u, float v) {
    h[0] = t;
    std::memcpy(&h[1], &prim, 4);
    h[2] = u;
    h[3] = v;
}

template <class F>
void parallel_chunks(uint32_t n, int threads, uint32_t chunk, F&& fn) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            uint32_t b = next.fetch_add(chunk);
            if (b >= n) break;
            uint32_t e = std::min(n, b + chunk);
            fn(b, e);
        }
    };
    if (threads == 1) { worker(); return; }
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
}

// ------------------------------------------------------------- BVH2 ------
// Box slab test made conservative the way Ize ("Robust BVH Ray Traversal",
// JCGT 2(2), 2013) prescribes: the far distance is scaled by 1 + 2*gamma(3).
constexpr float kFarScale = 1.0f + 2.0f * (3.0f * 5.9604645e-08f) / (1.0f - 3.0f * 5.9604645e-08f);

struct Node {
    float lo[3], hi[3];
    uint32_t first;   // internal: left child (right = first+1); leaf: first prim slot
    uint32_t count;   // 0 = internal
};

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    double area() const {
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

inline void safe_inv(const float* d, float* inv) {
    for (int k = 0; k < 3; ++k) {
        float dk = d[k];
        if (std::fabs(dk) < 1e-20f) dk = std::copysign(1e-20f, dk);
        inv[k] = 1.0f / dk;
    }
}

inline bool box_hit(const Node& n, const float* o, const float* inv, float tmin, float tmax,
                    float& tnear) {
    float t0x = (n.lo[0] - o[0]) * inv[0], t1x = (n.hi[0] - o[0]) * inv[0];
    float t0y = (n.lo[1] - o[1]) * inv[1], t1y = (n.hi[1] - o[1]) * inv[1];
    float t0z = (n.lo[2] - o[2]) * inv[2], t1z = (n.hi[2] - o[2]) * inv[2];
    float tn = std::max(std::max(std::min(t0x, t1x), std::min(t0y, t1y)),
                        std::max(std::min(t0z, t1z), tmin));
    float tf = std::min(std::min(std::max(t0x, t1x), std::max(t0y, t1y)),
                        std::min(std::max(t0z, t1z), tmax));
    tf = tf * kFarScale;
    tnear = tn;
    return tn <= tf;
}

}  // namespace

struct or_bvh {
    std::vector<Node> nodes;
    std::vector<float> tris;      // reordered soup, 9 floats per triangle
    std::vector<uint32_t> prim;   // original index of each reordered triangle
};

namespace {

struct BuildItem { uint32_t node, begin, end; };

void build_bvh(or_bvh& b, const float* tris, uint32_t n) {
    std::vector<Box> tb(n);
    std::vector<float> cen(3 * (size_t)n);
    float max_abs = 1.0f;
    for (uint32_t i = 0; i < n; ++i) {
        for (int v = 0; v < 3; ++v) {
            const float* p = tris + 9 * (size_t)i + 3 * v;
            tb[i].grow(p);
            for (int k = 0; k < 3; ++k) max_abs = std::max(max_abs, std::fabs(p[k]));
        }
        for (int k = 0; k < 3; ++k) cen[3 * (size_t)i + k] = 0.5f * (tb[i].lo[k] + tb[i].hi[k]);
    }
    // Node boxes are widened by 2^-16 of the largest coordinate so that the
    // box test never culls a primitive the watertight test would report
    // (its hit point may sit a few ulps outside the exact triangle).
    const float pad = std::ldexp(max_abs, -16);

    std::vector<uint32_t> idx(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    b.nodes.clear();
    b.nodes.reserve(n ? 2 * (size_t)n : 1);
    b.nodes.push_back(Node{});
    std::vector<BuildItem> stack;
    stack.push_back({0, 0, n});
    constexpr int kBins = 32;
    constexpr uint32_t kMaxLeaf = 8;
    while (!stack.empty()) {
        BuildItem it = stack.back();
        stack.pop_back();
        Box nb, cb;
        for (uint32_t i = it.begin; i < it.end; ++i) {
            nb.grow(tb[idx[i]]);
            cb.grow(&cen[3 * (size_t)idx[i]]);
        }
        Node& nd = b.nodes[it.node];
        for (int k = 0; k < 3; ++k) { nd.lo[k] = nb.lo[k] - pad; nd.hi[k] = nb.hi[k] + pad; }
        uint32_t cnt = it.end - it.begin;
        auto make_leaf = [&]() { b.nodes[it.node].first = it.begin; b.nodes[it.node].count = cnt; };
        if (cnt <= 2) { make_leaf(); continue; }
        // Binned SAH over the axis of largest centroid extent.
        int axis = 0;
        float ext[3];
        for (int k = 0; k < 3; ++k) ext[k] = cb.hi[k] - cb.lo[k];
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        uint32_t mid = it.begin;
        if (!(ext[axis] > 0.0f)) {
            if (cnt <= kMaxLeaf) { make_leaf(); continue; }
            mid = it.begin + cnt / 2;
        } else {
            Box bins[kBins];
            uint32_t bc[kBins] = {0};
            const float scale = (float)kBins / ext[axis];
            auto bin_of = [&](uint32_t t) {
                int bi = (int)((cen[3 * (size_t)t + axis] - cb.lo[axis]) * scale);
                return std::min(std::max(bi, 0), kBins - 1);
            };
            for (uint32_t i = it.begin; i < it.end; ++i) {
                int bi = bin_of(idx[i]);
                bins[bi].grow(tb[idx[i]]);
                bc[bi]++;
            }
            double rightA[kBins];
            uint32_t rightN[kBins];
            Box acc;
            uint32_t accn = 0;
            for (int i = kBins - 1; i > 0; --i) {
                acc.grow(bins[i]); accn += bc[i];
                rightA[i] = acc.area(); rightN[i] = accn;
            }
            Box lacc;
            uint32_t laccn = 0;
            double best = INFINITY;
            int best_split = -1;
            for (int i = 1; i < kBins; ++i) {
                lacc.grow(bins[i - 1]); laccn += bc[i - 1];
                if (laccn == 0 || rightN[i] == 0) continue;
                double c = lacc.area() * laccn + rightA[i] * rightN[i];
                if (c < best) { best = c; best_split = i; }
            }
            const double pa = nb.area();
            const double leaf_cost = (double)cnt;
            const double split_cost = 0.5 + (pa > 0 ? best / pa : (double)cnt);
            if (best_split < 0 || (cnt <= kMaxLeaf && leaf_cost <= split_cost)) {
                if (cnt <= kMaxLeaf) { make_leaf(); continue; }
                mid = it.begin + cnt / 2;
                std::nth_element(idx.begin() + it.begin, idx.begin() + mid, idx.begin() + it.end,
                                 [&](uint32_t a, uint32_t c) {
                                     return cen[3 * (size_t)a + axis] < cen[3 * (size_t)c + axis];
                                 });
            } else {
                auto pm = std::partition(idx.begin() + it.begin, idx.begin() + it.end,
                                         [&](uint32_t t) { return bin_of(t) < best_split; });
                mid = (uint32_t)(pm - idx.begin());
                if (mid == it.begin || mid == it.end) mid = it.begin + cnt / 2;
            }
        }
        uint32_t left = (uint32_t)b.nodes.size();
        b.nodes.push_back(Node{});
        b.nodes.push_back(Node{});
        b.nodes[it.node].first = left;
        b.nodes[it.node].count = 0;
        stack.push_back({left + 1, mid, it.end});
        stack.push_back({left, it.begin, mid});
    }
    b.tris.resize(9 * (size_t)n);
    b.prim = idx;
    for (uint32_t i = 0; i < n; ++i)
        std::memcpy(&b.tris[9 * (size_t)i], tris + 9 * (size_t)idx[i], 9 * sizeof(float));
}

// Closest hit through the oracle BVH; returns nodes visited through *visits.
inline void bvh_closest_one(const or_bvh& b, const float* ray, float* hit) {
    if (!ray_valid(ray)) { write_miss(hit); return; }
    WRay r = make_wray(ray);
    float inv[3];
    safe_inv(r.d, inv);
    float best_t = r.tmax;
    uint32_t best_p = 0xFFFFFFFFu;
    float bu = 0, bv = 0;
    uint32_t stack[256];
    int sp = 0;
    float tn0;
    if (b.nodes.empty() || !box_hit(b.nodes[0], r.o, inv, r.tmin, best_t, tn0)) {
        write_miss(hit);
        return;
    }
    uint32_t cur = 0;
    for (;;) {
        const Node& n = b.nodes[cur];
        if (n.count) {
            for (uint32_t i = 0; i < n.count; ++i) {
                const float* tv = &b.tris[9 * (size_t)(n.first + i)];
                float t, u, v;
                if (wt_hit(r, tv, tv + 3, tv + 6, best_t, t, u, v)) {
                    uint32_t p = b.prim[n.first + i];
                    if (t < best_t || p < best_p) { best_t = t; best_p = p; bu = u; bv = v; }
                }
            }
        } else {
            float ta, tb2;
            bool ha = box_hit(b.nodes[n.first], r.o, inv, r.tmin, best_t, ta);
            bool hb = box_hit(b.nodes[n.first + 1], r.o, inv, r.tmin, best_t, tb2);
            if (ha && hb) {
                uint32_t nearn = n.first, farn = n.first + 1;
                if (tb2 < ta) std::swap(nearn, farn);
                stack[sp++] = farn;
                cur = nearn;
                continue;
            }
            if (ha) { cur = n.first; continue; }
            if (hb) { cur = n.first + 1; continue; }
        }
        // pop; re-test the box against the shrunk interval
        bool found = false;
        while (sp > 0) {
            uint32_t c = stack[--sp];
            float tt;
            if (box_hit(b.nodes[c], r.o, inv, r.tmin, best_t, tt)) { cur = c; found = true; break; }
        }
        if (!found) break;
    }
    if (best_p == 0xFFFFFFFFu) write_miss(hit);
    else write_hit(hit, best_t, best_p, bu, bv);
}

// A second primitive (not `skip`) hit with t <= tmax_incl, through the BVH.
inline bool bvh_other_hit(const or_bvh& b, const WRay& r, float tmax_incl, uint32_t skip) {
    float inv[3];
    safe_inv(r.d, inv);
    uint32_t stack[256];
    int sp = 0;
    if (b.nodes.empty()) return false;
    stack[sp++] = 0;
    while (sp > 0) {
        uint32_t c = stack[--sp];
        const Node& n = b.nodes[c];
        float tt;
        if (!box_hit(n, r.o, inv, r.tmin, tmax_incl, tt)) continue;
        if (n.count) {
            for (uint32_t i = 0; i < n.count; ++i) {
                if (b.prim[n.first + i] == skip) continue;
                const float* tv = &b.tris[9 * (size_t)(n.first + i)];
                float t, u, v;
                if (wt_hit(r, tv, tv + 3, tv + 6, tmax_incl, t, u, v)) return true;
            }
        } else {
            stack[sp++] = n.first + 1;
            stack[sp++] = n.first;
        }
    }
    return false;
}

inline bool bvh_occluded_one(const or_bvh& b, const float* ray) {
    if (!ray_valid(ray)) return false;
    WRay r = make_wray(ray);
    float inv[3];
    safe_inv(r.d, inv);
    uint32_t stack[256];
    int sp = 0;
    if (b.nodes.empty()) return false;
    stack[sp++] = 0;
    while (sp > 0) {
        uint32_t c = stack[--sp];
        const Node& n = b.nodes[c];
        float tt;
        if (!box_hit(n, r.o, inv, r.tmin, r.tmax, tt)) continue;
        if (n.count) {
            for (uint32_t i = 0; i < n.count; ++i) {
                const float* tv = &b.tris[9 * (size_t)(n.first + i)];
                float t, u, v;
                if (wt_hit(r, tv, tv + 3, tv + 6, r.tmax, t, u, v)) return true;
            }
        } else {
            stack[sp++] = n.first + 1;
            stack[sp++] = n.first;
        }
    }
    return false;
}

}  // namespace

// ================================================================ C API ====
namespace {
// Value of a binary16 bit pattern holding a non-negative integer (BVH8 planes).
double half_int(uint16_t h) {
    const int e = (h >> 10) & 0x1F, m = h & 0x3FF;
    if (e == 0) return std::ldexp((double)m, -24);
    return std::ldexp(1.0 + m / 1024.0, e - 15);
}
}  // namespace

extern "C" {

void or_pcg32_u32(uint64_t initstate, uint64_t initseq, uint32_t* out, int n) {
    Pcg32 p;
    p.seed(initstate, initseq);
    for (int i = 0; i < n; ++i) out[i] = p.next();
}

void or_pcg32_f32(uint64_t initstate, uint64_t initseq, float* out, int n) {
    Pcg32 p;
    p.seed(initstate, initseq);
    for (int i = 0; i < n; ++i) out[i] = p.nextf();
}

int or_intersect_triangle(const float* ray8, const float* v0, const float* v1,
                          const float* v2, float* t, float* u, float* v) {
    if (!ray_valid(ray8)) return 0;
    WRay r = make_wray(ray8);
    return wt_hit(r, v0, v1, v2, r.tmax, *t, *u, *v) ? 1 : 0;
}

void or_closest_bruteforce(const float* tris, uint32_t n_tris, const float* rays,
                           uint32_t n_rays, float* hits, int threads) {
    parallel_chunks(n_rays, threads, 64, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            if (!ray_valid(rays + 8 * (size_t)i)) { write_miss(hits + 4 * (size_t)i); continue; }
            WRay r = make_wray(rays + 8 * (size_t)i);
            float best_t = r.tmax, bu = 0, bv = 0;
            uint32_t best_p = 0xFFFFFFFFu;
            for (uint32_t p = 0; p < n_tris; ++p) {
                const float* tv = tris + 9 * (size_t)p;
                float t, u, v;
                if (wt_hit(r, tv, tv + 3, tv + 6, best_t, t, u, v)) {
                    if (t < best_t || p < best_p) { best_t = t; best_p = p; bu = u; bv = v; }
                }
            }
            float* h = hits + 4 * (size_t)i;
            if (best_p == 0xFFFFFFFFu) write_miss(h);
            else write_hit(h, best_t, best_p, bu, bv);
        }
    });
}

// Two-level closest hit: each instance tests its mesh with the ray moved to
// object space by the instance's world-to-object matrix (Embree instances:
// rtcSetGeometryTransform on a subscene, src/core/accel.cpp:304-313, placed by
// EmbreeAccel::add_instances, accel.cpp:317-346). Brute force over instances
// and triangles; the world-space t is kept (the direction is transformed
// unnormalised).
void or_instanced_closest(const float* mesh_tris, const uint32_t* mesh_first, const uint32_t* mesh_count,
                          const float* inst_w2o, const uint32_t* inst_mesh, const uint32_t* inst_base,
                          uint32_t n_inst, const float* rays, uint32_t n_rays, float* hits, int threads) {
    parallel_chunks(n_rays, threads, 16, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            const float* wr = rays + 8 * (size_t)i;
            if (!ray_valid(wr)) { write_miss(hits + 4 * (size_t)i); continue; }
            float best_t = wr[7], bu = 0, bv = 0;
            uint32_t best_p = 0xFFFFFFFFu;
            for (uint32_t k = 0; k < n_inst; ++k) {
                const float* m = inst_w2o + 12 * (size_t)k;
                static const float eye[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
                float orr[8];
                if (std::memcmp(m, eye, sizeof(eye)) == 0) {
                    std::memcpy(orr, wr, sizeof(orr));   // identity placement: the ray as given
                } else {
                    for (int a = 0; a < 3; ++a) {
                        const float* row = m + 4 * a;
                        orr[a] = ((row[0] * wr[0] + row[1] * wr[1]) + row[2] * wr[2]) + row[3];
                        orr[4 + a] = (row[0] * wr[4] + row[1] * wr[5]) + row[2] * wr[6];
                    }
                    orr[3] = wr[3];
                    orr[7] = wr[7];
                }
                if (!ray_valid(orr)) continue;
                WRay r = make_wray(orr);
                const uint32_t mesh = inst_mesh[k];
                for (uint32_t p = 0; p < mesh_count[mesh]; ++p) {
                    const float* tv = mesh_tris + 9 * (size_t)(mesh_first[mesh] + p);
                    float t, u, v;
                    if (wt_hit(r, tv, tv + 3, tv + 6, best_t, t, u, v)) {
                        const uint32_t gp = inst_base[k] + p;
                        if (t < best_t || gp < best_p) { best_t = t; best_p = gp; bu = u; bv = v; }
                    }
                }
            }
            float* h = hits + 4 * (size_t)i;
            if (best_p == 0xFFFFFFFFu) write_miss(h);
            else write_hit(h, best_t, best_p, bu, bv);
        }
    });
}

uint32_t or_count_near_ties(const float* tris, uint32_t n_tris, const float* rays,
                            uint32_t n_rays, int threads) {
    std::atomic<uint32_t> ties{0};
    parallel_chunks(n_rays, threads, 64, [&](uint32_t b, uint32_t e) {
        uint32_t local = 0;
        for (uint32_t i = b; i < e; ++i) {
            if (!ray_valid(rays + 8 * (size_t)i)) continue;
            WRay r = make_wray(rays + 8 * (size_t)i);
            float best_t = r.tmax;
            uint32_t best_p = 0xFFFFFFFFu;
            for (uint32_t p = 0; p < n_tris; ++p) {
                const float* tv = tris + 9 * (size_t)p;
                float t, u, v;
                if (wt_hit(r, tv, tv + 3, tv + 6, best_t, t, u, v))
                    if (t < best_t || p < best_p) { best_t = t; best_p = p; }
            }
            if (best_p == 0xFFFFFFFFu) continue;
            const float lim = std::nextafter(best_t, INFINITY);
            for (uint32_t p = 0; p < n_tris; ++p) {
                if (p == best_p) continue;
                const float* tv = tris + 9 * (size_t)p;
                float t, u, v;
                if (wt_hit(r, tv, tv + 3, tv + 6, lim, t, u, v)) { local++; break; }
            }
        }
        ties += local;
    });
    return ties.load();
}

uint32_t or_bvh_count_near_ties(const or_bvh* b, const float* rays, uint32_t n_rays, const float* hits,
                                int threads) {
    std::atomic<uint32_t> ties{0};
    parallel_chunks(n_rays, threads, 256, [&](uint32_t lo, uint32_t hi) {
        uint32_t local = 0;
        for (uint32_t i = lo; i < hi; ++i) {
            uint32_t p;
            std::memcpy(&p, &hits[4 * (size_t)i + 1], 4);
            if (p == 0xFFFFFFFFu || !ray_valid(rays + 8 * (size_t)i)) continue;
            const WRay r = make_wray(rays + 8 * (size_t)i);
            // no primitive lies before the closest t, so any other hit up to
            // one ulp beyond it is a near tie
            if (bvh_other_hit(*b, r, std::nextafter(hits[4 * (size_t)i], INFINITY), p)) local++;
        }
        ties += local;
    });
    return ties.load();
}

or_bvh* or_bvh_build(const float* tris, uint32_t n_tris) {
    or_bvh* b = new or_bvh();
    build_bvh(*b, tris, n_tris);
    return b;
}

void or_bvh_free(or_bvh* b) { delete b; }

uint32_t or_bvh_node_count(const or_bvh* b) { return (uint32_t)b->nodes.size(); }

void or_bvh_closest(const or_bvh* b, const float* rays, uint32_t n_rays, float* hits,
                    int threads) {
    parallel_chunks(n_rays, threads, 256, [&](uint32_t s, uint32_t e) {
        for (uint32_t i = s; i < e; ++i)
            bvh_closest_one(*b, rays + 8 * (size_t)i, hits + 4 * (size_t)i);
    });
}

void or_bvh_occluded(const or_bvh* b, const float* rays, uint32_t n_rays, uint8_t* out,
                     int threads) {
    parallel_chunks(n_rays, threads, 256, [&](uint32_t s, uint32_t e) {
        for (uint32_t i = s; i < e; ++i) out[i] = bvh_occluded_one(*b, rays + 8 * (size_t)i) ? 1 : 0;
    });
}

double or_bvh_closest_timed(const or_bvh* b, const float* rays, uint32_t n_rays, float* hits,
                            int threads) {
    auto t0 = std::chrono::steady_clock::now();
    or_bvh_closest(b, rays, n_rays, hits, threads);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

int or_hardware_threads(void) { return (int)std::max(1u, std::thread::hardware_concurrency()); }

void or_bvh8_closest(const void* nodes_v, uint64_t n_nodes, const void* tris_v, uint64_t n_tris,
                     const float* rays, uint32_t n_rays, float* hits, uint64_t* nodes_visited,
                     uint64_t* tris_tested, int threads) {
    // Node record: p[3] f32 | e[3] i8, imask u8 | child_base u32 | prim_base u32
    //              | meta[8] u8 | q[6][8] binary16 integers (lo.xyz, hi.xyz planes).
    // Visit order of the HIP kernel (kazen_proto_amd/csrc/trace.hip traverse):
    // hit leaf children at once, the nearest hit inner child next, the other
    // hit inner children as one group taken in octant order and dropped when
    // the second-nearest entry distance exceeds the closest hit.  Child boxes
    // are tested exactly in FP64 (the kernel's FP32 test is conservative), so
    // the counts can differ slightly from the kernel's.
    const uint8_t* nodes = (const uint8_t*)nodes_v;
    const float* tris = (const float*)tris_v;
    std::atomic<uint64_t> tot_nodes{0}, tot_tris{0};
    parallel_chunks(n_rays, threads, 256, [&](uint32_t s, uint32_t e) {
        uint64_t ln = 0, lt = 0;
        struct Group { uint32_t base, imask, pm; float bound; };
        std::vector<Group> stack;
        for (uint32_t i = s; i < e; ++i) {
            const float* ray = rays + 8 * (size_t)i;
            if (!ray_valid(ray)) { write_miss(hits + 4 * (size_t)i); continue; }
            WRay r = make_wray(ray);
            uint32_t oct = 0;
            for (int k = 0; k < 3; ++k)
                if (std::signbit(r.d[k])) oct |= 1u << k;
            float best_t = r.tmax, bu = 0, bv = 0;
            uint32_t best_p = 0xFFFFFFFFu;
            stack.clear();
            uint32_t cur = 0;
            for (;;) {
                if (cur != 0xFFFFFFFFu && cur < n_nodes) {
                    ++ln;
                    const uint8_t* nd = nodes + 128 * (size_t)cur;
                    cur = 0xFFFFFFFFu;
                    float p[3];
                    std::memcpy(p, nd, 12);
                    int ex[3] = {(int8_t)nd[12], (int8_t)nd[13], (int8_t)nd[14]};
                    const uint32_t imask = nd[15];
                    uint32_t child_base, prim_base;
                    std::memcpy(&child_base, nd + 16, 4);
                    std::memcpy(&prim_base, nd + 20, 4);
                    const uint8_t* meta = nd + 24;
                    uint16_t q[6][8];
                    std::memcpy(q, nd + 32, 96);
                    uint32_t hitm = 0, k1 = 0xFFFFFFFFu, k2 = 0xFFFFFFFFu;
                    for (int sl = 0; sl < 8; ++sl) {
                        if (meta[sl] == 0) continue;
                        double tn0 = (double)r.tmin, tf0 = (double)best_t;
                        for (int k = 0; k < 3; ++k) {
                            const double sc = std::ldexp(1.0, ex[k]);
                            const double lo = (double)p[k] + half_int(q[k][sl]) * sc;
                            const double hi = (double)p[k] + half_int(q[3 + k][sl]) * sc;
                            const double d = (double)r.d[k], o = (double)r.o[k];
                            if (d == 0.0) {
                                if (o < lo || o > hi) { tn0 = 1.0; tf0 = 0.0; }
                                continue;
                            }
                            const double t0 = (lo - o) / d, t1 = (hi - o) / d;
                            tn0 = std::max(tn0, std::min(t0, t1));
                            tf0 = std::min(tf0, std::max(t0, t1));
                        }
                        if (!(tn0 <= tf0)) continue;
                        hitm |= 1u << sl;
                        if ((imask >> sl) & 1u) {
                            const float tnf = (float)tn0;
                            uint32_t bits;
                            std::memcpy(&bits, &tnf, 4);
                            const uint32_t key = (tnf > 0.0f ? (bits & 0xFFFFFFF8u) : 0u) | (uint32_t)sl;
                            k2 = std::min(k2, std::max(k1, key));
                            k1 = std::min(k1, key);
                        }
                    }
                    // leaf children in slot order
                    for (int sl = 0; sl < 8; ++sl) {
                        if (!((hitm >> sl) & 1u) || ((imask >> sl) & 1u)) continue;
                        const uint32_t first = prim_base + (meta[sl] & 31u), count = meta[sl] >> 5;
                        for (uint32_t k = 0; k < count; ++k) {
                            if (first + k >= n_tris) continue;
                            const float* tr = tris + 12 * (size_t)(first + k);
                            ++lt;
                            float t, u, v;
                            if (wt_hit(r, tr, tr + 4, tr + 8, best_t, t, u, v)) {
                                uint32_t pr;
                                std::memcpy(&pr, &tr[3], 4);
                                if (t < best_t || pr < best_p) { best_t = t; best_p = pr; bu = u; bv = v; }
                            }
                        }
                    }
                    const uint32_t inner = hitm & imask;
                    if (inner) {
                        const uint32_t s1 = k1 & 7u;
                        const uint32_t rest = inner & ~(1u << s1);
                        if (rest) {
                            uint32_t pm = 0;
                            for (uint32_t sl = 0; sl < 8; ++sl)
                                if ((rest >> sl) & 1u) pm |= 1u << (sl ^ oct);
                            uint32_t bb = (k2 < 8u) ? 0xFF800000u : (k2 & 0xFFFFFF00u);
                            float bound;
                            std::memcpy(&bound, &bb, 4);
                            stack.push_back({child_base, imask, pm, bound});
                        }
                        cur = child_base + (uint32_t)__builtin_popcount(imask & ((1u << s1) - 1u));
                    }
                }
                if (cur == 0xFFFFFFFFu) {
                    while (!stack.empty()) {
                        Group& g = stack.back();
                        if (g.bound > best_t) { stack.pop_back(); continue; }
                        const uint32_t rank = (uint32_t)__builtin_ctz(g.pm);
                        const uint32_t sl = rank ^ oct;
                        g.pm &= g.pm - 1u;
                        cur = g.base + (uint32_t)__builtin_popcount(g.imask & ((1u << sl) - 1u));
                        if (g.pm == 0) stack.pop_back();
                        break;
                    }
                    if (cur == 0xFFFFFFFFu) break;
                }
            }
            float* h = hits + 4 * (size_t)i;
            if (best_p == 0xFFFFFFFFu) write_miss(h);
            else write_hit(h, best_t, best_p, bu, bv);
        }
        tot_nodes += ln;
        tot_tris += lt;
    });
    if (nodes_visited) *nodes_visited = tot_nodes.load();
    if (tris_tested) *tris_tested = tot_tris.load();
}

}  // extern "C"
