// Host traversal of the BVH8 records (see host_trace.h).
//
// Built with contraction off for x86-64-v3 (FMA only where fmaf is written),
// so the triangle test and the instance transform round exactly as the HIP
// kernels in trace.hip do.  The node test uses the kernels' conservative
// quantised slab test; any conservative traversal finds the same closest hit
// because ties go to the lowest primitive id.
#include "host_trace.h"

#include <cmath>
#include <cstring>

namespace kp {
namespace {

struct Ray {
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
    float ix, iy, iz;
    float Sx, Sy, Sz;
    int kx, ky, kz;
};

inline float pick3(float a, float b, float c, int k) { return k == 0 ? a : (k == 1 ? b : c); }

inline float safe_rcp(float d) {
    const float dd = std::fabs(d) < 1e-20f ? std::copysign(1e-20f, d) : d;
    return 1.0f / dd;
}

// Same rule as the kernels: non-finite origin or direction, a zero
// direction or a NaN bound misses everything.
inline bool ray_valid(const float* a) {
    const bool fin = std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]) && std::isfinite(a[4]) &&
                     std::isfinite(a[5]) && std::isfinite(a[6]);
    const bool nz = a[4] != 0.0f || a[5] != 0.0f || a[6] != 0.0f;
    return fin && nz && !std::isnan(a[3]) && !std::isnan(a[7]);
}

inline void setup_ray(const float* a, Ray& r) {
    r.ox = a[0]; r.oy = a[1]; r.oz = a[2]; r.tmin = a[3];
    r.dx = a[4]; r.dy = a[5]; r.dz = a[6]; r.tmax = a[7];
    const float ax = std::fabs(r.dx), ay = std::fabs(r.dy), az = std::fabs(r.dz);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = (kz == 2) ? 0 : kz + 1;
    int ky = (kx == 2) ? 0 : kx + 1;
    const float dkz = pick3(r.dx, r.dy, r.dz, kz);
    if (dkz < 0.0f) { const int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = pick3(r.dx, r.dy, r.dz, kx) / dkz;
    r.Sy = pick3(r.dx, r.dy, r.dz, ky) / dkz;
    r.Sz = 1.0f / dkz;
    r.ix = safe_rcp(r.dx);
    r.iy = safe_rcp(r.dy);
    r.iz = safe_rcp(r.dz);
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), the kernels'
// operation order; zero edge functions are re-evaluated in double.
inline bool tri_test(const Ray& r, const float* a, const float* b, const float* c, float tmax_incl, float& t_out,
                     float& u_out, float& v_out) {
    const float A0 = a[0] - r.ox, A1 = a[1] - r.oy, A2 = a[2] - r.oz;
    const float B0 = b[0] - r.ox, B1 = b[1] - r.oy, B2 = b[2] - r.oz;
    const float C0 = c[0] - r.ox, C1 = c[1] - r.oy, C2 = c[2] - r.oz;
    const float Akz = pick3(A0, A1, A2, r.kz), Bkz = pick3(B0, B1, B2, r.kz), Ckz = pick3(C0, C1, C2, r.kz);
    const float Ax = pick3(A0, A1, A2, r.kx) - r.Sx * Akz;
    const float Ay = pick3(A0, A1, A2, r.ky) - r.Sy * Akz;
    const float Bx = pick3(B0, B1, B2, r.kx) - r.Sx * Bkz;
    const float By = pick3(B0, B1, B2, r.ky) - r.Sy * Bkz;
    const float Cx = pick3(C0, C1, C2, r.kx) - r.Sx * Ckz;
    const float Cy = pick3(C0, C1, C2, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float t = T / det;
    if (!(t >= r.tmin && t <= tmax_incl)) return false;
    t_out = t;
    u_out = V / det;
    v_out = W / det;
    return true;
}

inline float bits_f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline uint32_t f_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// 2^e as the kernels form it (exponent field e + 127).
inline float exp2i(int e) { return bits_f((uint32_t)(e + 127) << 23); }

// binary16 -> float, exact (the planes hold small integers).
inline float half_to_float(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 31u, m = h & 1023u;
    if (e == 0) {
        const float v = (float)m * 5.9604644775390625e-8f;   // m * 2^-24
        return s ? -v : v;
    }
    if (e == 31) return bits_f(s | 0x7F800000u | (m << 13));
    return bits_f(s | ((e + 112u) << 23) | (m << 13));
}

constexpr float kFarEps = 3.5762793e-7f;   // 2 gamma(3), Ize 2013

// Tests the 8 children of node nd against [r.tmin, tmax]: returns the hit
// slots, with each slot's entry distance in tn_out.
inline uint32_t visit(const KpNode8& nd, const Ray& r, float tmax, float* tn_out) {
    const bool nx = r.ix < 0.0f, ny = r.iy < 0.0f, nz = r.iz < 0.0f;
    const uint16_t* qnx = nd.q[nx ? 3 : 0];
    const uint16_t* qfx = nd.q[nx ? 0 : 3];
    const uint16_t* qny = nd.q[ny ? 4 : 1];
    const uint16_t* qfy = nd.q[ny ? 1 : 4];
    const uint16_t* qnz = nd.q[nz ? 5 : 2];
    const uint16_t* qfz = nd.q[nz ? 2 : 5];
    const float ax = exp2i(nd.ex) * r.ix, ay = exp2i(nd.ey) * r.iy, az = exp2i(nd.ez) * r.iz;
    const float bx = (nd.px - r.ox) * r.ix, by = (nd.py - r.oy) * r.iy, bz = (nd.pz - r.oz) * r.iz;
    uint32_t hit = 0u;
    for (int s = 0; s < 8; ++s) {
        if (nd.meta[s] == 0) continue;   // empty slot
        const float tnx = std::fmaf(half_to_float(qnx[s]), ax, bx);
        const float tfx = std::fmaf(half_to_float(qfx[s]), ax, bx);
        const float tny = std::fmaf(half_to_float(qny[s]), ay, by);
        const float tfy = std::fmaf(half_to_float(qfy[s]), ay, by);
        const float tnz = std::fmaf(half_to_float(qnz[s]), az, bz);
        const float tfz = std::fmaf(half_to_float(qfz[s]), az, bz);
        const float tn = std::fmax(std::fmax(tnx, tny), std::fmax(tnz, r.tmin));
        const float tf = std::fmin(std::fmin(tfx, tfy), std::fmin(tfz, tmax));
        if (tn <= std::fmaf(std::fabs(tf), kFarEps, tf)) {
            hit |= 1u << s;
            tn_out[s] = tn;
        }
    }
    return hit;
}

inline uint32_t inner_child(const KpNode8& nd, uint32_t slot) {
    return nd.child_base + (uint32_t)__builtin_popcount(nd.imask & ((1u << slot) - 1u));
}

// Node stack with inline storage; deeper traversals spill to the heap.
struct Stack {
    struct E {
        uint32_t node;
        float tn;
    };
    E buf[256];
    std::vector<E> more;
    uint32_t n = 0;
    void push(uint32_t node, float tn) {
        if (n < 256) buf[n] = E{node, tn};
        else more.push_back(E{node, tn});
        ++n;
    }
    E pop() {
        --n;
        if (n < 256) return buf[n];
        E e = more.back();
        more.pop_back();
        return e;
    }
};

struct Query {
    const HostBvh& b;
    bool any;
    float best_t;
    uint32_t best_p;
    float bu, bv;
    bool occl;
};

// World-space ray -> an instance's object space: o' = M o + m, d' = M d in
// FP32 with the kernels' operation order (no contraction).
inline void ray_to_object(const float* m, const float* a, float* o) {
    o[0] = ((m[0] * a[0] + m[1] * a[1]) + m[2] * a[2]) + m[3];
    o[1] = ((m[4] * a[0] + m[5] * a[1]) + m[6] * a[2]) + m[7];
    o[2] = ((m[8] * a[0] + m[9] * a[1]) + m[10] * a[2]) + m[11];
    o[3] = a[3];
    o[4] = (m[0] * a[4] + m[1] * a[5]) + m[2] * a[6];
    o[5] = (m[4] * a[4] + m[5] * a[5]) + m[6] * a[6];
    o[6] = (m[8] * a[4] + m[9] * a[5]) + m[10] * a[6];
    o[7] = a[7];
}

void walk(Query& q, const Ray& r, uint32_t root, uint32_t prim_off, bool top, const float* world8);

// Tests the leaf children `lm` of nd (triangles, or instances in the top level).
inline void leaves(Query& q, const Ray& r, const KpNode8& nd, uint32_t lm, const float* tn, uint32_t prim_off,
                   bool top, const float* world8) {
    while (lm) {
        const uint32_t s = (uint32_t)__builtin_ctz(lm);
        lm &= lm - 1u;
        const uint32_t meta = nd.meta[s];
        const uint32_t first = nd.prim_base + (meta & 31u), count = meta >> 5;
        for (uint32_t i = 0; i < count; ++i) {
            if (top) {
                if (!q.any && tn[s] > q.best_t) break;
                const uint32_t inst = q.b.tlas_inst[first + i];
                const float* rec = &q.b.insts[16 * (size_t)inst];
                uint32_t root = f_bits(rec[12]), base = f_bits(rec[13]), flags = f_bits(rec[14]);
                if (flags & 1u) {   // identity placement: the ray as given
                    walk(q, r, root, base, false, world8);
                } else {
                    float o8[8];
                    ray_to_object(rec, world8, o8);
                    if (!ray_valid(o8)) continue;
                    Ray ro;
                    setup_ray(o8, ro);
                    walk(q, ro, root, base, false, world8);
                }
                if (q.any && q.occl) return;
                continue;
            }
            const float* tp = &q.b.tris[12 * (size_t)(first + i)];
            float t, u, v;
            if (tri_test(r, tp, tp + 4, tp + 8, q.best_t, t, u, v)) {
                if (q.any) {
                    q.occl = true;
                    return;
                }
                const uint32_t pr = f_bits(tp[3]) + prim_off;
                if (t < q.best_t || pr < q.best_p) {
                    q.best_t = t;
                    q.best_p = pr;
                    q.bu = u;
                    q.bv = v;
                }
            }
        }
    }
}

void walk(Query& q, const Ray& r, uint32_t root, uint32_t prim_off, bool top, const float* world8) {
    Stack st;
    st.push(root, -INFINITY);
    while (st.n) {
        const Stack::E e = st.pop();
        if (!q.any && e.tn > q.best_t) continue;
        const KpNode8& nd = q.b.nodes[e.node];
        float tn[8];
        const uint32_t hit = visit(nd, r, q.best_t, tn);
        const uint32_t lm = hit & ~(uint32_t)nd.imask, im = hit & (uint32_t)nd.imask;
        if (lm) {
            leaves(q, r, nd, lm, tn, prim_off, top, world8);
            if (q.any && q.occl) return;
        }
        // inner children, nearest popped first
        uint32_t order[8], k = 0;
        for (uint32_t m = im; m; m &= m - 1u) order[k++] = (uint32_t)__builtin_ctz(m);
        for (uint32_t i = 1; i < k; ++i)
            for (uint32_t j = i; j > 0 && tn[order[j]] > tn[order[j - 1]]; --j) {
                const uint32_t t = order[j];
                order[j] = order[j - 1];
                order[j - 1] = t;
            }
        for (uint32_t i = 0; i < k; ++i) st.push(inner_child(nd, order[i]), tn[order[i]]);
    }
}

void query(Query& q, const float* ray8) {
    if (!ray_valid(ray8)) return;
    Ray r;
    setup_ray(ray8, r);
    q.best_t = r.tmax;
    if (!q.b.two_level) {
        walk(q, r, 0u, 0u, false, ray8);
        return;
    }
    if (q.b.world_root != KP_INVALID) {
        walk(q, r, q.b.world_root, 0u, false, ray8);
        if (q.any && q.occl) return;
    }
    walk(q, r, 0u, 0u, true, ray8);
}

}  // namespace

void host_closest(const HostBvh& b, const float* rays, uint32_t n, float* hits) {
    for (uint32_t i = 0; i < n; ++i) {
        Query q{b, false, 0.0f, KP_INVALID, 0.0f, 0.0f, false};
        if (!b.empty()) query(q, rays + 8 * (size_t)i);
        float* h = hits + 4 * (size_t)i;
        if (q.best_p == KP_INVALID) {
            h[0] = INFINITY;
            h[1] = bits_f(KP_INVALID);
            h[2] = 0.0f;
            h[3] = 0.0f;
        } else {
            h[0] = q.best_t;
            h[1] = bits_f(q.best_p);
            h[2] = q.bu;
            h[3] = q.bv;
        }
    }
}

void host_occluded(const HostBvh& b, const float* rays, uint32_t n, uint32_t* out) {
    for (uint32_t i = 0; i < n; ++i) {
        Query q{b, true, 0.0f, KP_INVALID, 0.0f, 0.0f, false};
        if (!b.empty()) query(q, rays + 8 * (size_t)i);
        out[i] = q.occl ? 1u : 0u;
    }
}

}  // namespace kp
