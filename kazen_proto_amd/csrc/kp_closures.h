// Closures of the wavefront integrator (device code, y-up local frame).
//
// Reference closures (file:line in kazen_proto):
//   Lambert      Diffuse::eval/sample            src/shading/bsdfs.cpp:27-41
//   conductor    KpPrincipleSpecularReflection   src/shading/bsdfs.cpp:688-730
//                + Schlick F0                     src/shading/fresnel.h:185-187
//   dielectric   KpGlass (Walter et al. 2007)     src/shading/bsdfs.cpp:441-554,
//                                                 bsdfs.h:542-599
//   gloss        KpGloss (dielectric Fresnel)     src/shading/bsdfs.cpp:359-439
//   microfacet   GGXDist / BeckmannDist / MicrofacetInterface, anisotropic
//                (stretched / projected roughness) src/shading/microfacet.h:60-291
//   Fresnel      fresnel_refl_dielectric          src/shading/fresnel.h:74-131
//   composite    CompositeClosure (<= 8 weighted lobes) src/shading/bsdf.h:132-196,
//                                                 bsdf.cpp:17-74
// All closures return f without the cosine; the integrator applies |cos| once.
// The VNDF samplers are exact (GGX: Heitz 2018; Beckmann: Jakob 2014 slope
// inversion), so the reported pdf is the density actually sampled, which the
// MIS weights require.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kp {

#define KP_PI 3.14159265358979323846f
#define KP_INV_PI 0.318309886183790671538f

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) {
    const float l = length(a);
    return {a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ __forceinline__ float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }

// Shading frame with the reference's construction (src/base/utils.h:183-210):
// local (x, y, z) = (s, n, t) components.
struct Frame {
    f3 s, t, n;
    __device__ __forceinline__ explicit Frame(f3 nn) : n(nn) {
        if (fabsf(n.x) > fabsf(n.y)) {
            const float inv = 1.0f / sqrtf(n.x * n.x + n.z * n.z);
            t = mk(n.z * inv, 0.0f, -n.x * inv);
        } else {
            const float inv = 1.0f / sqrtf(n.y * n.y + n.z * n.z);
            t = mk(0.0f, n.z * inv, -n.y * inv);
        }
        s = cross(t, n);
    }
    __device__ __forceinline__ f3 to_local(f3 v) const { return mk(dot(v, s), dot(v, n), dot(v, t)); }
    __device__ __forceinline__ f3 to_world(f3 v) const { return s * v.x + n * v.y + t * v.z; }
};

// One closure lobe (48 bytes): type (0 Lambert, 1 Schlick conductor, 2 rough
// dielectric, 3 gloss), distribution (0 GGX, 1 Beckmann), closure weight w,
// colour c (Lambert albedo or conductor F0), roughness ax along the frame's x
// axis and ay along z, eta = n_inside / n_outside.
struct Lobe {
    int32_t type, dist;
    float w[3];
    float c[3];
    float ax, ay, eta;
    uint32_t pad;
};

// A material (32 bytes): lobes [first, first + n) of the lobe table, flags
// (bit 0: a lobe transmits, so the shading normal keeps its orientation),
// the smallest microfacet roughness (guiding eligibility; +inf: none) and the
// emitted radiance.
#define KP_MAX_LOBES 8
#define KP_MATF_TRANSMIT 1u
struct MatRec {
    uint32_t first, n;
    uint32_t flags;
    float guide_alpha;
    float Le[3];
    float pad;
};

// ------------------------------------------------------------ microfacet --
// D(m) with roughness ax along x and ay along z (Heitz 2014 eq. 87, the
// stretched_roughness of microfacet.h:60-76); ax == ay takes the isotropic form.
__device__ __forceinline__ float mf_D(int dist, float ax, float ay, f3 m) {
    const float c = m.y;
    if (c <= 0.0f) return 0.0f;
    const float c2 = c * c;
    const float e = ax == ay ? ((1.0f - c2) / c2) / (ax * ax) : (m.x * m.x / (ax * ax) + m.z * m.z / (ay * ay)) / c2;
    if (dist == 0) {
        const float k = 1.0f + e;
        return 1.0f / (KP_PI * ax * ay * c2 * c2 * k * k);
    }
    return expf(-e) / (KP_PI * ax * ay * c2 * c2);
}

// Smith Lambda with the projected roughness (Heitz 2014 eq. 80,
// microfacet.h:78-93); exact Beckmann erf form.
__device__ __forceinline__ float mf_lambda(int dist, float ax, float ay, f3 w) {
    const float c = w.y;
    const float s2 = fmaxf(0.0f, 1.0f - c * c);
    if (s2 == 0.0f || c == 0.0f) return 0.0f;
    const float at = ax == ay ? ax * (sqrtf(s2) / fabsf(c))
                              : sqrtf(w.x * w.x * ax * ax + w.z * w.z * ay * ay) / fabsf(c);
    if (dist == 0) return 0.5f * (-1.0f + sqrtf(1.0f + at * at));
    const float a = 1.0f / at;
    return 0.5f * (erff(a) - 1.0f) + expf(-a * a) / (2.0f * a * 1.7724538509055160f);
}

__device__ __forceinline__ float mf_G1(int dist, float ax, float ay, f3 w) {
    return 1.0f / (1.0f + mf_lambda(dist, ax, ay, w));
}

__device__ __forceinline__ float mf_G2(int dist, float ax, float ay, f3 wo, f3 wi) {
    return 1.0f / (1.0f + mf_lambda(dist, ax, ay, wo) + mf_lambda(dist, ax, ay, wi));
}

__device__ __forceinline__ float mf_pdf_vndf(int dist, float ax, float ay, f3 wo, f3 m) {
    const float c = fabsf(wo.y);
    if (c == 0.0f) return 0.0f;
    return mf_G1(dist, ax, ay, wo) * fabsf(dot(wo, m)) * mf_D(dist, ax, ay, m) / c;
}

// Beckmann visible-slope sampling at alpha = 1 (W. Jakob, "An Improved
// Visible Normal Sampling Routine for the Beckmann Distribution", 2014).
__device__ inline void beckmann_sample11(float cos_t, float u1, float u2, float& sx, float& sy) {
    if (cos_t > 0.9999f) {
        const float r = sqrtf(-logf(1.0f - u1));
        float sp, cp;
        sincosf(2.0f * KP_PI * u2, &sp, &cp);
        sx = r * cp;
        sy = r * sp;
        return;
    }
    const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    const float tan_t = sin_t / cos_t, cot_t = 1.0f / tan_t;
    float a = -1.0f, c = erff(cot_t);
    const float sample_x = fmaxf(u1, 1e-6f);
    const float theta = acosf(cos_t);
    const float fit = 1.0f + theta * (-0.876f + theta * (0.4265f - 0.0594f * theta));
    float b = c - (1.0f + c) * powf(1.0f - sample_x, fit);
    const float inv_sqrt_pi = 0.56418958354775628f;
    const float norm = 1.0f / (1.0f + c + inv_sqrt_pi * tan_t * expf(-cot_t * cot_t));
    for (int it = 0; it < 12; ++it) {
        if (!(b >= a && b <= c)) b = 0.5f * (a + c);
        const float ie = erfinvf(b);
        const float value = norm * (1.0f + b + inv_sqrt_pi * tan_t * expf(-ie * ie)) - sample_x;
        const float deriv = norm * (1.0f - ie * tan_t);
        if (fabsf(value) < 1e-6f) break;
        if (value > 0.0f) c = b;
        else a = b;
        b -= value / deriv;
    }
    sx = erfinvf(b);
    sy = erfinvf(2.0f * fmaxf(u2, 1e-6f) - 1.0f);
}

// Visible normal from wo (returned in the upper hemisphere); anisotropic by
// stretching x with ax and z with ay.
__device__ inline f3 mf_sample_vndf(int dist, float ax, float ay, f3 wo, float u1, float u2) {
    f3 w = wo.y < 0.0f ? -wo : wo;
    if (dist == 0) {
        // GGX: E. Heitz, "Sampling the GGX Distribution of Visible Normals",
        // JCGT 7(4), 2018, in the y-up frame.
        const f3 vh = normalize(mk(ax * w.x, w.y, ay * w.z));
        const float lensq = vh.x * vh.x + vh.z * vh.z;
        const f3 t1 = lensq > 0.0f ? mk(vh.z, 0.0f, -vh.x) * (1.0f / sqrtf(lensq)) : mk(1.0f, 0.0f, 0.0f);
        const f3 t2 = cross(vh, t1);
        const float r = sqrtf(u1);
        float sp, cp;
        sincosf(2.0f * KP_PI * u2, &sp, &cp);
        const float p1 = r * cp;
        float p2 = r * sp;
        const float s = 0.5f * (1.0f + vh.y);
        p2 = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1)) + s * p2;
        const f3 nh = t1 * p1 + t2 * p2 + vh * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2));
        return normalize(mk(ax * nh.x, fmaxf(1e-7f, nh.y), ay * nh.z));
    }
    // Beckmann (Heitz & d'Eon 2014 sec. 4): stretch, sample, rotate, unstretch.
    const f3 ws = normalize(mk(ax * w.x, w.y, ay * w.z));
    float phi = atan2f(ws.z, ws.x);
    if (ws.x == 0.0f && ws.z == 0.0f) phi = 0.0f;
    float sx, sy;
    beckmann_sample11(ws.y, u1, u2, sx, sy);
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const float rx = (cp * sx - sp * sy) * ax;
    const float rz = (sp * sx + cp * sy) * ay;
    return normalize(mk(-rx, 1.0f, -rz));
}

__device__ __forceinline__ float fresnel_dielectric(float cos_i, float eta) {
    if (cos_i < 0.0f) {
        eta = 1.0f / eta;
        cos_i = -cos_i;
    }
    const float sin2_t = (1.0f - cos_i * cos_i) / (eta * eta);
    if (sin2_t >= 1.0f) return 1.0f;
    const float cos_t = sqrtf(fmaxf(0.0f, 1.0f - sin2_t));
    const float rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    const float rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (rp * rp + rs * rs);
}

__device__ __forceinline__ f3 reflect(f3 wo, f3 m) { return 2.0f * dot(wo, m) * m - wo; }

__device__ __forceinline__ bool refract(f3 wo, f3 m, float eta_rel, f3& wi) {
    const float c = dot(wo, m);
    const float inv = 1.0f / eta_rel;
    const float sin2_i = fmaxf(0.0f, 1.0f - c * c);
    const float sin2_t = sin2_i * inv * inv;
    if (sin2_t >= 1.0f) return false;
    const float cos_t = sqrtf(1.0f - sin2_t);
    wi = (-wo) * inv + m * (c * inv - cos_t);
    return true;
}

__device__ __forceinline__ f3 schlick(const float* F0, float c) {
    const float k = 1.0f - fminf(1.0f, fabsf(c));
    const float k5 = k * k * k * k * k;
    return mk(F0[0] + (1.0f - F0[0]) * k5, F0[1] + (1.0f - F0[1]) * k5, F0[2] + (1.0f - F0[2]) * k5);
}

// f (no cosine, no lobe weight) and solid-angle pdf of sampling wi given wo.
__device__ inline void lobe_eval(const Lobe& m, f3 wo, f3 wi, f3& f, float& pdf) {
    f = mk(0.0f, 0.0f, 0.0f);
    pdf = 0.0f;
    if (m.type == 0) {
        if (wo.y <= 0.0f || wi.y <= 0.0f) return;
        f = mk(m.c[0], m.c[1], m.c[2]) * KP_INV_PI;
        pdf = wi.y * KP_INV_PI;
        return;
    }
    if (m.type == 1 || m.type == 3) {
        if (wo.y <= 0.0f || wi.y <= 0.0f) return;
        const f3 h = normalize(wo + wi);
        const float D = mf_D(m.dist, m.ax, m.ay, h);
        const float G = mf_G2(m.dist, m.ax, m.ay, wo, wi);
        const float k = D * G / (4.0f * wo.y * wi.y);
        if (m.type == 1) {
            f = schlick(m.c, dot(wo, h)) * k;
        } else {
            const float F = fresnel_dielectric(dot(wo, h), m.eta);
            f = mk(F * k, F * k, F * k);
        }
        pdf = mf_G1(m.dist, m.ax, m.ay, wo) * D / (4.0f * wo.y);
        return;
    }
    const float co = wo.y, ci = wi.y;
    if (co == 0.0f || ci == 0.0f) return;
    const bool refl = co * ci > 0.0f;
    float etap = 1.0f;
    if (!refl) etap = co > 0.0f ? m.eta : 1.0f / m.eta;
    f3 wm = wi * etap + wo;
    if (dot(wm, wm) == 0.0f) return;
    wm = normalize(wm);
    if (wm.y < 0.0f) wm = -wm;
    if (dot(wm, wi) * ci < 0.0f || dot(wm, wo) * co < 0.0f) return;
    const float F = fresnel_dielectric(dot(wo, wm), m.eta);
    const float D = mf_D(m.dist, m.ax, m.ay, wm);
    const float G = mf_G2(m.dist, m.ax, m.ay, wo, wi);
    const float pvn = mf_pdf_vndf(m.dist, m.ax, m.ay, wo, wm);
    if (refl) {
        const float v = D * G * F / fabsf(4.0f * ci * co);
        f = mk(v, v, v);
        pdf = pvn / (4.0f * fabsf(dot(wo, wm))) * F;
    } else {
        const float dwi = dot(wi, wm), dwo = dot(wo, wm);
        float denom = dwi + dwo / etap;
        denom = denom * denom;
        float v = D * (1.0f - F) * G * fabsf(dwi * dwo / (denom * ci * co));
        // radiance BTDF: (n_view / n_light)^2 (Walter 2007 eq. 21 with eta_o on
        // the camera side); the reference's eval_refraction (bsdfs.h:566-583)
        // carries the adjoint eta^2 instead
        v /= etap * etap;
        f = mk(v, v, v);
        pdf = pvn * fabsf(dwi) / denom * (1.0f - F);
    }
}

// Samples wi from wo for one lobe (u5: reflection / refraction, u6, u7:
// direction); f and pdf are the lobe's own.
__device__ inline bool lobe_sample(const Lobe& m, f3 wo, float u5, float u6, float u7, f3& wi, f3& f,
                                   float& pdf) {
    f = mk(0.0f, 0.0f, 0.0f);
    pdf = 0.0f;
    if (m.type == 0) {
        if (wo.y <= 0.0f) return false;
        const float r = sqrtf(u7);
        float sp, cp;
        sincosf(2.0f * KP_PI * u6, &sp, &cp);
        wi = mk(r * cp, sqrtf(fmaxf(0.0f, 1.0f - u7)), r * sp);
        if (wi.y <= 0.0f) return false;
        f = mk(m.c[0], m.c[1], m.c[2]) * KP_INV_PI;
        pdf = wi.y * KP_INV_PI;
        return true;
    }
    if (m.type == 1 || m.type == 3) {
        if (wo.y <= 0.0f) return false;
        const f3 mm = mf_sample_vndf(m.dist, m.ax, m.ay, wo, u6, u7);
        wi = reflect(wo, mm);
        if (wi.y <= 0.0f) return false;
        lobe_eval(m, wo, wi, f, pdf);
        return pdf > 0.0f;
    }
    if (wo.y == 0.0f) return false;
    const f3 mm = mf_sample_vndf(m.dist, m.ax, m.ay, wo, u6, u7);
    const float F = fresnel_dielectric(dot(wo, mm), m.eta);
    if (u5 < F) {
        wi = reflect(wo, mm);
        if (wo.y * wi.y <= 0.0f) return false;
    } else {
        const float etap = wo.y > 0.0f ? m.eta : 1.0f / m.eta;
        const f3 mside = wo.y > 0.0f ? mm : -mm;
        if (!refract(wo, mside, etap, wi)) return false;
        if (wo.y * wi.y >= 0.0f || wi.y == 0.0f) return false;
    }
    lobe_eval(m, wo, wi, f, pdf);
    return pdf > 0.0f;
}

// A material at one vertex: its lobes and the lobe-pick probabilities
// p_i = dot(w_i, beta) / sum_j dot(w_j, beta) (CompositeClosure::compute_pdfs,
// bsdf.h:157-176, given the path throughput as integrator.cpp:425 gives it),
// uniform when every product is 0; one lobe: p = 1.  p_i is formed on the fly
// from the lobe's weight and inv_sum.
struct Bsdf {
    const Lobe* lobes;
    uint32_t n;
    f3 beta;
    float inv_sum;   // < 0: uniform pick
};

__device__ __forceinline__ Bsdf bsdf_at(const MatRec& m, const Lobe* lobes, f3 beta) {
    Bsdf b;
    b.lobes = lobes + m.first;
    b.n = m.n;
    b.beta = beta;
    b.inv_sum = 1.0f;
    if (b.n > 1) {
        float sum = 0.0f;
        for (uint32_t i = 0; i < b.n; ++i) {
            const Lobe& l = b.lobes[i];
            sum += fmaxf(0.0f, l.w[0] * beta.x + l.w[1] * beta.y + l.w[2] * beta.z);
        }
        b.inv_sum = (sum > 0.0f && isfinite(sum)) ? 1.0f / sum : -1.0f;
    }
    return b;
}

__device__ __forceinline__ float lobe_prob(const Bsdf& b, const Lobe& l) {
    if (b.inv_sum < 0.0f) return 1.0f / (float)b.n;
    return fmaxf(0.0f, l.w[0] * b.beta.x + l.w[1] * b.beta.y + l.w[2] * b.beta.z) * b.inv_sum;
}

// Whole material: f = sum_i w_i f_i, pdf = sum_i p_i pdf_i.
__device__ inline void bsdf_eval(const Bsdf& b, f3 wo, f3 wi, f3& f, float& pdf) {
    if (b.n == 1) {
        const Lobe l = b.lobes[0];
        f3 fl;
        lobe_eval(l, wo, wi, fl, pdf);
        f = mul(mk(l.w[0], l.w[1], l.w[2]), fl);
        return;
    }
    f = mk(0.0f, 0.0f, 0.0f);
    pdf = 0.0f;
    for (uint32_t i = 0; i < b.n; ++i) {
        const Lobe l = b.lobes[i];
        f3 fl;
        float pl;
        lobe_eval(l, wo, wi, fl, pl);
        f = f + mul(mk(l.w[0], l.w[1], l.w[2]), fl);
        pdf = pdf + lobe_prob(b, l) * pl;
    }
}

// Picks lobe i with probability p_i by u5, re-uses u5 rescaled inside the
// lobe, samples it and evaluates the whole material at the sampled direction.
__device__ inline bool bsdf_sample(const Bsdf& b, f3 wo, float u5, float u6, float u7, f3& wi, f3& f,
                                   float& pdf) {
    if (b.n == 1) {
        const Lobe l = b.lobes[0];
        f3 fl;
        const bool ok = lobe_sample(l, wo, u5, u6, u7, wi, fl, pdf);
        f = mul(mk(l.w[0], l.w[1], l.w[2]), fl);
        return ok;
    }
    f = mk(0.0f, 0.0f, 0.0f);
    pdf = 0.0f;
    uint32_t pick = b.n;
    float lo = 0.0f, pp = 0.0f;
    for (uint32_t i = 0; i < b.n; ++i) {
        const float p = lobe_prob(b, b.lobes[i]);
        if (p > 0.0f && (u5 < lo + p || i == b.n - 1)) {
            pick = i;
            pp = p;
            break;
        }
        lo += p;
    }
    if (pick == b.n) return false;
    const float ur = fminf((u5 - lo) / pp, 0.99999994f);
    f3 fl;
    float pl;
    if (!lobe_sample(b.lobes[pick], wo, fmaxf(ur, 0.0f), u6, u7, wi, fl, pl)) return false;
    bsdf_eval(b, wo, wi, f, pdf);
    return pdf > 0.0f;
}

}  // namespace kp
