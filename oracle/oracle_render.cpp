// CPU oracle: closures, area lights and the path integrator (FP64 film).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never the product.
//
// Reference algorithm restated (file:line into kazen_proto):
//   - local shading frame, y-up (cos theta = w.y): src/base/utils.h:83,183-210;
//     OSL globals use N as the frame normal, src/shading/renderservices.cpp:238,251
//   - Lambert: src/shading/bsdfs.cpp:27-41 (eval), sampling.h:102-108 (cosine warp)
//   - microfacet D / Lambda / G1 / G / VNDF pdf: src/shading/microfacet.h:95-291
//   - GGX conductor with Schlick F0: KpPrincipleSpecularReflection,
//     src/shading/bsdfs.cpp:688-730, fresnel.h:185-187
//   - rough dielectric (Walter 2007): KpGlass, src/shading/bsdfs.cpp:441-554,
//     bsdfs.h:542-599
//   - dielectric Fresnel: src/shading/fresnel.h:74-131
//   - area light sampling: GeometryLight::sample/pdf, src/core/light.cpp:52-85;
//     TriangleMesh::sample, src/core/shape.cpp:591-613; DiscrectPDF,
//     src/base/dpdf.h:8-58 (lower_bound on the area CDF)
//   - integrator: PathIntegrator::Li, src/core/integrator.cpp:362-561 (emission
//     with MIS, NEE with the power heuristic, Russian roulette with
//     q = min(max(beta)*0.95^2, 0.99) from depth 3, BSDF continuation)
//   - light pick: Integrator::get_random_light, integrator.cpp:25-34
//   - render loop: Renderer::render, src/core/renderer.cpp:18-90
// Deviations from the reference (its defects, replaced by the intended math):
//   D1 beta *= f*cos/pdf (integrator.cpp:503-505 forgets the pdf);
//   D2 cosine-weighted Lambert sampling (bsdfs.cpp:36 uses a uniform-theta warp);
//   D3/D4 KpGloss sample returns D*G*F/(4 cos cos) and its pdf carries the
//      1/(4|wo.m|) Jacobian (bsdfs.cpp:397-439);
//   D5 Beckmann VNDF sampling uses a real erfinv (microfacet.h:160-179 uses 1/erf);
//   D6 VNDF sampling in the stretched configuration (microfacet.h:196-225);
//   D7 explicit eta = n_inside/n_outside with side handling (fresnel.h:120-131);
//   D8 one depth increment per bounce (integrator.cpp:420,555);
//   D9 emission MIS uses the pdf of the light actually hit (integrator.cpp:452-461);
//   D10 all randomness from the per-path PCG32 stream (sampling.h:15-23);
//   D11 occlusion = any hit in [tmin,tmax] (hitable.h:28);
//   D16 a composite closure samples lobe i with the probabilities its
//       compute_pdfs forms (dot(w_i, beta) normalised, bsdf.h:157-176, with
//       beta the path throughput as integrator.cpp:425 passes it); the
//       reference draws the lobe uniformly but divides by those probabilities
//       (bsdf.cpp:32,38).  f and pdf are the one-sample-MIS sums over all
//       lobes, as bsdf.cpp:40-52 forms them;
//   D17 explicit pinhole camera with vertical fov (camera.h:67-96);
//   D18 ray offsetting after Waechter & Binder, Ray Tracing Gems ch. 6, the
//       method intersection.h:78-91 cites (its implementation has an
//       operator-precedence defect and is unused);
//   D19 every closure returns f without the cosine (bsdfs.cpp:31-32 vs 394);
//   D20 the rough-dielectric refraction value is the radiance BTDF: it carries
//       (n_view / n_light)^2, i.e. Walter 2007 eq. 21 with eta_o on the
//       camera side.  eval_refraction (bsdfs.h:566-583) multiplies by eta^2
//       with eta = n_light / n_view, the adjoint (importance) form, which in a
//       camera path tracer over-weights every refraction into the camera's
//       medium by eta^2 (2.25 at eta 1.5).  Pinned by the smooth-limit albedo
//       R + (1 - R)/eta^2 from outside, R + (1 - R) eta^2 from inside, and the
//       rough albedo written over microfacet normals;
//   D21 the shading-point offset of post_hit (shape.cpp:615-654, Hanika 2021)
//       is formed with world-space vertex normals and offsets; the reference
//       mixes object-space normals with object-space offsets, which is the
//       same for identity transforms and undefined otherwise.
//   Beckmann Lambda uses the exact erf form rather than the rational fit of
//   microfacet.h:134-141, so the VNDF pdf equals the sampled density; the
//   conductor does not clamp D<=4 or f<=1 (bsdfs.cpp:706,716).
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>


// Trace entry points of the oracle BVH (oracle_trace.cpp).
extern "C" {
void or_bvh_closest(const or_bvh* b, const float* rays, uint32_t n_rays, float* hits, int threads);
}

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr double kPiD = 3.14159265358979323846;

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {
    float l = length(a);
    return {a.x / l, a.y / l, a.z / l};
}
inline V3 load3(const float* p) { return {p[0], p[1], p[2]}; }
inline float maxc(V3 a) { return std::max(a.x, std::max(a.y, a.z)); }
inline bool is_zero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

// ----------------------------------------------------------------- RNG ----
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
    float nextf() {
        uint32_t bits = (next() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &bits, 4);
        return f - 1.0f;
    }
};

// Per-path stream: state from (seed, sample index), sequence = pixel index.
inline uint64_t path_initstate(uint32_t seed, uint32_t sample) {
    return ((uint64_t)sample << 32) ^ (uint64_t)seed ^ 0x9E3779B97F4A7C15ull;
}

// --------------------------------------------------------------- frame ----
// Same construction as the reference Frame (utils.h:183-210): to_local maps
// (s, n, t) to (x, y, z), so cos theta = y.
struct Frame {
    V3 s, t, n;
    explicit Frame(V3 nn) : n(nn) {
        if (std::fabs(n.x) > std::fabs(n.y)) {
            float inv = 1.0f / std::sqrt(n.x * n.x + n.z * n.z);
            t = v3(n.z * inv, 0.0f, -n.x * inv);
        } else {
            float inv = 1.0f / std::sqrt(n.y * n.y + n.z * n.z);
            t = v3(0.0f, n.z * inv, -n.y * inv);
        }
        s = cross(t, n);
    }
    V3 to_local(V3 v) const { return v3(dot(v, s), dot(v, n), dot(v, t)); }
    V3 to_world(V3 v) const { return s * v.x + n * v.y + t * v.z; }
};

// ---------------------------------------------------------- microfacet ----
// erfinv after M. Giles, "Approximating the erfinv function" (GPU Computing
// Gems Jade, 2011), single-precision polynomial, refined by two Newton steps
// on erf in double.
double erfinv_d(double x) {
    if (x <= -1.0) return -INFINITY;
    if (x >= 1.0) return INFINITY;
    double w = -std::log((1.0 - x) * (1.0 + x)), p;
    if (w < 5.0) {
        w = w - 2.5;
        p = 2.81022636e-08;
        p = 3.43273939e-07 + p * w;
        p = -3.5233877e-06 + p * w;
        p = -4.39150654e-06 + p * w;
        p = 0.00021858087 + p * w;
        p = -0.00125372503 + p * w;
        p = -0.00417768164 + p * w;
        p = 0.246640727 + p * w;
        p = 1.50140941 + p * w;
    } else {
        w = std::sqrt(w) - 3.0;
        p = -0.000200214257;
        p = 0.000100950558 + p * w;
        p = 0.00134934322 + p * w;
        p = -0.00367342844 + p * w;
        p = 0.00573950773 + p * w;
        p = -0.0076224613 + p * w;
        p = 0.00943887047 + p * w;
        p = 1.00167406 + p * w;
        p = 2.83297682 + p * w;
    }
    double y = p * x;
    for (int i = 0; i < 2; ++i) {
        double err = std::erf(y) - x;
        y -= err / (2.0 / std::sqrt(kPiD) * std::exp(-y * y));
    }
    return y;
}

// Microfacet distribution with roughness ax along the frame's x axis (the
// tangent) and ay along z (the bitangent), y-up: the stretched form of Heitz
// 2014 eq. 87 as microfacet.h:60-76,267-286 uses it (stretched_roughness:
// m.x with xalpha, m.z with yalpha); ax == ay takes the isotropic form.
double mf_D(int dist, double ax, double ay, V3 m) {
    double c = m.y;
    if (c <= 0.0) return 0.0;
    double c2 = c * c;
    double e;   // tan^2(theta) * stretched roughness
    if (ax == ay) {
        e = ((1.0 - c2) / c2) / (ax * ax);
    } else {
        e = ((double)m.x * m.x / (ax * ax) + (double)m.z * m.z / (ay * ay)) / c2;
    }
    if (dist == 0) {
        double k = 1.0 + e;
        return 1.0 / (kPiD * ax * ay * c2 * c2 * k * k);
    }
    return std::exp(-e) / (kPiD * ax * ay * c2 * c2);
}

// Smith Lambda with the projected roughness of Heitz 2014 eq. 80
// (microfacet.h:78-93,243-257); exact forms (GGX: microfacet.h:101-104;
// Beckmann: the erf form that microfacet.h:134-141 approximates).
double mf_lambda(int dist, double ax, double ay, V3 w) {
    double c = w.y;
    double c2 = c * c;
    double s2 = std::max(0.0, 1.0 - c2);
    if (s2 == 0.0 || c == 0.0) return 0.0;
    double at;   // projected alpha * tan(theta)
    if (ax == ay) {
        at = ax * (std::sqrt(s2) / std::fabs(c));
    } else {
        at = std::sqrt((double)w.x * w.x * ax * ax + (double)w.z * w.z * ay * ay) / std::fabs(c);
    }
    if (dist == 0) return 0.5 * (-1.0 + std::sqrt(1.0 + at * at));
    double a = 1.0 / at;
    return 0.5 * (std::erf(a) - 1.0) + std::exp(-a * a) / (2.0 * a * std::sqrt(kPiD));
}

double mf_G1(int dist, double ax, double ay, V3 w) { return 1.0 / (1.0 + mf_lambda(dist, ax, ay, w)); }
double mf_G2(int dist, double ax, double ay, V3 wo, V3 wi) {
    return 1.0 / (1.0 + mf_lambda(dist, ax, ay, wo) + mf_lambda(dist, ax, ay, wi));
}

// pdf of m under the distribution of visible normals seen from wo (Heitz
// 2014, eq. 2); wo may lie in either hemisphere.
double mf_pdf_vndf(int dist, double ax, double ay, V3 wo, V3 m) {
    double c = std::fabs((double)wo.y);
    if (c == 0.0) return 0.0;
    double wm = std::fabs((double)dot(wo, m));
    return mf_G1(dist, ax, ay, wo) * wm * mf_D(dist, ax, ay, m) / c;
}

// Visible-slope sampling for the Beckmann distribution at alpha = 1:
// W. Jakob, "An Improved Visible Normal Sampling Routine for the Beckmann
// Distribution" (2014).  Returns the slopes in a frame whose incident
// direction has azimuth 0.
void beckmann_sample11(double cos_t, double u1, double u2, double& sx, double& sy) {
    if (cos_t > 0.9999) {
        double r = std::sqrt(-std::log(1.0 - u1));
        double phi = 2.0 * kPiD * u2;
        sx = r * std::cos(phi);
        sy = r * std::sin(phi);
        return;
    }
    double sin_t = std::sqrt(std::max(0.0, 1.0 - cos_t * cos_t));
    double tan_t = sin_t / cos_t, cot_t = 1.0 / tan_t;
    double a = -1.0, c = std::erf(cot_t);
    double sample_x = std::max(u1, 1e-6);
    double theta = std::acos(cos_t);
    double fit = 1.0 + theta * (-0.876 + theta * (0.4265 - 0.0594 * theta));
    double b = c - (1.0 + c) * std::pow(1.0 - sample_x, fit);
    const double inv_sqrt_pi = 1.0 / std::sqrt(kPiD);
    double norm = 1.0 / (1.0 + c + inv_sqrt_pi * tan_t * std::exp(-cot_t * cot_t));
    for (int it = 0; it < 20; ++it) {
        if (!(b >= a && b <= c)) b = 0.5 * (a + c);
        double ie = erfinv_d(b);
        double value = norm * (1.0 + b + inv_sqrt_pi * tan_t * std::exp(-ie * ie)) - sample_x;
        double deriv = norm * (1.0 - ie * tan_t);
        if (std::fabs(value) < 1e-7) break;
        if (value > 0) c = b; else a = b;
        b -= value / deriv;
    }
    sx = erfinv_d(b);
    sy = erfinv_d(2.0 * std::max(u2, 1e-6) - 1.0);
}

// Sample a visible normal seen from wo (upper-hemisphere normal returned);
// anisotropic by stretching x with ax and z with ay.
V3 mf_sample_vndf(int dist, float ax, float ay, V3 wo, float u1, float u2) {
    V3 w = wo;
    if (w.y < 0.0f) w = -w;
    if (dist == 0) {
        // GGX: E. Heitz, "Sampling the GGX Distribution of Visible Normals",
        // JCGT 7(4), 2018, written for the y-up frame.
        V3 vh = normalize(v3(ax * w.x, w.y, ay * w.z));
        float lensq = vh.x * vh.x + vh.z * vh.z;
        V3 t1 = lensq > 0.0f ? v3(vh.z, 0.0f, -vh.x) * (1.0f / std::sqrt(lensq)) : v3(1.0f, 0.0f, 0.0f);
        V3 t2 = cross(vh, t1);
        float r = std::sqrt(u1);
        float phi = 2.0f * kPi * u2;
        float p1 = r * std::cos(phi);
        float p2 = r * std::sin(phi);
        float s = 0.5f * (1.0f + vh.y);
        p2 = (1.0f - s) * std::sqrt(std::max(0.0f, 1.0f - p1 * p1)) + s * p2;
        V3 nh = t1 * p1 + t2 * p2 + vh * std::sqrt(std::max(0.0f, 1.0f - p1 * p1 - p2 * p2));
        return normalize(v3(ax * nh.x, std::max(1e-7f, nh.y), ay * nh.z));
    }
    // Beckmann: stretch, sample slopes, rotate, unstretch (Heitz & d'Eon,
    // "Importance Sampling Microfacet-Based BSDFs using the Distribution of
    // Visible Normals", EGSR 2014, sec. 4).
    V3 ws = normalize(v3(ax * w.x, w.y, ay * w.z));
    double cos_t = ws.y;
    double phi = std::atan2((double)ws.z, (double)ws.x);
    if (ws.x == 0.0f && ws.z == 0.0f) phi = 0.0;
    double sx, sy;
    beckmann_sample11(cos_t, u1, u2, sx, sy);
    double cp = std::cos(phi), sp = std::sin(phi);
    double rx = cp * sx - sp * sy;
    double rz = sp * sx + cp * sy;
    rx *= ax;
    rz *= ay;
    return normalize(v3((float)-rx, 1.0f, (float)-rz));
}

// Unpolarised dielectric Fresnel reflectance; eta = n_t / n_i for a ray
// arriving from the side wo lies on (cos_i > 0 = outside).
double fresnel_dielectric(double cos_i, double eta) {
    if (cos_i < 0.0) {
        eta = 1.0 / eta;
        cos_i = -cos_i;
    }
    double sin2_t = (1.0 - cos_i * cos_i) / (eta * eta);
    if (sin2_t >= 1.0) return 1.0;
    double cos_t = std::sqrt(std::max(0.0, 1.0 - sin2_t));
    double rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    double rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5 * (rp * rp + rs * rs);
}

V3 reflect(V3 wo, V3 m) { return 2.0f * dot(wo, m) * m - wo; }

// Refract wo through the microfacet m (m on the same side as wo); eta_rel =
// n_t/n_i for that side.  Returns false on total internal reflection.
bool refract(V3 wo, V3 m, float eta_rel, V3& wi) {
    float c = dot(wo, m);
    float inv = 1.0f / eta_rel;
    float sin2_i = std::max(0.0f, 1.0f - c * c);
    float sin2_t = sin2_i * inv * inv;
    if (sin2_t >= 1.0f) return false;
    float cos_t = std::sqrt(1.0f - sin2_t);
    wi = (-wo) * inv + m * (c * inv - cos_t);
    return true;
}

// One closure lobe: type (0 Lambert, 1 Schlick conductor, 2 rough dielectric,
// 3 KpGloss), distribution, closure weight w, colour c (Lambert albedo or
// conductor F0), roughness (ax, ay) and eta.
struct Lobe {
    int type, dist;
    V3 w, c;
    float ax, ay, eta;
};

// A material: up to 8 weighted lobes (CompositeClosure, bsdf.h:132-196) and
// the emitted radiance; `dielectric` = a lobe transmits (the shading normal is
// then not turned toward the viewer); guide_alpha = the smallest microfacet
// roughness (guiding eligibility).
struct Mat {
    std::vector<Lobe> lobes;
    V3 Le;
    bool dielectric = false;
    float guide_alpha = INFINITY;
};

Lobe load_lobe(const or_material& m, V3 w) {
    Lobe l;
    l.type = m.type;
    l.dist = m.dist;
    l.w = w;
    l.c = m.type == 1 ? load3(m.F0) : load3(m.albedo);
    l.ax = m.alpha;
    l.ay = m.alpha_y > 0.0f ? m.alpha_y : m.alpha;
    l.eta = m.eta;
    return l;
}

void finish_mat(Mat& r) {
    r.dielectric = false;
    r.guide_alpha = INFINITY;
    for (const Lobe& l : r.lobes) {
        if (l.type == 2) r.dielectric = true;
        if (l.type != 0) r.guide_alpha = std::min(r.guide_alpha, std::min(l.ax, l.ay));
    }
}

Mat load_mat(const or_material& m) {
    Mat r;
    r.lobes.push_back(load_lobe(m, v3(1.0f, 1.0f, 1.0f)));
    r.Le = load3(m.Le);
    finish_mat(r);
    return r;
}

Mat load_composite(const or_lobe* lobes, uint32_t n, const float* Le) {
    Mat r;
    for (uint32_t i = 0; i < n; ++i) r.lobes.push_back(load_lobe(lobes[i].lobe, load3(lobes[i].weight)));
    r.Le = Le ? load3(Le) : v3(0, 0, 0);
    finish_mat(r);
    return r;
}

V3 schlick(V3 F0, float c) {
    float k = 1.0f - std::min(1.0f, std::fabs(c));
    float k5 = k * k * k * k * k;
    return v3(F0.x + (1.0f - F0.x) * k5, F0.y + (1.0f - F0.y) * k5, F0.z + (1.0f - F0.z) * k5);
}

// f (no cosine) and solid-angle pdf of wi given wo for one lobe, in the
// local frame, without the lobe weight.
void lobe_eval(const Lobe& m, V3 wo, V3 wi, V3& f, float& pdf) {
    f = v3(0, 0, 0);
    pdf = 0.0f;
    if (m.type == 0) {
        if (wo.y <= 0.0f || wi.y <= 0.0f) return;
        f = m.c * (1.0f / kPi);
        pdf = wi.y / kPi;
        return;
    }
    if (m.type == 1 || m.type == 3) {
        // microfacet reflection: Schlick conductor (bsdfs.cpp:688-730) or
        // KpGloss with dielectric Fresnel F(eta, m.wo) (bsdfs.cpp:359-395)
        if (wo.y <= 0.0f || wi.y <= 0.0f) return;
        V3 h = normalize(wo + wi);
        double D = mf_D(m.dist, m.ax, m.ay, h);
        double G = mf_G2(m.dist, m.ax, m.ay, wo, wi);
        float k = (float)(D * G / (4.0 * (double)wo.y * (double)wi.y));
        if (m.type == 1) {
            f = schlick(m.c, dot(wo, h)) * k;
        } else {
            float F = (float)fresnel_dielectric(dot(wo, h), m.eta);
            f = v3(F * k, F * k, F * k);
        }
        pdf = (float)(mf_G1(m.dist, m.ax, m.ay, wo) * D / (4.0 * (double)wo.y));
        return;
    }
    // rough dielectric, Walter et al. 2007 in the form of pbrt-v4's DielectricBxDF
    float co = wo.y, ci = wi.y;
    if (co == 0.0f || ci == 0.0f) return;
    bool refl = co * ci > 0.0f;
    float etap = 1.0f;
    if (!refl) etap = co > 0.0f ? m.eta : 1.0f / m.eta;
    V3 wm = wi * etap + wo;
    if (dot(wm, wm) == 0.0f) return;
    wm = normalize(wm);
    if (wm.y < 0.0f) wm = -wm;
    if (dot(wm, wi) * ci < 0.0f || dot(wm, wo) * co < 0.0f) return;   // back-facing microfacet
    double F = fresnel_dielectric(dot(wo, wm), m.eta);
    double D = mf_D(m.dist, m.ax, m.ay, wm);
    double G = mf_G2(m.dist, m.ax, m.ay, wo, wi);
    double pvn = mf_pdf_vndf(m.dist, m.ax, m.ay, wo, wm);
    if (refl) {
        double val = D * G * F / std::fabs(4.0 * (double)ci * (double)co);
        f = v3((float)val, (float)val, (float)val);
        pdf = (float)(pvn / (4.0 * std::fabs((double)dot(wo, wm))) * F);
    } else {
        double dwi = dot(wi, wm), dwo = dot(wo, wm);
        double denom = (dwi + dwo / etap);
        denom = denom * denom;
        double val = D * (1.0 - F) * G * std::fabs(dwi * dwo / (denom * (double)ci * (double)co));
        val /= (double)etap * (double)etap;   // radiance transport (D20)
        f = v3((float)val, (float)val, (float)val);
        pdf = (float)(pvn * std::fabs(dwi) / denom * (1.0 - F));
    }
}

// Samples wi for one lobe; u5 picks reflection/refraction, (u6, u7) the
// direction.  f and pdf are the lobe's own (no weight).
bool lobe_sample(const Lobe& m, V3 wo, float u5, float u6, float u7, V3& wi, V3& f, float& pdf) {
    f = v3(0, 0, 0);
    pdf = 0.0f;
    if (m.type == 0) {
        if (wo.y <= 0.0f) return false;
        float r = std::sqrt(u7);
        float phi = 2.0f * kPi * u6;
        wi = v3(r * std::cos(phi), std::sqrt(std::max(0.0f, 1.0f - u7)), r * std::sin(phi));
        if (wi.y <= 0.0f) return false;
        f = m.c * (1.0f / kPi);
        pdf = wi.y / kPi;
        return true;
    }
    if (m.type == 1 || m.type == 3) {
        if (wo.y <= 0.0f) return false;
        V3 mm = mf_sample_vndf(m.dist, m.ax, m.ay, wo, u6, u7);
        wi = reflect(wo, mm);
        if (wi.y <= 0.0f) return false;
        lobe_eval(m, wo, wi, f, pdf);
        return pdf > 0.0f;
    }
    if (wo.y == 0.0f) return false;
    V3 mm = mf_sample_vndf(m.dist, m.ax, m.ay, wo, u6, u7);
    double F = fresnel_dielectric(dot(wo, mm), m.eta);
    if (u5 < F) {
        wi = reflect(wo, mm);
        if (wo.y * wi.y <= 0.0f) return false;
    } else {
        float etap = wo.y > 0.0f ? m.eta : 1.0f / m.eta;
        V3 mside = wo.y > 0.0f ? mm : -mm;
        if (!refract(wo, mside, etap, wi)) return false;
        if (wo.y * wi.y >= 0.0f || wi.y == 0.0f) return false;
    }
    lobe_eval(m, wo, wi, f, pdf);
    return pdf > 0.0f;
}

// Lobe selection probabilities at a vertex with throughput beta:
// p_i = dot(w_i, beta) / sum_j dot(w_j, beta) (CompositeClosure::compute_pdfs,
// bsdf.h:157-176, with the reference's integrator passing the throughput,
// integrator.cpp:425); uniform if every product is 0.  One lobe: p = 1.
struct LobeProbs {
    float p[or_max_lobes];
};

LobeProbs lobe_probs(const Mat& m, V3 beta) {
    LobeProbs r;
    const uint32_t n = (uint32_t)m.lobes.size();
    if (n == 1) {
        r.p[0] = 1.0f;
        return r;
    }
    float sum = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        r.p[i] = std::max(0.0f, dot(m.lobes[i].w, beta));
        sum += r.p[i];
    }
    if (!(sum > 0.0f) || !std::isfinite(sum)) {
        for (uint32_t i = 0; i < n; ++i) r.p[i] = 1.0f / (float)n;
    } else {
        const float inv = 1.0f / sum;
        for (uint32_t i = 0; i < n; ++i) r.p[i] *= inv;
    }
    return r;
}

// Whole material: f = sum_i w_i f_i, pdf = sum_i p_i pdf_i (the one-sample
// MIS combination of the lobes, bsdf.cpp:17-74; deviation D16: the lobe is
// picked with p_i, not uniformly).
void bsdf_eval(const Mat& m, const LobeProbs& lp, V3 wo, V3 wi, V3& f, float& pdf) {
    const uint32_t n = (uint32_t)m.lobes.size();
    if (n == 1) {
        V3 fl;
        lobe_eval(m.lobes[0], wo, wi, fl, pdf);
        f = mul(m.lobes[0].w, fl);
        return;
    }
    f = v3(0, 0, 0);
    pdf = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        V3 fl;
        float pl;
        lobe_eval(m.lobes[i], wo, wi, fl, pl);
        f = f + mul(m.lobes[i].w, fl);
        pdf = pdf + lp.p[i] * pl;
    }
}

// Picks lobe i with probability p_i by u5 and re-uses u5 rescaled to
// [0, 1) inside the lobe, samples it, and evaluates the whole material at
// the sampled direction.
bool bsdf_sample(const Mat& m, const LobeProbs& lp, V3 wo, float u5, float u6, float u7, V3& wi, V3& f,
                 float& pdf) {
    const uint32_t n = (uint32_t)m.lobes.size();
    if (n == 1) {
        V3 fl;
        const bool ok = lobe_sample(m.lobes[0], wo, u5, u6, u7, wi, fl, pdf);
        f = mul(m.lobes[0].w, fl);
        return ok;
    }
    f = v3(0, 0, 0);
    pdf = 0.0f;
    uint32_t pick = n;
    float lo = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        if (lp.p[i] > 0.0f && (u5 < lo + lp.p[i] || i == n - 1)) {
            pick = i;
            break;
        }
        lo += lp.p[i];
    }
    if (pick == n) return false;
    const float ur = std::min((u5 - lo) / lp.p[pick], 0.99999994f);
    V3 fl;
    float pl;
    if (!lobe_sample(m.lobes[pick], wo, std::max(ur, 0.0f), u6, u7, wi, fl, pl)) return false;
    bsdf_eval(m, lp, wo, wi, f, pdf);
    return pdf > 0.0f;
}

// --------------------------------------------------------- ray offset ----
// Waechter & Binder, "A Fast and Robust Method for Avoiding
// Self-Intersection", Ray Tracing Gems ch. 6 (constants as in
// intersection.h:78-80).
inline float int_as_float(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }
inline int32_t float_as_int(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }
V3 offset_ray(V3 p, V3 n) {
    const float origin = 1.0f / 32.0f, fscale = 1.0f / 65536.0f, iscale = 256.0f;
    int32_t ox = (int32_t)(iscale * n.x), oy = (int32_t)(iscale * n.y), oz = (int32_t)(iscale * n.z);
    V3 pi = v3(int_as_float(float_as_int(p.x) + (p.x < 0 ? -ox : ox)),
               int_as_float(float_as_int(p.y) + (p.y < 0 ? -oy : oy)),
               int_as_float(float_as_int(p.z) + (p.z < 0 ? -oz : oz)));
    return v3(std::fabs(p.x) < origin ? p.x + fscale * n.x : pi.x,
              std::fabs(p.y) < origin ? p.y + fscale * n.y : pi.y,
              std::fabs(p.z) < origin ? p.z + fscale * n.z : pi.z);
}

}  // namespace

// ----------------------------------------------------------------- scene --
struct or_scene {
    std::vector<float> tris;
    std::vector<float> norms;        // 9 per triangle (vertex normals), empty = none
    std::vector<uint8_t> has_norms;  // per triangle
    std::vector<int32_t> tri_mat, tri_light;
    std::vector<Mat> mats;
    struct Light {
        std::vector<uint32_t> tris;
        std::vector<float> cdf;   // size n+1, cdf[0] = 0, cdf[n] = 1
        float area = 0.0f;
        V3 Le{0, 0, 0};
    };
    std::vector<Light> lights;
    or_bvh* bvh = nullptr;
};

extern "C" or_bvh* or_bvh_build(const float* tris, uint32_t n_tris);
extern "C" void or_bvh_free(or_bvh* b);

namespace {

float tri_area(const float* t) {
    V3 a = load3(t), b = load3(t + 3), c = load3(t + 6);
    return 0.5f * length(cross(b - a, c - a));
}

// DiscrectPDF::sample (dpdf.h:32-37): first cdf entry >= u, minus one.
uint32_t cdf_sample(const std::vector<float>& cdf, float u) {
    size_t pos = (size_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    size_t idx = pos > 0 ? pos - 1 : 0;
    return (uint32_t)std::min(idx, cdf.size() - 2);
}

struct Camera {
    V3 eye, fwd, right, upv;
    float tan_half, aspect;
    uint32_t w, h;
};

Camera make_camera(const or_render_desc& d) {
    Camera c;
    c.eye = load3(d.eye);
    c.fwd = normalize(load3(d.lookat) - c.eye);
    c.right = normalize(cross(c.fwd, load3(d.up)));
    c.upv = cross(c.right, c.fwd);
    c.tan_half = d.tan_half_fov;
    c.w = d.width;
    c.h = d.height;
    c.aspect = (float)d.width / (float)d.height;
    return c;
}

V3 camera_dir(const Camera& c, float px, float py) {
    float sx = px / (float)c.w, sy = py / (float)c.h;
    float nx = (2.0f * sx - 1.0f) * c.tan_half * c.aspect;
    float ny = (1.0f - 2.0f * sy) * c.tan_half;
    V3 d = c.fwd + c.right * nx + c.upv * ny;
    return normalize(d);
}

struct Hit {
    float t, u, v;
    uint32_t prim;
};

Hit trace(const or_scene& s, V3 o, V3 d, float tmax) {
    float ray[8] = {o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, tmax};
    float h[4];
    or_bvh_closest(s.bvh, ray, 1, h, 1);
    Hit r;
    r.t = h[0];
    std::memcpy(&r.prim, &h[1], 4);
    r.u = h[2];
    r.v = h[3];
    return r;
}

// Guided sampling set-up (see trace_path).
struct Guide {
    const float* params;
    V3 bmin, binv;
    float frac, min_alpha;
    bool sample;   // sample from the network (else BSDF only, 12 draws per vertex)
    uint32_t max_depth;   // guided vertices: depth <= max_depth (0: any depth)
};

V3 trace_path(const or_scene& s, const or_render_desc& d, Pcg32& rng, V3 o, V3 dir,
              const Guide* g = nullptr);

// TriangleMesh::post_hit (shape.cpp:615-654) with vertex normals nv (9
// floats): the interpolated shading normal ns, and the shading point P moved
// off the flat triangle by the method of Hanika, "Hacking the Shadow
// Terminator" (2021): each vertex's offset to P, projected onto the tangent
// plane of its normal where P lies below it, blended barycentrically.  World
// space throughout (the reference dots object-space normals with object-space
// offsets, identical for identity transforms: D21).
void shading_point(V3 v0, V3 v1, V3 v2, const float* nv, float u, float v, V3& P, V3& ns) {
    const float b0 = 1.0f - u - v;
    V3 n0 = load3(nv), n1 = load3(nv + 3), n2 = load3(nv + 6);
    ns = normalize(n0 * b0 + n1 * u + n2 * v);
    V3 tu = P - v0, tw = P - v1, tx = P - v2;
    tu = tu - n0 * std::min(0.0f, dot(tu, n0));
    tw = tw - n1 * std::min(0.0f, dot(tw, n1));
    tx = tx - n2 * std::min(0.0f, dot(tx, n2));
    P = P + (tu * b0 + tw * u + tx * v);
}

V3 path_radiance(const or_scene& s, const or_render_desc& d, const Camera& cam, uint32_t px,
                 uint32_t py, uint32_t sample, const Guide* g = nullptr) {
    Pcg32 rng;
    rng.seed(path_initstate(d.seed, sample), (uint64_t)py * d.width + px);
    float ux = rng.nextf(), uy = rng.nextf();
    V3 o = cam.eye;
    V3 dir = camera_dir(cam, (float)px + ux, (float)py + uy);
    return trace_path(s, d, rng, o, dir, g);
}

// PathIntegrator::Li restated: radiance along the ray (o, dir).
//
// With a guide (no reference counterpart) every vertex draws 12 numbers: the
// 8 of the plain protocol, then u8 (strategy), u9..u11 (network lobe and
// direction).  At a guided vertex (Lambert, or conductor with alpha >=
// min_alpha) the direction comes from the BSDF (u5..u7) if u8 < frac, else
// from the network's mixture (u9..u11); either way the sample is weighted by
// the one-sample MIS density frac * p_bsdf + (1 - frac) * q, and that density
// also stands in for p_bsdf in the next-event and emission power heuristics.
V3 trace_path(const or_scene& s, const or_render_desc& d, Pcg32& rng, V3 o, V3 dir, const Guide* g) {
    V3 beta = v3(1, 1, 1), L = v3(0, 0, 0);
    float prev_pdf = 0.0f;
    const uint32_t nl = (uint32_t)s.lights.size();
    const float eta_rr = 0.95f;
    for (uint32_t depth = 1;; ++depth) {
        Hit h = trace(s, o, dir, INFINITY);
        if (h.prim == 0xFFFFFFFFu) break;
        const float* tv = &s.tris[9 * (size_t)h.prim];
        V3 v0 = load3(tv), v1 = load3(tv + 3), v2 = load3(tv + 6);
        V3 ng = normalize(cross(v1 - v0, v2 - v0));
        float b0 = 1.0f - h.u - h.v;
        V3 P = v0 * b0 + v1 * h.u + v2 * h.v;
        V3 ns = ng;
        if (!s.has_norms.empty() && s.has_norms[h.prim])
            shading_point(v0, v1, v2, &s.norms[9 * (size_t)h.prim], h.u, h.v, P, ns);
        const Mat& m = s.mats[s.tri_mat[h.prim]];
        float cos_hit = dot(dir, ng);
        if (!is_zero(m.Le) && cos_hit < 0.0f) {
            float w = 1.0f;
            if (depth > 1) {
                int lid = s.tri_light[h.prim];
                if (lid >= 0) {
                    const auto& lt = s.lights[lid];
                    float pl = (1.0f / (float)nl) * (1.0f / lt.area) * (h.t * h.t) / (-cos_hit);
                    w = prev_pdf * prev_pdf / (prev_pdf * prev_pdf + pl * pl);
                }
            }
            L = L + mul(beta, m.Le) * w;
        }
        if (depth > d.max_depth) break;
        float u[12];
        const int draws = g ? 12 : 8;
        for (int i = 0; i < draws; ++i) u[i] = rng.nextf();
        V3 wo_w = -dir;
        // two-sided closures see the side the path arrives on; a transmitting
        // material keeps the normals' orientation (inside / outside)
        const V3 n = (!m.dielectric && dot(wo_w, ng) < 0.0f) ? -ns : ns;
        Frame fr(n);
        V3 wo = fr.to_local(wo_w);
        const LobeProbs lp = lobe_probs(m, beta);
        const bool guided = g && g->sample && !m.dielectric && m.guide_alpha >= g->min_alpha &&
                            (g->max_depth == 0 || depth <= g->max_depth);
        double gout[32];
        if (guided) {
            const float in6[6] = {(P.x - g->bmin.x) * g->binv.x, (P.y - g->bmin.y) * g->binv.y,
                                  (P.z - g->bmin.z) * g->binv.z, n.x, n.y, n.z};
            or_mlp_forward(g->params, 1, in6, 1, gout);
        }
        auto guide_pdf = [&](V3 w) {
            const float wd[3] = {w.x, w.y, w.z};
            return (float)or_vmf_pdf(gout, wd);
        };
        // next-event estimation (GeometryLight::sample, light.cpp:52-77)
        if (nl > 0) {
            uint32_t li = std::min((uint32_t)(u[0] * (float)nl), nl - 1);
            const auto& lt = s.lights[li];
            uint32_t lt_i = lt.tris[cdf_sample(lt.cdf, u[1])];
            const float* lv = &s.tris[9 * (size_t)lt_i];
            V3 a = load3(lv), b = load3(lv + 3), c = load3(lv + 6);
            float su0 = std::sqrt(u[2]);
            float bu = 1.0f - su0, bv = u[3] * su0;
            V3 pl = a * bu + b * bv + c * (1.0f - bu - bv);
            V3 nl_w = normalize(cross(b - a, c - a));
            V3 org = offset_ray(P, dot(pl - P, ng) > 0.0f ? ng : -ng);
            V3 dl = pl - org;
            float dist = length(dl);
            if (dist > 0.0f) {
                V3 wi_w = dl * (1.0f / dist);
                float cos_l = -dot(wi_w, nl_w);
                if (cos_l > 0.0f) {
                    V3 wi = fr.to_local(wi_w);
                    V3 f;
                    float pdf_b;
                    bsdf_eval(m, lp, wo, wi, f, pdf_b);
                    if (!is_zero(f)) {
                        float pdf_l = (1.0f / (float)nl) * (1.0f / lt.area) * (dist * dist) / cos_l;
                        if (guided) pdf_b = g->frac * pdf_b + (1.0f - g->frac) * guide_pdf(wi_w);
                        float w = pdf_l * pdf_l / (pdf_l * pdf_l + pdf_b * pdf_b);
                        V3 contrib = mul(mul(beta, f), lt.Le) * (std::fabs(wi.y) * w / pdf_l);
                        float sray[8] = {org.x, org.y, org.z, 0.0f, wi_w.x, wi_w.y, wi_w.z, dist * 0.9999f};
                        float sh[4];
                        or_bvh_closest(s.bvh, sray, 1, sh, 1);
                        uint32_t sp;
                        std::memcpy(&sp, &sh[1], 4);
                        if (sp == 0xFFFFFFFFu) L = L + contrib;
                    }
                }
            }
        }
        // Russian roulette (integrator.cpp:491-497)
        if (depth >= d.rr_depth) {
            float q = std::min(maxc(beta) * eta_rr * eta_rr, 0.99f);
            if (!(u[4] < q)) break;
            beta = beta * (1.0f / q);
        }
        V3 wi, f;
        float pdf;
        if (guided && !(u[8] < g->frac)) {
            const float u3[3] = {u[9], u[10], u[11]};
            float wd[3];
            double q;
            or_vmf_sample(gout, u3, wd, &q);
            const V3 ww = v3(wd[0], wd[1], wd[2]);
            wi = fr.to_local(ww);
            float pdf_bsdf;
            bsdf_eval(m, lp, wo, wi, f, pdf_bsdf);
            pdf = g->frac * pdf_bsdf + (1.0f - g->frac) * guide_pdf(ww);
            if (is_zero(f) || !(pdf > 0.0f)) break;
        } else {
            if (!bsdf_sample(m, lp, wo, u[5], u[6], u[7], wi, f, pdf)) break;
            if (guided) pdf = g->frac * pdf + (1.0f - g->frac) * guide_pdf(fr.to_world(wi));
        }
        beta = mul(beta, f) * (std::fabs(wi.y) / pdf);
        if (is_zero(beta)) break;
        prev_pdf = pdf;
        V3 wi_w = fr.to_world(wi);
        o = offset_ray(P, dot(wi_w, ng) > 0.0f ? ng : -ng);
        dir = wi_w;
    }
    return L;
}

}  // namespace

extern "C" {

void or_bsdf_eval(const or_material* m, const float* wo, const float* wi, float* f, float* pdf) {
    Mat mm = load_mat(*m);
    V3 ff;
    bsdf_eval(mm, lobe_probs(mm, v3(1, 1, 1)), load3(wo), load3(wi), ff, *pdf);
    f[0] = ff.x; f[1] = ff.y; f[2] = ff.z;
}

int or_bsdf_sample(const or_material* m, const float* wo, const float* u3, float* wi, float* f,
                   float* pdf) {
    Mat mm = load_mat(*m);
    V3 w, ff;
    bool ok = bsdf_sample(mm, lobe_probs(mm, v3(1, 1, 1)), load3(wo), u3[0], u3[1], u3[2], w, ff, *pdf);
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    f[0] = ff.x; f[1] = ff.y; f[2] = ff.z;
    return ok ? 1 : 0;
}

double or_mf_D(int dist, float ax, float ay, const float* m) { return mf_D(dist, ax, ay, load3(m)); }
double or_mf_lambda(int dist, float ax, float ay, const float* w) { return mf_lambda(dist, ax, ay, load3(w)); }
void or_mf_sample_vndf(int dist, float ax, float ay, const float* wo, const float* u2, float* m) {
    V3 r = mf_sample_vndf(dist, ax, ay, load3(wo), u2[0], u2[1]);
    m[0] = r.x; m[1] = r.y; m[2] = r.z;
}
double or_fresnel_dielectric(double cos_i, double eta) { return fresnel_dielectric(cos_i, eta); }

void or_shading_point(const float* tri9, const float* normals9, float u, float v, float* P, float* ns) {
    const V3 v0 = load3(tri9), v1 = load3(tri9 + 3), v2 = load3(tri9 + 6);
    V3 p = v0 * (1.0f - u - v) + v1 * u + v2 * v, n;
    shading_point(v0, v1, v2, normals9, u, v, p, n);
    P[0] = p.x; P[1] = p.y; P[2] = p.z;
    ns[0] = n.x; ns[1] = n.y; ns[2] = n.z;
}
double or_erfinv(double x) { return erfinv_d(x); }

void or_bsdf_eval_batch(const or_material* m, const float* wo, const float* wi, int n, float* f,
                        float* pdf) {
    or_composite_eval_batch(nullptr, m, 1, nullptr, wo, wi, n, f, pdf);
}

void or_bsdf_sample_batch(const or_material* m, const float* wo, const float* u3, int n, float* wi,
                          float* f, float* pdf, int32_t* ok) {
    or_composite_sample_batch(nullptr, m, 1, nullptr, wo, u3, n, wi, f, pdf, ok);
}

static Mat composite_or_single(const or_lobe* lobes, const or_material* single, uint32_t n) {
    return lobes ? load_composite(lobes, n, nullptr) : load_mat(*single);
}

void or_composite_eval_batch(const or_lobe* lobes, const or_material* single, uint32_t n_lobes,
                             const float* beta, const float* wo, const float* wi, int n, float* f, float* pdf) {
    const Mat mm = composite_or_single(lobes, single, n_lobes);
    const LobeProbs lp = lobe_probs(mm, beta ? load3(beta) : v3(1, 1, 1));
    for (int i = 0; i < n; ++i) {
        V3 ff;
        bsdf_eval(mm, lp, load3(wo + 3 * i), load3(wi + 3 * i), ff, pdf[i]);
        f[3 * i] = ff.x; f[3 * i + 1] = ff.y; f[3 * i + 2] = ff.z;
    }
}

void or_composite_sample_batch(const or_lobe* lobes, const or_material* single, uint32_t n_lobes,
                               const float* beta, const float* wo, const float* u3, int n, float* wi,
                               float* f, float* pdf, int32_t* ok) {
    const Mat mm = composite_or_single(lobes, single, n_lobes);
    const LobeProbs lp = lobe_probs(mm, beta ? load3(beta) : v3(1, 1, 1));
    for (int i = 0; i < n; ++i) {
        V3 w, ff;
        ok[i] = bsdf_sample(mm, lp, load3(wo + 3 * i), u3[3 * i], u3[3 * i + 1], u3[3 * i + 2], w, ff, pdf[i]);
        wi[3 * i] = w.x; wi[3 * i + 1] = w.y; wi[3 * i + 2] = w.z;
        f[3 * i] = ff.x; f[3 * i + 1] = ff.y; f[3 * i + 2] = ff.z;
    }
}

void or_mf_eval_batch(int dist, float ax, float ay, const float* w, int n, double* D, double* lambda) {
    for (int i = 0; i < n; ++i) {
        V3 v = load3(w + 3 * i);
        D[i] = mf_D(dist, ax, ay, v);
        lambda[i] = mf_lambda(dist, ax, ay, v);
    }
}

or_scene* or_scene_create(const float* tris, uint32_t n_tris, const int32_t* tri_material,
                          const int32_t* tri_light, const or_material* materials,
                          uint32_t n_materials) {
    or_scene* s = new or_scene();
    s->tris.assign(tris, tris + 9 * (size_t)n_tris);
    s->tri_mat.assign(tri_material, tri_material + n_tris);
    s->tri_light.assign(tri_light, tri_light + n_tris);
    for (uint32_t i = 0; i < n_materials; ++i) s->mats.push_back(load_mat(materials[i]));
    int32_t nl = -1;
    for (uint32_t i = 0; i < n_tris; ++i) nl = std::max(nl, tri_light[i]);
    s->lights.resize((size_t)(nl + 1));
    for (uint32_t i = 0; i < n_tris; ++i)
        if (tri_light[i] >= 0) s->lights[tri_light[i]].tris.push_back(i);
    for (auto& l : s->lights) {
        // setup_dpdf (shape.cpp:700-709): area CDF, normalised by its sum.
        l.cdf.assign(1, 0.0f);
        for (uint32_t t : l.tris) l.cdf.push_back(l.cdf.back() + tri_area(&s->tris[9 * (size_t)t]));
        float sum = l.cdf.back();
        l.area = sum;
        if (sum > 0.0f) {
            float inv = 1.0f / sum;
            for (size_t i = 1; i < l.cdf.size(); ++i) l.cdf[i] *= inv;
            l.cdf.back() = 1.0f;
        }
        if (!l.tris.empty()) l.Le = s->mats[s->tri_mat[l.tris[0]]].Le;
    }
    s->bvh = or_bvh_build(tris, n_tris);
    return s;
}

int or_scene_set_composite(or_scene* s, uint32_t material, const or_lobe* lobes, uint32_t n_lobes,
                           const float* Le) {
    if (!s || material >= s->mats.size() || !lobes || n_lobes == 0 || n_lobes > or_max_lobes) return 1;
    s->mats[material] = load_composite(lobes, n_lobes, Le);
    for (auto& l : s->lights)
        if (!l.tris.empty() && (uint32_t)s->tri_mat[l.tris[0]] == material) l.Le = s->mats[material].Le;
    return 0;
}

int or_scene_set_normals(or_scene* s, const float* normals9, const uint8_t* has) {
    if (!s || !normals9 || !has) return 1;
    const size_t n = s->tri_mat.size();
    s->norms.assign(normals9, normals9 + 9 * n);
    s->has_norms.assign(has, has + n);
    return 0;
}

void or_scene_free(or_scene* s) {
    if (!s) return;
    or_bvh_free(s->bvh);
    delete s;
}

static Guide make_guide(const or_guide_cfg* c) {
    Guide g;
    g.params = c->params;
    g.bmin = load3(c->bmin);
    g.binv = load3(c->binv);
    g.frac = c->bsdf_fraction;
    g.min_alpha = c->min_alpha;
    g.sample = c->sample != 0;
    g.max_depth = c->max_depth;
    return g;
}

int or_integrate_rays(const or_scene* s, const float* rays, uint32_t n, const or_render_desc* d,
                      float* out, int threads) {
    return or_integrate_rays_guided(s, rays, n, d, nullptr, out, threads);
}

int or_integrate_rays_guided(const or_scene* s, const float* rays, uint32_t n, const or_render_desc* d,
                             const or_guide_cfg* gc, float* out, int threads) {
    if (!s || !d || (!rays && n)) return 1;
    Guide gd{};
    const Guide* g = nullptr;
    if (gc) {
        gd = make_guide(gc);
        g = &gd;
    }
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= n) break;
            Pcg32 rng;
            rng.seed(path_initstate(d->seed, d->spp_offset), i);
            const float* r = rays + 8 * (size_t)i;
            V3 L = trace_path(*s, *d, rng, load3(r), load3(r + 4), g);
            out[3 * (size_t)i] = L.x;
            out[3 * (size_t)i + 1] = L.y;
            out[3 * (size_t)i + 2] = L.z;
        }
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    return 0;
}

int or_render_pixels(const or_scene* s, const or_render_desc* d, const uint32_t* pixels, uint32_t n,
                     double* out, int threads) {
    if (!s || !d || (n && (!pixels || !out)) || d->width == 0 || d->height == 0) return 1;
    Camera cam = make_camera(*d);
    const uint32_t npix = d->width * d->height;
    for (uint32_t i = 0; i < n; ++i)
        if (pixels[i] >= npix) return 1;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= n) break;
            const uint32_t p = pixels[i], px = p % d->width, py = p / d->width;
            double acc[3] = {0, 0, 0};
            for (uint32_t si = 0; si < d->spp; ++si) {
                V3 L = path_radiance(*s, *d, cam, px, py, d->spp_offset + si, nullptr);
                acc[0] += L.x; acc[1] += L.y; acc[2] += L.z;
            }
            for (int k = 0; k < 3; ++k) out[3 * (size_t)i + k] = acc[k] / (double)d->spp;
        }
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    return 0;
}

int or_render(const or_scene* s, const or_render_desc* d, double* out, float* sample0,
              int threads) {
    return or_render_guided(s, d, nullptr, out, sample0, threads);
}

int or_render_guided(const or_scene* s, const or_render_desc* d, const or_guide_cfg* gc, double* out,
                     float* sample0, int threads) {
    if (!s || !d || !out || d->width == 0 || d->height == 0) return 1;
    Guide gd{};
    const Guide* g = nullptr;
    if (gc) {
        gd = make_guide(gc);
        g = &gd;
    }
    Camera cam = make_camera(*d);
    const uint32_t npix = d->width * d->height;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            uint32_t p = next.fetch_add(1);
            if (p >= npix) break;
            uint32_t px = p % d->width, py = p / d->width;
            double acc[3] = {0, 0, 0};
            for (uint32_t si = 0; si < d->spp; ++si) {
                V3 L = path_radiance(*s, *d, cam, px, py, d->spp_offset + si, g);
                if (si == 0 && sample0) {
                    sample0[3 * (size_t)p] = L.x;
                    sample0[3 * (size_t)p + 1] = L.y;
                    sample0[3 * (size_t)p + 2] = L.z;
                }
                acc[0] += L.x; acc[1] += L.y; acc[2] += L.z;
            }
            for (int k = 0; k < 3; ++k) out[3 * (size_t)p + k] = acc[k] / (double)d->spp;
        }
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    return 0;
}

}  // extern "C"
