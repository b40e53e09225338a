// Reference-shaped value types for the C++ adapters: Vec3f/Vec2f/Vec3i,
// Ray, Intersection and the hitable interface.
//
// Mirrors (names and semantics) of kazen_proto:
//   Ray                 src/core/ray.h:10-45 (direction normalised by the
//                       constructor, tmin defaults to epsilon = 1e-4 from
//                       src/base/basic_types.h:9, tmax to FLT_MAX)
//   Intersection        src/core/intersection.h:24-70 (the fields a hit
//                       fills: P, N, shading_normal, tangent, bitangent,
//                       dpdu, dpdv, bary, uv, ray_t, backface, is_light,
//                       light_id, geom_id, prim_id, shader_name)
//   HitableInterface    src/core/hitable.h:9-37 (intersect x2, occluded x2)
#pragma once
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <string>

namespace kazen {

template <typename T>
constexpr T epsilon = static_cast<T>(1e-4);

struct Vec3f {
    float v[3] = {0.0f, 0.0f, 0.0f};
    Vec3f() = default;
    explicit Vec3f(float a) : v{a, a, a} {}
    Vec3f(float x, float y, float z) : v{x, y, z} {}
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    Vec3f operator+(const Vec3f& o) const { return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]}; }
    Vec3f operator-(const Vec3f& o) const { return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]}; }
    Vec3f operator-() const { return {-v[0], -v[1], -v[2]}; }
    Vec3f operator*(float s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
    Vec3f operator*(const Vec3f& o) const { return {v[0] * o.v[0], v[1] * o.v[1], v[2] * o.v[2]}; }
    Vec3f operator/(float s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
    Vec3f& operator+=(const Vec3f& o) { v[0] += o.v[0]; v[1] += o.v[1]; v[2] += o.v[2]; return *this; }
};
inline Vec3f operator*(float s, const Vec3f& a) { return a * s; }
inline float dot(const Vec3f& a, const Vec3f& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline Vec3f cross(const Vec3f& a, const Vec3f& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline float length(const Vec3f& a) { return std::sqrt(dot(a, a)); }
inline float length_squared(const Vec3f& a) { return dot(a, a); }
inline Vec3f normalize(const Vec3f& a) { return a / length(a); }

struct Vec2f {
    float v[2] = {0.0f, 0.0f};
    Vec2f() = default;
    Vec2f(float x, float y) : v{x, y} {}
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
};

struct Vec3i {
    int v[3] = {0, 0, 0};
    Vec3i() = default;
    Vec3i(int a, int b, int c) : v{a, b, c} {}
    int& operator[](int i) { return v[i]; }
    int operator[](int i) const { return v[i]; }
};

using RGBSpectrum = Vec3f;

// Row-major 4x4 (only the upper 3x4 is used: object -> world).
struct Mat4f {
    float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
};

class Ray {
public:
    Ray(Vec3f o, Vec3f d, const float tmin_ = epsilon<float>, const float tmax_ = FLT_MAX,
        const float time_ = 0.0f)
        : origin(o), direction(normalize(d)), tmin(tmin_), tmax(tmax_), time(time_) {
        origin_dx = origin;
        origin_dy = origin;
    }
    Vec3f at(const float t) const { return origin + direction * t; }

    Vec3f origin, origin_dx, origin_dy;
    Vec3f direction, direction_dx, direction_dy;
    float tmin, tmax, time;
};

struct Intersection {
    Vec3f P;                 // hit point; with vertex normals, moved off the flat triangle
    Vec3f N;                 // geometric normal (unit, winding order)
    Vec3f shading_normal;    // interpolated vertex normal, or N
    Vec3f tangent, bitangent;   // world (v1 - v0) direction, cross(tangent, N)
    Vec3f dpdu, dpdv;        // from the mesh uvs, else tangent / bitangent
    Vec3f bary;              // (1-u-v, u, v)
    Vec2f uv;                // interpolated mesh uv, else the barycentric (u, v)
    float ray_t = FLT_MAX;
    bool backface = false;
    bool is_light = false;
    uint32_t light_id = 0;
    size_t geom_id = 0;      // mesh index, in the order meshes were added
    size_t prim_id = 0;      // triangle index within the mesh
    std::string shader_name;
};

class HitableInterface {
public:
    virtual ~HitableInterface() = default;
    virtual bool intersect(const Ray& r, Intersection& isect) const = 0;
    virtual bool intersect(const Ray& r, float& t) const = 0;

    // Same contract as hitable.h:15-23: blocked between p1 and p2.
    bool occluded(const Vec3f& p1, const Vec3f& p2) const {
        const Vec3f dir = normalize(p2 - p1);
        Ray r(p1 + dir * 0.00001f, dir);
        float t;
        return intersect(r, t) && t < length(p2 - p1);
    }
    // hitable.h:25-33 intends "hit something other than the light": restated as
    // any hit whose mesh differs from dest_geom_id.
    bool occluded(const Ray& r, size_t& dest_geom_id, Vec3f& light_normal) const {
        Intersection isect;
        if (intersect(r, isect) && isect.geom_id != dest_geom_id) return true;
        light_normal = isect.N;
        return false;
    }
};

}  // namespace kazen
