// Scalar and vector arithmetic shared by the host builder and the gfx950 kernels.
//
// Every operation is a plain IEEE-754 binary32 operation evaluated in the order the reference
// writes it (BastianBlokland/sharpray src/math.cs, brdf.cs); the library is compiled with
// -ffp-contract=off and correctly rounded division / square root, so each result rounds exactly
// like RyuJIT's scalar SSE code. Min/Max follow .NET 8 MathF (NaN-propagating, -0 < +0).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SR_HD __host__ __device__ __forceinline__

namespace sr {

constexpr float kPi = 3.14159265358979323846f;  // MathF.PI
constexpr float kFloatMax = 3.40282347e+38f;     // float.MaxValue

SR_HD float net_max(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == b) return __builtin_signbit(a) ? b : a;
    return a > b ? a : b;
}
SR_HD float net_min(float a, float b) {
    if (a != a) return a;
    if (b != b) return b;
    if (a == b) return __builtin_signbit(a) ? a : b;
    return a < b ? a : b;
}
SR_HD float net_clamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

struct vec2 { float x, y; };
struct vec3 { float x, y, z; };
struct vec4 { float x, y, z, w; };
struct quat { float x, y, z, w; };

SR_HD vec2 mk2(float x, float y) { vec2 r; r.x = x; r.y = y; return r; }
SR_HD vec3 mk3(float x, float y, float z) { vec3 r; r.x = x; r.y = y; r.z = z; return r; }
SR_HD vec3 splat3(float s) { return mk3(s, s, s); }
SR_HD vec4 mk4(float x, float y, float z, float w) { vec4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }

SR_HD vec2 operator+(vec2 a, vec2 b) { return mk2(a.x + b.x, a.y + b.y); }
SR_HD vec2 operator*(vec2 a, float s) { return mk2(a.x * s, a.y * s); }
SR_HD vec3 operator-(vec3 a) { return mk3(-a.x, -a.y, -a.z); }
SR_HD vec3 operator+(vec3 a, vec3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
SR_HD vec3 operator-(vec3 a, vec3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
SR_HD vec3 operator*(vec3 a, vec3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
SR_HD vec3 operator*(vec3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
SR_HD vec3 operator*(float s, vec3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
SR_HD vec3 operator/(vec3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
SR_HD vec3 operator/(vec3 a, vec3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
SR_HD vec4 operator+(vec4 a, vec4 b) { return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
SR_HD vec4 operator*(vec4 a, float s) { return mk4(a.x * s, a.y * s, a.z * s, a.w * s); }
SR_HD vec3 xyz(vec4 v) { return mk3(v.x, v.y, v.z); }
SR_HD float axis_of(vec3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

SR_HD float dot3(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SR_HD vec3 cross3(vec3 a, vec3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
SR_HD float len_sqr(vec3 v) { return dot3(v, v); }
SR_HD vec3 normalized(vec3 v) { return v / sqrtf(len_sqr(v)); }
SR_HD vec3 normalized_or(vec3 v, vec3 fallback) {
    float m = len_sqr(v);
    return m >= 1e-12f ? v / sqrtf(m) : fallback;
}
SR_HD vec3 up3() { return mk3(0.0f, 1.0f, 0.0f); }
SR_HD vec3 right3() { return mk3(1.0f, 0.0f, 0.0f); }

// Colour helpers (math.cs:114-123, 176-184).
SR_HD float luminance(vec3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
SR_HD float max_comp(vec3 c) { return net_max(c.x, net_max(c.y, c.z)); }
SR_HD vec3 clamp_lum(vec3 c, float mx) {
    float l = luminance(c);
    return l > mx ? c * (mx / l) : c;
}
SR_HD vec3 lerp3(vec3 a, vec3 b, float t) {
    return mk3(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t);
}

// Quat * Vec3 (math.cs:609-617).
SR_HD vec3 rotate(quat q, vec3 v) {
    vec3 axis = mk3(q.x, q.y, q.z);
    float sqrMag = len_sqr(axis);
    vec3 a = axis * (dot3(axis, v) * 2.0f);
    vec3 b = v * (q.w * q.w - sqrMag);
    vec3 c = cross3(axis, v) * (q.w * 2.0f);
    return a + b + c;
}
SR_HD quat conj(quat q) { quat r; r.x = -q.x; r.y = -q.y; r.z = -q.z; r.w = q.w; return r; }

// Ray with the clamped reciprocal direction of math.cs:846-857.
struct ray3 { vec3 o, d, inv; };
SR_HD float safe_inv(float d) {
    const float mc = 1e-8f;
    return 1.0f / (fabsf(d) >= mc ? d : copysignf(mc, d));
}
SR_HD ray3 mk_ray(vec3 o, vec3 d) {
    ray3 r;
    r.o = o;
    r.d = d;
    r.inv = mk3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    return r;
}
SR_HD vec3 at(const ray3& r, float t) { return r.o + r.d * t; }

// Axis-aligned box (math.cs:936-1004).
struct box3 { vec3 mn, mx; };
SR_HD box3 box_empty() { box3 b; b.mn = splat3(kFloatMax); b.mx = splat3(-kFloatMax); return b; }
SR_HD vec3 box_mid(const box3& b) { return (b.mn + b.mx) * 0.5f; }
SR_HD vec3 box_ext(const box3& b) { return b.mx - b.mn; }
SR_HD float box_area(const box3& b) {
    vec3 s = box_ext(b);
    return 2.0f * (s.x * s.y + s.x * s.z + s.y * s.z);
}
SR_HD bool box_is_empty(const box3& b) { return b.mn.x > b.mx.x || b.mn.y > b.mx.y || b.mn.z > b.mx.z; }
SR_HD void box_grow(box3& b, vec3 p) {
    b.mn = mk3(net_min(b.mn.x, p.x), net_min(b.mn.y, p.y), net_min(b.mn.z, p.z));
    b.mx = mk3(net_max(b.mx.x, p.x), net_max(b.mx.y, p.y), net_max(b.mx.z, p.z));
}
SR_HD void box_grow(box3& b, const box3& o) {
    b.mn = mk3(net_min(b.mn.x, o.mn.x), net_min(b.mn.y, o.mn.y), net_min(b.mn.z, o.mn.z));
    b.mx = mk3(net_max(b.mx.x, o.mx.x), net_max(b.mx.y, o.mx.y), net_max(b.mx.z, o.mx.z));
}

}  // namespace sr
