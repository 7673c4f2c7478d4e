// SharpRay CPU oracle: a restatement of the reference's per-pixel path-tracing loop.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP product is compared against;
// it is loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
// nothing else. The product library never links, loads or calls it.
//
// Every function cites the reference C# (BastianBlokland/sharpray, src/*.cs) line range it
// restates. Arithmetic follows the reference expression by expression (same association, same
// IEEE single-precision operations, no FMA contraction: build with -ffp-contract=off), RNG draws
// are sequenced in C#'s left-to-right order, and MathF.Min/Max keep .NET 8's NaN-propagating,
// -0 < +0 semantics. Transcendentals go to the host libm (glibc), which is what .NET's MathF
// calls on Linux; the exact bits of those calls are the one part that no reference test pins
// ("parity unpinned": no reference test or fixture holds their outputs).
//
// Compiled twice: SR_REAL=float (the parity oracle) and SR_REAL=double (the FP64 oracle used to
// judge converged images). In the double build the BVH and all shading run in double precision.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#ifndef SR_REAL
#define SR_REAL float
#endif
typedef SR_REAL real;

namespace {

const real kPi = (real)3.14159265358979323846;
const real kMaxValue = (real)FLT_MAX;
const real kInf = (real)INFINITY;

// ---- .NET 8 MathF semantics --------------------------------------------------------------------
// MathF.Max / MathF.Min are IEEE 754-2019 maximum / minimum: NaN in -> NaN out, +0 > -0.
inline real nmax(real a, real b) {
    if (std::isnan(a)) return a;
    if (std::isnan(b)) return b;
    if (a == b) return std::signbit(a) ? b : a;
    return a > b ? a : b;
}
inline real nmin(real a, real b) {
    if (std::isnan(a)) return a;
    if (std::isnan(b)) return b;
    if (a == b) return std::signbit(a) ? a : b;
    return a < b ? a : b;
}
// MathF.Pow goes to libm's pow; the exponent passes through a volatile so the compiler cannot
// rewrite pow(x, 2) as x * x (which differs from libm pow on exact-midpoint squares).
inline real libm_pow(real x, real y) {
    volatile real vy = y;
    return std::pow(x, (real)vy);
}
// Math.Clamp(float, float, float): NaN passes through.
inline real nclamp(real v, real lo, real hi) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

// ---- Vectors (math.cs:246-560) -----------------------------------------------------------------
struct V2 { real x, y; };
struct V3 { real x, y, z; };
struct V4 { real x, y, z, w; };

inline V2 v2(real x, real y) { return V2{x, y}; }
inline V3 v3(real x, real y, real z) { return V3{x, y, z}; }
inline V3 v3s(real s) { return V3{s, s, s}; }
inline V2 operator+(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
inline V2 operator*(V2 a, real s) { return v2(a.x * s, a.y * s); }
inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(V3 a, real s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(real s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, real s) { return v3(a.x / s, a.y / s, a.z / s); }
inline V3 operator/(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline V4 operator+(V4 a, V4 b) { return V4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline V4 operator*(V4 a, real s) { return V4{a.x * s, a.y * s, a.z * s, a.w * s}; }
inline real comp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

inline real dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline real mag_sqr(V3 v) { return dot(v, v); }
// Vec3.Normalize, math.cs:370-375.
inline V3 normalize(V3 v) { return v / std::sqrt(mag_sqr(v)); }
// Vec3.NormalizeOr, math.cs:377-381.
inline V3 normalize_or(V3 v, V3 fb) {
    real m = mag_sqr(v);
    return m >= 1e-12f ? v / std::sqrt(m) : fb;
}
const V3 kUp = {0, 1, 0};
const V3 kRight = {1, 0, 0};

// Color (math.cs:88-239) shares V3's layout.
typedef V3 Color;
inline real luminance(Color c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
inline real max_component(Color c) { return nmax(c.x, nmax(c.y, c.z)); }
inline Color clamp_luminance(Color c, real mx) {
    real l = luminance(c);
    return l > mx ? c * (mx / l) : c;
}
inline Color lerp(Color a, Color b, real t) {
    return v3(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t);
}
const Color kWhite = {1, 1, 1};
const Color kBlack = {0, 0, 0};

// Vec3.Perp, math.cs:383-388.
inline V3 perp(V3 v) {
    V3 up = std::fabs(v.y) < 0.9f ? kUp : kRight;
    return normalize(cross(up, v));
}
// Vec3.EquirectUv, math.cs:390-398.
inline V2 equirect_uv(V3 d) {
    real u = std::atan2(d.z, d.x) / (kPi * 2.0f) + 0.5f;
    real v = 0.5f - std::asin(nclamp(d.y, -1.0f, 1.0f)) / kPi;
    return v2(u, v);
}
// Vec3.FromEquirectUv, math.cs:482-489.
inline V3 from_equirect_uv(V2 uv) {
    real phi = (uv.x - 0.5f) * 2.0f * kPi;
    real sinTheta = std::sin(kPi * uv.y);
    real y = std::cos(kPi * uv.y);
    return v3(sinTheta * std::cos(phi), y, sinTheta * std::sin(phi));
}
// Vec3.FromTangentSpace, math.cs:492-498.
inline V3 from_tangent_space(V3 dir, V4 tangent, V3 normal) {
    V3 tan = v3(tangent.x, tangent.y, tangent.z);
    V3 bitan = tangent.w * cross(normal, tan);
    return dir.x * tan + dir.y * bitan + dir.z * normal;
}
// Vec3.Reflect, math.cs:437-441.
inline V3 reflect(V3 v, V3 n) { return v - (real)2 * dot(v, n) * n; }

// ---- Quat / Transform (math.cs:564-840) --------------------------------------------------------
struct Quat { real x, y, z, w; };
inline Quat qinv(Quat q) { return Quat{-q.x, -q.y, -q.z, q.w}; }
// Quat * Vec3, math.cs:609-617.
inline V3 qrot(Quat q, V3 v) {
    V3 axis = v3(q.x, q.y, q.z);
    real sqrMag = mag_sqr(axis);
    V3 a = axis * (dot(axis, v) * (real)2);
    V3 b = v * (q.w * q.w - sqrMag);
    V3 c = cross(axis, v) * (q.w * (real)2);
    return a + b + c;
}
struct Transform { V3 pos; Quat rot; V3 scale; };

// ---- RNG (math.cs:1665-1743) -------------------------------------------------------------------
struct Rng {
    uint32_t s0, s1, s2, s3, s4;
    static uint64_t split_mix64(uint64_t& state) {
        uint64_t r = (state += 0x9e3779b97f4a7c15ULL);
        r = (r ^ (r >> 30)) * 0xbf58476d1ce4e5b9ULL;
        r = (r ^ (r >> 27)) * 0x94d049bb133111ebULL;
        return r ^ (r >> 31);
    }
    void seed(uint64_t seed) {
        uint64_t v1 = split_mix64(seed);
        uint64_t v2v = split_mix64(seed);
        s0 = (uint32_t)v1; s1 = (uint32_t)(v1 >> 32);
        s2 = (uint32_t)v2v; s3 = (uint32_t)(v2v >> 32);
        s4 = 0;
    }
    void seed_xy(uint32_t a, uint32_t b) { seed(((uint64_t)a << 32) | b); counter = false; }
    // Counter policy (no reference counterpart): draw k of sample `sample` of pixel `pixel` is word
    // k mod 4 of Philox4x32-10(counter = {sample, k / 4, 0, 0}, key = {seed, pixel}).
    bool counter = false;
    uint32_t key0 = 0, key1 = 0, sample = 0, draw = 0;
    void seed_counter(uint32_t seedv, uint32_t pixel, uint32_t sampleIndex) {
        counter = true; key0 = seedv; key1 = pixel; sample = sampleIndex; draw = 0;
    }
    // Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of two 32x32->64 multiplies
    // with Weyl-incremented keys.
    static void philox(uint32_t ctr[4], uint32_t k0, uint32_t k1) {
        const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
        for (int round = 0; round < 10; ++round) {
            uint64_t a = (uint64_t)M0 * ctr[0];
            uint64_t b = (uint64_t)M1 * ctr[2];
            uint32_t x0 = (uint32_t)(b >> 32) ^ ctr[1] ^ k0;
            uint32_t x1 = (uint32_t)b;
            uint32_t x2 = (uint32_t)(a >> 32) ^ ctr[3] ^ k1;
            uint32_t x3 = (uint32_t)a;
            ctr[0] = x0; ctr[1] = x1; ctr[2] = x2; ctr[3] = x3;
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
    }
    uint32_t next_uint() {
        if (counter) {
            uint32_t c[4] = {sample, draw / 4u, 0u, 0u};
            philox(c, key0, key1);
            return c[draw++ % 4u];
        }
        uint32_t t = s3;
        uint32_t s = s0;
        s3 = s2; s2 = s1; s1 = s;
        t ^= t >> 2;
        t ^= t << 1;
        t ^= s ^ (s << 4);
        s0 = t;
        s4 += 362437u;
        return t + s4;
    }
    float next_float() {
        const float toFloat = 1.0f / (4294967296.0f + 512.0f);
        return (float)next_uint() * toFloat;
    }
    real next_exponential() {
        float u;
        do { u = next_float(); } while (u <= 1e-8f);
        return -std::log((real)u);
    }
    void next_gauss(real& ga, real& gb) {
        float a;
        do { a = next_float(); } while (a <= 1e-8f);
        real b = kPi * 2.0f * (real)next_float();
        real mag = std::sqrt(-2.0f * std::log((real)a));
        ga = mag * std::cos(b);
        gb = mag * std::sin(b);
    }
    // Vec2.Rand, math.cs:283: X is drawn before Y.
    V2 rand2() {
        real x = next_float();
        real y = next_float();
        return v2(x, y);
    }
};

// Vec3.RandOnSphere, math.cs:443-454.
inline V3 rand_on_sphere(Rng& rng) {
    while (true) {
        real x, y, z, unused;
        rng.next_gauss(x, y);
        rng.next_gauss(z, unused);
        real magSqr = x * x + y * y + z * z;
        if (magSqr > 1e-6f) return v3(x, y, z) / std::sqrt(magSqr);
    }
}

// ---- Ray (math.cs:842-862) ---------------------------------------------------------------------
struct Ray { V3 o, d, inv; };
inline Ray make_ray(V3 o, V3 d) {
    const real mc = 1e-8f;
    Ray r;
    r.o = o;
    r.d = d;
    r.inv = v3((real)1 / (std::fabs(d.x) >= mc ? d.x : std::copysign(mc, d.x)),
               (real)1 / (std::fabs(d.y) >= mc ? d.y : std::copysign(mc, d.y)),
               (real)1 / (std::fabs(d.z) >= mc ? d.z : std::copysign(mc, d.z)));
    return r;
}
inline V3 ray_at(const Ray& r, real t) { return r.o + r.d * t; }

// Transform.TransformRayInv, math.cs:816-822 (TransformPointInv / TransformVectorInv, :811-814).
inline Ray transform_ray_inv(const Transform& tr, const Ray& ray, real& scaleOut) {
    Quat inv = qinv(tr.rot);
    V3 origin = qrot(inv, ray.o - tr.pos) / tr.scale;
    V3 dirRaw = qrot(inv, ray.d) / tr.scale;
    real scale = std::sqrt(mag_sqr(dirRaw));
    scaleOut = scale;
    return make_ray(origin, dirRaw / scale);
}

// ---- AABox (math.cs:936-1080) ------------------------------------------------------------------
struct AABox { V3 mn, mx; };
inline AABox box_inverted() { return AABox{v3s(kMaxValue), v3s(-kMaxValue)}; }
inline V3 box_center(const AABox& b) { return (b.mn + b.mx) * 0.5f; }
inline V3 box_size(const AABox& b) { return b.mx - b.mn; }
inline real box_sa(const AABox& b) {
    V3 s = box_size(b);
    return 2.0f * (s.x * s.y + s.x * s.z + s.y * s.z);
}
inline bool box_inverted_p(const AABox& b) { return b.mn.x > b.mx.x || b.mn.y > b.mx.y || b.mn.z > b.mx.z; }
inline void box_encap(AABox& b, V3 p) {
    b.mn = v3(nmin(b.mn.x, p.x), nmin(b.mn.y, p.y), nmin(b.mn.z, p.z));
    b.mx = v3(nmax(b.mx.x, p.x), nmax(b.mx.y, p.y), nmax(b.mx.z, p.z));
}
inline void box_encap(AABox& b, const AABox& o) {
    b.mn = v3(nmin(b.mn.x, o.mn.x), nmin(b.mn.y, o.mn.y), nmin(b.mn.z, o.mn.z));
    b.mx = v3(nmax(b.mx.x, o.mx.x), nmax(b.mx.y, o.mx.y), nmax(b.mx.z, o.mx.z));
}
inline AABox box_from_center(V3 c, V3 size) { return AABox{c - size * 0.5f, c + size * 0.5f}; }
// AABox.IntersectDist, math.cs:983-1004.
inline bool box_intersect_dist(const AABox& b, const Ray& r, real& tOut) {
    real t1 = (b.mn.x - r.o.x) * r.inv.x;
    real t2 = (b.mx.x - r.o.x) * r.inv.x;
    real t3 = (b.mn.y - r.o.y) * r.inv.y;
    real t4 = (b.mx.y - r.o.y) * r.inv.y;
    real t5 = (b.mn.z - r.o.z) * r.inv.z;
    real t6 = (b.mx.z - r.o.z) * r.inv.z;
    real minA = nmin(t1, t2), minB = nmin(t3, t4), minC = nmin(t5, t6);
    real maxA = nmax(t1, t2), maxB = nmax(t3, t4), maxC = nmax(t5, t6);
    real tMin = nmax(nmax(minA, minB), minC);
    real tMax = nmin(nmin(maxA, maxB), maxC);
    if (tMax < 0.0f || tMin > tMax) return false;
    tOut = nmax(0.0f, tMin);
    return true;
}

// ---- Shape hits (math.cs:22-61) ----------------------------------------------------------------
struct ShapeHit { real dist; V3 norm, normGeo; V4 tan; V2 uv; };
struct LeanHit { real dist; V2 uv; };

// ---- TriangleLean (math.cs:1334-1489) ----------------------------------------------------------
struct TriLean { V3 a, ab, ac; };
inline TriLean tri_lean(V3 pa, V3 pb, V3 pc) { return TriLean{pa, pb - pa, pc - pa}; }
inline V3 tri_normal(const TriLean& t) { return normalize_or(cross(t.ab, t.ac), kUp); }
inline AABox tri_bounds(const TriLean& t) {
    V3 b = t.a + t.ab, c = t.a + t.ac;
    return AABox{v3(nmin(t.a.x, nmin(b.x, c.x)), nmin(t.a.y, nmin(b.y, c.y)), nmin(t.a.z, nmin(b.z, c.z))),
                 v3(nmax(t.a.x, nmax(b.x, c.x)), nmax(t.a.y, nmax(b.y, c.y)), nmax(t.a.z, nmax(b.z, c.z)))};
}
// TriangleLean.Intersect (Moller-Trumbore), math.cs:1407-1435.
inline bool tri_intersect(const TriLean& t, const Ray& ray, LeanHit& out) {
    V3 h = cross(ray.d, t.ac);
    real det = dot(t.ab, h);
    if (std::fabs(det) <= 1e-7f) return false;
    real invDet = (real)1 / det;
    const real edgeEps = 1e-7f;
    V3 ao = ray.o - t.a;
    real u = dot(ao, h) * invDet;
    if (u < -edgeEps || u > (1.0f + edgeEps)) return false;
    V3 q = cross(ao, t.ab);
    real v = dot(ray.d, q) * invDet;
    if (v < -edgeEps || (u + v) > (1.0f + edgeEps)) return false;
    real tt = dot(t.ac, q) * invDet;
    if (tt < 0.0f) return false;
    out.dist = tt;
    out.uv = v2(u, v);
    return true;
}

// Triangle (math.cs:1491-1580): lean shape plus per-vertex normals, tangents and uvs.
struct Triangle {
    TriLean lean;
    V3 nA, nB, nC;
    V4 tA, tB, tC;
    V2 uvA, uvB, uvC;
};
static_assert(sizeof(Triangle) == 36 * sizeof(real), "Triangle layout");

// TriangleLean.Inflate, math.cs:1437-1459.
inline ShapeHit tri_inflate(const Triangle& tri, const LeanHit& lh) {
    real u = lh.uv.x, v = lh.uv.y;
    real w = 1.0f - u - v;
    V3 normInterp = tri.nA * w + tri.nB * u + tri.nC * v;
    V3 norm = mag_sqr(normInterp) >= 1e-12f ? normalize(normInterp) : tri_normal(tri.lean);
    V3 normFace = tri_normal(tri.lean);
    V4 tanInterp = tri.tA * w + tri.tB * u + tri.tC * v;
    V3 tx = v3(tanInterp.x, tanInterp.y, tanInterp.z);
    V3 tanXyz = normalize_or(tx - dot(tx, norm) * norm, kRight);
    ShapeHit hit;
    hit.dist = lh.dist;
    hit.norm = norm;
    hit.normGeo = normFace;
    hit.tan = V4{tanXyz.x, tanXyz.y, tanXyz.z, tanInterp.w >= 0.0f ? 1.0f : -1.0f};
    hit.uv = tri.uvA * w + tri.uvB * u + tri.uvC * v;
    return hit;
}

// Triangle.ComputeTangent, math.cs:1553-1579.
inline V4 compute_tangent(V3 pa, V3 pb, V3 pc, V2 ua, V2 ub, V2 uc, V3 norm) {
    V3 pd1 = pb - pa, pd2 = pc - pa;
    V2 ud1 = v2(ub.x - ua.x, ub.y - ua.y), ud2 = v2(uc.x - ua.x, uc.y - ua.y);
    real f = ud1.x * ud2.y - ud2.x * ud1.y;
    if (std::fabs(f) < 1e-8f) {
        V3 tg = normalize_or(pd1 - dot(pd1, norm) * norm, kRight);
        return V4{tg.x, tg.y, tg.z, 1.0f};
    }
    real invF = (real)1 / f;
    V3 tangent = (ud2.y * pd1 - ud1.y * pd2) * invF;
    V3 bitangent = (-ud2.x * pd1 + ud1.x * pd2) * invF;
    tangent = normalize_or(tangent - dot(tangent, norm) * norm, kRight);
    real w = dot(cross(norm, tangent), bitangent) < 0.0f ? -1.0f : 1.0f;
    return V4{tangent.x, tangent.y, tangent.z, w};
}
// Triangle(posA..C, uvA..C), math.cs:1508-1516: face normal at every vertex.
inline Triangle make_triangle_uv(V3 pa, V3 pb, V3 pc, V2 ua, V2 ub, V2 uc) {
    Triangle t;
    t.lean = tri_lean(pa, pb, pc);
    V3 n = tri_normal(t.lean);
    t.nA = t.nB = t.nC = n;
    t.tA = t.tB = t.tC = compute_tangent(pa, pb, pc, ua, ub, uc, n);
    t.uvA = ua; t.uvB = ub; t.uvC = uc;
    return t;
}
// Triangle(posA..C, normA..C, uvA..C), math.cs:1518-1531.
inline Triangle make_triangle_full(V3 pa, V3 pb, V3 pc, V3 na, V3 nb, V3 nc, V2 ua, V2 ub, V2 uc) {
    Triangle t;
    t.lean = tri_lean(pa, pb, pc);
    t.nA = na; t.nB = nb; t.nC = nc;
    t.uvA = ua; t.uvB = ub; t.uvC = uc;
    V3 geoNorm = normalize_or(cross(pb - pa, pc - pa), kUp);
    V3 avgNorm = normalize_or((na + nb + nc) / 3.0f, geoNorm);
    t.tA = t.tB = t.tC = compute_tangent(pa, pb, pc, ua, ub, uc, avgNorm);
    return t;
}

// ---- Counters (counters.cs:24-79, the subset bumped on the render path) ------------------------
enum CounterType {
    C_Pixel = 0, C_Sample, C_SampleBounce, C_SampleHit, C_SampleMiss, C_SampleTerminate,
    C_SampleFogScatter, C_SampleFogEscape, C_SamplePixelConverged, C_SamplePixelMaxed,
    C_SceneTrace, C_SceneOcclude, C_MeshIntersect, C_BvhIntersectNode, C_BvhIntersectShape,
    C_FilterSample, C_DenoiseEarlyOut,
    // Min / max-reduced (counters.cs:224-233): kept as their reference values, (long)(v * 1e4) for
    // the float ones; initial values long.MaxValue (min) and long.MinValue (max).
    C_SampleCountMin, C_SampleCountMax, C_SampleVarianceMax,
    C_DenoiseRejectMissing, C_DenoiseRejectNormal, C_DenoiseRejectDepth, C_DenoiseRejectLum,
    C_DenoiseWeightMax, C_DenoiseMaxLum, C_DenoiseMaxLumBoost, C_Count
};
struct Counters { int64_t v[C_Count]; };
inline bool counter_is_min(int k) { return k == C_SampleCountMin; }
inline bool counter_is_max(int k) {
    return k == C_SampleCountMax || k == C_SampleVarianceMax || k == C_DenoiseWeightMax || k == C_DenoiseMaxLum ||
           k == C_DenoiseMaxLumBoost;
}
inline void counters_init(Counters& c) {
    for (int k = 0; k != C_Count; ++k) c.v[k] = counter_is_min(k) ? INT64_MAX : (counter_is_max(k) ? INT64_MIN : 0);
}
inline void counters_merge(int64_t* into, const Counters& c) {
    for (int k = 0; k != C_Count; ++k) {
        if (counter_is_min(k)) into[k] = std::min(into[k], c.v[k]);
        else if (counter_is_max(k)) into[k] = std::max(into[k], c.v[k]);
        else into[k] += c.v[k];
    }
}
// Counters.BumpMax(float) (counters.cs:214-219): (long)(value * 1e4f), truncating; NaN and values
// out of the long range give long.MinValue, as x64's conversion does.
inline int64_t decimal_scaled(float v) {
    const float s = v * 1e4f;
    if (!(s == s) || s >= 9.2233720e18f || s < -9.2233720e18f) return INT64_MIN;
    return (int64_t)s;
}

// ---- Generic binned-SAH BVH (bvh.cs:16-477) ----------------------------------------------------
struct BvhNode { AABox bounds; int32_t child; int32_t shapeCount; };

struct Bvh {
    int binCount = 8;
    real costTraverse = 1.0f, costIntersect = 1.0f;
    std::vector<AABox> shapeBounds;  // Shape.Bounds() per shape (a pure function of the shape)
    std::vector<BvhNode> nodes;
    std::vector<int32_t> items;
    int nodeCount = 0;
    bool degenerate = false;         // a split produced an empty side (bvh.cs:448 assert)

    real sah_leaf(real sa, int count) const { return costIntersect * sa * (real)count; }
    real sah_split(real psa, real lsa, int lc, real rsa, int rc) const {
        return costTraverse * psa + sah_leaf(lsa, lc) + sah_leaf(rsa, rc);
    }
    // Bvh ctor, bvh.cs:58-75.
    void build(int bins, real cT, real cI) {
        binCount = bins; costTraverse = cT; costIntersect = cI;
        int n = (int)shapeBounds.size();
        nodes.assign(std::max(n * 2, 1), BvhNode{});
        items.resize(n);
        for (int i = 0; i != n; ++i) items[i] = i;
        nodeCount = 0;
        if (n > 0) subdivide(insert_root(), 0);
    }
    // InsertRoot, bvh.cs:285-296.
    int insert_root() {
        int index = nodeCount++;
        BvhNode& nd = nodes[index];
        nd.bounds = box_inverted();
        nd.child = 0;
        nd.shapeCount = (int)shapeBounds.size();
        for (int i = 0; i != nd.shapeCount; ++i) box_encap(nd.bounds, shapeBounds[items[i]]);
        return index;
    }
    // Insert, bvh.cs:303-313.
    int insert(int begin, int count) {
        int index = nodeCount++;
        BvhNode& nd = nodes[index];
        nd.bounds = box_inverted();
        nd.child = begin;
        nd.shapeCount = count;
        for (int i = 0; i != count; ++i) box_encap(nd.bounds, shapeBounds[items[begin + i]]);
        return index;
    }
    // CentroidBounds, bvh.cs:461-470.
    AABox centroid_bounds(int begin, int count) const {
        AABox b = box_inverted();
        for (int i = 0; i != count; ++i) box_encap(b, box_center(shapeBounds[items[begin + i]]));
        return b;
    }
    // SplitPick, bvh.cs:321-395.
    bool split_pick(int nodeIdx, int& axisOut, real& posOut) const {
        const BvhNode& nd = nodes[nodeIdx];
        int begin = nd.child, count = nd.shapeCount;
        AABox cb = centroid_bounds(begin, count);
        real parentSa = box_sa(nd.bounds);
        real bestCost = sah_leaf(parentSa, count);
        int bestAxis = -1;
        real bestPos = 0.0f;
        std::vector<AABox> binBounds(binCount);
        std::vector<int> binCnt(binCount);
        std::vector<real> leftSa(binCount - 1);
        std::vector<int> leftCount(binCount - 1);
        for (int axis = 0; axis != 3; ++axis) {
            real cmin = comp(cb.mn, axis), cmax = comp(cb.mx, axis);
            if (cmax - cmin < 1e-6f) continue;
            for (int b = 0; b != binCount; ++b) { binBounds[b] = box_inverted(); binCnt[b] = 0; }
            real scale = (real)binCount / (cmax - cmin);
            for (int i = 0; i != count; ++i) {
                const AABox& sb = shapeBounds[items[begin + i]];
                int bin = std::min((int)((comp(box_center(sb), axis) - cmin) * scale), binCount - 1);
                box_encap(binBounds[bin], sb);
                ++binCnt[bin];
            }
            AABox lb = box_inverted();
            int lc = 0;
            for (int b = 0; b != binCount - 1; ++b) {
                box_encap(lb, binBounds[b]);
                lc += binCnt[b];
                leftSa[b] = box_inverted_p(lb) ? 0.0f : box_sa(lb);
                leftCount[b] = lc;
            }
            AABox rb = box_inverted();
            int rc = 0;
            for (int b = binCount - 1; b != 0; --b) {
                box_encap(rb, binBounds[b]);
                rc += binCnt[b];
                real rightSa = box_inverted_p(rb) ? 0.0f : box_sa(rb);
                real cost = sah_split(parentSa, leftSa[b - 1], leftCount[b - 1], rightSa, rc);
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = axis;
                    bestPos = cmin + (real)b * (cmax - cmin) / (real)binCount;
                }
            }
        }
        if (bestAxis < 0) return false;
        axisOut = bestAxis;
        posOut = bestPos;
        return true;
    }
    // Partition, bvh.cs:401-426.
    int partition(int nodeIdx, int axis, real pos) {
        const BvhNode& nd = nodes[nodeIdx];
        int left = nd.child, right = nd.child + nd.shapeCount - 1;
        while (true) {
            const AABox& lb = shapeBounds[items[left]];
            real center = (comp(lb.mn, axis) + comp(lb.mx, axis)) * 0.5f;
            if (center < pos) {
                ++left;
                if (left > right) break;
            } else {
                if (left == right) break;
                std::swap(items[left], items[right]);
                --right;
            }
        }
        return left;
    }
    // Subdivide, bvh.cs:432-459 (recursive: child A's subtree is allocated before child B's).
    void subdivide(int nodeIdx, int depth) {
        if (depth >= 64) return;
        int axis;
        real pos;
        if (!split_pick(nodeIdx, axis, pos)) return;
        int pidx = partition(nodeIdx, axis, pos);
        int child = nodes[nodeIdx].child, cnt = nodes[nodeIdx].shapeCount;
        int countA = pidx - child;
        int countB = cnt - countA;
        if (countA <= 0 || countB <= 0) degenerate = true;
        int childA = insert(child, countA);
        int childB = insert(pidx, countB);
        (void)childB;
        nodes[nodeIdx].child = childA;
        nodes[nodeIdx].shapeCount = 0;
        subdivide(childA, depth + 1);
        subdivide(childA + 1, depth + 1);
    }
    AABox root_bounds() const { return nodeCount > 0 ? nodes[0].bounds : box_inverted(); }

    // Bvh.Intersect, bvh.cs:165-216. `test(idx, best, out)` returns true on a hit with dist.
    template <class Hit, class Fn>
    bool intersect(const Ray& ray, Counters* ctr, Fn test, Hit& bestHit, int& bestIdx) const {
        if (nodeCount == 0) return false;
        struct E { int idx; real t; };
        E queue[128];
        int qc = 1;
        queue[0] = E{0, 0.0f};
        bool found = false;
        real bestDist = kInf;
        while (qc > 0) {
            E e = queue[--qc];
            if (e.t >= bestDist) continue;
            const BvhNode& nd = nodes[e.idx];
            if (nd.shapeCount == 0) {
                if (ctr) ctr->v[C_BvhIntersectNode] += 2;
                real tA = 0, tB = 0;
                bool hA = box_intersect_dist(nodes[nd.child].bounds, ray, tA);
                bool hB = box_intersect_dist(nodes[nd.child + 1].bounds, ray, tB);
                bool pA = hA && tA <= bestDist, pB = hB && tB <= bestDist;
                if (pA) queue[qc++] = E{nd.child, tA};
                if (pB) queue[qc++] = E{nd.child + 1, tB};
                if (pA && pB && tA < tB) std::swap(queue[qc - 2], queue[qc - 1]);
            } else {
                if (ctr) ctr->v[C_BvhIntersectShape] += nd.shapeCount;
                for (int i = 0; i != nd.shapeCount; ++i) {
                    int idx = items[nd.child + i];
                    Hit h{};
                    if (test(idx, h) && h.dist < bestDist) {
                        bestDist = h.dist;
                        bestHit = h;
                        bestIdx = idx;
                        found = true;
                    }
                }
            }
        }
        return found;
    }

    // GetStats, bvh.cs:79-125.
    struct Stats { int nodeCount, depthMax, leafCount, leafMin, leafMax, shapeCount, depthWeighted; real sah; };
    Stats stats_node(int idx, int depth, real rootSa) const {
        const BvhNode& nd = nodes[idx];
        real sa = box_sa(nd.bounds);
        real ratio = rootSa > 0.0f ? sa / rootSa : 0.0f;
        if (nd.shapeCount > 0)
            return Stats{1, depth, 1, nd.shapeCount, nd.shapeCount, nd.shapeCount, depth * nd.shapeCount,
                         sah_leaf(ratio, nd.shapeCount)};
        Stats a = stats_node(nd.child, depth + 1, rootSa);
        Stats b = stats_node(nd.child + 1, depth + 1, rootSa);
        Stats r{a.nodeCount + b.nodeCount, std::max(a.depthMax, b.depthMax), a.leafCount + b.leafCount,
                std::min(a.leafMin, b.leafMin), std::max(a.leafMax, b.leafMax), a.shapeCount + b.shapeCount,
                a.depthWeighted + b.depthWeighted, a.sah + b.sah};
        r.nodeCount += 1;
        r.sah += costTraverse * ratio;
        return r;
    }
};

// ---- Mesh (mesh.cs:1-172) ----------------------------------------------------------------------
struct Mesh {
    std::vector<Triangle> tris;  // lean part + TriangleAttributes (mesh.cs:6-11) in one record
    Bvh bvh;
    // Mesh ctor, mesh.cs:18-46: sahCostIntersect 0.5, the given bin count (reference default 8).
    void build(int bins) {
        bvh.shapeBounds.resize(tris.size());
        for (size_t i = 0; i != tris.size(); ++i) bvh.shapeBounds[i] = tri_bounds(tris[i].lean);
        bvh.build(bins, 1.0f, 0.5f);
    }
    // Mesh.Intersect, mesh.cs:51-62.
    bool intersect(const Ray& ray, Counters* ctr, ShapeHit& out, int& triIdx) const {
        if (ctr) ctr->v[C_MeshIntersect]++;
        LeanHit best{};
        int idx = -1;
        auto test = [&](int i, LeanHit& h) { return tri_intersect(tris[i].lean, ray, h); };
        if (!bvh.intersect(ray, ctr, test, best, idx)) return false;
        out = tri_inflate(tris[idx], best);
        triIdx = idx;
        return true;
    }
};

// ---- Texture (texture.cs:6-179) ----------------------------------------------------------------
struct Texture {
    uint32_t w = 0, h = 0;
    int filter = 0;  // 0 bilinear, 1 point
    V2 tiling{1, 1};
    std::vector<Color> texels;
    Color get(int x, int y) const { return texels[(uint32_t)y * w + (uint32_t)x]; }
    // Texture.Sample, texture.cs:39-69.
    Color sample(V2 coord) const {
        real u = coord.x * tiling.x; u -= std::floor(u);
        real v = coord.y * tiling.y; v -= std::floor(v);
        if (filter == 1) {
            uint32_t x = (uint32_t)(u * (real)w) % w;
            uint32_t y = (uint32_t)(v * (real)h) % h;
            return texels[y * w + x];
        }
        real fx = u * (real)(w - 1);
        real fy = v * (real)(h - 1);
        uint32_t x0 = (uint32_t)fx, y0 = (uint32_t)fy;
        uint32_t x1 = (x0 + 1) % w, y1 = (y0 + 1) % h;
        real tx = fx - (real)x0, ty = fy - (real)y0;
        Color c1 = texels[y0 * w + x0], c2 = texels[y0 * w + x1];
        Color c3 = texels[y1 * w + x0], c4 = texels[y1 * w + x1];
        return lerp(lerp(c1, c2, tx), lerp(c3, c4, tx), ty);
    }
    // Texture.SampleNormal, texture.cs:71-75.
    V3 sample_normal(V2 coord) const {
        Color c = sample(coord);
        return normalize_or(v3(c.x * 2.0f - 1.0f, -(c.y * 2.0f - 1.0f), c.z * 2.0f - 1.0f), kUp);
    }
};

// ---- Cdf1 / Cdf2 (cdf.cs:7-98) -----------------------------------------------------------------
struct Cdf1 {
    std::vector<real> cdf;
    real total = 0;
    void build(const real* w, int n) {
        cdf.assign(n + 1, 0.0f);
        for (int i = 0; i != n; ++i) cdf[i + 1] = cdf[i] + w[i];
        total = cdf[n];
        if (total > 0.0f)
            for (int i = 1; i != n + 1; ++i) cdf[i] /= total;
    }
    uint32_t count() const { return (uint32_t)cdf.size() - 1; }
    uint32_t sample(real norm) const {
        uint32_t lo = 0, hi = count();
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (cdf[mid + 1] <= norm) lo = mid + 1;
            else hi = mid;
        }
        return std::min(lo, count() - 1);
    }
};
struct Cdf2 {
    Cdf1 marginal;
    std::vector<Cdf1> cond;
    real total() const { return marginal.total; }
};

// ---- Materials / sky / fog (scene.cs:6-505) ----------------------------------------------------
struct Material {
    Color albedo;
    real roughness, metallic, transparency, ior;
    Color radiance;
    int colorTex = -1, roughTex = -1, metalTex = -1, normalTex = -1;
};

struct Fog {  // height fog, scene.cs:438-505
    real density; Color color; real anisotropy; real heightFalloff;
    real density_at(real y) const { return density * std::exp(-heightFalloff * y); }
    // OpticalDepth, scene.cs:452-468.
    real optical_depth(const Ray& r, real start, real end) const {
        real hd = r.d.y;
        real hs = r.o.y + start * hd;
        real he = r.o.y + end * hd;
        real ds = density_at(hs);
        real de = density_at(he);
        if (std::fabs(hd) < 1e-6f) return ds * (end - start);
        return (ds - de) / (heightFalloff * hd);
    }
    real transmittance(const Ray& r, real start, real end) const { return std::exp(-optical_depth(r, start, end)); }
    // ScatterDistance, scene.cs:470-496.
    bool scatter_distance(const Ray& r, real start, real end, Rng& rng, real& tOut) const {
        real hd = r.d.y;
        real ho = r.o.y;
        real hs = ho + start * hd;
        real u = rng.next_exponential();
        real t;
        if (std::fabs(hd) < 1e-6f) {
            real ds = density_at(hs);
            if (ds <= 0.0f) return false;
            t = start + u / ds;
        } else {
            real rhs = std::exp(-heightFalloff * hs) - u * heightFalloff * hd / density;
            if (rhs <= 0.0f) return false;
            real yt = -std::log(rhs) / heightFalloff;
            t = (yt - ho) / hd;
        }
        if (t < end) { tOut = t; return true; }
        return false;
    }
};

// Bounded media (no reference counterpart): a homogeneous box (kind 1) or a box holding a
// piecewise-constant grid of half-precision densities (kind 2; sigma = scale * voxel value).
// Free flight: kind 1 inverts T = exp(-sigma d) with one exponential draw; kind 2 is delta
// tracking (tentative collisions at the majorant rate, accepted with probability
// sigma / majorant, at most 4096 tentative steps). Transmittance: exp of the exact optical
// depth, summed voxel by voxel along the ray (next boundary x, then y, then z on ties).
inline real half_bits(uint16_t h) {
    int e = (h >> 10) & 0x1f, m = h & 0x3ff;
    real v = e == 0 ? std::ldexp((real)m, -24) : (e == 31 ? (m ? (real)NAN : kInf) : std::ldexp((real)(m | 0x400), e - 25));
    return (h & 0x8000) ? -v : v;
}
struct Medium {
    int kind = 0;
    real sigma = 0;
    V3 bmin, bmax, cell, cellInv;
    int n[3] = {0, 0, 0};
    std::vector<real> grid;  // decoded values
    real majorant = 0;
    void init(int k, real s, V3 lo, V3 hi, const uint32_t* res, const uint16_t* bits) {
        kind = k; sigma = s; bmin = lo; bmax = hi;
        if (kind == 2) {
            for (int i = 0; i != 3; ++i) n[i] = (int)res[i];
            cell = v3((hi.x - lo.x) / (real)n[0], (hi.y - lo.y) / (real)n[1], (hi.z - lo.z) / (real)n[2]);
            cellInv = v3((real)1 / cell.x, (real)1 / cell.y, (real)1 / cell.z);
            size_t cnt = (size_t)n[0] * n[1] * n[2];
            grid.resize(cnt);
            real mx = 0;
            for (size_t i = 0; i != cnt; ++i) { grid[i] = half_bits(bits[i]); mx = std::max(mx, grid[i]); }
            majorant = sigma * mx;
        }
    }
    bool span(const Ray& r, real start, real end, real& a, real& b) const {
        real t1 = (bmin.x - r.o.x) * r.inv.x, t2 = (bmax.x - r.o.x) * r.inv.x;
        real t3 = (bmin.y - r.o.y) * r.inv.y, t4 = (bmax.y - r.o.y) * r.inv.y;
        real t5 = (bmin.z - r.o.z) * r.inv.z, t6 = (bmax.z - r.o.z) * r.inv.z;
        real lo = nmax(nmax(nmin(t1, t2), nmin(t3, t4)), nmin(t5, t6));
        real hi = nmin(nmin(nmax(t1, t2), nmax(t3, t4)), nmax(t5, t6));
        a = nmax(lo, start);
        b = nmin(hi, end);
        return a < b;
    }
    static int coord(real p, real lo, real inv, int cnt) {
        int i = (int)std::floor((p - lo) * inv);
        return i < 0 ? 0 : (i >= cnt ? cnt - 1 : i);
    }
    real sigma_cell(int x, int y, int z) const { return sigma * grid[((size_t)z * n[1] + y) * n[0] + x]; }
    real sigma_at(V3 p) const {
        return sigma_cell(coord(p.x, bmin.x, cellInv.x, n[0]), coord(p.y, bmin.y, cellInv.y, n[1]),
                          coord(p.z, bmin.z, cellInv.z, n[2]));
    }
    real grid_depth(const Ray& r, real a, real b) const {
        V3 p = ray_at(r, a);
        int c[3] = {coord(p.x, bmin.x, cellInv.x, n[0]), coord(p.y, bmin.y, cellInv.y, n[1]),
                    coord(p.z, bmin.z, cellInv.z, n[2])};
        real o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z}, inv[3] = {r.inv.x, r.inv.y, r.inv.z};
        real lo[3] = {bmin.x, bmin.y, bmin.z}, cs[3] = {cell.x, cell.y, cell.z};
        real next[3], delta[3];
        int step[3];
        for (int k = 0; k != 3; ++k) {
            step[k] = d[k] > 0 ? 1 : -1;
            real edge = lo[k] + (real)(c[k] + (d[k] > 0 ? 1 : 0)) * cs[k];
            next[k] = d[k] == 0 ? kInf : (edge - o[k]) * inv[k];
            delta[k] = d[k] == 0 ? kInf : cs[k] * std::fabs(inv[k]);
        }
        real t = a, tau = 0;
        for (int it = 0, lim = n[0] + n[1] + n[2] + 3; it != lim; ++it) {
            int k = next[0] <= next[1] ? (next[0] <= next[2] ? 0 : 2) : (next[1] <= next[2] ? 1 : 2);
            real tEnd = nmin(next[k], b);
            tau += sigma_cell(c[0], c[1], c[2]) * nmax((real)0, tEnd - t);
            if (tEnd >= b) break;
            t = tEnd;
            c[k] += step[k];
            if (c[k] < 0 || c[k] >= n[k]) break;
            next[k] += delta[k];
        }
        return tau;
    }
    real optical_depth(const Ray& r, real start, real end) const {
        real a, b;
        if (!span(r, start, end, a, b)) return 0;
        return kind == 1 ? sigma * (b - a) : grid_depth(r, a, b);
    }
    bool scatter_distance(const Ray& r, real start, real end, Rng& rng, real& tOut) const {
        real a, b;
        if (!span(r, start, end, a, b)) return false;
        if (kind == 1) {
            if (!(sigma > 0)) return false;
            real t = a + (real)rng.next_exponential() / sigma;
            if (t < b) { tOut = t; return true; }
            return false;
        }
        if (!(majorant > 0)) return false;
        real t = a;
        for (int it = 0; it != 4096; ++it) {
            t += (real)rng.next_exponential() / majorant;
            if (!(t < b)) return false;
            real s = sigma_at(ray_at(r, t));
            if ((real)rng.next_float() < s / majorant) { tOut = t; return true; }
        }
        return false;
    }
};

struct SampleLight { V3 dir; real pdf; int texel = -1; };

// Quat.AngleAxis, math.cs:621-626.
inline Quat angle_axis(real angle, V3 axis) {
    V3 vec = axis * std::sin(angle * 0.5f);
    return Quat{vec.x, vec.y, vec.z, std::cos(angle * 0.5f)};
}
// Mat4.RotateLook (math.cs:742-758) + Quat.FromMat4 (math.cs:631-655): only the 3x3 part matters.
inline Quat quat_look(V3 forward, V3 upRef) {
    V3 c0, c1, c2;
    if (mag_sqr(forward) <= 1e-6f || mag_sqr(upRef) <= 1e-6f) {
        c0 = v3(1, 0, 0); c1 = v3(0, 1, 0); c2 = v3(0, 0, 1);
    } else {
        V3 fwd = normalize(forward);
        V3 right = cross(upRef, fwd);
        right = mag_sqr(right) > 1e-6f ? normalize(right) : v3(1, 0, 0);
        V3 up = cross(fwd, right);
        c0 = right; c1 = up; c2 = fwd;
    }
    real trace = c0.x + c1.y + c2.z;
    if (trace > 1e-6f) {
        real s = std::sqrt(trace + 1) * 2;
        return Quat{(c1.z - c2.y) / s, (c2.x - c0.z) / s, (c0.y - c1.x) / s, s * 0.25f};
    }
    if (c0.x > c1.y && c0.x > c2.z) {
        real s = std::sqrt(1 + c0.x - c1.y - c2.z) * 2;
        return Quat{s * 0.25f, (c1.x + c0.y) / s, (c2.x + c0.z) / s, (c1.z - c2.y) / s};
    }
    if (c1.y > c2.z) {
        real s = std::sqrt(1 + c1.y - c0.x - c2.z) * 2;
        return Quat{(c1.x + c0.y) / s, s * 0.25f, (c2.y + c1.z) / s, (c2.x - c0.z) / s};
    }
    real s = std::sqrt(1 + c2.z - c0.x - c1.y) * 2;
    return Quat{(c2.x + c0.z) / s, (c2.y + c1.z) / s, s * 0.25f, (c0.y - c1.x) / s};
}

// Walker/Vose alias table over the environment's texel weights (no reference counterpart; the
// reference samples the same discrete distribution with Cdf2, cdf.cs:62-98). Built in double:
// scaled_i = w_i * n / sum(w); the worklists are stacks filled in index order; a small entry s takes
// the last large entry l as its alias and l keeps (scaled_l + scaled_s) - 1. Leftovers keep prob 1.
struct AliasTable {
    std::vector<float> prob;
    std::vector<int32_t> alias;
    void build(const std::vector<real>& w) {
        size_t n = w.size();
        double total = 0;
        for (size_t i = 0; i != n; ++i) total += (double)w[i];
        prob.assign(n, 1.0f);
        alias.resize(n);
        for (size_t i = 0; i != n; ++i) alias[i] = (int32_t)i;
        if (!(total > 0)) return;
        std::vector<double> scaled(n);
        std::vector<int32_t> small, large;
        for (size_t i = 0; i != n; ++i) {
            scaled[i] = (double)w[i] * (double)n / total;
            (scaled[i] < 1.0 ? small : large).push_back((int32_t)i);
        }
        while (!small.empty() && !large.empty()) {
            int32_t a = small.back(); small.pop_back();
            int32_t b = large.back(); large.pop_back();
            prob[a] = (float)scaled[a];
            alias[a] = b;
            scaled[b] = (scaled[b] + scaled[a]) - 1.0;
            (scaled[b] < 1.0 ? small : large).push_back(b);
        }
    }
    // Two uniforms (u.x picks the column, u.y accepts it or takes its alias).
    uint32_t sample(real ux, real uy) const {
        uint32_t n = (uint32_t)prob.size();
        uint32_t i = (uint32_t)(float)((float)ux * (float)n);
        if (i >= n) i = n - 1;
        return (float)uy < prob[i] ? i : (uint32_t)alias[i];
    }
};

struct Sky {
    int kind = 0;  // 0 none, 1 texture, 2 procedural
    int sampler = 0;  // texture sky light sampling: 0 Cdf2 (reference), 1 alias table
    // SkyTexture, scene.cs:379-435.
    Texture tex;
    Cdf2 cdf;
    AliasTable aliasTab;
    real pdfScale = 0, multiplier = 1, yaw = 0;
    Quat rot{0, 0, 0, 1}, rotInv{0, 0, 0, 1};
    // SkyProcedural / SunProcedural, scene.cs:281-377.
    Color top, middle, bottom, sunRadiance;
    V3 sunDir;
    real sunAngle = 0, sunAngleCos = 1, sunPdfDenom = 0;
    Quat sunRot{0, 0, 0, 1};

    void init_texture() {
        int W = (int)tex.w, H = (int)tex.h;
        cdf.cond.resize(H);
        std::vector<real> rowW(H), colW(W), all((size_t)W * H);
        for (int y = 0; y != H; ++y) {
            for (int x = 0; x != W; ++x) {
                colW[x] = luminance(tex.get(x, y)) * std::sin(kPi * ((real)y + 0.5f) / (real)H);
                all[(size_t)y * W + x] = colW[x];
            }
            cdf.cond[y].build(colW.data(), W);
            rowW[y] = cdf.cond[y].total;
        }
        aliasTab.build(all);
        cdf.marginal.build(rowW.data(), H);
        pdfScale = (real)(W * H) / (cdf.total() * 2.0f * kPi * kPi);
        rot = angle_axis(yaw, kUp);
        rotInv = qinv(rot);
    }
    void init_procedural() {
        sunAngleCos = std::cos(sunAngle);
        sunRot = quat_look(sunDir, kUp);
        sunPdfDenom = kPi * (1.0f - sunAngleCos);
    }
    Color radiance(V3 dir) const {
        if (kind == 1) return tex.sample(equirect_uv(qrot(rotInv, dir))) * multiplier;
        if (kind == 2) {
            const real bias = 0.0001f;
            real topBlend = 1.0f - libm_pow(nmin(1.0f, 1.0f + bias - dir.y), (real)4.0f);
            real bottomBlend = 1.0f - libm_pow(nmin(1.0f, 1.0f + bias + dir.y), (real)40.0f);
            real middleBlend = 1.0f - topBlend - bottomBlend;
            Color ambient = top * topBlend + middle * middleBlend + bottom * bottomBlend;
            real blend = nmax(0.0f, (dot(dir, sunDir) - sunAngleCos) / (1.0f - sunAngleCos));
            return ambient + sunRadiance * blend;
        }
        return kBlack;
    }
    // LightDirRand: SkyTexture scene.cs:408-416, SunProcedural scene.cs:315-324.
    bool light_dir_rand(Rng& rng, SampleLight& out) const {
        if (kind == 1) {
            if (cdf.total() <= 0.0f) return false;
            V2 c = rng.rand2();
            int x, y;
            if (sampler == 1) {
                uint32_t t = aliasTab.sample(c.x, c.y);
                x = (int)(t % tex.w);
                y = (int)(t / tex.w);
            } else {
                y = (int)cdf.marginal.sample(c.y);
                x = (int)cdf.cond[y].sample(c.x);
            }
            V2 uv = v2(((real)x + 0.5f) / (real)tex.w, ((real)y + 0.5f) / (real)tex.h);
            V3 dir = qrot(rot, from_equirect_uv(uv));
            real pdf = luminance(tex.get(x, y)) * pdfScale;
            if (!(pdf > 0.0f)) return false;
            out = SampleLight{dir, pdf, y * (int)tex.w + x};
            return true;
        }
        if (kind == 2) {
            real phi = kPi * 2.0f * (real)rng.next_float();
            real blend = std::sqrt((real)rng.next_float());
            real z = sunAngleCos + (1.0f - sunAngleCos) * blend;
            real sinT = std::sqrt(nmax(0.0f, 1.0f - z * z));
            V3 dir = qrot(sunRot, v3(std::cos(phi) * sinT, std::sin(phi) * sinT, z));
            out = SampleLight{dir, blend / sunPdfDenom};
            return true;
        }
        return false;
    }
    // Texel whose luminance LightDir's pdf reads for direction dir (texture sky).
    int light_texel(V3 dir) const {
        V2 uv = equirect_uv(qrot(rotInv, dir));
        int cx = ((int)(uv.x * (real)(int)tex.w)) % (int)tex.w;
        int cy = ((int)(uv.y * (real)(int)tex.h)) % (int)tex.h;
        return cy * (int)tex.w + cx;
    }
    // Derivative of a texel-sky light pdf p = lum(texel) * W * H / (total * 2 pi^2) with respect to
    // the texels, scaled by s: the texel's own luminance directly, every texel through total (the
    // latter is gathered in gTotal and spread by apply_total_grad).
    void pdf_grad(int texel, real pdf, real s, real* g, real& gTotal) const {
        if (kind != 1 || texel < 0 || s == 0) return;
        g[texel * 3] += s * (real)0.2126f * pdfScale;
        g[texel * 3 + 1] += s * (real)0.7152f * pdfScale;
        g[texel * 3 + 2] += s * (real)0.0722f * pdfScale;
        gTotal += s * (-pdf / cdf.total());
    }
    void apply_total_grad(real gTotal, real* g) const {
        if (kind != 1 || gTotal == 0) return;
        for (uint32_t y = 0; y != tex.h; ++y) {
            real sy = std::sin(kPi * ((real)y + 0.5f) / (real)tex.h);
            for (uint32_t x = 0; x != tex.w; ++x) {
                size_t t = (size_t)y * tex.w + x;
                g[t * 3] += gTotal * (real)0.2126f * sy;
                g[t * 3 + 1] += gTotal * (real)0.7152f * sy;
                g[t * 3 + 2] += gTotal * (real)0.0722f * sy;
            }
        }
    }
    // Adds s_c * multiplier * (bilinear weight of texel t) to g[t * 3 + c] for the texels that
    // Radiance(dir) interpolates: the derivative of Radiance with respect to the sky texels.
    void radiance_grad(V3 dir, Color s, real* g) const {
        if (kind != 1) return;
        V2 c = equirect_uv(qrot(rotInv, dir));
        real u = c.x * tex.tiling.x; u -= std::floor(u);
        real v = c.y * tex.tiling.y; v -= std::floor(v);
        uint32_t idx[4];
        real w[4];
        if (tex.filter == 1) {
            idx[0] = ((uint32_t)(v * (real)tex.h) % tex.h) * tex.w + (uint32_t)(u * (real)tex.w) % tex.w;
            w[0] = 1;
            idx[1] = idx[2] = idx[3] = idx[0];
            w[1] = w[2] = w[3] = 0;
        } else {
            real fx = u * (real)(tex.w - 1), fy = v * (real)(tex.h - 1);
            uint32_t x0 = (uint32_t)fx, y0 = (uint32_t)fy;
            uint32_t x1 = (x0 + 1) % tex.w, y1 = (y0 + 1) % tex.h;
            real tx = fx - (real)x0, ty = fy - (real)y0;
            idx[0] = y0 * tex.w + x0; w[0] = (1 - tx) * (1 - ty);
            idx[1] = y0 * tex.w + x1; w[1] = tx * (1 - ty);
            idx[2] = y1 * tex.w + x0; w[2] = (1 - tx) * ty;
            idx[3] = y1 * tex.w + x1; w[3] = tx * ty;
        }
        for (int i = 0; i != 4; ++i) {
            real k = w[i] * multiplier;
            g[idx[i] * 3] += s.x * k;
            g[idx[i] * 3 + 1] += s.y * k;
            g[idx[i] * 3 + 2] += s.z * k;
        }
    }
    // LightDir (MIS pdf of a direction): SkyTexture scene.cs:418-426, SunProcedural scene.cs:306-313.
    bool light_dir(V3 dir, real& pdfOut) const {
        if (kind == 1) {
            if (cdf.total() <= 0.0f) return false;
            V2 uv = equirect_uv(qrot(rotInv, dir));
            int cx = ((int)(uv.x * (real)(int)tex.w)) % (int)tex.w;
            int cy = ((int)(uv.y * (real)(int)tex.h)) % (int)tex.h;
            real pdf = luminance(tex.get(cx, cy)) * pdfScale;
            if (!(pdf > 0.0f)) return false;
            pdfOut = pdf;
            return true;
        }
        if (kind == 2) {
            real cosTheta = dot(dir, sunDir);
            if (cosTheta < sunAngleCos) return false;
            real blend = (cosTheta - sunAngleCos) / (1.0f - sunAngleCos);
            pdfOut = blend / sunPdfDenom;
            return true;
        }
        return false;
    }
};

// ---- BRDF (brdf.cs:7-123) ----------------------------------------------------------------------
inline Color base_reflectivity(real ior, Color albedo, real metallic) {
    real f0 = libm_pow((ior - 1.0f) / (ior + 1.0f), (real)2.0f);
    return lerp(v3s(f0), albedo, metallic);
}
inline Color fresnel(real nDotV, Color f0) {
    real x = 1.0f - nDotV;
    real x2 = x * x;
    real x5 = x2 * x2 * x;
    return f0 + (kWhite - f0) * x5;
}
inline real ggx_d(real nDotH, real r2) {
    real alpha2 = r2 * r2;
    real nh2 = nDotH * nDotH;
    real d = (1.0f - nh2) + nh2 * alpha2;
    return alpha2 / (kPi * d * d);
}
inline real smith_g1(real nDotX, real r2) {
    real k = r2 / 2.0f;
    return nDotX / (nDotX * (1.0f - k) + k);
}
inline V3 ggx_sample_local(real r2, V2 u) {
    real cosTheta = std::sqrt((1.0f - u.x) / (u.x * (r2 * r2 - 1.0f) + 1.0f));
    real sinTheta = std::sqrt(nmax(0.0f, 1.0f - cosTheta * cosTheta));
    real phi = 2.0f * kPi * u.y;
    return v3(sinTheta * std::cos(phi), sinTheta * std::sin(phi), cosTheta);
}
inline real power_heuristic(real p1, real p2) {
    real a = p1 * p1, b = p2 * p2;
    return a + b > 0.0f ? a / (a + b) : 0.0f;
}
inline real hg_pdf(real rDotL, real g) {
    real g2 = g * g;
    real d = 1.0f + g2 - 2.0f * g * rDotL;
    return (1.0f - g2) / (4.0f * kPi * d * std::sqrt(d));
}
inline bool refract(V3 inc, V3 n, real eta, V3& out) {
    real cosI = -dot(inc, n);
    real sin2T = eta * eta * (1.0f - cosI * cosI);
    if (sin2T >= 1.0f) return false;
    real cosT = std::sqrt(1.0f - sin2T);
    out = eta * inc + (eta * cosI - cosT) * n;
    return true;
}
inline Color beer_law(Color albedo, real d) {
    return v3(libm_pow(nmax(albedo.x, 0.0f), d), libm_pow(nmax(albedo.y, 0.0f), d), libm_pow(nmax(albedo.z, 0.0f), d));
}
inline V3 hg_scatter_dir(V3 forward, real g, Rng& rng) {
    V2 u = rng.rand2();
    real cosTheta;
    if (std::fabs(g) < 1e-3f) {
        cosTheta = 1.0f - 2.0f * u.x;
    } else {
        real s = (1.0f - g * g) / (1.0f - g + 2.0f * g * u.x);
        cosTheta = (1.0f + g * g - s * s) / (2.0f * g);
    }
    cosTheta = nclamp(cosTheta, -1.0f, 1.0f);
    real sinTheta = std::sqrt(1.0f - cosTheta * cosTheta);
    real phi = 2.0f * kPi * u.y;
    V3 tangent = perp(forward);
    V3 bitangent = cross(forward, tangent);
    return cosTheta * forward + sinTheta * std::cos(phi) * tangent + sinTheta * std::sin(phi) * bitangent;
}

// ---- Surface (scene.cs:15-152) -----------------------------------------------------------------
struct Scatter { V3 dir; Color weight; bool hasPdf; real pdf; bool refracted; };
struct Surface {
    Color radiance, albedo;
    real roughness, metallic, transparency, ior;
    V3 normal, normalGeo;
    V4 tangent;
    V2 uv;

    V3 clamp_to_geo_side(V3 dir, V3 sideHint) const {
        real d = dot(dir, normalGeo);
        if (d * dot(sideHint, normalGeo) < 0.0f) dir = normalize_or(dir - 2.0f * d * normalGeo, sideHint);
        return dir;
    }
    real specular_probability(V3 viewDir) const {
        real nDotV = nmax(1e-4f, dot(normal, viewDir));
        real f = luminance(fresnel(nDotV, base_reflectivity(ior, albedo, metallic)));
        return nclamp(nmax(f, metallic), 0.001f, 0.999f);
    }
    V3 specular_dir(V3 inc, Rng& rng) const {
        real r2 = nmax(roughness * roughness, 1e-4f);
        V3 h = from_tangent_space(ggx_sample_local(r2, rng.rand2()), tangent, normal);
        return reflect(inc, h);
    }
    Color opaque_eval(V3 viewDir, V3 lightDir) const {
        real nDotL = dot(normal, lightDir);
        if (nDotL <= 0.0f) return kBlack;
        real nDotV = nmax(1e-4f, dot(normal, viewDir));
        real r2 = nmax(roughness * roughness, 1e-4f);
        Color f0 = base_reflectivity(ior, albedo, metallic);
        V3 h = normalize_or(viewDir + lightDir, normal);
        real nDotH = dot(normal, h);
        real hDotV = dot(h, viewDir);
        Color f = fresnel(hDotV, f0);
        real g = smith_g1(nDotV, r2) * smith_g1(nDotL, r2);
        Color specular = ggx_d(nDotH, r2) * f * g / (4.0f * nDotV);
        Color diffuse = (kWhite - fresnel(nDotV, f0)) * albedo * (1.0f - metallic) * (nDotL / kPi);
        return (specular + diffuse) * (1.0f - transparency);
    }
    // Partial derivatives of OpaqueEval with respect to this surface's albedo (per channel: the
    // Jacobian is diagonal) and roughness. For the adjoint renderer; no reference counterpart.
    void opaque_eval_grad(V3 viewDir, V3 lightDir, Color& dAlbedo, Color& dRough) const {
        dAlbedo = kBlack;
        dRough = kBlack;
        real nDotL = dot(normal, lightDir);
        if (nDotL <= 0.0f) return;
        real nDotV = nmax(1e-4f, dot(normal, viewDir));
        real r2 = nmax(roughness * roughness, 1e-4f);
        Color f0 = base_reflectivity(ior, albedo, metallic);
        V3 h = normalize_or(viewDir + lightDir, normal);
        real nDotH = dot(normal, h);
        real hDotV = dot(h, viewDir);
        real xh = 1 - hDotV, xv = 1 - nDotV;
        real xh5 = xh * xh * xh * xh * xh, xv5 = xv * xv * xv * xv * xv;
        Color fh = fresnel(hDotV, f0), fv = fresnel(nDotV, f0);
        real g1v = smith_g1(nDotV, r2), g1l = smith_g1(nDotL, r2);
        real dist = ggx_d(nDotH, r2);
        real keep = 1 - transparency;
        // d F / d albedo = metallic * (1 - x^5) for both Fresnel terms (F0 = lerp(f0, albedo, metallic)).
        real specA = dist * g1v * g1l / (4 * nDotV) * metallic * (1 - xh5);
        real diffK = (1 - metallic) * (nDotL / kPi);
        dAlbedo = v3(specA + diffK * ((1 - fv.x) - albedo.x * metallic * (1 - xv5)),
                     specA + diffK * ((1 - fv.y) - albedo.y * metallic * (1 - xv5)),
                     specA + diffK * ((1 - fv.z) - albedo.z * metallic * (1 - xv5))) * keep;
        if (roughness * roughness > (real)1e-4f) {  // r2 = max(r^2, 1e-4): no derivative while clamped
            real alpha2 = r2 * r2, nh2 = nDotH * nDotH;
            real dd = (1 - nh2) + nh2 * alpha2;
            real dD_dAlpha2 = 1 / (kPi * dd * dd) - 2 * alpha2 * nh2 / (kPi * dd * dd * dd);
            real dD = dD_dAlpha2 * 2 * r2;
            real k = r2 / 2;
            auto dG1dk = [&](real x) { real den = x * (1 - k) + k; return -x * (1 - x) / (den * den); };
            real dG = 0.5f * (dG1dk(nDotV) * g1l + g1v * dG1dk(nDotL));
            real dSpec = (dD * g1v * g1l + dist * dG) / (4 * nDotV) * (2 * roughness);
            dRough = fh * dSpec * keep;
        }
    }
    // d GgxD(nDotH, max(r^2, 1e-4)) / d r.
    real ggx_d_dr(real nDotH) const {
        if (!(roughness * roughness > (real)1e-4f)) return 0;
        real r2 = roughness * roughness, alpha2 = r2 * r2, nh2 = nDotH * nDotH;
        real dd = (1 - nh2) + nh2 * alpha2;
        return (1 / (kPi * dd * dd) - 2 * alpha2 * nh2 / (kPi * dd * dd * dd)) * 2 * r2 * 2 * roughness;
    }
    // Density of the specular candidate direction d (GGX half vector, reflected view) at any d,
    // and its derivative with respect to roughness.
    void spec_candidate_pdf(V3 viewDir, V3 d, real& pdf, real& dPdf) const {
        pdf = dPdf = 0;
        V3 hv = viewDir + d;
        if (mag_sqr(hv) <= 1e-12f) return;
        V3 h = normalize(hv);
        real nh = dot(normal, h);
        if (nh < 0) { h = -h; nh = -nh; }
        real vh = std::fabs(dot(h, viewDir));
        if (!(vh > 1e-6f)) return;
        real r2 = nmax(roughness * roughness, 1e-4f);
        pdf = ggx_d(nh, r2) * nh / (4 * vh);
        dPdf = ggx_d_dr(nh) * nh / (4 * vh);
    }
    // d SpecularProbability / d albedo: sp = clamp(max(lum F(nDotV), metallic), .001, .999) moves
    // with albedo only where lum F is the maximum and inside the clamp.
    Color spec_prob_grad(V3 viewDir) const {
        real nDotV = nmax(1e-4f, dot(normal, viewDir));
        real lumF = luminance(fresnel(nDotV, base_reflectivity(ior, albedo, metallic)));
        if (!(lumF > metallic && lumF > 0.001f && lumF < 0.999f)) return kBlack;
        real xv = 1 - nDotV, base = metallic * (1 - xv * xv * xv * xv * xv);
        return v3(0.2126f * base, 0.7152f * base, 0.0722f * base);
    }
    // d OpaquePdf(viewDir, l, sp(albedo)) / d albedo (per channel) and / d roughness.
    void opaque_pdf_grad(V3 viewDir, V3 l, real sp, Color& dA, real& dR) const {
        dA = kBlack;
        dR = 0;
        real nDotL = dot(normal, l);
        if (nDotL <= 0.0f) return;
        V3 h = normalize_or(viewDir + l, normal);
        real nh = nmax(0.0f, dot(normal, h));
        real hv = nmax(1e-6f, dot(h, viewDir));
        real r2 = nmax(roughness * roughness, 1e-4f);
        real keep = 1 - transparency;
        real pSpec = ggx_d(nh, r2) * nh / (4 * hv), pDiff = nDotL / kPi;
        dR = sp * ggx_d_dr(nh) * nh / (4 * hv) * keep;
        dA = spec_prob_grad(viewDir) * ((pSpec - pDiff) * keep);
    }
    // Score correction of a scatter direction l drawn by OpaqueScatter: d/dθ [log q(l) - log p(l)],
    // q being the density of the procedure (lobe choice, GGX reflection or cosine lobe, then the
    // ClampToGeoSide mirror, which folds the density at the mirror image of l onto l) and p the
    // OpaquePdf the estimator divides by. Zero when q = p. sA: per albedo channel, sR: roughness.
    void scatter_score(V3 viewDir, V3 l, real sp, Color& sA, real& sR) const {
        sA = kBlack;
        sR = 0;
        real p = opaque_pdf_sp(viewDir, l, sp);
        V3 hh = normalize_or(viewDir + l, normal);
        real nh = nmax(0.0f, dot(normal, hh));
        real hv = nmax(1e-6f, dot(hh, viewDir));
        real r2 = nmax(roughness * roughness, 1e-4f);
        real pSpec = ggx_d(nh, r2) * nh / (4 * hv), pDiff = dot(normal, l) / kPi;
        real dpSpec = ggx_d_dr(nh) * nh / (4 * hv);
        if (dot(normal, l) <= 0) { pSpec = pDiff = dpSpec = 0; }
        V3 m = l - 2.0f * dot(l, normalGeo) * normalGeo;
        real mSpec = 0, dmSpec = 0, mDiff = 0;
        if (dot(m, normalGeo) < 0) {
            spec_candidate_pdf(viewDir, m, mSpec, dmSpec);
            mDiff = nmax(0.0f, dot(normal, m)) / kPi;
        }
        real mm = sp * mSpec + (1 - sp) * mDiff;
        real q = p + mm;
        if (!(q > 0)) return;
        bool pLive = p >= 1e-6f;  // below that the estimator divides by the constant 1e-6
        real dpR = sp * dpSpec, dmR = sp * dmSpec;
        sR = (dpR + dmR) / q - (pLive ? dpR / p : 0);
        real dpS = pSpec - pDiff, dmS = mSpec - mDiff;
        sA = spec_prob_grad(viewDir) * ((dpS + dmS) / q - (pLive ? dpS / p : 0));
    }
    real opaque_pdf_sp(V3 viewDir, V3 lightDir, real sp) const {
        real nDotL = dot(normal, lightDir);
        if (nDotL <= 0.0f) return 0.0f;
        V3 h = normalize_or(viewDir + lightDir, normal);
        real nDotH = nmax(0.0f, dot(normal, h));
        real hDotV = nmax(1e-6f, dot(h, viewDir));
        real r2 = nmax(roughness * roughness, 1e-4f);
        real pdfSpec = ggx_d(nDotH, r2) * nDotH / (4.0f * hDotV);
        real pdfDiff = nDotL / kPi;
        return (sp * pdfSpec + (1.0f - sp) * pdfDiff) * (1.0f - transparency);
    }
    real opaque_pdf(V3 viewDir, V3 lightDir) const { return opaque_pdf_sp(viewDir, lightDir, specular_probability(viewDir)); }
    Scatter opaque_scatter(V3 inc, V3 viewDir, Rng& rng) const {
        real sp = specular_probability(viewDir);
        V3 dir;
        if ((real)rng.next_float() < sp) dir = specular_dir(inc, rng);
        else dir = normalize_or(normal + rand_on_sphere(rng), normal);
        dir = clamp_to_geo_side(dir, normalGeo);
        real pdf = opaque_pdf_sp(viewDir, dir, sp);
        return Scatter{dir, opaque_eval(viewDir, dir) / nmax(pdf, 1e-6f), true, pdf, false};
    }
    Scatter transparent_scatter(V3 inc, V3 viewDir, Rng& rng) const {
        bool entering = dot(inc, normalGeo) < 0.0f;
        V3 refractNormal = entering ? normalGeo : -normalGeo;
        real iorRatio = entering ? (real)1 / ior : ior;
        real r2 = nmax(roughness * roughness, 1e-4f);
        V3 h = from_tangent_space(ggx_sample_local(r2, rng.rand2()), tangent, normal);
        if (!entering) h = -h;
        if (dot(inc, h) > 0.0f) h = refractNormal;
        real hDotV = nmax(1e-4f, dot(h, viewDir));
        real f = luminance(fresnel(hDotV, base_reflectivity(ior, albedo, metallic)));
        V3 rdir;
        if ((real)rng.next_float() >= f && refract(inc, h, iorRatio, rdir))
            return Scatter{clamp_to_geo_side(rdir, -refractNormal), kWhite, false, 0, true};
        return Scatter{clamp_to_geo_side(reflect(inc, h), refractNormal), kWhite, false, 0, false};
    }
    Scatter scatter(V3 inc, V3 viewDir, Rng& rng) const {
        if ((real)rng.next_float() < transparency) return transparent_scatter(inc, viewDir, rng);
        return opaque_scatter(inc, viewDir, rng);
    }
};

// ---- Objects and scene (scene.cs:204-270, 507-811) ---------------------------------------------
struct Object {
    Transform trans;
    Material mat;
    int kind;   // 0 mesh, 1 sphere, 2 aabox
    int mesh;   // mesh index
    V3 sphereCenter; real sphereRadius;
    AABox box;
    AABox bounds;
};

struct Fragment {
    Color radiance, radianceFog;
    Color albedo;  // primary surface albedo (valid with hasDepth); no reference counterpart
    bool hasNormal; V3 normal;
    bool hasUv; V2 uv;
    bool hasDepth; real depth;
};

struct Scene {
    std::vector<Mesh> meshes;
    std::vector<Texture> textures;
    std::vector<Object> objects;
    int skyTexId = -1;
    Sky sky;
    bool hasFog = false;
    Fog fog{};        // height fog, or color / anisotropy of a bounded medium
    Medium medium;    // kind != 0: bounded medium in place of the height fog
    real fog_depth(const Ray& r, real start, real end) const {
        return medium.kind ? medium.optical_depth(r, start, end) : fog.optical_depth(r, start, end);
    }
    bool fog_scatter(const Ray& r, real start, real end, Rng& rng, real& t) const {
        return medium.kind ? medium.scatter_distance(r, start, end, rng, t) : fog.scatter_distance(r, start, end, rng, t);
    }
    Bvh bvh;
    bool built = false;

    // Sphere.Intersect, math.cs:1210-1234.
    static bool sphere_intersect(V3 center, real radius, const Ray& ray, ShapeHit& out) {
        V3 toCenter = center - ray.o;
        real tcd = mag_sqr(toCenter);
        real a = dot(toCenter, ray.d);
        real disc = radius * radius - tcd + a * a;
        if (disc < 0.0f) return false;
        real f = std::sqrt(disc);
        real t = tcd < radius * radius ? a + f : a - f;
        if (t < 0.0f) return false;
        V3 norm = normalize(ray_at(ray, t) - center);
        V3 tg = normalize_or(v3(-norm.z, 0.0f, norm.x), kRight);
        real texU = 0.5f + std::atan2(norm.z, norm.x) / (2.0f * kPi);
        real texV = 0.5f - std::asin(nclamp(norm.y, -1.0f, 1.0f)) / kPi;
        out = ShapeHit{t, norm, norm, V4{tg.x, tg.y, tg.z, 1.0f}, v2(texU, texV)};
        return true;
    }
    // AABox.Intersect, math.cs:1006-1059.
    static bool aabox_intersect(const AABox& b, const Ray& ray, ShapeHit& out) {
        real t1 = (b.mn.x - ray.o.x) * ray.inv.x;
        real t2 = (b.mx.x - ray.o.x) * ray.inv.x;
        real t3 = (b.mn.y - ray.o.y) * ray.inv.y;
        real t4 = (b.mx.y - ray.o.y) * ray.inv.y;
        real t5 = (b.mn.z - ray.o.z) * ray.inv.z;
        real t6 = (b.mx.z - ray.o.z) * ray.inv.z;
        real minA = nmin(t1, t2), minB = nmin(t3, t4), minC = nmin(t5, t6);
        real maxA = nmax(t1, t2), maxB = nmax(t3, t4), maxC = nmax(t5, t6);
        real tMin = nmax(nmax(minA, minB), minC);
        real tMax = nmin(nmin(maxA, maxB), maxC);
        if (tMax < 0.0f || tMin > tMax) return false;
        bool inside = tMin < 0.0f;
        real t = inside ? tMax : tMin;
        V3 hit = ray_at(ray, t);
        V3 norm; V4 tan; V2 face;
        if (minA >= minB && minA >= minC) {
            bool neg = t1 <= t2;
            norm = neg ? v3(-1, 0, 0) : v3(1, 0, 0);
            tan = V4{0, 1, 0, neg ? -1.0f : 1.0f};
            face = v2((hit.y - b.mn.y) / (b.mx.y - b.mn.y), (hit.z - b.mn.z) / (b.mx.z - b.mn.z));
        } else if (minB >= minA && minB >= minC) {
            bool neg = t3 <= t4;
            norm = neg ? v3(0, -1, 0) : v3(0, 1, 0);
            tan = V4{1, 0, 0, neg ? 1.0f : -1.0f};
            face = v2((hit.x - b.mn.x) / (b.mx.x - b.mn.x), (hit.z - b.mn.z) / (b.mx.z - b.mn.z));
        } else {
            bool neg = t5 <= t6;
            norm = neg ? v3(0, 0, -1) : v3(0, 0, 1);
            tan = V4{1, 0, 0, neg ? -1.0f : 1.0f};
            face = v2((hit.x - b.mn.x) / (b.mx.x - b.mn.x), (hit.y - b.mn.y) / (b.mx.y - b.mn.y));
        }
        V3 n = inside ? -norm : norm;
        out = ShapeHit{t, n, n, tan, face};
        return true;
    }
    // Object local bounds: Mesh.Bounds (mesh.cs:48), Sphere.Bounds (math.cs:1192-1196), AABox.
    AABox local_bounds(const Object& o) const {
        if (o.kind == 0) return meshes[o.mesh].bvh.root_bounds();
        if (o.kind == 1) {
            V3 r = v3s(o.sphereRadius);
            return AABox{o.sphereCenter - r, o.sphereCenter + r};
        }
        return o.box;
    }
    // Object ctor bounds: AABox.Transform (math.cs:974-975) then Box.Bounds (math.cs:1101-1114).
    AABox object_bounds(const Object& o) const {
        AABox lb = local_bounds(o);
        const Transform& tr = o.trans;
        V3 center = tr.pos + qrot(tr.rot, box_center(lb) * tr.scale);
        AABox local = box_from_center(center, box_size(lb) * tr.scale);
        V3 c = box_center(local);
        V3 h = box_size(local) * 0.5f;
        V3 ax = qrot(tr.rot, v3(h.x, 0, 0));
        V3 ay = qrot(tr.rot, v3(0, h.y, 0));
        V3 az = qrot(tr.rot, v3(0, 0, h.z));
        V3 ext = v3(std::fabs(ax.x) + std::fabs(ay.x) + std::fabs(az.x),
                    std::fabs(ax.y) + std::fabs(ay.y) + std::fabs(az.y),
                    std::fabs(ax.z) + std::fabs(ay.z) + std::fabs(az.z));
        return box_from_center(c, ext * 2.0f);
    }
    // Object.Intersect -> ShapeExtensions.Intersect, scene.cs:227; math.cs:67-79.
    bool object_intersect(const Object& o, const Ray& ray, Counters* ctr, ShapeHit& out, int& prim) const {
        real scale;
        Ray local = transform_ray_inv(o.trans, ray, scale);
        ShapeHit h;
        prim = -1;
        bool hit;
        if (o.kind == 0) hit = meshes[o.mesh].intersect(local, ctr, h, prim);
        else if (o.kind == 1) hit = sphere_intersect(o.sphereCenter, o.sphereRadius, local, h);
        else hit = aabox_intersect(o.box, local, h);
        if (!hit) return false;
        const Transform& tr = o.trans;
        out.norm = normalize(qrot(tr.rot, h.norm / tr.scale));
        out.normGeo = normalize(qrot(tr.rot, h.normGeo / tr.scale));
        V3 wt = normalize(qrot(tr.rot, v3(h.tan.x, h.tan.y, h.tan.z)));
        out.tan = V4{wt.x, wt.y, wt.z, h.tan.w};
        out.uv = h.uv;
        out.dist = h.dist / scale;
        return true;
    }
    // Scene.Build, scene.cs:517-541.
    bool build() {
        if (built) return true;
        built = true;
        if (sky.kind == 0) return false;
        bvh.shapeBounds.resize(objects.size());
        for (size_t i = 0; i != objects.size(); ++i) {
            objects[i].bounds = object_bounds(objects[i]);
            bvh.shapeBounds[i] = objects[i].bounds;
        }
        bvh.build(32, 1.0f, 10.0f);
        return true;
    }
    struct SceneHit { real dist; ShapeHit sh; int prim; };
    bool scene_intersect(const Ray& ray, Counters* ctr, ShapeHit& hit, int& objIdx, int& prim) const {
        struct H { real dist; ShapeHit sh; int prim; };
        H best{};
        auto test = [&](int i, H& h) {
            if (!object_intersect(objects[i], ray, ctr, h.sh, h.prim)) return false;
            h.dist = h.sh.dist;
            return true;
        };
        if (!bvh.intersect(ray, ctr, test, best, objIdx)) return false;
        hit = best.sh;
        prim = best.prim;
        return true;
    }
    Color sample_albedo(const Material& m, V2 uv) const {
        Color t = m.colorTex >= 0 ? textures[m.colorTex].sample(uv) : kWhite;
        return t * m.albedo;
    }
    // ShadowTransmittance, scene.cs:576-600.
    Color shadow_transmittance(Ray ray, Counters* ctr) const {
        Color tr = kWhite;
        do {
            if (ctr) ctr->v[C_SceneOcclude]++;
            ShapeHit hit;
            int idx, prim;
            if (!scene_intersect(ray, ctr, hit, idx, prim)) return tr;
            const Material& mat = objects[idx].mat;
            if (mat.transparency <= 0.0f) return kBlack;
            Color albedo = sample_albedo(mat, hit.uv);
            real nDotV = std::fabs(dot(hit.normGeo, ray.d));
            Color f = fresnel(nDotV, base_reflectivity(mat.ior, albedo, mat.metallic));
            tr = tr * (albedo * mat.transparency * (kWhite - f));
            ray = make_ray(ray_at(ray, hit.dist + nmax(1e-4f, hit.dist * 1e-4f)), ray.d);
        } while (max_component(tr) > 1e-4f);
        return tr;
    }
    // Scene.Trace, scene.cs:602-625.
    bool trace(const Ray& ray, Counters* ctr, Surface& surf, real& dist, int* objOut, int* primOut) const {
        if (ctr) ctr->v[C_SceneTrace]++;
        ShapeHit hit;
        int idx, prim;
        if (!scene_intersect(ray, ctr, hit, idx, prim)) return false;
        if (objOut) *objOut = idx;
        if (primOut) *primOut = prim;
        const Material& mat = objects[idx].mat;
        Color albedo = sample_albedo(mat, hit.uv);
        real roughness = (mat.roughTex >= 0 ? textures[mat.roughTex].sample(hit.uv).x : 1.0f) * mat.roughness;
        real metallic = (mat.metalTex >= 0 ? textures[mat.metalTex].sample(hit.uv).x : 1.0f) * mat.metallic;
        V3 normal = hit.norm;
        if (mat.normalTex >= 0)
            normal = normalize_or(from_tangent_space(textures[mat.normalTex].sample_normal(hit.uv), hit.tan, hit.norm), hit.norm);
        V3 tx = v3(hit.tan.x, hit.tan.y, hit.tan.z);
        V3 tdir = normalize_or(tx - dot(tx, normal) * normal, tx);
        surf = Surface{mat.radiance, albedo, roughness, metallic, mat.transparency, mat.ior,
                       normal, hit.normGeo, V4{tdir.x, tdir.y, tdir.z, hit.tan.w}, hit.uv};
        dist = hit.dist;
        return true;
    }
    // SampleSkyLight, scene.cs:758-772.
    bool sample_sky_light(V3 pos, Rng& rng, Counters* ctr, SampleLight& light, Color& tr) const {
        if (!sky.light_dir_rand(rng, light) || light.pdf <= 0.0f) return false;
        Ray shadow = make_ray(pos, light.dir);
        tr = shadow_transmittance(shadow, ctr);
        if (max_component(tr) <= 0.0f) return false;
        if (hasFog) tr = tr * std::exp(-fog_depth(shadow, 0.0f, kInf));
        return true;
    }
    // SampleSkyDirect, scene.cs:774-794.
    Color sample_sky_direct(const Surface& s, V3 hitPos, V3 viewDir, Rng& rng, Counters* ctr) const {
        SampleLight light;
        Color tr;
        if (!sample_sky_light(hitPos, rng, ctr, light, tr)) return kBlack;
        V3 n = dot(s.normal, viewDir) >= 0.0f ? s.normal : -s.normal;
        if (dot(n, light.dir) <= 0.0f) return kBlack;
        V3 ng = dot(s.normalGeo, viewDir) >= 0.0f ? s.normalGeo : -s.normalGeo;
        if (dot(ng, light.dir) <= 0.0f) return kBlack;
        Color refl = s.opaque_eval(viewDir, light.dir);
        real surfPdf = s.opaque_pdf(viewDir, light.dir);
        real w = power_heuristic(light.pdf, surfPdf);
        return sky.radiance(light.dir) * tr * refl * w / light.pdf;
    }
    // SampleSkyDirectFog, scene.cs:796-810.
    Color sample_sky_direct_fog(V3 pos, V3 rayDir, Rng& rng, Counters* ctr) const {
        SampleLight light;
        Color tr;
        if (!sample_sky_light(pos, rng, ctr, light, tr)) return kBlack;
        real rDotL = dot(rayDir, light.dir);
        real fogPdf = hg_pdf(rDotL, fog.anisotropy);
        real w = power_heuristic(light.pdf, fogPdf);
        return sky.radiance(light.dir) * tr * fogPdf * w / light.pdf;
    }
    // Adjoint of one camera path (no reference counterpart): adds to gMat (4 per object: albedo
    // rgb, roughness) and gEnv (3 per sky texel) the gradient of sum_c delta_c * L_c, L being this
    // path's radiance as Sample computes it, i.e. of the estimator's expectation: a scatter's BSDF
    // factor scales every contribution after it, so its gradient is delta * (radiance gathered
    // downstream) * (dBSDF / BSDF + d log q - d log p), q being the density of the directions
    // OpaqueScatter draws (ClampToGeoSide folds mirrored ones) and p the OpaquePdf divided by; the
    // MIS weights are differentiated through both pdfs. Russian-roulette survival, the light pmf
    // (whose score cancels the 1/pdf derivative) and the indirect clamp's scale are held fixed.
    // The per-vertex contributions are kept and summed from the end of the path. Opaque, no fog.
    void sample_adjoint(Ray ray, Rng& rng, uint32_t bounces, real indirectClamp, Color delta, real* gMat,
                        real* gEnv, real& gTotal) const {
        Color energy = kWhite;
        bool hasScatterPdf = false;
        real scatterPdf = 0;
        int prevObj = -1;          // material of the scatter that produced the current ray
        Color prevDpA = kBlack;    // its d pdf / d albedo, d pdf / d roughness (texture-scaled)
        real prevDpR = 0;
        std::vector<int> vObj;
        std::vector<Color> vRa, vRr, vC;
        std::vector<bool> vScat;
        for (uint32_t i = 0; i != bounces + 1; ++i) {
            Surface surf;
            real dist = 0;
            int obj = -1, prim = -1;
            bool hit = trace(ray, nullptr, surf, dist, &obj, &prim);
            if (!hit) {
                real mis = 1.0f;
                Color skyD = sky.radiance(ray.d);
                if (hasScatterPdf) {
                    real lp = 0.0f;
                    if (!sky.light_dir(ray.d, lp)) lp = 0.0f;
                    mis = power_heuristic(scatterPdf, lp);
                    // d w / d pb of pb^2 / (pb^2 + pl^2), pb being the previous scatter's pdf.
                    real pb = scatterPdf, den = pb * pb + lp * lp;
                    real dw = den > 0 ? 2 * pb * lp * lp / (den * den) : 0;
                    // The weight also moves with the sky pdf lp (a function of the texels).
                    if (lp > 0 && den > 0) {
                        Color sw = delta * skyD * energy;
                        sky.pdf_grad(sky.light_texel(ray.d), lp, (sw.x + sw.y + sw.z) * (-2 * pb * pb * lp / (den * den)),
                                     gEnv, gTotal);
                    }
                    if (prevObj >= 0 && dw != 0) {
                        Color s = delta * skyD * energy * dw;
                        real sum = s.x + s.y + s.z;
                        gMat[prevObj * 4] += sum * prevDpA.x;
                        gMat[prevObj * 4 + 1] += sum * prevDpA.y;
                        gMat[prevObj * 4 + 2] += sum * prevDpA.z;
                        gMat[prevObj * 4 + 3] += sum * prevDpR;
                    }
                }
                Color c = skyD * mis * energy;
                sky.radiance_grad(ray.d, delta * (energy * mis), gEnv);
                vObj.push_back(-1); vRa.push_back(kBlack); vRr.push_back(kBlack); vC.push_back(c); vScat.push_back(false);
                break;
            }
            const Material& mat = objects[obj].mat;
            Color tA = mat.colorTex >= 0 ? textures[mat.colorTex].sample(surf.uv) : kWhite;
            real tR = mat.roughTex >= 0 ? textures[mat.roughTex].sample(surf.uv).x : 1.0f;
            V3 viewDir = -ray.d;
            real distBias = nmax(1e-4f, dist * 1e-4f);
            V3 hitPos = ray_at(ray, dist - distBias);
            Color c = surf.radiance * energy;
            SampleLight light;
            Color tr;
            if (sample_sky_light(hitPos, rng, nullptr, light, tr)) {
                V3 n = dot(surf.normal, viewDir) >= 0.0f ? surf.normal : -surf.normal;
                V3 ng = dot(surf.normalGeo, viewDir) >= 0.0f ? surf.normalGeo : -surf.normalGeo;
                if (dot(n, light.dir) > 0.0f && dot(ng, light.dir) > 0.0f) {
                    Color refl = surf.opaque_eval(viewDir, light.dir);
                    real sp = surf.specular_probability(viewDir);
                    real pb = surf.opaque_pdf_sp(viewDir, light.dir, sp);
                    real w = power_heuristic(light.pdf, pb);
                    Color skyL = sky.radiance(light.dir);
                    c = c + skyL * tr * refl * w / light.pdf * energy;
                    Color dA, dR, dpA;
                    real dpR;
                    surf.opaque_eval_grad(viewDir, light.dir, dA, dR);
                    surf.opaque_pdf_grad(viewDir, light.dir, sp, dpA, dpR);
                    // d w / d pb of the power heuristic pl^2 / (pl^2 + pb^2).
                    real pl2 = light.pdf * light.pdf, den = pl2 + pb * pb;
                    real dw = den > 0 ? -2 * pl2 * pb / (den * den) : 0;
                    if (den > 0) {
                        Color sw = delta * skyL * tr * refl * (1 / light.pdf) * energy;
                        sky.pdf_grad(light.texel, light.pdf, (sw.x + sw.y + sw.z) * (2 * light.pdf * pb * pb / (den * den)),
                                     gEnv, gTotal);
                    }
                    Color common = delta * skyL * tr * (1 / light.pdf) * energy;
                    Color gA = dA * w + refl * dpA * dw, gR = dR * w + refl * (dpR * dw);
                    gMat[obj * 4] += common.x * gA.x * tA.x;
                    gMat[obj * 4 + 1] += common.y * gA.y * tA.y;
                    gMat[obj * 4 + 2] += common.z * gA.z * tA.z;
                    gMat[obj * 4 + 3] += (common.x * gR.x + common.y * gR.y + common.z * gR.z) * tR;
                    sky.radiance_grad(light.dir, delta * tr * refl * (w / light.pdf) * energy, gEnv);
                }
            }
            vObj.push_back(obj); vC.push_back(c);
            if (russian_roulette(energy, i, rng, nullptr)) {
                vRa.push_back(kBlack); vRr.push_back(kBlack); vScat.push_back(false);
                break;
            }
            Scatter sc = surf.scatter(ray.d, viewDir, rng);
            Color f = surf.opaque_eval(viewDir, sc.dir);
            Color dA, dR, sA;
            real sR;
            surf.opaque_eval_grad(viewDir, sc.dir, dA, dR);
            surf.scatter_score(viewDir, sc.dir, surf.specular_probability(viewDir), sA, sR);
            auto ratio = [](real d, real v) { return v > 0 ? d / v : (real)0; };
            vRa.push_back(v3((ratio(dA.x, f.x) + sA.x) * tA.x, (ratio(dA.y, f.y) + sA.y) * tA.y,
                             (ratio(dA.z, f.z) + sA.z) * tA.z));
            vRr.push_back(v3((ratio(dR.x, f.x) + sR) * tR, (ratio(dR.y, f.y) + sR) * tR, (ratio(dR.z, f.z) + sR) * tR));
            vScat.push_back(true);
            energy = energy * sc.weight;
            energy = clamp_luminance(energy, indirectClamp);
            hasScatterPdf = sc.hasPdf;
            scatterPdf = sc.pdf;
            {
                Color dpA;
                real dpR;
                surf.opaque_pdf_grad(viewDir, sc.dir, surf.specular_probability(viewDir), dpA, dpR);
                prevObj = obj;
                prevDpA = dpA * tA;
                prevDpR = dpR * tR;
            }
            ray = make_ray(ray_at(ray, dist) + sc.dir * distBias, sc.dir);
        }
        Color after = kBlack;  // radiance gathered after the current vertex's scatter
        for (size_t j = vC.size(); j-- > 0;) {
            if (vScat[j]) {
                Color s = delta * after;
                int o = vObj[j];
                gMat[o * 4] += s.x * vRa[j].x;
                gMat[o * 4 + 1] += s.y * vRa[j].y;
                gMat[o * 4 + 2] += s.z * vRa[j].z;
                gMat[o * 4 + 3] += s.x * vRr[j].x + s.y * vRr[j].y + s.z * vRr[j].z;
            }
            after = after + vC[j];
        }
    }
    // RussianRoulette, scene.cs:744-756.
    static bool russian_roulette(Color& energy, uint32_t bounce, Rng& rng, Counters* ctr) {
        if (bounce < 3) return false;
        real survive = nmin(1.0f, max_component(energy));
        if ((real)rng.next_float() >= survive) {
            if (ctr) ctr->v[C_SampleTerminate]++;
            return true;
        }
        energy = energy / survive;
        return false;
    }
    // Scene.Sample, scene.cs:627-741.
    Fragment sample(Ray ray, Rng& rng, uint32_t bounces, real indirectClamp, Counters* ctr) const {
        if (ctr) ctr->v[C_Sample]++;
        Color radiance = kBlack, radianceFog = kBlack, energy = kWhite;
        Fragment fr{};
        bool hasScatterPdf = false;
        real scatterPdf = 0;
        bool hasMedium = false;
        Color mediumAlbedo = kBlack;
        for (uint32_t i = 0; i != bounces + 1; ++i) {
            if (ctr) ctr->v[C_SampleBounce]++;
            bool primary = i == 0;
            Surface surf;
            real dist = 0;
            bool hit = trace(ray, ctr, surf, dist, nullptr, nullptr);
            real scatterDist = 0;
            if (hasFog && fog_scatter(ray, 0.0f, hit ? dist : kInf, rng, scatterDist)) {
                if (ctr) ctr->v[C_SampleFogScatter]++;
                energy = energy * fog.color;
                Color sr = sample_sky_direct_fog(ray_at(ray, scatterDist), ray.d, rng, ctr) * energy;
                if (primary) radianceFog = radianceFog + sr;
                else radiance = radiance + sr;
                if (russian_roulette(energy, i, rng, ctr)) break;
                V3 sd = hg_scatter_dir(ray.d, fog.anisotropy, rng);
                hasScatterPdf = true;
                scatterPdf = hg_pdf(dot(ray.d, sd), fog.anisotropy);
                ray = make_ray(ray_at(ray, scatterDist), sd);
            } else if (hit) {
                if (ctr) {
                    ctr->v[C_SampleHit]++;
                    if (hasFog) ctr->v[C_SampleFogEscape]++;
                }
                V3 viewDir = -ray.d;
                real distBias = nmax(1e-4f, dist * 1e-4f);
                V3 hitPos = ray_at(ray, dist - distBias);
                if (hasMedium) energy = energy * beer_law(mediumAlbedo, dist);
                radiance = radiance + surf.radiance * energy;
                if (i == 0) {
                    fr.hasNormal = true; fr.normal = surf.normal;
                    fr.hasUv = true; fr.uv = surf.uv;
                    fr.hasDepth = true; fr.depth = dist;
                    fr.albedo = surf.albedo;
                }
                radiance = radiance + sample_sky_direct(surf, hitPos, viewDir, rng, ctr) * energy;
                if (russian_roulette(energy, i, rng, ctr)) break;
                Scatter sc = surf.scatter(ray.d, viewDir, rng);
                energy = energy * sc.weight;
                energy = clamp_luminance(energy, indirectClamp);
                bool enter = dot(sc.dir, surf.normalGeo) < 0.0f;
                if (sc.refracted) {
                    hasMedium = enter;
                    if (enter) mediumAlbedo = surf.albedo;
                } else if (!enter) {
                    hasMedium = false;
                }
                hasScatterPdf = sc.hasPdf;
                scatterPdf = sc.pdf;
                ray = make_ray(ray_at(ray, dist) + sc.dir * distBias, sc.dir);
            } else {
                if (ctr) {
                    ctr->v[C_SampleMiss]++;
                    if (hasFog) ctr->v[C_SampleFogEscape]++;
                }
                real mis = 1.0f;
                if (hasScatterPdf) {
                    real lp = 0.0f;
                    if (!sky.light_dir(ray.d, lp)) lp = 0.0f;
                    mis = power_heuristic(scatterPdf, lp);
                }
                radiance = radiance + sky.radiance(ray.d) * mis * energy;
                break;
            }
        }
        fr.radiance = radiance;
        fr.radianceFog = radianceFog;
        return fr;
    }
};

// View (math.cs:1582-1663).
struct View {
    Transform trans;
    real fov, near_;
    // View.Ray, math.cs:1613-1636. tanHalf = MathF.Tan(Fov * 0.5) is recomputed per call in the
    // reference; it is a pure function of Fov, so hoisting it does not change any bit.
    Ray ray(V2 sp, real aspect, real tanFov) const {
        real ndcX = sp.x * 2.0f - 1.0f;
        real ndcY = -sp.y * 2.0f + 1.0f;
        real th, tv;
        if (aspect >= 1.0f) { th = tanFov; tv = th / aspect; }
        else { tv = tanFov; th = tv * aspect; }
        V3 localDir = v3(ndcX * th, ndcY * tv, 1.0f);
        V3 origin = trans.pos + qrot(trans.rot, v3(0.0f, 0.0f, near_));
        V3 dir = qrot(trans.rot, normalize(localDir));
        return make_ray(origin, dir);
    }
};

struct RenderParams {
    uint32_t width, height, minSamples, maxSamples, bounces;
    real varianceThreshold, indirectClamp;
    int policy;     // 0 reference xorwow stream per pixel, 1 counter (Philox) per sample
    uint32_t seed;  // counter policy key
};

// Renderer.RelativeStdErr, renderer.cs:232-237.
inline real relative_std_err(real sum, real sumSqr, uint32_t n) {
    real mean = sum / (real)n;
    real variance = nmax(0.0f, sumSqr / (real)n - mean * mean);
    return std::sqrt(variance / (real)n) / nmax(mean, 1e-4f);
}

struct PixelOut {
    Color radiance, radianceFog;
    Color albedo;  // mean primary-hit albedo over samples with a primary hit (0 if none)
    V3 normal;
    V2 uv;
    real depth;
    uint32_t samples;
    real variance;
};

// Renderer.Render, renderer.cs:153-229.
PixelOut render_pixel(const Scene& sc, const View& view, const RenderParams& p, real tanFov, uint32_t x, uint32_t y, Counters* ctr) {
    if (ctr) ctr->v[C_Pixel]++;
    Rng rng;
    rng.seed_xy(x, y);
    real aspect = (real)((float)p.width / (float)p.height);
    Color radianceSum = kBlack, radianceFogSum = kBlack;
    real lumSum = 0, lumSumSqr = 0, lumSumSurf = 0, lumSumSurfSqr = 0;
    V3 normalSum = v3s(0), normalFallback = v3s(0);
    uint32_t normalCount = 0;
    bool hasUv = false;
    V2 uv = v2(0, 0);
    real depthSum = 0;
    uint32_t depthCount = 0;
    Color albedoSum = kBlack;
    uint32_t sampleIndex = 0;
    for (; sampleIndex != p.maxSamples; ++sampleIndex) {
        if (p.policy == 1) rng.seed_counter(p.seed, y * p.width + x, sampleIndex);
        real jx = (real)rng.next_float();
        real jy = (real)rng.next_float();
        V2 pos = v2(((real)x + jx) / (real)p.width, ((real)y + jy) / (real)p.height);
        Ray ray = view.ray(pos, aspect, tanFov);
        Fragment fr = sc.sample(ray, rng, p.bounces, p.indirectClamp, ctr);
        radianceSum = radianceSum + fr.radiance;
        radianceFogSum = radianceFogSum + fr.radianceFog;
        real lum = luminance(fr.radiance) + luminance(fr.radianceFog);
        lumSum += lum;
        lumSumSqr += lum * lum;
        real lumSurf = luminance(fr.radiance);
        lumSumSurf += lumSurf;
        lumSumSurfSqr += lumSurf * lumSurf;
        if (fr.hasNormal) {
            normalSum = normalSum + fr.normal;
            normalFallback = fr.normal;
            normalCount++;
        }
        if (fr.hasDepth) {
            depthSum += fr.depth;
            depthCount++;
            albedoSum = albedoSum + fr.albedo;
        }
        if (!hasUv && fr.hasUv) { hasUv = true; uv = fr.uv; }
        if (sampleIndex >= p.minSamples - 1) {
            if (relative_std_err(lumSum, lumSumSqr, sampleIndex + 1) < p.varianceThreshold) break;
        }
    }
    uint32_t sampleCount = sampleIndex + 1;
    PixelOut o;
    o.variance = relative_std_err(lumSumSurf, lumSumSurfSqr, sampleCount);
    o.radiance = radianceSum / (real)sampleCount;
    o.radianceFog = radianceFogSum / (real)sampleCount;
    o.normal = normalCount > 0 ? normalize_or(normalSum, normalFallback) : v3s(0);
    o.uv = hasUv ? uv : v2(0, 0);
    o.depth = depthCount > 0 ? depthSum / (real)depthCount : kInf;
    o.albedo = depthCount > 0 ? albedoSum / (real)depthCount : kBlack;
    o.samples = sampleCount;
    if (ctr) {
        ctr->v[sampleCount < p.maxSamples ? C_SamplePixelConverged : C_SamplePixelMaxed]++;
        ctr->v[C_SampleCountMin] = std::min(ctr->v[C_SampleCountMin], (int64_t)sampleCount);
        ctr->v[C_SampleCountMax] = std::max(ctr->v[C_SampleCountMax], (int64_t)sampleCount);
        ctr->v[C_SampleVarianceMax] = std::max(ctr->v[C_SampleVarianceMax], decimal_scaled((float)o.variance));
    }
    return o;
}

struct Denoise {
    real radius, strength, strengthMax, lumBoost, lumLimitInv, normalLimitInv, depthLimitInv;
    const float* albedo = nullptr;  // optional albedo guide (no reference counterpart)
    real albedoLimitInv = 0;
};

// Compositor.Filter, compose.cs:101-204 (ctor constants compose.cs:42-44).
Color filter_pixel(const Denoise& dn, const float* radiance, const float* normals, const float* depth,
                   const float* variance, int W, int H, int cx, int cy, Counters* ctr) {
    real radiusPixels = dn.radius * std::sqrt((real)(W * H));
    real radiusPixelsInv = (real)1 / (radiusPixels * radiusPixels * 2.0f);
    int kr = (int)std::ceil(radiusPixels * 3.0f);
    int ci = cy * W + cx;
    Color cr = v3(radiance[ci * 3], radiance[ci * 3 + 1], radiance[ci * 3 + 2]);
    V3 cn = v3(normals[ci * 3], normals[ci * 3 + 1], normals[ci * 3 + 2]);
    real cd = depth[ci];
    bool hasCN = mag_sqr(cn) > 0.0f;
    bool hasCD = !std::isinf(cd);
    real lum = luminance(cr);
    real lb = 1.0f + lum * dn.lumBoost;
    real dw = nmin((real)variance[ci] * lb * dn.strength, dn.strengthMax);
    if (dw < 1e-4f) {
        if (ctr) ctr->v[C_DenoiseEarlyOut]++;
        return cr;
    }
    if (ctr) {  // compose.cs:134-136
        ctr->v[C_DenoiseWeightMax] = std::max(ctr->v[C_DenoiseWeightMax], decimal_scaled((float)dw));
        ctr->v[C_DenoiseMaxLum] = std::max(ctr->v[C_DenoiseMaxLum], decimal_scaled((float)lum));
        ctr->v[C_DenoiseMaxLumBoost] = std::max(ctr->v[C_DenoiseMaxLumBoost], decimal_scaled((float)lb));
    }
    real wSum = 1.0f;
    Color rSum = cr;
    for (int ky = -kr; ky <= kr; ++ky) {
        int ky2 = ky * ky;
        for (int kx = -kr; kx <= kr; ++kx) {
            if (kx == 0 && ky == 0) continue;
            real kd2 = (real)(kx * kx + ky2);
            if (kd2 > (real)(kr * kr)) continue;
            int nx = cx + kx, ny = cy + ky;
            if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
            int ni = ny * W + nx;
            Color nr = v3(radiance[ni * 3], radiance[ni * 3 + 1], radiance[ni * 3 + 2]);
            V3 nn = v3(normals[ni * 3], normals[ni * 3 + 1], normals[ni * 3 + 2]);
            real nd = depth[ni];
            bool hasNN = mag_sqr(nn) > 0.0f;
            bool hasND = !std::isinf(nd);
            if (hasCN != hasNN || hasCD != hasND) {
                if (ctr) ctr->v[C_DenoiseRejectMissing]++;
                continue;
            }
            real w = dw * std::exp(-kd2 * radiusPixelsInv);
            if (hasCN && w > 0.0f) {
                V3 delta = cn - nn;
                w *= std::exp(-mag_sqr(delta) * dn.normalLimitInv);
                if (ctr && w < 1e-4f) ctr->v[C_DenoiseRejectNormal]++;
            }
            if (hasCD && w > 0.0f) {
                real dd = cd - nd;
                w *= std::exp(-(dd * dd) * dn.depthLimitInv);
                if (ctr && w < 1e-4f) ctr->v[C_DenoiseRejectDepth]++;
            }
            // Albedo guide (an extension): after the depth term, on surface pixels only.
            if (dn.albedo && hasCN && w > 0.0f) {
                V3 da = v3(dn.albedo[ci * 3], dn.albedo[ci * 3 + 1], dn.albedo[ci * 3 + 2]) -
                        v3(dn.albedo[ni * 3], dn.albedo[ni * 3 + 1], dn.albedo[ni * 3 + 2]);
                w *= std::exp(-mag_sqr(da) * dn.albedoLimitInv);
            }
            if (w > 0.0f) {
                real ld = nmax(0.0f, luminance(nr) - lum);
                w *= std::exp(-(ld * ld) * dn.lumLimitInv);
                if (ctr && w < 1e-4f) ctr->v[C_DenoiseRejectLum]++;
            }
            wSum += w;
            rSum = rSum + nr * w;
            if (ctr) ctr->v[C_FilterSample]++;
        }
    }
    return rSum / wSum;
}

inline V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }
inline Transform ld_transform(const float* t) {
    return Transform{v3(t[0], t[1], t[2]), Quat{t[3], t[4], t[5], t[6]}, v3(t[7], t[8], t[9])};
}
inline Material ld_material(const float* m, const int32_t* tex) {
    Material mat;
    mat.albedo = ld3(m);
    mat.roughness = m[3]; mat.metallic = m[4]; mat.transparency = m[5]; mat.ior = m[6];
    mat.radiance = ld3(m + 7);
    if (tex) { mat.colorTex = tex[0]; mat.roughTex = tex[1]; mat.metalTex = tex[2]; mat.normalTex = tex[3]; }
    return mat;
}
inline Triangle ld_triangle(const float* f) {
    Triangle t;
    real* d = (real*)&t;
    for (int i = 0; i != 36; ++i) d[i] = (real)f[i];
    return t;
}
inline void st_triangle(const Triangle& t, float* f) {
    const real* d = (const real*)&t;
    for (int i = 0; i != 36; ++i) f[i] = (float)d[i];
}

struct Vec3Key { int64_t x, y, z; bool operator==(const Vec3Key& o) const { return x == o.x && y == o.y && z == o.z; } };
struct Vec3KeyHash { size_t operator()(const Vec3Key& k) const { return (size_t)(k.x * 73856093LL ^ k.y * 19349663LL ^ k.z * 83492791LL); } };

}  // namespace

// =================================================================================================
// C ABI used by the tests (ctypes). Float arrays in the reference's blittable layouts.
// =================================================================================================
extern "C" {

int oc_real_size() { return (int)sizeof(real); }

// Rng (math.cs:1665-1743): writes `n` successive NextUInt values for seed (a<<32|b).
void oc_rng_uints(uint32_t a, uint32_t b, uint32_t n, uint32_t* out) {
    Rng r; r.seed_xy(a, b);
    for (uint32_t i = 0; i != n; ++i) out[i] = r.next_uint();
}
void oc_rng_floats(uint32_t a, uint32_t b, uint32_t n, float* out) {
    Rng r; r.seed_xy(a, b);
    for (uint32_t i = 0; i != n; ++i) out[i] = r.next_float();
}
// Philox4x32-10 block for known-answer tests: ctr[4] in, out[4].
void oc_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    Rng::philox(c, k0, k1);
    for (int i = 0; i != 4; ++i) out[i] = c[i];
}
uint64_t oc_splitmix64(uint64_t state, uint32_t steps) {
    uint64_t v = 0;
    for (uint32_t i = 0; i != steps; ++i) v = Rng::split_mix64(state);
    return v;
}
void oc_rng_sphere(uint32_t a, uint32_t b, uint32_t n, float* out) {
    Rng r; r.seed_xy(a, b);
    for (uint32_t i = 0; i != n; ++i) {
        V3 v = rand_on_sphere(r);
        out[i * 3] = (float)v.x; out[i * 3 + 1] = (float)v.y; out[i * 3 + 2] = (float)v.z;
    }
}

// TriangleLean.Intersect for a batch of rays against one triangle; out = (hit, t, u, v).
void oc_tri_intersect(const float* tri9, const float* rays6, uint32_t n, float* out4) {
    TriLean t{ld3(tri9), ld3(tri9 + 3), ld3(tri9 + 6)};
    for (uint32_t i = 0; i != n; ++i) {
        Ray r = make_ray(ld3(rays6 + i * 6), ld3(rays6 + i * 6 + 3));
        LeanHit h{};
        bool ok = tri_intersect(t, r, h);
        out4[i * 4] = ok ? 1.0f : 0.0f;
        out4[i * 4 + 1] = (float)h.dist; out4[i * 4 + 2] = (float)h.uv.x; out4[i * 4 + 3] = (float)h.uv.y;
    }
}
// AABox.IntersectDist; out = (hit, t).
void oc_box_intersect(const float* box6, const float* rays6, uint32_t n, float* out2) {
    AABox b{ld3(box6), ld3(box6 + 3)};
    for (uint32_t i = 0; i != n; ++i) {
        Ray r = make_ray(ld3(rays6 + i * 6), ld3(rays6 + i * 6 + 3));
        real t = 0;
        bool ok = box_intersect_dist(b, r, t);
        out2[i * 2] = ok ? 1.0f : 0.0f;
        out2[i * 2 + 1] = (float)t;
    }
}

// Mesh.Intersect through the mesh BVH (bvh.cs:165-216) for a batch of rays; idx = -1 on a miss.
void oc_mesh_intersect(void* s, int mesh, const float* rays6, uint32_t n, int32_t* idxOut, float* tOut) {
    const Mesh& m = ((Scene*)s)->meshes[mesh];
    for (uint32_t i = 0; i != n; ++i) {
        Ray r = make_ray(ld3(rays6 + i * 6), ld3(rays6 + i * 6 + 3));
        ShapeHit h{};
        int idx = -1;
        if (m.intersect(r, nullptr, h, idx)) { idxOut[i] = idx; tOut[i] = (float)h.dist; }
        else { idxOut[i] = -1; tOut[i] = INFINITY; }
    }
}

// Brdf functions evaluated elementwise (brdf.cs): which 0 GgxD, 1 SmithG1, 2 PowerHeuristic, 3 HgPdf.
void oc_brdf_scalar(int which, const float* a, const float* b, uint32_t n, float* out) {
    for (uint32_t i = 0; i != n; ++i) {
        real r = 0;
        if (which == 0) r = ggx_d(a[i], b[i]);
        else if (which == 1) r = smith_g1(a[i], b[i]);
        else if (which == 2) r = power_heuristic(a[i], b[i]);
        else if (which == 3) r = hg_pdf(a[i], b[i]);
        out[i] = (float)r;
    }
}
// GgxSampleLocal (brdf.cs:45-51) for given (u.x,u.y) pairs.
void oc_ggx_sample(float r2, const float* u2, uint32_t n, float* out3) {
    for (uint32_t i = 0; i != n; ++i) {
        V3 h = ggx_sample_local(r2, v2(u2[i * 2], u2[i * 2 + 1]));
        out3[i * 3] = (float)h.x; out3[i * 3 + 1] = (float)h.y; out3[i * 3 + 2] = (float)h.z;
    }
}
// HgScatterDir (brdf.cs:96-122) from seed (a,b): n directions around `forward`.
void oc_hg_scatter(const float* fwd, float g, uint32_t a, uint32_t b, uint32_t n, float* out3) {
    Rng r; r.seed_xy(a, b);
    for (uint32_t i = 0; i != n; ++i) {
        V3 d = hg_scatter_dir(ld3(fwd), g, r);
        out3[i * 3] = (float)d.x; out3[i * 3 + 1] = (float)d.y; out3[i * 3 + 2] = (float)d.z;
    }
}
// Cdf1 (cdf.cs:7-57): builds from weights, writes normalised cdf (n+1) and samples for `u`.
float oc_cdf1(const float* w, uint32_t n, float* cdfOut, const float* u, uint32_t m, uint32_t* idxOut) {
    std::vector<real> ww(w, w + n);
    Cdf1 c; c.build(ww.data(), (int)n);
    for (uint32_t i = 0; i != n + 1; ++i) cdfOut[i] = (float)c.cdf[i];
    for (uint32_t i = 0; i != m; ++i) idxOut[i] = c.sample(u[i]);
    return (float)c.total;
}

// Triangle ctors (math.cs:1508-1531) + Mesh.ComputeSmoothNormals (mesh.cs:99-123).
// pos: n*9; nrm: n*9 or null; uv: n*6 or null (null -> the OBJ defaults (0,0),(1,0),(0,1)).
// smooth != 0 and nrm == null applies ComputeSmoothNormals. out: n*36 (Triangle layout).
void oc_triangles_make(const float* pos, const float* nrm, const float* uv, uint32_t n, int smooth, float* out) {
    std::vector<Triangle> tris(n);
    for (uint32_t i = 0; i != n; ++i) {
        const float* p = pos + i * 9;
        V2 ua = uv ? v2(uv[i * 6], uv[i * 6 + 1]) : v2(0, 0);
        V2 ub = uv ? v2(uv[i * 6 + 2], uv[i * 6 + 3]) : v2(1, 0);
        V2 uc = uv ? v2(uv[i * 6 + 4], uv[i * 6 + 5]) : v2(0, 1);
        if (nrm)
            tris[i] = make_triangle_full(ld3(p), ld3(p + 3), ld3(p + 6), ld3(nrm + i * 9), ld3(nrm + i * 9 + 3),
                                         ld3(nrm + i * 9 + 6), ua, ub, uc);
        else
            tris[i] = make_triangle_uv(ld3(p), ld3(p + 3), ld3(p + 6), ua, ub, uc);
    }
    if (smooth && !nrm) {
        const real invEps = 1.0f / 1e-6f;
        auto key = [&](V3 v) {
            return Vec3Key{(int64_t)std::nearbyint(v.x * invEps), (int64_t)std::nearbyint(v.y * invEps),
                           (int64_t)std::nearbyint(v.z * invEps)};
        };
        std::unordered_map<Vec3Key, V3, Vec3KeyHash> acc;
        acc.reserve(n * 2);
        for (uint32_t i = 0; i != n; ++i) {
            const TriLean& L = tris[i].lean;
            V3 c = cross(L.ab, L.ac);
            V3 nw = tri_normal(L) * (std::sqrt(mag_sqr(c)) * 0.5f);
            V3 pa = L.a, pb = L.a + L.ab, pc = L.a + L.ac;
            for (V3 p : {pa, pb, pc}) {
                auto it = acc.find(key(p));
                if (it == acc.end()) acc.emplace(key(p), v3s(0) + nw);
                else it->second = it->second + nw;
            }
        }
        for (uint32_t i = 0; i != n; ++i) {
            Triangle& t = tris[i];
            V3 pa = t.lean.a, pb = t.lean.a + t.lean.ab, pc = t.lean.a + t.lean.ac;
            V3 fn = tri_normal(t.lean);
            V3 na = normalize_or(acc[key(pa)], fn);
            V3 nb = normalize_or(acc[key(pb)], fn);
            V3 nc = normalize_or(acc[key(pc)], fn);
            t = make_triangle_full(pa, pb, pc, na, nb, nc, t.uvA, t.uvB, t.uvC);
        }
    }
    for (uint32_t i = 0; i != n; ++i) st_triangle(tris[i], out + i * 36);
}

// ---- scene handle ----
void* oc_scene_new() { return new Scene(); }
void oc_scene_free(void* s) { delete (Scene*)s; }

// Texture from float RGB (texels w*h*3); filter 0 bilinear / 1 point. Returns the texture id.
int oc_scene_add_texture(void* s, const float* rgb, uint32_t w, uint32_t h, int filter, float tilX, float tilY) {
    Scene* sc = (Scene*)s;
    Texture t;
    t.w = w; t.h = h; t.filter = filter; t.tiling = v2(tilX, tilY);
    t.texels.resize((size_t)w * h);
    for (size_t i = 0; i != (size_t)w * h; ++i) t.texels[i] = ld3(rgb + i * 3);
    sc->textures.push_back(std::move(t));
    return (int)sc->textures.size() - 1;
}
// Mesh from Triangle records (n*36 floats); bins = mesh BVH split bins. Returns mesh id or -1 if
// the build hit a degenerate split.
int oc_scene_add_mesh(void* s, const float* tris36, uint32_t n, int bins) {
    Scene* sc = (Scene*)s;
    Mesh m;
    m.tris.resize(n);
    for (uint32_t i = 0; i != n; ++i) m.tris[i] = ld_triangle(tris36 + i * 36);
    m.build(bins);
    sc->meshes.push_back(std::move(m));
    return sc->meshes.back().bvh.degenerate ? -1 : (int)sc->meshes.size() - 1;
}
// Object (scene.cs:214-223). transform: pos3 quat4 scale3. material: albedo3 rough metal transp ior
// radiance3. tex4: color/rough/metal/normal texture ids (-1 none) or null. kind 0 mesh (shape =
// mesh id), 1 sphere (shapeParams = center3 radius), 2 aabox (shapeParams = min3 max3).
int oc_scene_add_object(void* s, const float* transform, const float* material, const int32_t* tex4, int kind,
                        int shape, const float* shapeParams) {
    Scene* sc = (Scene*)s;
    Object o{};
    o.trans = ld_transform(transform);
    o.mat = ld_material(material, tex4);
    o.kind = kind;
    o.mesh = shape;
    if (kind == 1) { o.sphereCenter = ld3(shapeParams); o.sphereRadius = shapeParams[3]; }
    if (kind == 2) { o.box = AABox{ld3(shapeParams), ld3(shapeParams + 3)}; }
    sc->objects.push_back(o);
    return (int)sc->objects.size() - 1;
}
void oc_scene_set_sky_texture(void* s, int texId, float multiplier, float yaw) {
    Scene* sc = (Scene*)s;
    sc->sky.kind = 1;
    sc->sky.tex = sc->textures[texId];
    sc->skyTexId = texId;
    sc->sky.multiplier = multiplier;
    sc->sky.yaw = yaw;
    sc->sky.init_texture();
}
// SkyProcedural (scene.cs:341-352): top3 middle3 bottom3 sunDir3 sunRadiance3 sunAngle.
void oc_scene_set_sky_procedural(void* s, const float* p16) {
    Scene* sc = (Scene*)s;
    sc->sky.kind = 2;
    sc->sky.top = ld3(p16); sc->sky.middle = ld3(p16 + 3); sc->sky.bottom = ld3(p16 + 6);
    sc->sky.sunDir = ld3(p16 + 9); sc->sky.sunRadiance = ld3(p16 + 12); sc->sky.sunAngle = p16[15];
    sc->sky.init_procedural();
}
void oc_scene_set_fog(void* s, float density, const float* color, float anisotropy, float heightFalloff) {
    Scene* sc = (Scene*)s;
    sc->hasFog = true;
    sc->fog = Fog{density, ld3(color), anisotropy, heightFalloff};
    sc->medium = Medium{};
}
// Bounded medium: kind 1 homogeneous, 2 grid (res3, half-precision bits, x fastest).
void oc_scene_set_fog_volume(void* s, int kind, float sigma, const float* color, float anisotropy, const float* bmin,
                             const float* bmax, const uint32_t* res, const uint16_t* density) {
    Scene* sc = (Scene*)s;
    sc->hasFog = true;
    sc->fog = Fog{sigma, ld3(color), anisotropy, 0};
    sc->medium = Medium{};
    sc->medium.init(kind, sigma, ld3(bmin), ld3(bmax), res, density);
}
// Medium queries for pinning tests: optical depth of [start, end) segments of rays, and
// free-flight draws from seed (a, b) (t, or -1 when the ray leaves the medium unscattered).
void oc_medium_optical_depth(void* s, const float* rays6, const float* se2, uint32_t n, float* out) {
    Scene* sc = (Scene*)s;
    for (uint32_t i = 0; i != n; ++i) {
        Ray r = make_ray(ld3(rays6 + i * 6), ld3(rays6 + i * 6 + 3));
        out[i] = (float)sc->fog_depth(r, se2[i * 2], se2[i * 2 + 1]);
    }
}
void oc_medium_scatter(void* s, const float* ray6, float start, float end, uint32_t a, uint32_t b, uint32_t n,
                       float* out) {
    Scene* sc = (Scene*)s;
    Ray r = make_ray(ld3(ray6), ld3(ray6 + 3));
    Rng rng; rng.seed_xy(a, b);
    for (uint32_t i = 0; i != n; ++i) {
        real t = 0;
        out[i] = sc->fog_scatter(r, start, end, rng, t) ? (float)t : -1.0f;
    }
}
int oc_scene_build(void* s) { return ((Scene*)s)->build() ? 0 : -1; }

// BVH export. level 0: the scene BVH over objects; level 1: mesh `mesh`. nodes: nodeCount*8 floats
// (min3 max3, child and shapeCount as int bit patterns), items: shape count.
int oc_bvh_export(void* s, int level, int mesh, float* nodes, int32_t* items) {
    Scene* sc = (Scene*)s;
    const Bvh& b = level == 0 ? sc->bvh : sc->meshes[mesh].bvh;
    if (nodes) {
        for (int i = 0; i != b.nodeCount; ++i) {
            const BvhNode& nd = b.nodes[i];
            float* o = nodes + i * 8;
            o[0] = (float)nd.bounds.mn.x; o[1] = (float)nd.bounds.mn.y; o[2] = (float)nd.bounds.mn.z;
            o[3] = (float)nd.bounds.mx.x; o[4] = (float)nd.bounds.mx.y; o[5] = (float)nd.bounds.mx.z;
            std::memcpy(o + 6, &nd.child, 4);
            std::memcpy(o + 7, &nd.shapeCount, 4);
        }
    }
    if (items) for (size_t i = 0; i != b.items.size(); ++i) items[i] = b.items[i];
    return b.nodeCount;
}
// Bvh.GetStats (bvh.cs:79-125): ints[7] = nodeCount depthMax leafCount leafMin leafMax shapes depthWeighted; returns sah.
float oc_bvh_stats(void* s, int level, int mesh, int32_t* ints) {
    Scene* sc = (Scene*)s;
    const Bvh& b = level == 0 ? sc->bvh : sc->meshes[mesh].bvh;
    if (b.nodeCount == 0) { std::memset(ints, 0, 7 * 4); return 0; }
    Bvh::Stats st = b.stats_node(0, 0, box_sa(b.nodes[0].bounds));
    ints[0] = st.nodeCount; ints[1] = st.depthMax; ints[2] = st.leafCount; ints[3] = st.leafMin;
    ints[4] = st.leafMax; ints[5] = st.shapeCount; ints[6] = st.depthWeighted;
    return (float)st.sah;
}

// View: pos3 quat4 fov near (9 floats). tanFov = MathF.Tan(Fov * 0.5f) is evaluated here (libm).
static View ld_view(const float* v) {
    View vw;
    vw.trans = Transform{v3(v[0], v[1], v[2]), Quat{v[3], v[4], v[5], v[6]}, v3s(1)};
    vw.fov = v[7];
    vw.near_ = v[8];
    return vw;
}

// Sample-0 primary hit per pixel (renderer.cs:156,175 + math.cs:1613-1636 + scene.cs:602-625).
// Rays for pixels [0,W) x [y0,y1). objOut/primOut = -1 on a miss; tOut = +inf on a miss.
void oc_trace_primary(void* s, const float* view9, uint32_t W, uint32_t H, uint32_t y0, uint32_t y1,
                      int32_t* objOut, int32_t* primOut, float* tOut, float* normalOut) {
    Scene* sc = (Scene*)s;
    View vw = ld_view(view9);
    real tanFov = std::tan(vw.fov * 0.5f);
    real aspect = (real)((float)W / (float)H);
    for (uint32_t y = y0; y < y1; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            size_t o = (size_t)(y - y0) * W + x;
            Rng rng; rng.seed_xy(x, y);
            real jx = (real)rng.next_float();
            real jy = (real)rng.next_float();
            V2 pos = v2(((real)x + jx) / (real)W, ((real)y + jy) / (real)H);
            Ray ray = vw.ray(pos, aspect, tanFov);
            Surface surf;
            real dist = 0;
            int obj = -1, prim = -1;
            if (sc->trace(ray, nullptr, surf, dist, &obj, &prim)) {
                objOut[o] = obj; primOut[o] = prim; tOut[o] = (float)dist;
                if (normalOut) { normalOut[o * 3] = (float)surf.normal.x; normalOut[o * 3 + 1] = (float)surf.normal.y; normalOut[o * 3 + 2] = (float)surf.normal.z; }
            } else {
                objOut[o] = -1; primOut[o] = -1; tOut[o] = INFINITY;
                if (normalOut) { normalOut[o * 3] = 0; normalOut[o * 3 + 1] = 0; normalOut[o * 3 + 2] = 0; }
            }
        }
}

// Renderer (renderer.cs:40-229): renders rows [y0,y1) of a W x H image with `threads` workers
// claiming 32x32-pixel blocks through an atomic counter (renderer.cs:107-151). Output arrays hold
// (y1-y0)*W pixels in the AOV layouts of renderer.cs:14-20. params: minS maxS bounces (u32) and
// varianceThreshold indirectClamp (f32) in rp. counters: C_Count int64 (may be null).
int oc_render(void* s, const float* view9, uint32_t W, uint32_t H, uint32_t minS, uint32_t maxS, float threshold,
              uint32_t bounces, float indirectClamp, uint32_t y0, uint32_t y1, int threads,
              float* radiance, float* radianceFog, float* normals, float* uvs, float* depth, uint32_t* samples,
              float* variance, int64_t* counters, int policy, uint32_t seed, int lightSampler, float* albedo) {
    Scene* sc = (Scene*)s;
    if (!sc->built) return -1;
    sc->sky.sampler = lightSampler;
    View vw = ld_view(view9);
    RenderParams p{W, H, minS, maxS, bounces, (real)threshold, (real)indirectClamp, policy, seed};
    real tanFov = std::tan(vw.fov * 0.5f);
    const uint32_t bs = 32;
    uint32_t rows = y1 - y0;
    uint32_t bx = (W + bs - 1) / bs, by = (rows + bs - 1) / bs, total = bx * by;
    std::atomic<uint32_t> next{0};
    std::vector<Counters> ctrs(std::max(threads, 1));
    for (auto& c : ctrs) counters_init(c);
    auto worker = [&](int wi) {
        Counters* ctr = counters ? &ctrs[wi] : nullptr;
        while (true) {
            uint32_t blk = next.fetch_add(1);
            if (blk >= total) break;
            uint32_t xmin = (blk % bx) * bs, ymin = y0 + (blk / bx) * bs;
            uint32_t xmax = std::min(xmin + bs, W), ymax = std::min(ymin + bs, y1);
            for (uint32_t y = ymin; y != ymax; ++y)
                for (uint32_t x = xmin; x != xmax; ++x) {
                    PixelOut po = render_pixel(*sc, vw, p, tanFov, x, y, ctr);
                    size_t o = (size_t)(y - y0) * W + x;
                    radiance[o * 3] = (float)po.radiance.x; radiance[o * 3 + 1] = (float)po.radiance.y; radiance[o * 3 + 2] = (float)po.radiance.z;
                    radianceFog[o * 3] = (float)po.radianceFog.x; radianceFog[o * 3 + 1] = (float)po.radianceFog.y; radianceFog[o * 3 + 2] = (float)po.radianceFog.z;
                    normals[o * 3] = (float)po.normal.x; normals[o * 3 + 1] = (float)po.normal.y; normals[o * 3 + 2] = (float)po.normal.z;
                    uvs[o * 2] = (float)po.uv.x; uvs[o * 2 + 1] = (float)po.uv.y;
                    depth[o] = (float)po.depth;
                    samples[o] = po.samples;
                    variance[o] = (float)po.variance;
                    if (albedo) {
                        albedo[o * 3] = (float)po.albedo.x; albedo[o * 3 + 1] = (float)po.albedo.y;
                        albedo[o * 3 + 2] = (float)po.albedo.z;
                    }
                }
        }
    };
    int nt = std::max(threads, 1);
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(worker, i);
    worker(0);
    for (auto& t : pool) t.join();
    if (counters) {
        Counters init;
        counters_init(init);
        std::memcpy(counters, init.v, sizeof(init.v));
        for (auto& c : ctrs) counters_merge(counters, c);
    }
    return 0;
}

// Compositor.Filter over a whole image. params: radius strength strengthMax lumBoost lumLimit
// normalLimit depthLimit (the Compositor ctor arguments, compose.cs:22-34).
void oc_filter(const float* params7, const float* radiance, const float* normals, const float* depth,
               const float* variance, uint32_t W, uint32_t H, int threads, float* out, int64_t* filterCounters,
               const float* albedo, float albedoLimit) {
    Denoise dn;
    dn.albedo = albedo;
    if (albedo) dn.albedoLimitInv = 1.0f / ((real)albedoLimit * (real)albedoLimit * 2.0f);
    dn.radius = params7[0]; dn.strength = params7[1]; dn.strengthMax = params7[2]; dn.lumBoost = params7[3];
    dn.lumLimitInv = 1.0f / ((real)params7[4] * (real)params7[4] * 2.0f);
    dn.normalLimitInv = 1.0f / ((real)params7[5] * (real)params7[5] * 2.0f);
    dn.depthLimitInv = 1.0f / ((real)params7[6] * (real)params7[6] * 2.0f);
    std::atomic<uint32_t> next{0};
    int nt = std::max(threads, 1);
    std::vector<Counters> ctrs(nt);
    for (auto& c : ctrs) counters_init(c);
    auto worker = [&](int wi) {
        while (true) {
            uint32_t y = next.fetch_add(1);
            if (y >= H) break;
            for (uint32_t x = 0; x != W; ++x) {
                Color c = filter_pixel(dn, radiance, normals, depth, variance, (int)W, (int)H, (int)x, (int)y, &ctrs[wi]);
                size_t o = (size_t)y * W + x;
                out[o * 3] = (float)c.x; out[o * 3 + 1] = (float)c.y; out[o * 3 + 2] = (float)c.z;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(worker, i);
    worker(0);
    for (auto& t : pool) t.join();
    if (filterCounters) {  // Counters of the whole image in C_Count layout
        Counters init;
        counters_init(init);
        std::memcpy(filterCounters, init.v, sizeof(init.v));
        for (auto& c : ctrs) counters_merge(filterCounters, c);
    }
}

// ---- Compose tail ----------------------------------------------------------------------------
// Compositor.FilterFog, compose.cs:206-245: circular Gaussian over the fog radiance with the centre
// weight fixed at 1 and every neighbour weighted strength * exp(-d^2 / (2 r^2)).
void oc_filter_fog(float radius, float strength, const float* fog, uint32_t W, uint32_t H, float* out) {
    const real radiusPixels = (real)radius * std::sqrt((real)(W * H));
    const real inv = (real)1 / (radiusPixels * radiusPixels * (real)2);
    const int kr = (int)std::ceil(radiusPixels * (real)3);
    for (int cy = 0; cy < (int)H; ++cy)
        for (int cx = 0; cx < (int)W; ++cx) {
            const size_t ci = (size_t)cy * W + cx;
            real wsum = 1;
            Color sum = v3(fog[ci * 3], fog[ci * 3 + 1], fog[ci * 3 + 2]);
            for (int ky = -kr; ky <= kr; ++ky) {
                const int ky2 = ky * ky;
                for (int kx = -kr; kx <= kr; ++kx) {
                    if (kx == 0 && ky == 0) continue;
                    const real kd2 = (real)(kx * kx + ky2);
                    if (kd2 > (real)(kr * kr)) continue;
                    const int nx = cx + kx, ny = cy + ky;
                    if (nx < 0 || nx >= (int)W || ny < 0 || ny >= (int)H) continue;
                    const real w = (real)strength * std::exp(-kd2 * inv);
                    const size_t ni = (size_t)ny * W + nx;
                    wsum += w;
                    sum = sum + v3(fog[ni * 3], fog[ni * 3 + 1], fog[ni * 3 + 2]) * w;
                }
            }
            const Color r = sum / wsum;
            out[ci * 3] = (float)r.x; out[ci * 3 + 1] = (float)r.y; out[ci * 3 + 2] = (float)r.z;
        }
}

// Color.ToPixel, math.cs:127-136: Clamp01 (Math.Clamp leaves NaN), the sRGB power approximation,
// MathF.Max against 0 and the byte conversion (a NaN converts to 0 through the int32 route).
static uint8_t to_srgb_byte(real c) {
    c = c < (real)0 ? (real)0 : (c > (real)1 ? (real)1 : c);
    const real e = nmax((real)1.055f * std::pow(c, (real)0.416666667f) - (real)0.055f, (real)0);
    const real v = e * (real)255 + (real)0.5f;
    return v == v ? (uint8_t)(int32_t)v : (uint8_t)0;
}

// Compositor.Tonemap, compose.cs:247-266 (0 = Reinhard, 1 = LinearSmooth), then ToPixel, for
// main + fog per pixel (Compose, compose.cs:92; Preview, compose.cs:62).
void oc_tonemap(const float* main, const float* fog, uint32_t n, int tonemapper, float exposure, uint8_t* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const Color rad = v3(main[i * 3], main[i * 3 + 1], main[i * 3 + 2]) + v3(fog[i * 3], fog[i * 3 + 1], fog[i * 3 + 2]);
        const Color x = rad * (real)exposure;
        Color t;
        if (tonemapper == 0) {
            t = x / (x + v3(1, 1, 1));
        } else {
            const real a = (real)1.8f, b = (real)1.4f, c = (real)0.5f, d = (real)1.5f;
            t = (x * (a * x + v3(b, b, b))) / (x * (a * x + v3(c, c, c)) + v3(d, d, d));
        }
        out[i * 3] = to_srgb_byte(t.x);
        out[i * 3 + 1] = to_srgb_byte(t.y);
        out[i * 3 + 2] = to_srgb_byte(t.z);
    }
}

// ---- Adjoint (no reference counterpart) -----------------------------------------------------
// Parameters: per object albedo rgb + roughness (4 floats), then the sky texture's texels (3 per
// texel, row-major) when the sky is a texture.
uint32_t oc_param_count(void* s) {
    Scene* sc = (Scene*)s;
    return (uint32_t)sc->objects.size() * 4 + (sc->sky.kind == 1 ? sc->sky.tex.w * sc->sky.tex.h * 3 : 0);
}
void oc_params_get(void* s, float* p) {
    Scene* sc = (Scene*)s;
    size_t k = 0;
    for (auto& o : sc->objects) {
        p[k++] = (float)o.mat.albedo.x; p[k++] = (float)o.mat.albedo.y; p[k++] = (float)o.mat.albedo.z;
        p[k++] = (float)o.mat.roughness;
    }
    if (sc->sky.kind == 1)
        for (auto& t : sc->sky.tex.texels) { p[k++] = (float)t.x; p[k++] = (float)t.y; p[k++] = (float)t.z; }
}
// Sets the parameters and rebuilds the sky's sampling tables (a material using the sky's texture
// sees the new texels too).
void oc_params_set(void* s, const float* p) {
    Scene* sc = (Scene*)s;
    size_t k = 0;
    for (auto& o : sc->objects) {
        o.mat.albedo = v3(p[k], p[k + 1], p[k + 2]);
        o.mat.roughness = p[k + 3];
        k += 4;
    }
    if (sc->sky.kind == 1) {
        for (auto& t : sc->sky.tex.texels) { t = v3(p[k], p[k + 1], p[k + 2]); k += 3; }
        if (sc->skyTexId >= 0) sc->textures[sc->skyTexId].texels = sc->sky.tex.texels;
        sc->sky.init_texture();
    }
}
// Gradient of sum_pixels sum_c dImage[p,c] * Radiance[p,c] for a fixed-spp render with the
// counter RNG policy (Radiance = sum of spp samples / (spp + 1), renderer.cs:173,215).
int oc_adjoint(void* s, const float* view9, uint32_t W, uint32_t H, uint32_t spp, uint32_t bounces, float clamp,
               uint32_t seed, const float* dImage, int threads, float* grad) {
    Scene* sc = (Scene*)s;
    if (!sc->built) return -1;
    for (auto& o : sc->objects)
        if (o.mat.transparency > 0.0f) return -2;
    if (sc->hasFog) return -2;
    View vw = ld_view(view9);
    real tanFov = std::tan(vw.fov * 0.5f);
    real aspect = (real)((float)W / (float)H);
    size_t nMat = sc->objects.size() * 4, n = oc_param_count(s);
    int nt = std::max(threads, 1);
    std::vector<std::vector<real>> acc(nt, std::vector<real>(n, 0));
    std::vector<real> accTotal(nt, 0);
    std::atomic<uint32_t> next{0};
    auto worker = [&](int wi) {
        real* gm = acc[wi].data();
        real* ge = gm + nMat;
        for (;;) {
            uint32_t y = next.fetch_add(1);
            if (y >= H) break;
            for (uint32_t x = 0; x != W; ++x) {
                size_t o = (size_t)y * W + x;
                Color delta = v3(dImage[o * 3], dImage[o * 3 + 1], dImage[o * 3 + 2]) / (real)(spp + 1);
                for (uint32_t smp = 0; smp != spp; ++smp) {
                    Rng rng;
                    rng.seed_counter(seed, y * W + x, smp);
                    real jx = (real)rng.next_float();
                    real jy = (real)rng.next_float();
                    V2 pos = v2(((real)x + jx) / (real)W, ((real)y + jy) / (real)H);
                    sc->sample_adjoint(vw.ray(pos, aspect, tanFov), rng, bounces, (real)clamp, delta, gm, ge,
                                       accTotal[wi]);
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nt; ++i) pool.emplace_back(worker, i);
    worker(0);
    for (auto& t : pool) t.join();
    real total = 0;
    for (int t = 0; t != nt; ++t) total += accTotal[t];
    std::vector<real> sum(n, 0);
    for (size_t i = 0; i != n; ++i)
        for (int t = 0; t != nt; ++t) sum[i] += acc[t][i];
    sc->sky.apply_total_grad(total, sum.data() + nMat);
    for (size_t i = 0; i != n; ++i) grad[i] = (float)sum[i];
    return 0;
}

// Sky queries for pinning tests: radiance(dir) and the MIS pdf LightDir(dir); pdf = 0 when null.
void oc_sky_eval(void* s, const float* dirs, uint32_t n, float* rad3, float* pdf) {
    Scene* sc = (Scene*)s;
    for (uint32_t i = 0; i != n; ++i) {
        V3 d = ld3(dirs + i * 3);
        Color c = sc->sky.radiance(d);
        rad3[i * 3] = (float)c.x; rad3[i * 3 + 1] = (float)c.y; rad3[i * 3 + 2] = (float)c.z;
        real p = 0;
        pdf[i] = sc->sky.light_dir(d, p) ? (float)p : 0.0f;
    }
}
// Sky light samples from seed (a,b): dir3 + pdf (pdf = 0 when LightDirRand returns null);
// lightSampler 0 = Cdf2, 1 = alias table.
void oc_sky_sample(void* s, uint32_t a, uint32_t b, uint32_t n, int lightSampler, float* out4) {
    Scene* sc = (Scene*)s;
    sc->sky.sampler = lightSampler;
    Rng r; r.seed_xy(a, b);
    for (uint32_t i = 0; i != n; ++i) {
        SampleLight l{v3s(0), 0};
        if (!sc->sky.light_dir_rand(r, l)) l = SampleLight{v3s(0), 0};
        out4[i * 4] = (float)l.dir.x; out4[i * 4 + 1] = (float)l.dir.y; out4[i * 4 + 2] = (float)l.dir.z;
        out4[i * 4 + 3] = (float)l.pdf;
    }
}
// Surface BSDF queries: material10 (as oc_scene_add_object), normal3, normalGeo3, tangent4;
// for each (view3, light3) pair returns eval3 + pdf.
void oc_surface_eval(const float* material, const float* n3, const float* ng3, const float* t4,
                     const float* vl6, uint32_t cnt, float* out4) {
    Material m = ld_material(material, nullptr);
    Surface s{m.radiance, m.albedo, m.roughness, m.metallic, m.transparency, m.ior,
              ld3(n3), ld3(ng3), V4{t4[0], t4[1], t4[2], t4[3]}, v2(0, 0)};
    for (uint32_t i = 0; i != cnt; ++i) {
        V3 v = ld3(vl6 + i * 6), l = ld3(vl6 + i * 6 + 3);
        Color e = s.opaque_eval(v, l);
        out4[i * 4] = (float)e.x; out4[i * 4 + 1] = (float)e.y; out4[i * 4 + 2] = (float)e.z;
        out4[i * 4 + 3] = (float)s.opaque_pdf(v, l);
    }
}
// d OpaqueEval / d albedo (3, per channel) and d OpaqueEval / d roughness (3) per (view, light) pair.
void oc_surface_eval_grad(const float* material, const float* n3, const float* ng3, const float* t4,
                          const float* vl6, uint32_t cnt, float* out6) {
    Material m = ld_material(material, nullptr);
    Surface s{m.radiance, m.albedo, m.roughness, m.metallic, m.transparency, m.ior,
              ld3(n3), ld3(ng3), V4{t4[0], t4[1], t4[2], t4[3]}, v2(0, 0)};
    for (uint32_t i = 0; i != cnt; ++i) {
        Color dA, dR;
        s.opaque_eval_grad(ld3(vl6 + i * 6), ld3(vl6 + i * 6 + 3), dA, dR);
        float* o = out6 + i * 6;
        o[0] = (float)dA.x; o[1] = (float)dA.y; o[2] = (float)dA.z;
        o[3] = (float)dR.x; o[4] = (float)dR.y; o[5] = (float)dR.z;
    }
}
// Surface.Scatter samples from seed (a,b) for incoming direction inc3: dir3 weight3 pdf refracted.
void oc_surface_scatter(const float* material, const float* n3, const float* ng3, const float* t4,
                        const float* inc3, uint32_t a, uint32_t b, uint32_t cnt, float* out8) {
    Material m = ld_material(material, nullptr);
    Surface s{m.radiance, m.albedo, m.roughness, m.metallic, m.transparency, m.ior,
              ld3(n3), ld3(ng3), V4{t4[0], t4[1], t4[2], t4[3]}, v2(0, 0)};
    Rng r; r.seed_xy(a, b);
    V3 inc = ld3(inc3);
    for (uint32_t i = 0; i != cnt; ++i) {
        Scatter sc = s.scatter(inc, -inc, r);
        float* o = out8 + i * 8;
        o[0] = (float)sc.dir.x; o[1] = (float)sc.dir.y; o[2] = (float)sc.dir.z;
        o[3] = (float)sc.weight.x; o[4] = (float)sc.weight.y; o[5] = (float)sc.weight.z;
        o[6] = sc.hasPdf ? (float)sc.pdf : -1.0f; o[7] = sc.refracted ? 1.0f : 0.0f;
    }
}
// Scene.Trace (scene.cs:602-625) of arbitrary rays (6 floats: origin, direction). Per ray 20
// floats: object (-1 on a miss), primitive, distance, shading normal 3, geometric normal 3, tangent
// 4, uv 2, albedo 3, roughness, metallic.
void oc_scene_trace(void* s, const float* rays6, uint32_t n, float* out20) {
    const Scene* sc = (const Scene*)s;
    for (uint32_t i = 0; i != n; ++i) {
        const float* r = rays6 + (size_t)i * 6;
        float* o = out20 + (size_t)i * 20;
        Ray ray = make_ray(v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]));
        Surface surf;
        real dist;
        int obj = -1, prim = -1;
        for (int k = 0; k < 20; ++k) o[k] = 0.0f;
        if (!sc->trace(ray, nullptr, surf, dist, &obj, &prim)) { o[0] = -1.0f; continue; }
        o[0] = (float)obj; o[1] = (float)prim; o[2] = (float)dist;
        o[3] = (float)surf.normal.x; o[4] = (float)surf.normal.y; o[5] = (float)surf.normal.z;
        o[6] = (float)surf.normalGeo.x; o[7] = (float)surf.normalGeo.y; o[8] = (float)surf.normalGeo.z;
        o[9] = (float)surf.tangent.x; o[10] = (float)surf.tangent.y; o[11] = (float)surf.tangent.z;
        o[12] = (float)surf.tangent.w; o[13] = (float)surf.uv.x; o[14] = (float)surf.uv.y;
        o[15] = (float)surf.albedo.x; o[16] = (float)surf.albedo.y; o[17] = (float)surf.albedo.z;
        o[18] = (float)surf.roughness; o[19] = (float)surf.metallic;
    }
}
// ShadowTransmittance (scene.cs:576-600) of rays: 3 floats each.
void oc_shadow_transmittance(void* s, const float* rays6, uint32_t n, float* out3) {
    const Scene* sc = (const Scene*)s;
    for (uint32_t i = 0; i != n; ++i) {
        const float* r = rays6 + (size_t)i * 6;
        Color t = sc->shadow_transmittance(make_ray(v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5])), nullptr);
        out3[i * 3] = (float)t.x; out3[i * 3 + 1] = (float)t.y; out3[i * 3 + 2] = (float)t.z;
    }
}
// Brdf.Refract (brdf.cs:76-92): per item (dir 3, 1 if refracted else 0).
void oc_refract(const float* inc3, const float* n3, float eta, uint32_t n, float* out4) {
    for (uint32_t i = 0; i != n; ++i) {
        V3 d;
        const bool ok = refract(ld3(inc3 + i * 3), ld3(n3 + i * 3), (real)eta, d);
        out4[i * 4] = ok ? (float)d.x : 0.0f; out4[i * 4 + 1] = ok ? (float)d.y : 0.0f;
        out4[i * 4 + 2] = ok ? (float)d.z : 0.0f; out4[i * 4 + 3] = ok ? 1.0f : 0.0f;
    }
}
// Scene.RussianRoulette (scene.cs:744-756) on n draws of the Rng(a, b) stream: per draw the energy
// after the decision (3) and 1 if terminated.
void oc_russian_roulette(const float* energy3, uint32_t bounce, uint32_t a, uint32_t b, uint32_t n, float* out4) {
    Rng rng;
    rng.seed_xy(a, b);
    for (uint32_t i = 0; i != n; ++i) {
        Color e = ld3(energy3);
        const bool term = Scene::russian_roulette(e, bounce, rng, nullptr);
        out4[i * 4] = (float)e.x; out4[i * 4 + 1] = (float)e.y; out4[i * 4 + 2] = (float)e.z;
        out4[i * 4 + 3] = term ? 1.0f : 0.0f;
    }
}
// Color.ClampLuminance (math.cs:120-124).
void oc_clamp_luminance(const float* c3, float mx, uint32_t n, float* out3) {
    for (uint32_t i = 0; i != n; ++i) {
        Color c = clamp_luminance(ld3(c3 + i * 3), (real)mx);
        out3[i * 3] = (float)c.x; out3[i * 3 + 1] = (float)c.y; out3[i * 3 + 2] = (float)c.z;
    }
}
// Brdf.BeerLawTransmittance (brdf.cs:87-92).
void oc_beer(const float* albedo3, float d, float* out3) {
    Color c = beer_law(ld3(albedo3), (real)d);
    out3[0] = (float)c.x; out3[1] = (float)c.y; out3[2] = (float)c.z;
}
// Renderer.RelativeStdErr (renderer.cs:232-237).
float oc_relative_std_err(float sum, float sumSqr, uint32_t n) {
    return (float)relative_std_err((real)sum, (real)sumSqr, n);
}
// Height fog (scene.cs:438-505): optical depth for segments [start,end) of rays.
void oc_fog_optical_depth(const float* fog6, const float* rays6, const float* se2, uint32_t n, float* out) {
    Fog f{fog6[0], ld3(fog6 + 1), fog6[4], fog6[5]};
    for (uint32_t i = 0; i != n; ++i) {
        Ray r = make_ray(ld3(rays6 + i * 6), ld3(rays6 + i * 6 + 3));
        out[i] = (float)f.optical_depth(r, se2[i * 2], se2[i * 2 + 1]);
    }
}
// Fog.ScatterDistance draws from seed (a,b) for one ray segment: t or -1 when the ray escapes.
void oc_fog_scatter(const float* fog6, const float* ray6, float start, float end, uint32_t a, uint32_t b,
                    uint32_t n, float* out) {
    Fog f{fog6[0], ld3(fog6 + 1), fog6[4], fog6[5]};
    Ray r = make_ray(ld3(ray6), ld3(ray6 + 3));
    Rng rng; rng.seed_xy(a, b);
    for (uint32_t i = 0; i != n; ++i) {
        real t = 0;
        out[i] = f.scatter_distance(r, start, end, rng, t) ? (float)t : -1.0f;
    }
}

}  // extern "C"
