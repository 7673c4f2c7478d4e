// Device-side scene layout and per-path arithmetic of the gfx950 wavefront tracer.
//
// Layout in HBM (all arrays owned by the context, uploaded once by sr_scene_build):
//  - nodes: every BVH (one per mesh plus the scene BVH over objects) in the reference's node
//    order, 32 B per node (bvh.cs:26-31), each BVH preceded by one pad node and starting at an
//    even index, so the sibling pair (child, child+1) of any interior node is one 64-byte aligned
//    record that a wave fetches with four 16-byte loads.
//  - tris: TriangleLean (PosA, AB, AC; 36 B) stored in BVH item order, so a leaf's triangles are
//    contiguous; slot_tri maps an item slot back to the mesh's triangle index.
//  - attrs: TriangleAttributes (mesh.cs:6-11; 27 floats) in the same item order, read once per
//    closest hit at the hit's item slot.
//  - objects: transform, material and shape per Object (scene.cs:204-223).
//  - texels / textures: linear RGB texels (12 B) and per-texture descriptors.
//  - cdf: the sky's marginal and conditional CDF tables (cdf.cs:62-98), then their bucket records.
//  - skyTab: per sky texel, the light sample's direction, pdf and radiance (k_sky_table).
#pragma once

#include "sr_math.h"

#define SR_DEV __device__ __forceinline__

namespace sr {

struct DevTexture {
    int32_t offset;  // first texel in the pool
    int32_t w, h;
    int32_t filter;  // 0 bilinear, 1 point
    float tilx, tily;
    int32_t pad[2];
};

struct DevObject {
    quat rot;
    vec3 pos;
    vec3 scale;
    int32_t kind;       // 0 mesh, 1 sphere, 2 box
    int32_t nodeBase;   // mesh BVH: index of its pad node in the node pool
    int32_t triBase;    // mesh: first triangle (slot and attribute arrays share the base)
    int32_t texColor, texRough, texMetal, texNormal;
    vec3 albedo;
    float roughness, metallic, transparency, ior;
    vec3 radiance;
    float f0;           // MathF.Pow((ior - 1) / (ior + 1), 2f) (brdf.cs:14), evaluated on the host
    float shape[6];     // sphere centre + radius, or box min + max
    int32_t xform;      // the first object with a bit-identical transform (rot, pos, scale)
};

struct DevSky {
    int32_t kind;  // 0 none, 1 texture, 2 procedural
    int32_t tex;
    float multiplier;
    quat rot, rotInv;
    int32_t w, h;
    int32_t cdfOff;  // marginal CDF (h+1 floats) followed by h conditional CDFs (w+1 floats each)
    int32_t guideOff;  // -1, or the CDFs' bucket records (float4 each, in the same pool; cdf_sample_guided)
    int32_t sampler;  // 0: row/column CDFs, 1: alias table (DevScene::alias)
    float total, pdfScale;
    vec3 top, middle, bottom, sunDir, sunRad;
    float sunCos, sunPdfDenom;
    quat sunRot;
};

struct DevFog {
    int32_t enabled;
    int32_t kind;  // 0 height fog (scene.cs:438-505), 1 homogeneous box, 2 fp16 density grid
    float density;  // height fog density / box and grid extinction scale
    vec3 color;
    float anisotropy;
    float heightFalloff;
    vec3 bmin, bmax;   // medium box (kinds 1, 2)
    vec3 cell, cellInv;  // grid voxel size and its reciprocal
    int32_t nx, ny, nz;
    float majorant;    // density * max grid value
    const uint16_t* grid;  // fp16 bits, x fastest
};

struct DevScene {
    const float4* nodes;
    const float* tris;
    const int32_t* slotTri;
    const float* attrs;
    const int32_t* sceneItems;
    const DevObject* objects;
    const float* texels;
    const DevTexture* textures;
    const float* cdf;
    const float4* alias;  // per texel: acceptance threshold, alias texel (int bits), pdf, alias pdf
    // Texture sky, per texel (or null): the direction toward its centre and its light pdf, then the
    // sky radiance seen in that direction (k_sky_table), two float4 per texel.
    const float4* skyTab;
    int32_t sceneBase;   // pad node of the scene BVH
    int32_t sceneNodes;  // node count of the scene BVH (0 = empty scene)
    int32_t anyTransparent;
    // Hit records carry (object << 24 | item slot) in one word (fewer than 128 objects, mesh item
    // slots below 2^24 - 1); otherwise the object alone, with the slot in its own array.
    int32_t packHit;
    DevSky sky;
    DevFog fog;
};

// ---- RNG policies ------------------------------------------------------------------------------
// Reference policy: the stateful xorwow stream of math.cs:1665-1714 seeded per pixel.
struct XorwowRng {
    uint32_t s0, s1, s2, s3, s4;
    SR_HD uint32_t next_uint() {
        uint32_t t = s3;
        const uint32_t s = s0;
        s3 = s2;
        s2 = s1;
        s1 = s;
        t ^= t >> 2;
        t ^= t << 1;
        t ^= s ^ (s << 4);
        s0 = t;
        s4 += 362437u;
        return t + s4;
    }
};

SR_HD uint64_t split_mix64(uint64_t& state) {
    uint64_t r = (state += 0x9e3779b97f4a7c15ULL);
    r = (r ^ (r >> 30)) * 0xbf58476d1ce4e5b9ULL;
    r = (r ^ (r >> 27)) * 0x94d049bb133111ebULL;
    return r ^ (r >> 31);
}
// Rng(uint seedA, uint seedB) (math.cs:1669-1680).
SR_HD XorwowRng xorwow_seed(uint32_t a, uint32_t b) {
    uint64_t st = ((uint64_t)a << 32) | b;
    const uint64_t v1 = split_mix64(st);
    const uint64_t v2 = split_mix64(st);
    XorwowRng r;
    r.s0 = (uint32_t)v1;
    r.s1 = (uint32_t)(v1 >> 32);
    r.s2 = (uint32_t)v2;
    r.s3 = (uint32_t)(v2 >> 32);
    r.s4 = 0;
    return r;
}

// Counter policy: Philox4x32-10 (Salmon et al., SC'11) keyed by (seed, pixel), counter
// (sample, draw / 4, 0, 0); draw k of a path is word k % 4 of block k / 4.
SR_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
// CACHE keeps the last block drawn from in registers (blk, for block index blkIdx), so four
// consecutive draws cost one Philox evaluation instead of four; the stream is the same. A path
// record keeps only (key, sample, draw): a reloaded generator starts with no block. The adjoint
// kernel, which runs at a tight register budget, draws without the cache.
template <bool CACHE>
struct CounterRngT {
    uint32_t key0, key1, sample, draw;
    uint32_t blkIdx = ~0u;
    uint32_t blk0 = 0, blk1 = 0, blk2 = 0, blk3 = 0;
    SR_HD uint32_t next_uint() {
        const uint32_t b = draw >> 2;
        if (!CACHE || b != blkIdx) {
            uint32_t c[4] = {sample, b, 0u, 0u};
            philox4x32_10(c, key0, key1);
            blk0 = c[0]; blk1 = c[1]; blk2 = c[2]; blk3 = c[3];
            blkIdx = b;
        }
        const uint32_t w = draw & 3u;
        const uint32_t v = w == 0u ? blk0 : w == 1u ? blk1 : w == 2u ? blk2 : blk3;
        ++draw;
        return v;
    }
};
using CounterRng = CounterRngT<true>;

template <class R>
SR_HD float next_float(R& r) {
    const float toFloat = 1.0f / (4294967296.0f + 512.0f);  // 0x2F7FFFFE (math.cs:1716-1720)
    return (float)r.next_uint() * toFloat;
}
template <class R>
SR_HD vec2 next_vec2(R& r) {  // Vec2.Rand (math.cs:283): X before Y
    const float x = next_float(r);
    const float y = next_float(r);
    return mk2(x, y);
}
template <class R>
SR_DEV float next_exponential(R& r) {  // math.cs:1722-1727
    float u;
    do { u = next_float(r); } while (u <= 1e-8f);
    return -logf(u);
}
template <class R>
SR_DEV void next_gauss(R& r, float& ga, float& gb) {  // math.cs:1729-1738
    float a;
    do { a = next_float(r); } while (a <= 1e-8f);
    const float b = kPi * 2.0f * next_float(r);
    const float mag = sqrtf(-2.0f * logf(a));
    ga = mag * cosf(b);
    gb = mag * sinf(b);
}
template <class R>
SR_DEV vec3 rand_on_sphere(R& r) {  // math.cs:443-454
    for (;;) {
        float x, y, z, unused;
        next_gauss(r, x, y);
        next_gauss(r, z, unused);
        const float m = x * x + y * y + z * z;
        if (m > 1e-6f) return mk3(x, y, z) / sqrtf(m);
    }
}

// ---- Geometry ----------------------------------------------------------------------------------
SR_DEV vec2 equirect_uv(vec3 d) {  // math.cs:390-398
    const float u = atan2f(d.z, d.x) / (kPi * 2.0f) + 0.5f;
    const float v = 0.5f - asinf(net_clamp(d.y, -1.0f, 1.0f)) / kPi;
    return mk2(u, v);
}
SR_DEV vec3 from_equirect_uv(vec2 uv) {  // math.cs:482-489
    const float phi = (uv.x - 0.5f) * 2.0f * kPi;
    const float st = sinf(kPi * uv.y);
    const float y = cosf(kPi * uv.y);
    return mk3(st * cosf(phi), y, st * sinf(phi));
}
SR_DEV vec3 from_tangent(vec3 d, vec4 t, vec3 n) {  // math.cs:492-498
    const vec3 tt = xyz(t);
    const vec3 bt = t.w * cross3(n, tt);
    return d.x * tt + d.y * bt + d.z * n;
}
SR_DEV vec3 reflect3(vec3 v, vec3 n) { return v - 2.0f * dot3(v, n) * n; }  // math.cs:437-441
SR_DEV vec3 perp3(vec3 v) {  // math.cs:383-388
    const vec3 up = fabsf(v.y) < 0.9f ? up3() : right3();
    return normalized(cross3(up, v));
}

// Slab test of AABox.IntersectDist (math.cs:983-1004) with the caller's bound: the box counts as hit
// when t = max(0, tmin) <= min(tmax, limit), which is the reference's test (tmax >= 0 and
// tmin <= tmax) followed by t <= limit in fewer instructions; hardware min / max return the other
// operand of a NaN, so both forms agree for NaN slab distances too. Inputs are finite, so hardware
// min/max agree with MathF's; a -0 entry distance only ever meets ordered comparisons.
SR_DEV bool slab(float mnx, float mny, float mnz, float mxx, float mxy, float mxz, const ray3& r, float limit, float& t) {
    const float t1 = (mnx - r.o.x) * r.inv.x;
    const float t2 = (mxx - r.o.x) * r.inv.x;
    const float t3 = (mny - r.o.y) * r.inv.y;
    const float t4 = (mxy - r.o.y) * r.inv.y;
    const float t5 = (mnz - r.o.z) * r.inv.z;
    const float t6 = (mxz - r.o.z) * r.inv.z;
    const float tmin = fmaxf(fmaxf(fminf(t1, t2), fminf(t3, t4)), fminf(t5, t6));
    const float tmax = fminf(fminf(fmaxf(t1, t2), fmaxf(t3, t4)), fmaxf(t5, t6));
    t = fmaxf(0.0f, tmin);
    return t <= fminf(tmax, limit);
}

// TriangleLean.Intersect (Moller-Trumbore, math.cs:1407-1435).
SR_DEV bool tri_hit(const float* T, const ray3& r, float& t, float& u, float& v) {
    const vec3 a = mk3(T[0], T[1], T[2]);
    const vec3 ab = mk3(T[3], T[4], T[5]);
    const vec3 ac = mk3(T[6], T[7], T[8]);
    const vec3 h = cross3(r.d, ac);
    const float det = dot3(ab, h);
    if (fabsf(det) <= 1e-7f) return false;
    const float invDet = 1.0f / det;
    const vec3 ao = r.o - a;
    const float uu = dot3(ao, h) * invDet;
    if (uu < -1e-7f || uu > (1.0f + 1e-7f)) return false;
    const vec3 q = cross3(ao, ab);
    const float vv = dot3(r.d, q) * invDet;
    if (vv < -1e-7f || (uu + vv) > (1.0f + 1e-7f)) return false;
    const float tt = dot3(ac, q) * invDet;
    if (tt < 0.0f) return false;
    t = tt;
    u = uu;
    v = vv;
    return true;
}

// Transform.TransformRayInv (math.cs:816-822).
SR_DEV ray3 to_object(const DevObject& ob, const ray3& r, float& scale) {
    const quat inv = conj(ob.rot);
    const vec3 o = rotate(inv, r.o - ob.pos) / ob.scale;
    const vec3 dr = rotate(inv, r.d) / ob.scale;
    scale = sqrtf(len_sqr(dr));
    return mk_ray(o, dr / scale);
}

struct ShapeHitD {
    float dist;
    vec3 norm, normGeo;
    vec4 tan;
    vec2 uv;
};

// Sphere.Intersect (math.cs:1210-1234).
SR_DEV bool sphere_hit(const float* s, const ray3& r, ShapeHitD& h, bool full) {
    const vec3 c = mk3(s[0], s[1], s[2]);
    const float rad = s[3];
    const vec3 tc = c - r.o;
    const float tcd = len_sqr(tc);
    const float a = dot3(tc, r.d);
    const float disc = rad * rad - tcd + a * a;
    if (disc < 0.0f) return false;
    const float f = sqrtf(disc);
    const float t = tcd < rad * rad ? a + f : a - f;
    if (t < 0.0f) return false;
    h.dist = t;
    if (full) {
        const vec3 n = normalized(at(r, t) - c);
        const vec3 tg = normalized_or(mk3(-n.z, 0.0f, n.x), right3());
        h.norm = n;
        h.normGeo = n;
        h.tan = mk4(tg.x, tg.y, tg.z, 1.0f);
        h.uv = mk2(0.5f + atan2f(n.z, n.x) / (2.0f * kPi), 0.5f - asinf(net_clamp(n.y, -1.0f, 1.0f)) / kPi);
    }
    return true;
}

// AABox.Intersect (math.cs:1006-1059), with MathF min/max kept exact for the face choice.
SR_DEV bool box_hit(const float* s, const ray3& r, ShapeHitD& h, bool full) {
    const vec3 mn = mk3(s[0], s[1], s[2]), mx = mk3(s[3], s[4], s[5]);
    const float t1 = (mn.x - r.o.x) * r.inv.x;
    const float t2 = (mx.x - r.o.x) * r.inv.x;
    const float t3 = (mn.y - r.o.y) * r.inv.y;
    const float t4 = (mx.y - r.o.y) * r.inv.y;
    const float t5 = (mn.z - r.o.z) * r.inv.z;
    const float t6 = (mx.z - r.o.z) * r.inv.z;
    const float minA = net_min(t1, t2), minB = net_min(t3, t4), minC = net_min(t5, t6);
    const float maxA = net_max(t1, t2), maxB = net_max(t3, t4), maxC = net_max(t5, t6);
    const float tMin = net_max(net_max(minA, minB), minC);
    const float tMax = net_min(net_min(maxA, maxB), maxC);
    if (tMax < 0.0f || tMin > tMax) return false;
    const bool inside = tMin < 0.0f;
    const float t = inside ? tMax : tMin;
    h.dist = t;
    if (full) {
        const vec3 p = at(r, t);
        vec3 n;
        vec4 tg;
        vec2 f;
        if (minA >= minB && minA >= minC) {
            const bool neg = t1 <= t2;
            n = neg ? mk3(-1, 0, 0) : mk3(1, 0, 0);
            tg = mk4(0, 1, 0, neg ? -1.0f : 1.0f);
            f = mk2((p.y - mn.y) / (mx.y - mn.y), (p.z - mn.z) / (mx.z - mn.z));
        } else if (minB >= minA && minB >= minC) {
            const bool neg = t3 <= t4;
            n = neg ? mk3(0, -1, 0) : mk3(0, 1, 0);
            tg = mk4(1, 0, 0, neg ? 1.0f : -1.0f);
            f = mk2((p.x - mn.x) / (mx.x - mn.x), (p.z - mn.z) / (mx.z - mn.z));
        } else {
            const bool neg = t5 <= t6;
            n = neg ? mk3(0, 0, -1) : mk3(0, 0, 1);
            tg = mk4(1, 0, 0, neg ? -1.0f : 1.0f);
            f = mk2((p.x - mn.x) / (mx.x - mn.x), (p.y - mn.y) / (mx.y - mn.y));
        }
        const vec3 nn = inside ? -n : n;
        h.norm = nn;
        h.normGeo = nn;
        h.tan = tg;
        h.uv = f;
    }
    return true;
}

// TriangleLean.Inflate (math.cs:1437-1459) from the lean triangle and its 27 attribute floats.
SR_DEV void inflate(const float* T, const float* A, float u, float v, ShapeHitD& h) {
    const float w = 1.0f - u - v;
    const vec3 nA = mk3(A[0], A[1], A[2]), nB = mk3(A[3], A[4], A[5]), nC = mk3(A[6], A[7], A[8]);
    const vec4 tA = mk4(A[9], A[10], A[11], A[12]), tB = mk4(A[13], A[14], A[15], A[16]);
    const vec4 tC = mk4(A[17], A[18], A[19], A[20]);
    const vec2 uA = mk2(A[21], A[22]), uB = mk2(A[23], A[24]), uC = mk2(A[25], A[26]);
    const vec3 face = normalized_or(cross3(mk3(T[3], T[4], T[5]), mk3(T[6], T[7], T[8])), up3());
    const vec3 ni = nA * w + nB * u + nC * v;
    const vec3 n = len_sqr(ni) >= 1e-12f ? normalized(ni) : face;
    const vec4 ti = tA * w + tB * u + tC * v;
    const vec3 tx = xyz(ti);
    const vec3 t = normalized_or(tx - dot3(tx, n) * n, right3());
    h.norm = n;
    h.normGeo = face;
    h.tan = mk4(t.x, t.y, t.z, ti.w >= 0.0f ? 1.0f : -1.0f);
    h.uv = uA * w + uB * u + uC * v;
}

// Texture.Sample (texture.cs:39-69).
SR_DEV vec3 tex_get(const DevScene& S, const DevTexture& t, uint32_t idx) {
    const float* p = S.texels + (size_t)(t.offset + (int32_t)idx) * 3;
    return mk3(p[0], p[1], p[2]);
}
SR_DEV vec3 tex_sample(const DevScene& S, int32_t id, vec2 c) {
    const DevTexture t = S.textures[id];
    float u = c.x * t.tilx;
    u -= floorf(u);
    float v = c.y * t.tily;
    v -= floorf(v);
    const uint32_t W = (uint32_t)t.w, H = (uint32_t)t.h;
    if (t.filter == 1) {
        const uint32_t x = (uint32_t)(u * (float)W) % W;
        const uint32_t y = (uint32_t)(v * (float)H) % H;
        return tex_get(S, t, y * W + x);
    }
    const float fx = u * (float)(W - 1);
    const float fy = v * (float)(H - 1);
    const uint32_t x0 = (uint32_t)fx, y0 = (uint32_t)fy;
    const uint32_t x1 = (x0 + 1) % W, y1 = (y0 + 1) % H;
    const float tx = fx - (float)x0, ty = fy - (float)y0;
    const vec3 c1 = tex_get(S, t, y0 * W + x0), c2 = tex_get(S, t, y0 * W + x1);
    const vec3 c3 = tex_get(S, t, y1 * W + x0), c4 = tex_get(S, t, y1 * W + x1);
    return lerp3(lerp3(c1, c2, tx), lerp3(c3, c4, tx), ty);
}

// ---- Sky (scene.cs:281-435) --------------------------------------------------------------------
SR_DEV vec3 sky_radiance(const DevScene& S, vec3 d) {
    const DevSky& k = S.sky;
    if (k.kind == 1) return tex_sample(S, k.tex, equirect_uv(rotate(k.rotInv, d))) * k.multiplier;
    if (k.kind == 2) {
        const float bias = 0.0001f;
        const float topBlend = 1.0f - powf(net_min(1.0f, 1.0f + bias - d.y), 4.0f);
        const float bottomBlend = 1.0f - powf(net_min(1.0f, 1.0f + bias + d.y), 40.0f);
        const float middleBlend = 1.0f - topBlend - bottomBlend;
        const vec3 amb = k.top * topBlend + k.middle * middleBlend + k.bottom * bottomBlend;
        const float blend = net_max(0.0f, (dot3(d, k.sunDir) - k.sunCos) / (1.0f - k.sunCos));
        return amb + k.sunRad * blend;
    }
    return splat3(0.0f);
}
// Cdf1.Sample (cdf.cs:36-54) over `n` entries of a table with n+1 values.
SR_DEV uint32_t cdf_sample(const float* cdf, uint32_t n, float x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid + 1] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo < n - 1 ? lo : n - 1;
}
// Bucket records of a CDF table with n + 1 entries: with m buckets (a power of two), bucket j covers
// x in [j / m, (j + 1) / m) and its 16-byte record holds the search's unclamped answers lo for
// x = j / m and hi for x = (j + 1) / m (n for the last bucket), then copies of cdf[lo + 1] and
// cdf[lo + 2] (0 past the table). The marginal CDF has kGuideMarg buckets, each conditional CDF
// kGuideCond: as many buckets as the 2048-texel rows of a 2048 x 1024 map have texels, so a bucket
// mostly holds at most two texel boundaries and one record read settles the search.
#ifndef SR_GUIDE_COND
#define SR_GUIDE_COND 2048
#endif
constexpr uint32_t kGuideMarg = 1024, kGuideCond = SR_GUIDE_COND;
// Cdf1.Sample with the search confined to x's bucket j = floor(x m): x lies in [j / m, (j + 1) / m)
// and the predicate cdf[i + 1] <= x falls from true to false once along i (the table does not
// decrease), so its first false index lies in [lo, hi]. The record's copies decide the first two
// candidates lo and lo + 1; a bucket holding more texel boundaries continues with a binary search
// of the table over the rest of [lo, hi]. Same index as the full search.
SR_DEV uint32_t cdf_sample_guided(const float* cdf, uint32_t n, float x, const float4* g, uint32_t m) {
    const float s = x * (float)m;
    const uint32_t j = s >= 0.0f ? (s < (float)m ? (uint32_t)s : m - 1) : 0u;
    const float4 r = g[j];
    uint32_t lo = __float_as_uint(r.x), hi = __float_as_uint(r.y);
    if (lo < hi && r.z <= x) {
        ++lo;
        if (lo < hi && r.w <= x) {
            ++lo;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (cdf[mid + 1] <= x) lo = mid + 1;
                else hi = mid;
            }
        }
    }
    return lo < n - 1 ? lo : n - 1;
}
// SkyTexture.LightDirRand's direction toward the centre of texel (x, y) (scene.cs:413) and the pdf
// of sampling that texel (scene.cs:414). k_sky_table tabulates both with these same functions.
SR_DEV vec3 sky_texel_dir(const DevSky& k, uint32_t x, uint32_t y) {
    const vec2 uv = mk2(((float)(int32_t)x + 0.5f) / (float)k.w, ((float)(int32_t)y + 0.5f) / (float)k.h);
    return rotate(k.rot, from_equirect_uv(uv));
}
SR_DEV float sky_texel_pdf(const DevScene& S, uint32_t texel) {
    const DevTexture t = S.textures[S.sky.tex];
    return luminance(tex_get(S, t, texel)) * S.sky.pdfScale;
}
// Sky radiance toward a light sample: the table's value when sky_light_rand read one (ok), else
// evaluated on demand.
struct SkyRad {
    vec3 v;
    bool ok;
};
SR_DEV vec3 sky_light_radiance(const DevScene& S, const SkyRad& r, vec3 dir) { return r.ok ? r.v : sky_radiance(S, dir); }

template <class R>
SR_DEV bool sky_light_rand(const DevScene& S, R& rng, vec3& dir, float& pdf, uint32_t* texelOut = nullptr,
                           SkyRad* radOut = nullptr) {
    if (radOut) radOut->ok = false;
    const DevSky& k = S.sky;
    if (k.kind == 1) {
        if (k.total <= 0.0f) return false;
        const vec2 c = next_vec2(rng);
        uint32_t x, y;
        if (k.sampler == 1) {
            // One 16-byte entry decides the texel and carries both candidates' pdfs.
            const uint32_t n = (uint32_t)k.w * (uint32_t)k.h;
            uint32_t i = (uint32_t)(c.x * (float)n);
            if (i >= n) i = n - 1;
            const float4 e = S.alias[i];
            const bool keep = c.y < e.x;
            const uint32_t texel = keep ? i : (uint32_t)__float_as_int(e.y);
            x = texel % (uint32_t)k.w;
            y = texel / (uint32_t)k.w;
            pdf = keep ? e.z : e.w;
        } else {
            const float* marg = S.cdf + k.cdfOff;
            if (k.guideOff >= 0) {
                const float4* G = (const float4*)(S.cdf + k.guideOff);
                y = cdf_sample_guided(marg, (uint32_t)k.h, c.y, G, kGuideMarg);
                const float* cond = marg + (k.h + 1) + (size_t)y * (k.w + 1);
                x = cdf_sample_guided(cond, (uint32_t)k.w, c.x, G + kGuideMarg + (size_t)y * kGuideCond, kGuideCond);
            } else {
                y = cdf_sample(marg, (uint32_t)k.h, c.y);
                const float* cond = marg + (k.h + 1) + (size_t)y * (k.w + 1);
                x = cdf_sample(cond, (uint32_t)k.w, c.x);
            }
        }
        const uint32_t texel = y * (uint32_t)k.w + x;
        if (S.skyTab) {
            // One 32-byte entry instead of the texel's pdf, four trigonometric functions for the
            // direction and the bilinear sky lookup along it (bit-identical: tabulated by the same code).
            const float4 e0 = S.skyTab[2 * (size_t)texel], e1 = S.skyTab[2 * (size_t)texel + 1];
            dir = mk3(e0.x, e0.y, e0.z);
            if (k.sampler != 1) pdf = e0.w;
            if (radOut) {
                radOut->v = mk3(e1.x, e1.y, e1.z);
                radOut->ok = true;
            }
        } else {
            if (k.sampler != 1) pdf = sky_texel_pdf(S, texel);
            dir = sky_texel_dir(k, x, y);
        }
        if (texelOut) *texelOut = texel;
        return pdf > 0.0f;
    }
    if (k.kind == 2) {
        const float phi = kPi * 2.0f * next_float(rng);
        const float blend = sqrtf(next_float(rng));
        const float z = k.sunCos + (1.0f - k.sunCos) * blend;
        const float sinT = sqrtf(net_max(0.0f, 1.0f - z * z));
        dir = rotate(k.sunRot, mk3(cosf(phi) * sinT, sinf(phi) * sinT, z));
        pdf = blend / k.sunPdfDenom;
        return true;
    }
    return false;
}
SR_DEV bool sky_light_pdf(const DevScene& S, vec3 d, float& pdf) {
    const DevSky& k = S.sky;
    if (k.kind == 1) {
        if (k.total <= 0.0f) return false;
        const vec2 uv = equirect_uv(rotate(k.rotInv, d));
        const int32_t cx = ((int32_t)(uv.x * (float)k.w)) % k.w;
        const int32_t cy = ((int32_t)(uv.y * (float)k.h)) % k.h;
        const DevTexture t = S.textures[k.tex];
        pdf = luminance(tex_get(S, t, (uint32_t)(cy * k.w + cx))) * k.pdfScale;
        return pdf > 0.0f;
    }
    if (k.kind == 2) {
        const float c = dot3(d, k.sunDir);
        if (c < k.sunCos) return false;
        pdf = ((c - k.sunCos) / (1.0f - k.sunCos)) / k.sunPdfDenom;
        return true;
    }
    return false;
}

// ---- Height fog (scene.cs:438-505) -------------------------------------------------------------
SR_DEV float fog_density(const DevFog& f, float y) { return f.density * expf(-f.heightFalloff * y); }
SR_DEV float fog_optical_depth(const DevFog& f, const ray3& r, float start, float end) {
    const float hd = r.d.y;
    const float hs = r.o.y + start * hd;
    const float he = r.o.y + end * hd;
    const float ds = fog_density(f, hs);
    const float de = fog_density(f, he);
    if (fabsf(hd) < 1e-6f) return ds * (end - start);
    return (ds - de) / (f.heightFalloff * hd);
}
template <class R>
SR_DEV bool fog_scatter(const DevFog& f, const ray3& r, float start, float end, R& rng, float& tOut) {
    const float hd = r.d.y;
    const float ho = r.o.y;
    const float hs = ho + start * hd;
    const float u = next_exponential(rng);
    float t;
    if (fabsf(hd) < 1e-6f) {
        const float ds = fog_density(f, hs);
        if (ds <= 0.0f) return false;
        t = start + u / ds;
    } else {
        const float rhs = expf(-f.heightFalloff * hs) - u * f.heightFalloff * hd / f.density;
        if (rhs <= 0.0f) return false;
        const float yt = -logf(rhs) / f.heightFalloff;
        t = (yt - ho) / hd;
    }
    if (t < end) {
        tOut = t;
        return true;
    }
    return false;
}

// ---- Bounded media (no reference counterpart) ---------------------------------------------------
// A homogeneous box or a box holding a piecewise-constant fp16 density grid (extinction
// sigma = density * value of the voxel containing the point). Free flight in the grid is drawn by
// delta tracking against the grid's majorant; transmittance is the exact optical depth summed
// voxel by voxel along the ray (3D DDA), so shadow rays need no random numbers.
constexpr float kInfF = __builtin_huge_valf();
constexpr int32_t kMaxDeltaSteps = 4096;  // bound on tentative collisions per free-flight draw
SR_DEV float half_bits_to_float(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// The part [a, b) of [start, end) inside the medium box.
SR_DEV bool medium_span(const DevFog& f, const ray3& r, float start, float end, float& a, float& b) {
    const float t1 = (f.bmin.x - r.o.x) * r.inv.x, t2 = (f.bmax.x - r.o.x) * r.inv.x;
    const float t3 = (f.bmin.y - r.o.y) * r.inv.y, t4 = (f.bmax.y - r.o.y) * r.inv.y;
    const float t5 = (f.bmin.z - r.o.z) * r.inv.z, t6 = (f.bmax.z - r.o.z) * r.inv.z;
    const float tmin = net_max(net_max(net_min(t1, t2), net_min(t3, t4)), net_min(t5, t6));
    const float tmax = net_min(net_min(net_max(t1, t2), net_max(t3, t4)), net_max(t5, t6));
    a = net_max(tmin, start);
    b = net_min(tmax, end);
    return a < b;
}
SR_DEV int32_t grid_coord(float p, float lo, float inv, int32_t n) {
    const int32_t i = (int32_t)floorf((p - lo) * inv);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
SR_DEV float grid_sigma(const DevFog& f, int32_t x, int32_t y, int32_t z) {
    return f.density * half_bits_to_float(f.grid[((size_t)z * f.ny + y) * f.nx + x]);
}
SR_DEV float grid_sigma_at(const DevFog& f, vec3 p) {
    return grid_sigma(f, grid_coord(p.x, f.bmin.x, f.cellInv.x, f.nx), grid_coord(p.y, f.bmin.y, f.cellInv.y, f.ny),
                      grid_coord(p.z, f.bmin.z, f.cellInv.z, f.nz));
}
// Optical depth of the grid over [a, b): walk the voxels the ray crosses (Amanatides-Woo), next
// boundary first on x, then y, then z on ties.
SR_DEV float grid_optical_depth(const DevFog& f, const ray3& r, float a, float b) {
    const vec3 p = at(r, a);
    int32_t c[3] = {grid_coord(p.x, f.bmin.x, f.cellInv.x, f.nx), grid_coord(p.y, f.bmin.y, f.cellInv.y, f.ny),
                    grid_coord(p.z, f.bmin.z, f.cellInv.z, f.nz)};
    const int32_t n[3] = {f.nx, f.ny, f.nz};
    const float o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z}, inv[3] = {r.inv.x, r.inv.y, r.inv.z};
    const float lo[3] = {f.bmin.x, f.bmin.y, f.bmin.z}, cs[3] = {f.cell.x, f.cell.y, f.cell.z};
    float tNext[3], tDelta[3];
    int32_t step[3];
    for (int k = 0; k < 3; ++k) {
        step[k] = d[k] > 0.0f ? 1 : -1;
        const float edge = lo[k] + (float)(c[k] + (d[k] > 0.0f ? 1 : 0)) * cs[k];
        tNext[k] = d[k] == 0.0f ? kInfF : (edge - o[k]) * inv[k];
        tDelta[k] = d[k] == 0.0f ? kInfF : cs[k] * fabsf(inv[k]);
    }
    float t = a, tau = 0.0f;
    for (int32_t it = 0, lim = f.nx + f.ny + f.nz + 3; it < lim; ++it) {
        const int k = tNext[0] <= tNext[1] ? (tNext[0] <= tNext[2] ? 0 : 2) : (tNext[1] <= tNext[2] ? 1 : 2);
        const float tEnd = net_min(tNext[k], b);
        tau += grid_sigma(f, c[0], c[1], c[2]) * net_max(0.0f, tEnd - t);
        if (tEnd >= b) break;
        t = tEnd;
        c[k] += step[k];
        if (c[k] < 0 || c[k] >= n[k]) break;
        tNext[k] += tDelta[k];
    }
    return tau;
}

// ---- Medium dispatch ----------------------------------------------------------------------------
SR_DEV float medium_optical_depth(const DevFog& f, const ray3& r, float start, float end) {
    if (f.kind == 0) return fog_optical_depth(f, r, start, end);
    float a, b;
    if (!medium_span(f, r, start, end, a, b)) return 0.0f;
    if (f.kind == 1) return f.density * (b - a);
    return grid_optical_depth(f, r, a, b);
}
template <class R>
SR_DEV bool medium_scatter(const DevFog& f, const ray3& r, float start, float end, R& rng, float& tOut) {
    if (f.kind == 0) return fog_scatter(f, r, start, end, rng, tOut);
    float a, b;
    if (!medium_span(f, r, start, end, a, b)) return false;
    if (f.kind == 1) {
        if (!(f.density > 0.0f)) return false;
        const float t = a + next_exponential(rng) / f.density;
        if (t < b) { tOut = t; return true; }
        return false;
    }
    // Delta tracking: tentative collisions at majorant rate, real with probability sigma / majorant.
    if (!(f.majorant > 0.0f)) return false;
    float t = a;
    for (int32_t it = 0; it < kMaxDeltaSteps; ++it) {
        t += next_exponential(rng) / f.majorant;
        if (!(t < b)) return false;
        const float s = grid_sigma_at(f, at(r, t));
        if (next_float(rng) < s / f.majorant) { tOut = t; return true; }
    }
    return false;
}

// ---- BRDF (brdf.cs) ----------------------------------------------------------------------------
SR_DEV vec3 base_refl(float f0, vec3 albedo, float metallic) { return lerp3(splat3(f0), albedo, metallic); }
SR_DEV vec3 fresnel3(float nDotV, vec3 f0) {
    const float x = 1.0f - nDotV;
    const float x2 = x * x;
    const float x5 = x2 * x2 * x;
    return f0 + (splat3(1.0f) - f0) * x5;
}
SR_DEV float ggx_d(float nDotH, float r2) {
    const float a2 = r2 * r2;
    const float nh2 = nDotH * nDotH;
    const float d = (1.0f - nh2) + nh2 * a2;
    return a2 / (kPi * d * d);
}
SR_DEV float smith_g1(float nDotX, float r2) {
    const float k = r2 / 2.0f;
    return nDotX / (nDotX * (1.0f - k) + k);
}
SR_DEV vec3 ggx_sample(float r2, vec2 u) {
    const float ct = sqrtf((1.0f - u.x) / (u.x * (r2 * r2 - 1.0f) + 1.0f));
    const float st = sqrtf(net_max(0.0f, 1.0f - ct * ct));
    const float phi = 2.0f * kPi * u.y;
    return mk3(st * cosf(phi), st * sinf(phi), ct);
}
SR_DEV float power_heuristic(float p1, float p2) {
    const float a = p1 * p1, b = p2 * p2;
    return a + b > 0.0f ? a / (a + b) : 0.0f;
}
SR_DEV float hg_pdf(float rDotL, float g) {
    const float g2 = g * g;
    const float d = 1.0f + g2 - 2.0f * g * rDotL;
    return (1.0f - g2) / (4.0f * kPi * d * sqrtf(d));
}
SR_DEV bool refract3(vec3 i, vec3 n, float eta, vec3& out) {
    const float cosI = -dot3(i, n);
    const float sin2T = eta * eta * (1.0f - cosI * cosI);
    if (sin2T >= 1.0f) return false;
    const float cosT = sqrtf(1.0f - sin2T);
    out = eta * i + (eta * cosI - cosT) * n;
    return true;
}
SR_DEV vec3 beer(vec3 a, float d) {
    return mk3(powf(net_max(a.x, 0.0f), d), powf(net_max(a.y, 0.0f), d), powf(net_max(a.z, 0.0f), d));
}
template <class R>
SR_DEV vec3 hg_scatter(vec3 fwd, float g, R& rng) {
    const vec2 u = next_vec2(rng);
    float ct;
    if (fabsf(g) < 1e-3f) {
        ct = 1.0f - 2.0f * u.x;
    } else {
        const float s = (1.0f - g * g) / (1.0f - g + 2.0f * g * u.x);
        ct = (1.0f + g * g - s * s) / (2.0f * g);
    }
    ct = net_clamp(ct, -1.0f, 1.0f);
    const float st = sqrtf(1.0f - ct * ct);
    const float phi = 2.0f * kPi * u.y;
    const vec3 t = perp3(fwd);
    const vec3 b = cross3(fwd, t);
    return ct * fwd + st * cosf(phi) * t + st * sinf(phi) * b;
}

// ---- Surface (scene.cs:15-152) -----------------------------------------------------------------
struct SurfaceD {
    vec3 radiance, albedo;
    float roughness, metallic, transparency, ior, f0;
    vec3 normal, normalGeo;
    vec4 tangent;
    vec2 uv;
};

SR_DEV vec3 clamp_geo_side(const SurfaceD& s, vec3 dir, vec3 hint) {
    const float d = dot3(dir, s.normalGeo);
    if (d * dot3(hint, s.normalGeo) < 0.0f) dir = normalized_or(dir - 2.0f * d * s.normalGeo, hint);
    return dir;
}
SR_DEV float spec_prob(const SurfaceD& s, vec3 v) {
    const float nDotV = net_max(1e-4f, dot3(s.normal, v));
    const float f = luminance(fresnel3(nDotV, base_refl(s.f0, s.albedo, s.metallic)));
    return net_clamp(net_max(f, s.metallic), 0.001f, 0.999f);
}
SR_DEV vec3 opaque_eval(const SurfaceD& s, vec3 v, vec3 l) {
    const float nDotL = dot3(s.normal, l);
    if (nDotL <= 0.0f) return splat3(0.0f);
    const float nDotV = net_max(1e-4f, dot3(s.normal, v));
    const float r2 = net_max(s.roughness * s.roughness, 1e-4f);
    const vec3 F0 = base_refl(s.f0, s.albedo, s.metallic);
    const vec3 h = normalized_or(v + l, s.normal);
    const float nDotH = dot3(s.normal, h);
    const float hDotV = dot3(h, v);
    const vec3 f = fresnel3(hDotV, F0);
    const float g = smith_g1(nDotV, r2) * smith_g1(nDotL, r2);
    const vec3 spec = ggx_d(nDotH, r2) * f * g / (4.0f * nDotV);
    const vec3 diff = (splat3(1.0f) - fresnel3(nDotV, F0)) * s.albedo * (1.0f - s.metallic) * (nDotL / kPi);
    return (spec + diff) * (1.0f - s.transparency);
}
SR_DEV float opaque_pdf(const SurfaceD& s, vec3 v, vec3 l, float sp) {
    const float nDotL = dot3(s.normal, l);
    if (nDotL <= 0.0f) return 0.0f;
    const vec3 h = normalized_or(v + l, s.normal);
    const float nDotH = net_max(0.0f, dot3(s.normal, h));
    const float hDotV = net_max(1e-6f, dot3(h, v));
    const float r2 = net_max(s.roughness * s.roughness, 1e-4f);
    const float pdfSpec = ggx_d(nDotH, r2) * nDotH / (4.0f * hDotV);
    const float pdfDiff = nDotL / kPi;
    return (sp * pdfSpec + (1.0f - sp) * pdfDiff) * (1.0f - s.transparency);
}

struct ScatterD {
    vec3 dir, weight;
    bool hasPdf;
    float pdf;
    bool refracted;
};

template <class R>
SR_DEV ScatterD surface_scatter(const SurfaceD& s, vec3 inc, vec3 view, R& rng) {
    ScatterD out;
    if (next_float(rng) < s.transparency) {
        // TransparentScatter (scene.cs:119-144).
        const bool entering = dot3(inc, s.normalGeo) < 0.0f;
        const vec3 rn = entering ? s.normalGeo : -s.normalGeo;
        const float eta = entering ? 1.0f / s.ior : s.ior;
        const float r2 = net_max(s.roughness * s.roughness, 1e-4f);
        vec3 h = from_tangent(ggx_sample(r2, next_vec2(rng)), s.tangent, s.normal);
        if (!entering) h = -h;
        if (dot3(inc, h) > 0.0f) h = rn;
        const float hDotV = net_max(1e-4f, dot3(h, view));
        const float f = luminance(fresnel3(hDotV, base_refl(s.f0, s.albedo, s.metallic)));
        vec3 rd;
        out.weight = splat3(1.0f);
        out.hasPdf = false;
        out.pdf = 0.0f;
        if (next_float(rng) >= f && refract3(inc, h, eta, rd)) {
            out.dir = clamp_geo_side(s, rd, -rn);
            out.refracted = true;
            return out;
        }
        out.dir = clamp_geo_side(s, reflect3(inc, h), rn);
        out.refracted = false;
        return out;
    }
    // OpaqueScatter (scene.cs:103-117).
    const float sp = spec_prob(s, view);
    vec3 dir;
    if (next_float(rng) < sp) {
        const float r2 = net_max(s.roughness * s.roughness, 1e-4f);
        const vec3 h = from_tangent(ggx_sample(r2, next_vec2(rng)), s.tangent, s.normal);
        dir = reflect3(inc, h);
    } else {
        dir = normalized_or(s.normal + rand_on_sphere(rng), s.normal);
    }
    dir = clamp_geo_side(s, dir, s.normalGeo);
    const float pdf = opaque_pdf(s, view, dir, sp);
    out.dir = dir;
    out.weight = opaque_eval(s, view, dir) / net_max(pdf, 1e-6f);
    out.hasPdf = true;
    out.pdf = pdf;
    out.refracted = false;
    return out;
}

}  // namespace sr
