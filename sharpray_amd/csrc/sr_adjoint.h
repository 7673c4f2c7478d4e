// Adjoint rendering: gradients of sum(dImage * Radiance) with respect to every object's albedo and
// roughness and the sky texture's texels. No counterpart in the reference; the forward estimator it
// differentiates is Scene.Sample (scene.cs:627-741) as rendered with the counter RNG policy.
//
// One lane replays one (pixel, sample) path with the same random numbers as the forward render and
// keeps, per vertex, the contribution it added and d log(BSDF weight) / d parameter; summing the
// contributions from the end of the path gives each scatter's downstream radiance, which that
// scatter's log-derivative scales (path-replay backpropagation with per-vertex storage). Because the
// reference's BSDF sampling is not an exact importance sampler (ClampToGeoSide folds the density
// of mirrored directions onto the upper side, so the drawn directions' density q differs from the
// OpaquePdf p the estimator divides by), each scatter also carries d/dθ [log q - log p] (a score
// term), and the power-heuristic MIS weights are differentiated through both pdfs (the BSDF's and
// the sky's, whose CDF total ties every texel in). Russian roulette (probability and 1/survive
// cancel), the light-sampling pmf (its score cancels the 1/pdf derivative) and the indirect
// clamp's rescale are held fixed. Opaque surfaces, no fog.
#pragma once

#include "sr_device.h"
#include "sr_trace.h"

namespace sr {

constexpr int kAdjMaxVertices = 17;  // bounces <= 16

// d GgxD(nDotH, max(r^2, 1e-4)) / d r.
SR_DEV float ggx_d_dr(float roughness, float nh) {
    if (!(roughness * roughness > 1e-4f)) return 0.0f;
    const float r2 = roughness * roughness, a2 = r2 * r2, nh2 = nh * nh;
    const float dd = (1.0f - nh2) + nh2 * a2;
    return (1.0f / (kPi * dd * dd) - 2.0f * a2 * nh2 / (kPi * dd * dd * dd)) * 2.0f * r2 * 2.0f * roughness;
}

// d OpaqueEval / d albedo (per channel) and d OpaqueEval / d roughness.
SR_DEV void opaque_eval_grad(const SurfaceD& s, vec3 v, vec3 l, vec3& dA, vec3& dR) {
    dA = splat3(0.0f);
    dR = splat3(0.0f);
    const float nDotL = dot3(s.normal, l);
    if (nDotL <= 0.0f) return;
    const float nDotV = net_max(1e-4f, dot3(s.normal, v));
    const float r2 = net_max(s.roughness * s.roughness, 1e-4f);
    const vec3 F0 = base_refl(s.f0, s.albedo, s.metallic);
    const vec3 h = normalized_or(v + l, s.normal);
    const float nDotH = dot3(s.normal, h);
    const float hDotV = dot3(h, v);
    const float xh = 1.0f - hDotV, xv = 1.0f - nDotV;
    const float xh5 = xh * xh * xh * xh * xh, xv5 = xv * xv * xv * xv * xv;
    const vec3 fh = fresnel3(hDotV, F0), fv = fresnel3(nDotV, F0);
    const float g1v = smith_g1(nDotV, r2), g1l = smith_g1(nDotL, r2);
    const float dist = ggx_d(nDotH, r2);
    const float keep = 1.0f - s.transparency;
    const float specA = dist * g1v * g1l / (4.0f * nDotV) * s.metallic * (1.0f - xh5);
    const float diffK = (1.0f - s.metallic) * (nDotL / kPi);
    const float m5 = s.metallic * (1.0f - xv5);
    dA = mk3(specA + diffK * ((1.0f - fv.x) - s.albedo.x * m5), specA + diffK * ((1.0f - fv.y) - s.albedo.y * m5),
             specA + diffK * ((1.0f - fv.z) - s.albedo.z * m5)) * keep;
    if (s.roughness * s.roughness > 1e-4f) {
        const float a2 = r2 * r2, nh2 = nDotH * nDotH;
        const float dd = (1.0f - nh2) + nh2 * a2;
        const float dD = (1.0f / (kPi * dd * dd) - 2.0f * a2 * nh2 / (kPi * dd * dd * dd)) * 2.0f * r2;
        const float k = r2 / 2.0f;
        const float denV = nDotV * (1.0f - k) + k, denL = nDotL * (1.0f - k) + k;
        const float dgV = -nDotV * (1.0f - nDotV) / (denV * denV), dgL = -nDotL * (1.0f - nDotL) / (denL * denL);
        const float dG = 0.5f * (dgV * g1l + g1v * dgL);
        const float dSpec = (dD * g1v * g1l + dist * dG) / (4.0f * nDotV) * (2.0f * s.roughness);
        dR = fh * dSpec * keep;
    }
}

// d SpecularProbability / d albedo (zero where the metallic term or the clamp decides).
SR_DEV vec3 spec_prob_grad(const SurfaceD& s, vec3 v) {
    const float nDotV = net_max(1e-4f, dot3(s.normal, v));
    const float lumF = luminance(fresnel3(nDotV, base_refl(s.f0, s.albedo, s.metallic)));
    if (!(lumF > s.metallic && lumF > 0.001f && lumF < 0.999f)) return splat3(0.0f);
    const float xv = 1.0f - nDotV;
    const float base = s.metallic * (1.0f - xv * xv * xv * xv * xv);
    return mk3(0.2126f * base, 0.7152f * base, 0.0722f * base);
}

// d OpaquePdf(v, l, sp(albedo)) / d albedo and / d roughness.
SR_DEV void opaque_pdf_grad(const SurfaceD& s, vec3 v, vec3 l, float sp, vec3& dA, float& dR) {
    dA = splat3(0.0f);
    dR = 0.0f;
    const float nDotL = dot3(s.normal, l);
    if (nDotL <= 0.0f) return;
    const vec3 h = normalized_or(v + l, s.normal);
    const float nh = net_max(0.0f, dot3(s.normal, h));
    const float hv = net_max(1e-6f, dot3(h, v));
    const float r2 = net_max(s.roughness * s.roughness, 1e-4f);
    const float keep = 1.0f - s.transparency;
    const float pSpec = ggx_d(nh, r2) * nh / (4.0f * hv), pDiff = nDotL / kPi;
    dR = sp * ggx_d_dr(s.roughness, nh) * nh / (4.0f * hv) * keep;
    dA = spec_prob_grad(s, v) * ((pSpec - pDiff) * keep);
}

// Density of the specular candidate direction d (GGX half vector reflecting v) and its roughness
// derivative, for any d (also below the surface).
SR_DEV void spec_candidate_pdf(const SurfaceD& s, vec3 v, vec3 d, float& pdf, float& dPdf) {
    pdf = dPdf = 0.0f;
    const vec3 hv = v + d;
    if (len_sqr(hv) <= 1e-12f) return;
    vec3 h = normalized(hv);
    float nh = dot3(s.normal, h);
    if (nh < 0.0f) { h = -h; nh = -nh; }
    const float vh = fabsf(dot3(h, v));
    if (!(vh > 1e-6f)) return;
    const float r2 = net_max(s.roughness * s.roughness, 1e-4f);
    pdf = ggx_d(nh, r2) * nh / (4.0f * vh);
    dPdf = ggx_d_dr(s.roughness, nh) * nh / (4.0f * vh);
}

// d/dθ [log q(l) - log p(l)] for a direction l drawn by OpaqueScatter (q: density of the drawing
// procedure including the mirror fold, p: the OpaquePdf divided by).
SR_DEV void scatter_score(const SurfaceD& s, vec3 v, vec3 l, float sp, vec3& sA, float& sR) {
    sA = splat3(0.0f);
    sR = 0.0f;
    const float p = opaque_pdf(s, v, l, sp);
    const vec3 hh = normalized_or(v + l, s.normal);
    const float nh = net_max(0.0f, dot3(s.normal, hh));
    const float hv = net_max(1e-6f, dot3(hh, v));
    const float r2 = net_max(s.roughness * s.roughness, 1e-4f);
    float pSpec = ggx_d(nh, r2) * nh / (4.0f * hv), pDiff = dot3(s.normal, l) / kPi;
    float dpSpec = ggx_d_dr(s.roughness, nh) * nh / (4.0f * hv);
    if (dot3(s.normal, l) <= 0.0f) pSpec = pDiff = dpSpec = 0.0f;
    const vec3 m = l - 2.0f * dot3(l, s.normalGeo) * s.normalGeo;
    float mSpec = 0.0f, dmSpec = 0.0f, mDiff = 0.0f;
    if (dot3(m, s.normalGeo) < 0.0f) {
        spec_candidate_pdf(s, v, m, mSpec, dmSpec);
        mDiff = net_max(0.0f, dot3(s.normal, m)) / kPi;
    }
    const float q = p + sp * mSpec + (1.0f - sp) * mDiff;
    if (!(q > 0.0f)) return;
    const bool pLive = p >= 1e-6f;
    const float dpR = sp * dpSpec, dmR = sp * dmSpec;
    sR = (dpR + dmR) / q - (pLive ? dpR / p : 0.0f);
    const float dpS = pSpec - pDiff, dmS = mSpec - mDiff;
    sA = spec_prob_grad(s, v) * ((dpS + dmS) / q - (pLive ? dpS / p : 0.0f));
}

// Adds s_c * multiplier * (bilinear weight) to g[texel * 3 + c] for the texels SkyTexture.Radiance
// interpolates for direction d.
SR_DEV void sky_radiance_grad(const DevScene& S, vec3 d, vec3 s, float* g) {
    const DevSky& k = S.sky;
    if (k.kind != 1) return;
    const DevTexture t = S.textures[k.tex];
    const vec2 c = equirect_uv(rotate(k.rotInv, d));
    float u = c.x * t.tilx;
    u -= floorf(u);
    float v = c.y * t.tily;
    v -= floorf(v);
    const uint32_t W = (uint32_t)t.w, H = (uint32_t)t.h;
    uint32_t idx[4];
    float w[4];
    if (t.filter == 1) {
        idx[0] = ((uint32_t)(v * (float)H) % H) * W + (uint32_t)(u * (float)W) % W;
        w[0] = 1.0f;
        idx[1] = idx[2] = idx[3] = idx[0];
        w[1] = w[2] = w[3] = 0.0f;
    } else {
        const float fx = u * (float)(W - 1), fy = v * (float)(H - 1);
        const uint32_t x0 = (uint32_t)fx, y0 = (uint32_t)fy;
        const uint32_t x1 = (x0 + 1) % W, y1 = (y0 + 1) % H;
        const float tx = fx - (float)x0, ty = fy - (float)y0;
        idx[0] = y0 * W + x0; w[0] = (1.0f - tx) * (1.0f - ty);
        idx[1] = y0 * W + x1; w[1] = tx * (1.0f - ty);
        idx[2] = y1 * W + x0; w[2] = (1.0f - tx) * ty;
        idx[3] = y1 * W + x1; w[3] = tx * ty;
    }
    for (int i = 0; i < 4; ++i) {
        const float kw = w[i] * k.multiplier;
        if (kw == 0.0f) continue;
        unsafeAtomicAdd(&g[idx[i] * 3], s.x * kw);
        unsafeAtomicAdd(&g[idx[i] * 3 + 1], s.y * kw);
        unsafeAtomicAdd(&g[idx[i] * 3 + 2], s.z * kw);
    }
}

// Texel LightDir's point lookup reads for direction d (texture sky).
SR_DEV uint32_t sky_light_texel(const DevScene& S, vec3 d) {
    const DevSky& k = S.sky;
    const vec2 uv = equirect_uv(rotate(k.rotInv, d));
    const int32_t cx = ((int32_t)(uv.x * (float)k.w)) % k.w;
    const int32_t cy = ((int32_t)(uv.y * (float)k.h)) % k.h;
    return (uint32_t)(cy * k.w + cx);
}

// Material-gradient accumulator: per-block LDS slots for up to kAdjLdsObjects objects, global
// atomics beyond; gTotal gathers the part of sky-pdf derivatives that runs through the CDF total.
constexpr int kAdjLdsObjects = 256;
struct MatGrad {
    float* lds;     // 4 per object
    float* global;  // 4 per object
    uint32_t nObj;
    float gTotal;
    SR_DEV void add(int32_t obj, int k, float v) {
        if (v == 0.0f) return;
        if (obj < kAdjLdsObjects) atomicAdd(&lds[obj * 4 + k], v);
        else unsafeAtomicAdd(&global[obj * 4 + k], v);
    }
    // Derivative of a texel-sky pdf p = lum(texel) * W * H / (total * 2 pi^2) scaled by s: the
    // texel's own luminance directly, every texel through total (spread once after the pass).
    SR_DEV void sky_pdf(const DevScene& S, uint32_t texel, float pdf, float s, float* gEnv) {
        if (S.sky.kind != 1 || s == 0.0f) return;
        const float k = s * S.sky.pdfScale;
        unsafeAtomicAdd(&gEnv[texel * 3], k * 0.2126f);
        unsafeAtomicAdd(&gEnv[texel * 3 + 1], k * 0.7152f);
        unsafeAtomicAdd(&gEnv[texel * 3 + 2], k * 0.0722f);
        gTotal += s * (-pdf / S.sky.total);
    }
};

// One path's adjoint (the forward steps are those of the shade kernel's surface and sky branches,
// with the shadow ray traced inline).
SR_DEV void adjoint_path(const DevScene& S, ray3 ray, CounterRngT<false>& rng, uint32_t bounces, float clampLum,
                         vec3 delta, MatGrad& gm, float* gEnv, uint2* lds) {
    vec3 energy = splat3(1.0f);
    bool hasScatterPdf = false;
    float scatterPdf = 0.0f;
    int32_t prevObj = -1;
    vec3 prevDpA = splat3(0.0f);
    float prevDpR = 0.0f;
    int32_t vObj[kAdjMaxVertices];
    vec3 vRa[kAdjMaxVertices], vRr[kAdjMaxVertices], vC[kAdjMaxVertices];
    int nv = 0;
    Tally tl = {0, 0, 0};
    for (uint32_t i = 0; i <= bounces; ++i) {
        HitRec h;
        h.t = kInf; h.u = 0.0f; h.v = 0.0f; h.obj = -1; h.slot = -1;
        const bool hit = scene_closest<false, false, 0>(S, ray, h, tl, lds);
        if (!hit) {
            float mis = 1.0f;
            const vec3 skyD = sky_radiance(S, ray.d);
            if (hasScatterPdf) {
                float lp = 0.0f;
                if (!sky_light_pdf(S, ray.d, lp)) lp = 0.0f;
                mis = power_heuristic(scatterPdf, lp);
                const float pb = scatterPdf, den = pb * pb + lp * lp;
                const float dw = den > 0.0f ? 2.0f * pb * lp * lp / (den * den) : 0.0f;
                if (lp > 0.0f && den > 0.0f) {
                    const vec3 sw = delta * skyD * energy;
                    gm.sky_pdf(S, sky_light_texel(S, ray.d), lp, (sw.x + sw.y + sw.z) * (-2.0f * pb * pb * lp / (den * den)),
                               gEnv);
                }
                if (prevObj >= 0 && dw != 0.0f) {
                    const vec3 s = delta * skyD * energy * dw;
                    const float sum = s.x + s.y + s.z;
                    gm.add(prevObj, 0, sum * prevDpA.x);
                    gm.add(prevObj, 1, sum * prevDpA.y);
                    gm.add(prevObj, 2, sum * prevDpA.z);
                    gm.add(prevObj, 3, sum * prevDpR);
                }
            }
            sky_radiance_grad(S, ray.d, delta * (energy * mis), gEnv);
            vObj[nv] = -1;
            vC[nv] = skyD * mis * energy;
            ++nv;
            break;
        }
        SurfaceD s;
        make_surface(S, ray, h, s);
        const DevObject& ob = S.objects[h.obj];
        const vec3 tA = ob.texColor >= 0 ? tex_sample(S, ob.texColor, s.uv) : splat3(1.0f);
        const float tR = ob.texRough >= 0 ? tex_sample(S, ob.texRough, s.uv).x : 1.0f;
        const float dist = h.t;
        const vec3 view = -ray.d;
        const float bias = net_max(1e-4f, dist * 1e-4f);
        const vec3 hitPos = at(ray, dist - bias);
        vec3 c = s.radiance * energy;
        vec3 ld;
        float lpdf;
        uint32_t ltex = 0;
        if (sky_light_rand(S, rng, ld, lpdf, &ltex) && lpdf > 0.0f) {
            const vec3 nn = dot3(s.normal, view) >= 0.0f ? s.normal : -s.normal;
            const vec3 ng = dot3(s.normalGeo, view) >= 0.0f ? s.normalGeo : -s.normalGeo;
            if (dot3(nn, ld) > 0.0f && dot3(ng, ld) > 0.0f && !scene_any<false, 0>(S, mk_ray(hitPos, ld), tl, lds)) {
                const vec3 skyL = sky_radiance(S, ld);
                const vec3 refl = opaque_eval(s, view, ld);
                const float sp = spec_prob(s, view);
                const float pb = opaque_pdf(s, view, ld, sp);
                const float w = power_heuristic(lpdf, pb);
                c = c + skyL * refl * w / lpdf * energy;
                vec3 dA, dR, dpA;
                float dpR;
                opaque_eval_grad(s, view, ld, dA, dR);
                opaque_pdf_grad(s, view, ld, sp, dpA, dpR);
                const float pl2 = lpdf * lpdf, den = pl2 + pb * pb;
                const float dw = den > 0.0f ? -2.0f * pl2 * pb / (den * den) : 0.0f;
                if (den > 0.0f) {
                    const vec3 sw = delta * skyL * refl * (1.0f / lpdf) * energy;
                    gm.sky_pdf(S, ltex, lpdf, (sw.x + sw.y + sw.z) * (2.0f * lpdf * pb * pb / (den * den)), gEnv);
                }
                const vec3 common = delta * skyL * (1.0f / lpdf) * energy;
                const vec3 gA = dA * w + refl * dpA * dw, gR = dR * w + refl * (dpR * dw);
                gm.add(h.obj, 0, common.x * gA.x * tA.x);
                gm.add(h.obj, 1, common.y * gA.y * tA.y);
                gm.add(h.obj, 2, common.z * gA.z * tA.z);
                gm.add(h.obj, 3, (common.x * gR.x + common.y * gR.y + common.z * gR.z) * tR);
                sky_radiance_grad(S, ld, delta * refl * (w / lpdf) * energy, gEnv);
            }
        }
        vObj[nv] = h.obj;
        vC[nv] = c;
        if (i >= 3) {
            const float survive = net_min(1.0f, max_comp(energy));
            if (next_float(rng) >= survive) {
                vRa[nv] = splat3(0.0f); vRr[nv] = splat3(0.0f);
                vObj[nv] = -1;  // no scatter at this vertex
                ++nv;
                break;
            }
            energy = energy / survive;
        }
        const float sp = spec_prob(s, view);
        const ScatterD sc = surface_scatter(s, ray.d, view, rng);
        const vec3 f = opaque_eval(s, view, sc.dir);
        vec3 dA, dR, sA;
        float sR;
        opaque_eval_grad(s, view, sc.dir, dA, dR);
        scatter_score(s, view, sc.dir, sp, sA, sR);
        auto ratio = [](float d, float v) { return v > 0.0f ? d / v : 0.0f; };
        vRa[nv] = mk3((ratio(dA.x, f.x) + sA.x) * tA.x, (ratio(dA.y, f.y) + sA.y) * tA.y,
                      (ratio(dA.z, f.z) + sA.z) * tA.z);
        vRr[nv] = mk3((ratio(dR.x, f.x) + sR) * tR, (ratio(dR.y, f.y) + sR) * tR, (ratio(dR.z, f.z) + sR) * tR);
        ++nv;
        energy = energy * sc.weight;
        energy = clamp_lum(energy, clampLum);
        hasScatterPdf = sc.hasPdf;
        scatterPdf = sc.pdf;
        vec3 dpA;
        float dpR;
        opaque_pdf_grad(s, view, sc.dir, sp, dpA, dpR);
        prevObj = h.obj;
        prevDpA = dpA * tA;
        prevDpR = dpR * tR;
        ray = mk_ray(at(ray, dist) + sc.dir * bias, sc.dir);
    }
    vec3 after = splat3(0.0f);
    for (int j = nv - 1; j >= 0; --j) {
        if (vObj[j] >= 0) {
            const vec3 s = delta * after;
            gm.add(vObj[j], 0, s.x * vRa[j].x);
            gm.add(vObj[j], 1, s.y * vRa[j].y);
            gm.add(vObj[j], 2, s.z * vRa[j].z);
            gm.add(vObj[j], 3, s.x * vRr[j].x + s.y * vRr[j].y + s.z * vRr[j].z);
        }
        after = after + vC[j];
    }
}

}  // namespace sr
