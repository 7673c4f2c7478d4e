// Two-level closest-hit and any-hit traversal of the reference's BVHs on gfx950, plus the
// reconstruction of the hit surface (Scene.Trace). Shared by the tracer's kernels.
#pragma once

#include "sr_device.h"

namespace sr {

constexpr float kInf = __builtin_huge_valf();
constexpr uint32_t kMaxChild = (1u << 24) - 1;
constexpr int kStack = 64;  // bvh.cs:436 caps depth at 64, so at most 64 deferred siblings

// Closest-hit record of a ray: world distance, barycentrics, object and item slot (-1 = miss).
struct HitRec {
    float t, u, v;
    int32_t obj, slot;
};

struct Tally {
    uint32_t box, shape, mesh;
};

// Eight LDS entries per lane: 4 KB per one-wave trace block, so 7 waves per SIMD still fit the
// CU's 160 KB (16 entries cost occupancy; 6 to 10 measure the same).
#ifndef SR_LDS_DEPTH
#define SR_LDS_DEPTH 8
#endif
constexpr int kTravBlock = 256;        // threads per block of the traversal kernels
constexpr int kLdsDepth = SR_LDS_DEPTH;  // stack entries per lane kept in LDS

// Mesh traversal stack: the first LD entries of each lane live in LDS (each wave owns LD rows of
// 64 entries, one column per lane, so a wave's accesses are bank-conflict free); deeper entries
// spill to private memory. The column pointer is a value of the traversal, not a member: stored
// beside the private arrays it lived in scratch and every LDS access became a generic (flat) load
// through a pointer reloaded from scratch.
template <int LD>
struct TravStack {
    uint32_t node[kStack - LD];
    float t[kStack - LD];
    SR_DEV void put(uint2* col, int i, uint32_t n, float tt) {
        if (LD > 0 && i < LD) col[i * 64] = make_uint2(n, __float_as_uint(tt));
        else { node[i - LD] = n; t[i - LD] = tt; }
    }
    SR_DEV void get(const uint2* col, int i, uint32_t& n, float& tt) const {
        if (LD > 0 && i < LD) {
            const uint2 e = col[i * 64];
            n = e.x;
            tt = __uint_as_float(e.y);
        } else {
            // The empty asm statements keep the compiler from merging this branch's loads with the
            // LDS load above into one generic (flat) load from a selected address, which would turn
            // every pop of an LDS entry into a flat load waiting on all outstanding memory operations.
            asm volatile("" ::: "memory");
            n = node[i - LD];
            tt = t[i - LD];
            asm volatile("" ::: "memory");
        }
    }
};
// Dynamic LDS of a traversal launch with `threads` threads per block.
constexpr size_t trav_lds_bytes(int threads) { return (size_t)kLdsDepth * (size_t)threads * sizeof(uint2); }

// A node's stack word (count << 24 | child), stored in its child field by the node pool.
SR_DEV uint32_t node_word(float wordBits) { return (uint32_t)__float_as_int(wordBits); }

// Closest hit in one mesh BVH (bvh.cs:165-216 + mesh.cs:51-62). Nearer child first, ties to
// child B; prune when the entry distance is >= the best hit; accept strictly closer hits.
// While-while form: interior steps repeat until the lane holds a leaf (or its stack runs dry),
// then one leaf step, so a wave runs its box tests and its triangle tests as separate batches
// instead of paying for both whenever any lane sits on a leaf. Each lane visits the same node
// sequence as the plain loop over (pop, test) (same hits, same counters).
template <bool COUNT, int LD>
SR_DEV bool mesh_closest(const DevScene& S, int32_t nodeBase, int32_t triBase, const ray3& r, float& bestT,
                         float& bu, float& bv, int32_t& bslot, Tally& tl, uint2* lds) {
    const float4* N = S.nodes;
    TravStack<LD> st;
    int sp = 0;
    bestT = kInf;
    bool found = false;
    const float4 root = N[(size_t)(nodeBase + 1) * 2 + 1];
    uint32_t cur = node_word(root.z);
    float curT = 0.0f;
    bool live = true;
    while (live) {
        while ((cur >> 24) == 0) {
            bool next = false;
            if (curT < bestT) {
                if (COUNT) tl.box += 2;
                const float4* P = N + (size_t)(nodeBase + 1 + (int32_t)(cur & kMaxChild)) * 2;
                const float4 a0 = P[0], a1 = P[1], b0 = P[2], b1 = P[3];
                float tA = 0.0f, tB = 0.0f;
                const bool hA = slab(a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, r, bestT, tA);
                const bool hB = slab(b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, r, bestT, tB);
                const uint32_t pA = node_word(a1.z), pB = node_word(b1.z);
                if (hA && hB) {
                    const bool aNear = tA < tB;
                    st.put(lds, sp, aNear ? pB : pA, aNear ? tB : tA);
                    ++sp;
                    cur = aNear ? pA : pB;
                    curT = aNear ? tA : tB;
                    next = true;
                } else if (hA) {
                    cur = pA; curT = tA; next = true;
                } else if (hB) {
                    cur = pB; curT = tB; next = true;
                }
            }
            if (!next) {
                if (sp == 0) { live = false; break; }
                --sp;
                st.get(lds, sp, cur, curT);
            }
        }
        if (!live) break;
        if (curT < bestT) {
            const uint32_t cnt = cur >> 24, child = cur & kMaxChild;
            if (COUNT) tl.shape += cnt;
            const float* T = S.tris + (size_t)(triBase + (int32_t)child) * 9;
            for (uint32_t i = 0; i < cnt; ++i, T += 9) {
                float t, u, v;
                if (tri_hit(T, r, t, u, v) && t < bestT) {
                    bestT = t; bu = u; bv = v;
                    bslot = (int32_t)(child + i);
                    found = true;
                }
            }
        }
        if (sp == 0) break;
        --sp;
        st.get(lds, sp, cur, curT);
    }
    return found;
}

// Whether any primitive of a mesh is hit (the same node visits as the closest-hit traversal up
// to its first accepted hit, which is where the existence of a hit is decided); while-while form.
template <bool COUNT, int LD>
SR_DEV bool mesh_any(const DevScene& S, int32_t nodeBase, int32_t triBase, const ray3& r, Tally& tl, uint2* lds) {
    const float4* N = S.nodes;
    TravStack<LD> st;
    int sp = 0;
    const float4 root = N[(size_t)(nodeBase + 1) * 2 + 1];
    uint32_t cur = node_word(root.z);
    float unused;
    for (;;) {
        while ((cur >> 24) == 0) {
            if (COUNT) tl.box += 2;
            const float4* P = N + (size_t)(nodeBase + 1 + (int32_t)(cur & kMaxChild)) * 2;
            const float4 a0 = P[0], a1 = P[1], b0 = P[2], b1 = P[3];
            float tA = 0.0f, tB = 0.0f;
            const bool hA = slab(a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, r, kInf, tA);
            const bool hB = slab(b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, r, kInf, tB);
            const uint32_t pA = node_word(a1.z), pB = node_word(b1.z);
            if (hA && hB) {
                const bool aNear = tA < tB;
                st.put(lds, sp++, aNear ? pB : pA, 0.0f);
                cur = aNear ? pA : pB;
            } else if (hA) {
                cur = pA;
            } else if (hB) {
                cur = pB;
            } else {
                if (sp == 0) return false;
                st.get(lds, --sp, cur, unused);
            }
        }
        const uint32_t cnt = cur >> 24, child = cur & kMaxChild;
        const float* T = S.tris + (size_t)(triBase + (int32_t)child) * 9;
        for (uint32_t i = 0; i < cnt; ++i, T += 9) {
            if (COUNT) tl.shape += 1;
            float t, u, v;
            if (tri_hit(T, r, t, u, v)) return true;
        }
        if (sp == 0) return false;
        st.get(lds, --sp, cur, unused);
    }
}

// Speculative form of mesh_closest (Aila and Laine's speculative while-while): a lane that reaches
// a leaf postpones it and keeps descending while other lanes of its wave still look for one, so the
// wave's triangle tests run with more lanes busy. Leaves are still intersected in the order the
// plain traversal reaches them, with the same best distance at each, so the hit is the same (a
// leaf the plain traversal would have pruned holds no strictly closer hit); the extra box tests
// come from descending before a postponed leaf has lowered the best distance, which is why
// counting renders use mesh_closest itself.
// `bound`: hits at or beyond it are not looked for (kInf for the whole mesh); the closest hit below
// it is the one the whole-mesh search finds, since every node holding it or an equal hit found
// before it enters below the bound.
template <int LD>
SR_DEV bool mesh_closest_spec(const DevScene& S, int32_t nodeBase, int32_t triBase, const ray3& r, float& bestT,
                              float& bu, float& bv, int32_t& bslot, uint2* lds, float bound = kInf) {
    const float4* N = S.nodes;
    TravStack<LD> st;
    int sp = 0;
    bestT = bound;
    bool found = false;
    const float4 root = N[(size_t)(nodeBase + 1) * 2 + 1];
    uint32_t cur = node_word(root.z);
    float curT = 0.0f;
    bool live = true;   // cur is a node still to visit
    uint32_t leaf = 0;  // postponed leaf (count << 24 | first item; 0 = none)
    float leafT = 0.0f;
    for (;;) {
        while (live) {
            if ((cur >> 24) != 0) {
                if (leaf != 0) break;  // a second leaf: intersect the postponed one first
                leaf = cur;
                leafT = curT;
                if (sp == 0) { live = false; break; }
                --sp;
                st.get(lds, sp, cur, curT);
            } else {
                bool next = false;
                if (curT < bestT) {
                    const float4* P = N + (size_t)(nodeBase + 1 + (int32_t)(cur & kMaxChild)) * 2;
                    const float4 a0 = P[0], a1 = P[1], b0 = P[2], b1 = P[3];
                    float tA = 0.0f, tB = 0.0f;
                    const bool hA = slab(a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, r, bestT, tA);
                    const bool hB = slab(b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, r, bestT, tB);
                    const uint32_t pA = node_word(a1.z), pB = node_word(b1.z);
                    // Nearer child first, ties to B, as selects: one push of the farther child
                    // when both are hit.
                    const bool goA = hA && (!hB || tA < tB);
                    if (hA && hB) {
                        st.put(lds, sp, goA ? pB : pA, goA ? tB : tA);
                        ++sp;
                    }
                    cur = goA ? pA : pB;
                    curT = goA ? tA : tB;
                    next = hA || hB;
                }
                if (!next) {
                    if (sp == 0) { live = false; break; }
                    --sp;
                    st.get(lds, sp, cur, curT);
                }
            }
            // Every lane still descending holds a leaf: intersect them together.
            if (__ballot(leaf != 0) == __ballot(1)) break;
        }
        if (leaf != 0) {
            if (leafT < bestT) {
                const uint32_t cnt = leaf >> 24, child = leaf & kMaxChild;
                const float* T = S.tris + (size_t)(triBase + (int32_t)child) * 9;
                for (uint32_t i = 0; i < cnt; ++i, T += 9) {
                    float t, u, v;
                    if (tri_hit(T, r, t, u, v) && t < bestT) {
                        bestT = t; bu = u; bv = v;
                        bslot = (int32_t)(child + i);
                        found = true;
                    }
                }
            }
            leaf = 0;
        } else if (!live) {
            break;
        }
    }
    return found;
}

// Speculative form of mesh_any, used outside counting renders: lanes holding a leaf keep
// descending while other lanes of the wave look for one. Whether a hit exists does not depend on
// the order of the leaves.
template <int LD>
SR_DEV bool mesh_any_spec(const DevScene& S, int32_t nodeBase, int32_t triBase, const ray3& r, uint2* lds) {
    const float4* N = S.nodes;
    TravStack<LD> st;
    int sp = 0;
    const float4 root = N[(size_t)(nodeBase + 1) * 2 + 1];
    uint32_t cur = node_word(root.z);
    float unused;
    bool live = true;
    uint32_t leaf = 0;
    for (;;) {
        while (live) {
            if ((cur >> 24) != 0) {
                if (leaf != 0) break;
                leaf = cur;
                if (sp == 0) { live = false; break; }
                st.get(lds, --sp, cur, unused);
            } else {
                const float4* P = N + (size_t)(nodeBase + 1 + (int32_t)(cur & kMaxChild)) * 2;
                const float4 a0 = P[0], a1 = P[1], b0 = P[2], b1 = P[3];
                float tA = 0.0f, tB = 0.0f;
                const bool hA = slab(a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, r, kInf, tA);
                const bool hB = slab(b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, r, kInf, tB);
                const uint32_t pA = node_word(a1.z), pB = node_word(b1.z);
                const bool goA = hA && (!hB || tA < tB);
                if (hA && hB) st.put(lds, sp++, goA ? pB : pA, 0.0f);
                if (hA || hB) {
                    cur = goA ? pA : pB;
                } else {
                    if (sp == 0) { live = false; break; }
                    st.get(lds, --sp, cur, unused);
                }
            }
            if (__ballot(leaf != 0) == __ballot(1)) break;
        }
        if (leaf != 0) {
            const uint32_t cnt = leaf >> 24, child = leaf & kMaxChild;
            const float* T = S.tris + (size_t)(triBase + (int32_t)child) * 9;
            for (uint32_t i = 0; i < cnt; ++i, T += 9) {
                float t, u, v;
                if (tri_hit(T, r, t, u, v)) return true;
            }
            leaf = 0;
        } else if (!live) {
            return false;
        }
    }
}

// The object-space ray of one object transform (Transform.TransformRayInv, math.cs:816-822) kept
// while a traversal visits objects: objects with a bit-identical transform (DevObject::xform) get
// the same ray, so it is computed once for all of them (the bench scene's nine objects share one).
struct ObjRay {
    int32_t xform = -1;
    ray3 r;
    float scale;
    SR_DEV const ray3& get(const DevObject& ob, const ray3& world) {
        if (ob.xform != xform) {
            r = to_object(ob, world, scale);
            xform = ob.xform;
        }
        return r;
    }
};

// Object.Intersect (scene.cs:227 -> math.cs:67-79): distance in world units only; the hit
// record keeps (u, v, slot) so shading can rebuild the full ShapeHit. `best` is the closest world
// distance found so far: outside counting renders the mesh search stops looking at object-space
// distances whose world distance would round to `best` or beyond (a bound 2^-18 above
// best * scale), which loses no hit the caller would accept (it takes d < best only).
template <bool COUNT, bool SPEC, int LD>
SR_DEV bool object_closest(const DevScene& S, const DevObject& ob, const ray3& r, ObjRay& orr, float best, float& dist,
                           float& u, float& v, int32_t& slot, Tally& tl, uint2* lds) {
    const ray3& lr = orr.get(ob, r);
    const float scale = orr.scale;
    float t = 0.0f;
    slot = -1;
    u = v = 0.0f;
    bool hit;
    if (ob.kind == 0) {
        if (COUNT) tl.mesh += 1;
        if (SPEC && !COUNT) {
            // best * scale in the normal range keeps its relative rounding error below 2^-24.
            const float bs = best * scale;
            const float bound = best < kInf && bs >= 1e-30f ? bs * (1.0f + 0x1p-18f) : kInf;
            hit = ob.nodeBase >= 0 && mesh_closest_spec<LD>(S, ob.nodeBase, ob.triBase, lr, t, u, v, slot, lds, bound);
        } else {
            hit = ob.nodeBase >= 0 && mesh_closest<COUNT, LD>(S, ob.nodeBase, ob.triBase, lr, t, u, v, slot, tl, lds);
        }
    } else {
        ShapeHitD h;
        hit = ob.kind == 1 ? sphere_hit(ob.shape, lr, h, false) : box_hit(ob.shape, lr, h, false);
        t = h.dist;
    }
    if (!hit) return false;
    dist = t / scale;
    return true;
}

template <bool COUNT, int LD>
SR_DEV bool object_any(const DevScene& S, const DevObject& ob, const ray3& r, ObjRay& orr, Tally& tl, uint2* lds) {
    const ray3& lr = orr.get(ob, r);
    if (ob.kind == 0) {
        if (COUNT) tl.mesh += 1;
        if (!COUNT) return ob.nodeBase >= 0 && mesh_any_spec<LD>(S, ob.nodeBase, ob.triBase, lr, lds);
        return ob.nodeBase >= 0 && mesh_any<COUNT, LD>(S, ob.nodeBase, ob.triBase, lr, tl, lds);
    }
    ShapeHitD h;
    return ob.kind == 1 ? sphere_hit(ob.shape, lr, h, false) : box_hit(ob.shape, lr, h, false);
}

// Scene BVH over objects (costI 10, 32 bins; scene.cs:528-533), two-level closest hit, in the
// while-while form of the mesh traversal: a wave descends the object BVH until every lane holds
// an object leaf, so the lanes enter their mesh traversals together.
// LD: mesh-stack entries per lane kept in the caller's LDS rows (0: all in private memory).
template <bool COUNT, bool SPEC = false, int LD = kLdsDepth>
SR_DEV bool scene_closest(const DevScene& S, const ray3& r, HitRec& out, Tally& tl, uint2* lds) {
    if (S.sceneNodes == 0) return false;
    const float4* N = S.nodes;
    uint32_t stk[kStack];
    float stt[kStack];
    int sp = 0;
    float best = kInf;
    bool found = false;
    const float4 root = N[(size_t)(S.sceneBase + 1) * 2 + 1];
    uint32_t cur = node_word(root.z);
    float curT = 0.0f;
    bool live = true;
    ObjRay orr;
    while (live) {
        while ((cur >> 24) == 0) {
            bool next = false;
            if (curT < best) {
                if (COUNT) tl.box += 2;
                const float4* P = N + (size_t)(S.sceneBase + 1 + (int32_t)(cur & kMaxChild)) * 2;
                const float4 a0 = P[0], a1 = P[1], b0 = P[2], b1 = P[3];
                float tA = 0.0f, tB = 0.0f;
                const bool hA = slab(a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, r, best, tA);
                const bool hB = slab(b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, r, best, tB);
                const uint32_t pA = node_word(a1.z), pB = node_word(b1.z);
                if (hA && hB) {
                    const bool aNear = tA < tB;
                    stk[sp] = aNear ? pB : pA;
                    stt[sp] = aNear ? tB : tA;
                    ++sp;
                    cur = aNear ? pA : pB;
                    curT = aNear ? tA : tB;
                    next = true;
                } else if (hA) {
                    cur = pA; curT = tA; next = true;
                } else if (hB) {
                    cur = pB; curT = tB; next = true;
                }
            }
            if (!next) {
                if (sp == 0) { live = false; break; }
                --sp;
                cur = stk[sp];
                curT = stt[sp];
            }
        }
        if (!live) break;
        if (curT < best) {
            const uint32_t cnt = cur >> 24, child = cur & kMaxChild;
            if (COUNT) tl.shape += cnt;
            for (uint32_t i = 0; i < cnt; ++i) {
                const int32_t oi = S.sceneItems[child + i];
                float d, u, v;
                int32_t slot;
                if (object_closest<COUNT, SPEC, LD>(S, S.objects[oi], r, orr, best, d, u, v, slot, tl, lds) && d < best) {
                    best = d;
                    out.t = d; out.u = u; out.v = v; out.obj = oi; out.slot = slot;
                    found = true;
                }
            }
        }
        if (sp == 0) break;
        --sp;
        cur = stk[sp];
        curT = stt[sp];
    }
    return found;
}

// Whether a ray reaches no surface at all: for scenes without transparent objects this decides
// ShadowTransmittance (scene.cs:576-600) exactly as its closest-hit trace does.
template <bool COUNT, int LD = kLdsDepth>
SR_DEV bool scene_any(const DevScene& S, const ray3& r, Tally& tl, uint2* lds) {
    if (S.sceneNodes == 0) return false;
    const float4* N = S.nodes;
    uint32_t stk[kStack];
    int sp = 0;
    const float4 root = N[(size_t)(S.sceneBase + 1) * 2 + 1];
    uint32_t cur = node_word(root.z);
    ObjRay orr;
    for (;;) {
        while ((cur >> 24) == 0) {
            if (COUNT) tl.box += 2;
            const float4* P = N + (size_t)(S.sceneBase + 1 + (int32_t)(cur & kMaxChild)) * 2;
            const float4 a0 = P[0], a1 = P[1], b0 = P[2], b1 = P[3];
            float tA = 0.0f, tB = 0.0f;
            const bool hA = slab(a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, r, kInf, tA);
            const bool hB = slab(b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, r, kInf, tB);
            const uint32_t pA = node_word(a1.z), pB = node_word(b1.z);
            if (hA && hB) {
                const bool aNear = tA < tB;
                stk[sp++] = aNear ? pB : pA;
                cur = aNear ? pA : pB;
            } else if (hA) {
                cur = pA;
            } else if (hB) {
                cur = pB;
            } else {
                if (sp == 0) return false;
                cur = stk[--sp];
            }
        }
        const uint32_t cnt = cur >> 24, child = cur & kMaxChild;
        for (uint32_t i = 0; i < cnt; ++i) {
            if (COUNT) tl.shape += 1;
            if (object_any<COUNT, LD>(S, S.objects[S.sceneItems[child + i]], r, orr, tl, lds)) return true;
        }
        if (sp == 0) return false;
        cur = stk[--sp];
    }
}
// Scene.Trace (scene.cs:602-625) from a hit record: world-space ShapeHit, material sampling,
// normal mapping and tangent re-orthogonalisation.
SR_DEV void make_surface(const DevScene& S, const ray3& r, const HitRec& h, SurfaceD& s) {
    const DevObject& ob = S.objects[h.obj];
    ShapeHitD sh;
    if (ob.kind == 0) {
        inflate(S.tris + (size_t)(ob.triBase + h.slot) * 9, S.attrs + (size_t)(ob.triBase + h.slot) * 27, h.u, h.v, sh);
    } else {
        float scale;
        const ray3 lr = to_object(ob, r, scale);
        if (ob.kind == 1) sphere_hit(ob.shape, lr, sh, true);
        else box_hit(ob.shape, lr, sh, true);
    }
    const vec3 wn = normalized(rotate(ob.rot, sh.norm / ob.scale));
    const vec3 wng = normalized(rotate(ob.rot, sh.normGeo / ob.scale));
    const vec3 wt = normalized(rotate(ob.rot, xyz(sh.tan)));
    const vec2 uv = sh.uv;
    const vec3 texC = ob.texColor >= 0 ? tex_sample(S, ob.texColor, uv) : splat3(1.0f);
    s.albedo = texC * ob.albedo;
    s.roughness = (ob.texRough >= 0 ? tex_sample(S, ob.texRough, uv).x : 1.0f) * ob.roughness;
    s.metallic = (ob.texMetal >= 0 ? tex_sample(S, ob.texMetal, uv).x : 1.0f) * ob.metallic;
    const vec4 htan = mk4(wt.x, wt.y, wt.z, sh.tan.w);
    vec3 n = wn;
    if (ob.texNormal >= 0) {
        const vec3 c = tex_sample(S, ob.texNormal, uv);
        const vec3 tn = normalized_or(mk3(c.x * 2.0f - 1.0f, -(c.y * 2.0f - 1.0f), c.z * 2.0f - 1.0f), up3());
        n = normalized_or(from_tangent(tn, htan, wn), wn);
    }
    const vec3 td = normalized_or(wt - dot3(wt, n) * n, wt);
    s.radiance = ob.radiance;
    s.transparency = ob.transparency;
    s.ior = ob.ior;
    s.f0 = ob.f0;
    s.normal = n;
    s.normalGeo = wng;
    s.tangent = mk4(td.x, td.y, td.z, sh.tan.w);
    s.uv = uv;
}

}  // namespace sr
