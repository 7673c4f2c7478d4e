// Host-side scene preparation. See sr_host.h.
#include "sr_host.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>

namespace sr {

box3 lean_bounds(const float* t) {
    vec3 a = mk3(t[0], t[1], t[2]);
    vec3 b = a + mk3(t[3], t[4], t[5]);
    vec3 c = a + mk3(t[6], t[7], t[8]);
    box3 r;
    r.mn = mk3(net_min(a.x, net_min(b.x, c.x)), net_min(a.y, net_min(b.y, c.y)), net_min(a.z, net_min(b.z, c.z)));
    r.mx = mk3(net_max(a.x, net_max(b.x, c.x)), net_max(a.y, net_max(b.y, c.y)), net_max(a.z, net_max(b.z, c.z)));
    return r;
}

namespace {

// Work-list builder. Each primitive's box and box centre are computed once; the recursion of
// bvh.cs:432-459 becomes an explicit LIFO (child B pushed under child A) so that nodes are
// allocated in the same depth-first pair order.
class SahBuilder {
public:
    SahBuilder(const box3* bounds, int n, int bins, float ct, float ci)
        : bb_(bounds), n_(n), bins_(bins), ct_(ct), ci_(ci), mid_(n), binBox_(bins), binCnt_(bins),
          leftArea_(bins), leftCnt_(bins) {
        for (int i = 0; i < n; ++i) mid_[i] = box_mid(bb_[i]);
    }

    BvhResult run() {
        BvhResult r;
        r.nodes.assign(std::max(n_ * 2, 1), RefNode{});
        r.items.resize(n_);
        std::iota(r.items.begin(), r.items.end(), 0);
        nodes_ = &r.nodes;
        items_ = &r.items;
        if (n_ == 0) return r;
        leaf(0, 0, n_);
        r.used = 1;
        struct Job { int node, depth; };
        std::vector<Job> jobs;
        jobs.push_back({0, 0});
        while (!jobs.empty()) {
            Job j = jobs.back();
            jobs.pop_back();
            r.max_depth = std::max(r.max_depth, j.depth);
            if (j.depth >= 64) continue;
            int axis;
            float pos;
            if (!pick(j.node, axis, pos)) continue;
            RefNode& nd = r.nodes[j.node];
            int begin = nd.child, cnt = nd.count;
            int split = partition(begin, cnt, axis, pos);
            int na = split - begin, nb = cnt - na;
            if (na <= 0 || nb <= 0) r.degenerate = true;
            int ca = r.used;
            leaf(ca, begin, na);
            leaf(ca + 1, split, nb);
            r.used += 2;
            r.nodes[j.node].child = ca;
            r.nodes[j.node].count = 0;
            jobs.push_back({ca + 1, j.depth + 1});
            jobs.push_back({ca, j.depth + 1});
        }
        return r;
    }

private:
    void leaf(int idx, int begin, int count) {
        RefNode& nd = (*nodes_)[idx];
        nd.bounds = box_empty();
        nd.child = begin;
        nd.count = count;
        const int32_t* it = items_->data() + begin;
        for (int i = 0; i < count; ++i) box_grow(nd.bounds, bb_[it[i]]);
    }

    float leaf_cost(float area, int count) const { return ci_ * area * (float)count; }

    // Best (axis, position) over bins_-1 candidate planes per axis (bvh.cs:321-395).
    bool pick(int idx, int& axisOut, float& posOut) {
        const RefNode& nd = (*nodes_)[idx];
        const int32_t* it = items_->data() + nd.child;
        const int cnt = nd.count;
        box3 cb = box_empty();
        for (int i = 0; i < cnt; ++i) box_grow(cb, mid_[it[i]]);
        const float parentArea = box_area(nd.bounds);
        float best = leaf_cost(parentArea, cnt);
        int bestAxis = -1;
        float bestPos = 0.0f;
        for (int axis = 0; axis < 3; ++axis) {
            const float lo = axis_of(cb.mn, axis), hi = axis_of(cb.mx, axis);
            if (hi - lo < 1e-6f) continue;
            std::fill(binBox_.begin(), binBox_.end(), box_empty());
            std::fill(binCnt_.begin(), binCnt_.end(), 0);
            const float scale = (float)bins_ / (hi - lo);
            for (int i = 0; i < cnt; ++i) {
                const int p = it[i];
                int b = (int)((axis_of(mid_[p], axis) - lo) * scale);
                if (b > bins_ - 1) b = bins_ - 1;
                box_grow(binBox_[b], bb_[p]);
                binCnt_[b] += 1;
            }
            box3 acc = box_empty();
            int accCnt = 0;
            for (int b = 0; b < bins_ - 1; ++b) {
                box_grow(acc, binBox_[b]);
                accCnt += binCnt_[b];
                leftArea_[b] = box_is_empty(acc) ? 0.0f : box_area(acc);
                leftCnt_[b] = accCnt;
            }
            acc = box_empty();
            accCnt = 0;
            for (int b = bins_ - 1; b > 0; --b) {
                box_grow(acc, binBox_[b]);
                accCnt += binCnt_[b];
                const float rightArea = box_is_empty(acc) ? 0.0f : box_area(acc);
                const float cost = ct_ * parentArea + leaf_cost(leftArea_[b - 1], leftCnt_[b - 1]) +
                                   leaf_cost(rightArea, accCnt);
                if (cost < best) {
                    best = cost;
                    bestAxis = axis;
                    bestPos = lo + (float)b * (hi - lo) / (float)bins_;
                }
            }
        }
        if (bestAxis < 0) return false;
        axisOut = bestAxis;
        posOut = bestPos;
        return true;
    }

    // Two-cursor swap partition of bvh.cs:401-426: left cursor advances over items whose box
    // centre lies below the plane, otherwise the item is swapped with the right cursor's.
    int partition(int begin, int cnt, int axis, float pos) {
        int32_t* it = items_->data();
        int l = begin, r = begin + cnt - 1;
        for (;;) {
            if (axis_of(mid_[it[l]], axis) < pos) {
                if (++l > r) break;
            } else {
                if (l == r) break;
                std::swap(it[l], it[r]);
                --r;
            }
        }
        return l;
    }

    const box3* bb_;
    int n_, bins_;
    float ct_, ci_;
    std::vector<vec3> mid_;
    std::vector<box3> binBox_;
    std::vector<int> binCnt_;
    std::vector<float> leftArea_;
    std::vector<int> leftCnt_;
    std::vector<RefNode>* nodes_ = nullptr;
    std::vector<int32_t>* items_ = nullptr;
};

vec3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

// Triangle.ComputeTangent (math.cs:1553-1579).
vec4 tangent_of(vec3 pa, vec3 pb, vec3 pc, vec2 ua, vec2 ub, vec2 uc, vec3 n) {
    const vec3 e1 = pb - pa, e2 = pc - pa;
    const float du1 = ub.x - ua.x, dv1 = ub.y - ua.y;
    const float du2 = uc.x - ua.x, dv2 = uc.y - ua.y;
    const float det = du1 * dv2 - du2 * dv1;
    if (fabsf(det) < 1e-8f) {
        vec3 t = normalized_or(e1 - dot3(e1, n) * n, right3());
        return mk4(t.x, t.y, t.z, 1.0f);
    }
    const float inv = 1.0f / det;
    vec3 t = (dv2 * e1 - dv1 * e2) * inv;
    const vec3 bt = (-du2 * e1 + du1 * e2) * inv;
    t = normalized_or(t - dot3(t, n) * n, right3());
    const float w = dot3(cross3(n, t), bt) < 0.0f ? -1.0f : 1.0f;
    return mk4(t.x, t.y, t.z, w);
}

void put3(float* o, vec3 v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
void put4(float* o, vec4 v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }

// Writes one Triangle record (36 floats: lean 9, normals 9, tangents 12, uvs 6).
void emit(float* o, vec3 pa, vec3 pb, vec3 pc, const vec3* nrm, vec2 ua, vec2 ub, vec2 uc) {
    const vec3 ab = pb - pa, ac = pc - pa;
    put3(o, pa);
    put3(o + 3, ab);
    put3(o + 6, ac);
    vec3 na, nb, nc;
    vec4 tg;
    if (nrm) {
        na = nrm[0]; nb = nrm[1]; nc = nrm[2];
        const vec3 geo = normalized_or(cross3(pb - pa, pc - pa), up3());
        const vec3 avg = normalized_or((na + nb + nc) / 3.0f, geo);
        tg = tangent_of(pa, pb, pc, ua, ub, uc, avg);
    } else {
        na = nb = nc = normalized_or(cross3(ab, ac), up3());
        tg = tangent_of(pa, pb, pc, ua, ub, uc, na);
    }
    put3(o + 9, na);
    put3(o + 12, nb);
    put3(o + 15, nc);
    put4(o + 18, tg);
    put4(o + 22, tg);
    put4(o + 26, tg);
    o[30] = ua.x; o[31] = ua.y; o[32] = ub.x; o[33] = ub.y; o[34] = uc.x; o[35] = uc.y;
}

}  // namespace

BvhResult build_bvh(const box3* bounds, int n, int bins, float ct, float ci) {
    SahBuilder b(bounds, n, bins, ct, ci);
    return b.run();
}

void build_triangles(const float* pos, const float* nrm, const float* uv, uint32_t n, bool smooth, float* out,
                     const uint8_t* hasNrm) {
    auto uvs = [&](uint32_t i, vec2& a, vec2& b, vec2& c) {
        if (uv) {
            a = mk2(uv[i * 6], uv[i * 6 + 1]);
            b = mk2(uv[i * 6 + 2], uv[i * 6 + 3]);
            c = mk2(uv[i * 6 + 4], uv[i * 6 + 5]);
        } else {
            a = mk2(0.0f, 0.0f); b = mk2(1.0f, 0.0f); c = mk2(0.0f, 1.0f);
        }
    };
    for (uint32_t i = 0; i < n; ++i) {
        vec2 a, b, c;
        uvs(i, a, b, c);
        vec3 nv[3];
        if (nrm) { nv[0] = ld3(nrm + i * 9); nv[1] = ld3(nrm + i * 9 + 3); nv[2] = ld3(nrm + i * 9 + 6); }
        const bool useN = nrm && (!hasNrm || hasNrm[i]);
        emit(out + (size_t)i * 36, ld3(pos + i * 9), ld3(pos + i * 9 + 3), ld3(pos + i * 9 + 6), useN ? nv : nullptr, a, b,
             c);
    }
    if (!smooth || nrm || n == 0) return;

    // Smooth normals: vertices whose positions round to the same 1e-6 grid cell share the sum of
    // the area-weighted face normals, accumulated in triangle order (A, B, C). Grouping is done by
    // sorting corner keys; ties keep corner order so each group's sum runs in the reference order.
    const float invEps = 1.0f / 1e-6f;
    const size_t corners = (size_t)n * 3;
    struct Key { int64_t x, y, z; };
    std::vector<Key> keys(corners);
    std::vector<vec3> cornerPos(corners);
    std::vector<vec3> weighted(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* o = out + (size_t)i * 36;
        const vec3 pa = ld3(o), ab = ld3(o + 3), ac = ld3(o + 6);
        const vec3 cr = cross3(ab, ac);
        weighted[i] = normalized_or(cr, up3()) * (sqrtf(len_sqr(cr)) * 0.5f);
        cornerPos[i * 3] = pa;
        cornerPos[i * 3 + 1] = pa + ab;
        cornerPos[i * 3 + 2] = pa + ac;
    }
    for (size_t k = 0; k < corners; ++k) {
        const vec3 p = cornerPos[k];
        keys[k] = Key{(int64_t)nearbyintf(p.x * invEps), (int64_t)nearbyintf(p.y * invEps), (int64_t)nearbyintf(p.z * invEps)};
    }
    std::vector<uint32_t> order(corners);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const Key &ka = keys[a], &kb = keys[b];
        if (ka.x != kb.x) return ka.x < kb.x;
        if (ka.y != kb.y) return ka.y < kb.y;
        return ka.z < kb.z;
    });
    std::vector<vec3> sum(corners);
    size_t g = 0;
    while (g < corners) {
        size_t e = g + 1;
        const Key& kg = keys[order[g]];
        while (e < corners && keys[order[e]].x == kg.x && keys[order[e]].y == kg.y && keys[order[e]].z == kg.z) ++e;
        vec3 s = splat3(0.0f);
        for (size_t k = g; k < e; ++k) s = s + weighted[order[k] / 3];
        for (size_t k = g; k < e; ++k) sum[order[k]] = s;
        g = e;
    }
    for (uint32_t i = 0; i < n; ++i) {
        float* o = out + (size_t)i * 36;
        const vec3 face = normalized_or(cross3(ld3(o + 3), ld3(o + 6)), up3());
        vec3 nv[3];
        for (int c = 0; c < 3; ++c) nv[c] = normalized_or(sum[(size_t)i * 3 + c], face);
        const vec2 ua = mk2(o[30], o[31]), ub = mk2(o[32], o[33]), uc = mk2(o[34], o[35]);
        emit(o, cornerPos[i * 3], cornerPos[i * 3 + 1], cornerPos[i * 3 + 2], nv, ua, ub, uc);
    }
}

void build_alias(const float* w, size_t n, float* prob, int32_t* alias) {
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += (double)w[i];
    for (size_t i = 0; i < n; ++i) { prob[i] = 1.0f; alias[i] = (int32_t)i; }
    if (!(sum > 0.0)) return;
    // Columns below the mean donate their deficit to one column above it; the worklists are used
    // as stacks in index order so the table is a pure function of the weights.
    std::vector<double> excess(n);
    std::vector<int32_t> under, over;
    under.reserve(n);
    over.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        excess[i] = (double)w[i] * (double)n / sum;
        if (excess[i] < 1.0) under.push_back((int32_t)i);
        else over.push_back((int32_t)i);
    }
    while (!under.empty() && !over.empty()) {
        const int32_t lo = under.back();
        under.pop_back();
        const int32_t hi = over.back();
        over.pop_back();
        prob[lo] = (float)excess[lo];
        alias[lo] = hi;
        excess[hi] = (excess[hi] + excess[lo]) - 1.0;
        if (excess[hi] < 1.0) under.push_back(hi);
        else over.push_back(hi);
    }
}

float half_to_float_host(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float v;
    if (e == 0) v = ldexpf((float)m, -24);                      // zero and subnormals
    else if (e == 31) v = m ? __builtin_nanf("") : __builtin_huge_valf();
    else v = ldexpf((float)(m | 0x400u), (int)e - 25);
    uint32_t bits;
    memcpy(&bits, &v, 4);
    bits |= sign;
    memcpy(&v, &bits, 4);
    return v;
}

float build_cdf1(const float* w, int n, float* cdf) {
    cdf[0] = 0.0f;
    for (int i = 0; i < n; ++i) cdf[i + 1] = cdf[i] + w[i];
    const float total = cdf[n];
    if (total > 0.0f)
        for (int i = 1; i <= n; ++i) cdf[i] /= total;
    return total;
}


namespace {
// GetStatsNode (bvh.cs:96-124): leaves give their own record; interior nodes merge A then B
// (NodeStats.Merge, bvh.cs:38-48), add themselves and their traversal cost.
BvhStatsHost stats_node(const BvhResult& b, int idx, int depth, float rootSa, float ct, float ci) {
    const RefNode& nd = b.nodes[idx];
    const float sa = box_area(nd.bounds);
    const float ratio = rootSa > 0.0f ? sa / rootSa : 0.0f;
    BvhStatsHost r;
    if (nd.count > 0) {
        r.nodeCount = 1;
        r.depthMax = depth;
        r.leafCount = 1;
        r.leafMin = r.leafMax = r.shapeCount = nd.count;
        r.depthWeighted = depth * nd.count;
        r.sah = ci * ratio * (float)nd.count;
        return r;
    }
    const BvhStatsHost a = stats_node(b, nd.child, depth + 1, rootSa, ct, ci);
    const BvhStatsHost c = stats_node(b, nd.child + 1, depth + 1, rootSa, ct, ci);
    r.nodeCount = a.nodeCount + c.nodeCount + 1;
    r.leafCount = a.leafCount + c.leafCount;
    r.leafMin = std::min(a.leafMin, c.leafMin);
    r.leafMax = std::max(a.leafMax, c.leafMax);
    r.shapeCount = a.shapeCount + c.shapeCount;
    r.depthMax = std::max(a.depthMax, c.depthMax);
    r.depthWeighted = a.depthWeighted + c.depthWeighted;
    r.sah = a.sah + c.sah;
    r.sah += ct * ratio;
    return r;
}
}  // namespace

BvhStatsHost bvh_stats(const BvhResult& b, float cost_traverse, float cost_intersect) {
    if (b.used == 0) return BvhStatsHost();
    return stats_node(b, 0, 0, box_area(b.nodes[0].bounds), cost_traverse, cost_intersect);
}

}  // namespace sr
