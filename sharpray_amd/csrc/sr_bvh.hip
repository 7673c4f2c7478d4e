// Device binned-SAH BVH builder producing the reference's tree bit for bit (bvh.cs:285-476): same
// node bounds, split candidates and costs, the same in-place two-cursor partition, and nodes
// numbered in the order the reference's depth-first Subdivide allocates its sibling pairs.
//
// Phase 1 (level by level, one 256-thread workgroup per node) takes nodes of more than kSubtreeMax
// items: node and centroid bounds by block reduction, the 3 x bins SAH histogram in LDS, the
// reference's sweeps on one lane, and the partition as a scatter. The two-cursor partition
// (bvh.cs:401-426) is order dependent but has a closed form over the node's item range [B, R]
// with p = B + (number of items left of the plane):
//   - a left item below p stays;
//   - the k-th right item below p (a "hole", k from 0) moves to q_k - 1, where q_0 = R + 1 and q_k
//     is the position of the k-th left item at or above p counted from R downwards;
//   - the m-th left item at or above p (m from 0, counted from R) fills hole m;
//   - a right item at or above p moves down one place, except the one at p, which goes to q_A - 1
//     (A = number of holes).
// Phase 2 (one wave per remaining node) builds each small subtree the way the reference does,
// depth first with the sequential partition, into a local node pool numbered like the reference.
// Phase 3 counts internal nodes bottom-up, assigns every internal node its depth-first rank r
// top-down (its children are nodes 1 + 2r and 2 + 2r) and writes the final node array.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "sr_host.h"
#include "sr_math.h"

namespace sr {
namespace {

constexpr int kSubtreeMax = 512;  // nodes of at most this many items are built by one wave
constexpr int kMaxBins = 64;
constexpr int kMaxDepth = 64;     // bvh.cs:436

struct Record {
    box3 bounds;
    int32_t begin, count, depth;
    int32_t firstChild;  // record index of child A (B follows), -1 for a leaf
    int32_t subtree;     // 1: built by the subtree pass
    int32_t poolBase, poolCount;
    int32_t ic;          // internal nodes in the subtree rooted here
    int32_t rank, fid;   // depth-first subdivision rank, final node index
};

struct Args {
    const box3* box;
    const vec3* mid;
    int32_t* items;
    int32_t* tmp;
    int32_t* scan;   // per position: left items before it within its node (phase 1)
    int32_t* hole;   // per node range: position of hole k at [begin + k]
    int32_t* qpos;   // per node range: position of the m-th left item from the top at [begin + m]
    Record* rec;
    RefNode* pool;
    int32_t* counters;  // [0] records, [1] next list, [2] subtree list, [3] pool, [4] degenerate, [5] max depth
    int32_t* nextList;
    int32_t* subList;
    int32_t bins;
    float ct, ci;
};

__device__ __forceinline__ uint32_t okey(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b >> 31) ? ~b : (b ^ 0x80000000u);
}
__device__ __forceinline__ float unkey(uint32_t k) {
    return __uint_as_float((k >> 31) ? (k ^ 0x80000000u) : ~k);
}

__global__ void k_prep(const float* lean, int n, box3* box, vec3* mid, int32_t* items) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* t = lean + (size_t)i * 9;
    const vec3 a = mk3(t[0], t[1], t[2]);
    const vec3 b = a + mk3(t[3], t[4], t[5]);
    const vec3 c = a + mk3(t[6], t[7], t[8]);
    box3 r;
    r.mn = mk3(net_min(a.x, net_min(b.x, c.x)), net_min(a.y, net_min(b.y, c.y)), net_min(a.z, net_min(b.z, c.z)));
    r.mx = mk3(net_max(a.x, net_max(b.x, c.x)), net_max(a.y, net_max(b.y, c.y)), net_max(a.z, net_max(b.z, c.z)));
    box[i] = r;
    mid[i] = box_mid(r);
    items[i] = i;
}

__device__ __forceinline__ float axis_of(vec3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// Lane 0's sweep over one axis's histogram (bvh.cs:360-392), updating the best split.
__device__ void sweep_axis(const uint32_t* bmn, const uint32_t* bmx, const int32_t* bcnt, int bins, int axis, float lo,
                           float hi, float parentArea, float ct, float ci, float* leftArea, int* leftCnt, float& best,
                           int& bestAxis, float& bestPos) {
    box3 acc = box_empty();
    int accCnt = 0;
    auto bin_box = [&](int b) {
        box3 r;
        r.mn = mk3(unkey(bmn[b * 3]), unkey(bmn[b * 3 + 1]), unkey(bmn[b * 3 + 2]));
        r.mx = mk3(unkey(bmx[b * 3]), unkey(bmx[b * 3 + 1]), unkey(bmx[b * 3 + 2]));
        return r;
    };
    for (int b = 0; b < bins - 1; ++b) {
        box_grow(acc, bin_box(b));
        accCnt += bcnt[b];
        leftArea[b] = box_is_empty(acc) ? 0.0f : box_area(acc);
        leftCnt[b] = accCnt;
    }
    acc = box_empty();
    accCnt = 0;
    for (int b = bins - 1; b > 0; --b) {
        box_grow(acc, bin_box(b));
        accCnt += bcnt[b];
        const float rightArea = box_is_empty(acc) ? 0.0f : box_area(acc);
        const float cost = ct * parentArea + ci * leftArea[b - 1] * (float)leftCnt[b - 1] + ci * rightArea * (float)accCnt;
        if (cost < best) {
            best = cost;
            bestAxis = axis;
            bestPos = lo + (float)b * (hi - lo) / (float)bins;
        }
    }
}

// ---- Phase 1: one workgroup per large node --------------------------------------------------------
__global__ void __launch_bounds__(256) k_large(Args A, const int32_t* list) {
    __shared__ uint32_t sMn[3 * kMaxBins * 3], sMx[3 * kMaxBins * 3];
    __shared__ int32_t sCnt[3 * kMaxBins];
    __shared__ uint32_t sRed[12];
    __shared__ float sLeftArea[kMaxBins];
    __shared__ int sLeftCnt[kMaxBins];
    __shared__ int sAxis, sA, sWave[4];
    __shared__ float sPos;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    Record& R = A.rec[list[blockIdx.x]];
    const int begin = R.begin, count = R.count, depth = R.depth;
    const int32_t* items = A.items + begin;

    // Node bounds (Insert, bvh.cs:303-313) and centroid bounds (bvh.cs:455-464).
    if (tid < 12) sRed[tid] = (tid % 6) < 3 ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    box3 nb = box_empty(), cb = box_empty();
    for (int i = tid; i < count; i += 256) {
        const int p = items[i];
        box_grow(nb, A.box[p]);
        box_grow(cb, A.mid[p]);
    }
    {
        const float v[12] = {nb.mn.x, nb.mn.y, nb.mn.z, nb.mx.x, nb.mx.y, nb.mx.z,
                             cb.mn.x, cb.mn.y, cb.mn.z, cb.mx.x, cb.mx.y, cb.mx.z};
        for (int k = 0; k < 12; ++k) {
            if ((k % 6) < 3) atomicMin(&sRed[k], okey(v[k]));
            else atomicMax(&sRed[k], okey(v[k]));
        }
    }
    __syncthreads();
    nb.mn = mk3(unkey(sRed[0]), unkey(sRed[1]), unkey(sRed[2]));
    nb.mx = mk3(unkey(sRed[3]), unkey(sRed[4]), unkey(sRed[5]));
    cb.mn = mk3(unkey(sRed[6]), unkey(sRed[7]), unkey(sRed[8]));
    cb.mx = mk3(unkey(sRed[9]), unkey(sRed[10]), unkey(sRed[11]));
    if (tid == 0) {
        R.bounds = nb;
        R.firstChild = -1;
        R.subtree = 0;
        R.ic = 0;
        atomicMax(&A.counters[5], depth);
    }
    if (depth >= kMaxDepth) return;

    // SAH histogram of all three axes (bvh.cs:339-358).
    const int bins = A.bins;
    float lo[3], hi[3], scale[3];
    bool valid[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = axis_of(cb.mn, a);
        hi[a] = axis_of(cb.mx, a);
        valid[a] = !(hi[a] - lo[a] < 1e-6f);
        scale[a] = (float)bins / (hi[a] - lo[a]);
    }
    for (int k = tid; k < 3 * bins * 3; k += 256) {
        sMn[k] = okey(kFloatMax);
        sMx[k] = okey(-kFloatMax);
    }
    for (int k = tid; k < 3 * bins; k += 256) sCnt[k] = 0;
    __syncthreads();
    for (int i = tid; i < count; i += 256) {
        const int p = items[i];
        const box3 b = A.box[p];
        const vec3 m = A.mid[p];
        for (int a = 0; a < 3; ++a) {
            if (!valid[a]) continue;
            int bin = (int)((axis_of(m, a) - lo[a]) * scale[a]);
            if (bin > bins - 1) bin = bins - 1;
            uint32_t* mn = sMn + (a * bins + bin) * 3;
            uint32_t* mx = sMx + (a * bins + bin) * 3;
            atomicMin(&mn[0], okey(b.mn.x));
            atomicMin(&mn[1], okey(b.mn.y));
            atomicMin(&mn[2], okey(b.mn.z));
            atomicMax(&mx[0], okey(b.mx.x));
            atomicMax(&mx[1], okey(b.mx.y));
            atomicMax(&mx[2], okey(b.mx.z));
            atomicAdd(&sCnt[a * bins + bin], 1);
        }
    }
    __syncthreads();
    if (tid == 0) {
        const float parentArea = box_area(nb);
        float best = A.ci * parentArea * (float)count;
        int bestAxis = -1;
        float bestPos = 0.0f;
        for (int a = 0; a < 3; ++a)
            if (valid[a])
                sweep_axis(sMn + a * bins * 3, sMx + a * bins * 3, sCnt + a * bins, bins, a, lo[a], hi[a], parentArea,
                           A.ct, A.ci, sLeftArea, sLeftCnt, best, bestAxis, bestPos);
        sAxis = bestAxis;
        sPos = bestPos;
    }
    __syncthreads();
    const int axis = sAxis;
    if (axis < 0) return;
    const float pos = sPos;

    // Partition as a scatter (closed form in the file comment). Pass 1: left items before every
    // position, kept in A.scan, and the node's total.
    int running = 0;
    for (int base = 0; base < count; base += 256) {
        const int i = base + tid;
        const bool left = i < count && axis_of(A.mid[items[i]], axis) < pos;
        const unsigned long long ball = __ballot(left);
        const int inWave = __popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) sWave[wave] = __popcll(ball);
        __syncthreads();
        int before = running;
        for (int w = 0; w < wave; ++w) before += sWave[w];
        if (i < count) A.scan[begin + i] = before + inWave;
        running += sWave[0] + sWave[1] + sWave[2] + sWave[3];
        __syncthreads();
    }
    const int nL = running;
    const int p = begin + nL, last = begin + count - 1;
    if (nL == 0 || nL == count) {  // a side would be empty (bvh.cs:448 asserts against it)
        if (tid == 0) A.counters[4] = 1;
        return;
    }
    // Pass 2: hole and q positions by rank.
    for (int i = tid; i < count; i += 256) {
        const int pos_i = begin + i;
        const bool left = axis_of(A.mid[items[i]], axis) < pos;
        const int lcb = A.scan[pos_i];
        if (pos_i < p && !left) A.hole[begin + (i - lcb)] = pos_i;
        if (pos_i >= p && left) A.qpos[begin + (nL - lcb - 1)] = pos_i;
        if (pos_i == p) sA = nL - lcb;
    }
    __syncthreads();
    const int holes = sA;
    // Pass 3: destinations.
    for (int i = tid; i < count; i += 256) {
        const int pos_i = begin + i;
        const int it = items[i];
        const bool left = axis_of(A.mid[it], axis) < pos;
        const int lcb = A.scan[pos_i];
        int dst;
        if (pos_i < p) {
            if (left) {
                dst = pos_i;
            } else {
                const int k = i - lcb;
                dst = k == 0 ? last : A.qpos[begin + k - 1] - 1;
            }
        } else if (left) {
            dst = A.hole[begin + (nL - lcb - 1)];
        } else if (pos_i == p) {
            dst = (holes == 0 ? last + 1 : A.qpos[begin + holes - 1]) - 1;
        } else {
            dst = pos_i - 1;
        }
        A.tmp[dst] = it;
    }
    __syncthreads();
    for (int i = tid; i < count; i += 256) A.items[begin + i] = A.tmp[begin + i];
    if (tid == 0) {
        const int c = atomicAdd(&A.counters[0], 2);
        R.firstChild = c;
        const int cnt[2] = {nL, count - nL};
        const int beg[2] = {begin, p};
        for (int k = 0; k < 2; ++k) {
            Record& C = A.rec[c + k];
            C.begin = beg[k];
            C.count = cnt[k];
            C.depth = depth + 1;
            C.firstChild = -1;
            C.subtree = 0;
            C.ic = 0;
            if (cnt[k] > kSubtreeMax) A.nextList[atomicAdd(&A.counters[1], 1)] = c + k;
            else A.subList[atomicAdd(&A.counters[2], 1)] = c + k;
        }
    }
}

// ---- Phase 2: one wave per small subtree ----------------------------------------------------------
__device__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = net_min(v, __shfl_xor(v, o));
    return v;
}
__device__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = net_max(v, __shfl_xor(v, o));
    return v;
}

// Insert (bvh.cs:303-313): a leaf over [begin, begin + count) with its bounds, by one wave.
__device__ void wave_leaf(const Args& A, RefNode* nd, int begin, int count, int lane) {
    box3 b = box_empty();
    for (int i = lane; i < count; i += 64) box_grow(b, A.box[A.items[begin + i]]);
    b.mn = mk3(wave_min(b.mn.x), wave_min(b.mn.y), wave_min(b.mn.z));
    b.mx = mk3(wave_max(b.mx.x), wave_max(b.mx.y), wave_max(b.mx.z));
    if (lane == 0) {
        nd->bounds = b;
        nd->child = begin;
        nd->count = count;
    }
}

__global__ void __launch_bounds__(256) k_subtree(Args A, int nSub) {
    __shared__ uint32_t sMn[4][kMaxBins * 3], sMx[4][kMaxBins * 3];
    __shared__ int32_t sCnt[4][kMaxBins];
    __shared__ float sLeftArea[4][kMaxBins];
    __shared__ int sLeftCnt[4][kMaxBins];
    __shared__ int sStack[4][2 * (kMaxDepth + 2)];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int s = blockIdx.x * 4 + wave;
    if (s >= nSub) return;  // whole wave
    Record& R = A.rec[A.subList[s]];
    const int rootCount = R.count;
    int base = 0;
    if (lane == 0) base = atomicAdd(&A.counters[3], 2 * rootCount - 1);
    base = __shfl(base, 0);
    RefNode* pool = A.pool + base;
    wave_leaf(A, &pool[0], R.begin, rootCount, lane);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    int used = 1, sp = 0, maxDepth = R.depth;
    if (lane == 0) {
        sStack[wave][0] = 0;
        sStack[wave][1] = R.depth;
    }
    sp = 1;
    const int bins = A.bins;
    while (sp > 0) {
        --sp;
        const int node = sStack[wave][2 * sp], depth = sStack[wave][2 * sp + 1];
        maxDepth = max(maxDepth, depth);
        if (depth >= kMaxDepth) continue;
        const RefNode nd = pool[node];
        const int begin = nd.child, count = nd.count;
        // Centroid bounds.
        box3 cb = box_empty();
        for (int i = lane; i < count; i += 64) box_grow(cb, A.mid[A.items[begin + i]]);
        cb.mn = mk3(wave_min(cb.mn.x), wave_min(cb.mn.y), wave_min(cb.mn.z));
        cb.mx = mk3(wave_max(cb.mx.x), wave_max(cb.mx.y), wave_max(cb.mx.z));
        const float parentArea = box_area(nd.bounds);
        float best = A.ci * parentArea * (float)count;
        int bestAxis = -1;
        float bestPos = 0.0f;
        for (int a = 0; a < 3; ++a) {
            const float lo = axis_of(cb.mn, a), hi = axis_of(cb.mx, a);
            if (hi - lo < 1e-6f) continue;
            for (int k = lane; k < bins * 3; k += 64) {
                sMn[wave][k] = okey(kFloatMax);
                sMx[wave][k] = okey(-kFloatMax);
            }
            for (int k = lane; k < bins; k += 64) sCnt[wave][k] = 0;
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            const float scale = (float)bins / (hi - lo);
            for (int i = lane; i < count; i += 64) {
                const int p = A.items[begin + i];
                int bin = (int)((axis_of(A.mid[p], a) - lo) * scale);
                if (bin > bins - 1) bin = bins - 1;
                const box3 b = A.box[p];
                atomicMin(&sMn[wave][bin * 3], okey(b.mn.x));
                atomicMin(&sMn[wave][bin * 3 + 1], okey(b.mn.y));
                atomicMin(&sMn[wave][bin * 3 + 2], okey(b.mn.z));
                atomicMax(&sMx[wave][bin * 3], okey(b.mx.x));
                atomicMax(&sMx[wave][bin * 3 + 1], okey(b.mx.y));
                atomicMax(&sMx[wave][bin * 3 + 2], okey(b.mx.z));
                atomicAdd(&sCnt[wave][bin], 1);
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == 0)
                sweep_axis(sMn[wave], sMx[wave], sCnt[wave], bins, a, lo, hi, parentArea, A.ct, A.ci, sLeftArea[wave],
                           sLeftCnt[wave], best, bestAxis, bestPos);
            __builtin_amdgcn_wave_barrier();
        }
        bestAxis = __shfl(bestAxis, 0);
        bestPos = __shfl(bestPos, 0);
        if (bestAxis < 0) continue;
        // The reference's two-cursor partition (bvh.cs:401-426), on lane 0.
        int split = 0;
        if (lane == 0) {
            int32_t* it = A.items;
            int l = begin, r = begin + count - 1;
            for (;;) {
                if (axis_of(A.mid[it[l]], bestAxis) < bestPos) {
                    if (++l > r) break;
                } else {
                    if (l == r) break;
                    const int t = it[l];
                    it[l] = it[r];
                    it[r] = t;
                    --r;
                }
            }
            split = l;
        }
        split = __shfl(split, 0);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int na = split - begin, nb = count - na;
        if (na <= 0 || nb <= 0) {
            if (lane == 0) A.counters[4] = 1;
            continue;
        }
        const int ca = used;
        wave_leaf(A, &pool[ca], begin, na, lane);
        wave_leaf(A, &pool[ca + 1], split, nb, lane);
        used += 2;
        if (lane == 0) {
            pool[node].child = ca;
            pool[node].count = 0;
            sStack[wave][2 * sp] = ca + 1;
            sStack[wave][2 * sp + 1] = depth + 1;
            sStack[wave][2 * sp + 2] = ca;
            sStack[wave][2 * sp + 3] = depth + 1;
        }
        sp += 2;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (lane == 0) {
        R.bounds = pool[0].bounds;
        R.firstChild = -1;
        R.subtree = 1;
        R.poolBase = base;
        R.poolCount = used;
        R.ic = (used - 1) / 2;
        atomicMax(&A.counters[5], maxDepth);
    }
}

// ---- Phase 3 -----------------------------------------------------------------------------------
__global__ void k_count(Record* rec, int lo, int hi) {
    const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    Record& R = rec[i];
    if (R.firstChild >= 0) R.ic = 1 + rec[R.firstChild].ic + rec[R.firstChild + 1].ic;
}

__global__ void k_rank(Record* rec, int lo, int hi) {
    const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    const Record& R = rec[i];
    if (R.firstChild < 0) return;
    Record& a = rec[R.firstChild];
    Record& b = rec[R.firstChild + 1];
    a.fid = 1 + 2 * R.rank;
    b.fid = 2 + 2 * R.rank;
    a.rank = R.rank + 1;
    b.rank = R.rank + 1 + a.ic;
}

__global__ void k_emit_records(const Record* rec, int n, RefNode* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Record& R = rec[i];
    if (R.subtree) return;
    RefNode nd;
    nd.bounds = R.bounds;
    if (R.firstChild >= 0) {
        nd.child = rec[R.firstChild].fid;
        nd.count = 0;
    } else {
        nd.child = R.begin;
        nd.count = R.count;
    }
    out[R.fid] = nd;
}

// A subtree's local node k: 0 is the subtree root (final index fid); pair (1 + 2q, 2 + 2q) was the
// q-th local subdivision, i.e. the (rank + q)-th overall.
__global__ void k_emit_subtree(const Record* rec, const int32_t* subList, int nSub, const RefNode* pool,
                               RefNode* out) {
    const int s = blockIdx.x;
    if (s >= nSub) return;
    const Record& R = rec[subList[s]];
    for (int k = threadIdx.x; k < R.poolCount; k += blockDim.x) {
        RefNode nd = pool[R.poolBase + k];
        if (nd.count == 0) nd.child = 1 + 2 * (R.rank + (nd.child - 1) / 2);
        out[k == 0 ? R.fid : 1 + 2 * (R.rank + (k - 1) / 2) + ((k - 1) & 1)] = nd;
    }
}

template <class T>
bool dalloc(T*& p, size_t n) {
    return hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess;
}

}  // namespace

// Builds the BVH of `n` TriangleLean records (9 floats, device memory) on the current device.
// Returns false on a HIP error (message in err); out.degenerate is set like the host builder.
bool device_build_bvh(const float* dLean, int n, int bins, float ct, float ci, BvhResult& out, hipStream_t s,
                      std::string& err) {
    out = BvhResult();
    out.items.resize(n);
    if (n == 0) {
        out.nodes.assign(1, RefNode{});
        return true;
    }
    if (bins < 2 || bins > kMaxBins) {
        err = "device BVH build supports 2..64 bins";
        return false;
    }
    Args A{};
    A.bins = bins;
    A.ct = ct;
    A.ci = ci;
    box3* dBox = nullptr;
    vec3* dMid = nullptr;
    int32_t *dItems = nullptr, *dTmp = nullptr, *dScan = nullptr, *dHole = nullptr, *dQ = nullptr, *dCtr = nullptr,
            *dNext = nullptr, *dCur = nullptr, *dSub = nullptr;
    Record* dRec = nullptr;
    RefNode *dPool = nullptr, *dOut = nullptr;
    const size_t cap = 2 * (size_t)n + 1;
    bool ok = dalloc(dBox, n) && dalloc(dMid, n) && dalloc(dItems, n) && dalloc(dTmp, n) && dalloc(dScan, n) &&
              dalloc(dHole, n) && dalloc(dQ, n) && dalloc(dCtr, 8) && dalloc(dNext, cap) && dalloc(dCur, cap) &&
              dalloc(dSub, cap) && dalloc(dRec, cap) && dalloc(dPool, cap) && dalloc(dOut, cap);
    hipError_t e = ok ? hipSuccess : hipErrorOutOfMemory;
    std::vector<int> levelLo, levelHi;
    int nSub = 0, nRec = 1;
    if (e == hipSuccess) {
        A.box = dBox;
        A.mid = dMid;
        A.items = dItems;
        A.tmp = dTmp;
        A.scan = dScan;
        A.hole = dHole;
        A.qpos = dQ;
        A.rec = dRec;
        A.pool = dPool;
        A.counters = dCtr;
        A.subList = dSub;
        hipLaunchKernelGGL(k_prep, dim3((n + 255) / 256), dim3(256), 0, s, dLean, n, dBox, dMid, dItems);
        Record root{};
        root.begin = 0;
        root.count = n;
        root.depth = 0;
        root.firstChild = -1;
        int32_t ctr[8] = {1, 0, n <= kSubtreeMax ? 1 : 0, 0, 0, 0, 0, 0};
        const int32_t zero = 0;
        e = hipStreamSynchronize(s);
        if (e == hipSuccess) e = hipMemcpy(dRec, &root, sizeof(root), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dCtr, ctr, sizeof(ctr), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(n <= kSubtreeMax ? dSub : dCur, &zero, 4, hipMemcpyHostToDevice);
        levelLo.push_back(0);
        levelHi.push_back(1);
        int nCur = n > kSubtreeMax ? 1 : 0;
        while (e == hipSuccess && nCur > 0) {
            e = hipMemsetAsync(dCtr + 1, 0, 4, s);
            if (e != hipSuccess) break;
            A.nextList = dNext;
            hipLaunchKernelGGL(k_large, dim3(nCur), dim3(256), 0, s, A, dCur);
            e = hipGetLastError();
            if (e != hipSuccess) break;
            e = hipMemcpyAsync(ctr, dCtr, sizeof(ctr), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) break;
            levelLo.push_back(nRec);
            levelHi.push_back(ctr[0]);
            nRec = ctr[0];
            nCur = ctr[1];
            std::swap(dCur, dNext);
        }
        nSub = ctr[2];
        if (e == hipSuccess && nSub > 0) {
            hipLaunchKernelGGL(k_subtree, dim3((nSub + 3) / 4), dim3(256), 0, s, A, nSub);
            e = hipGetLastError();
        }
        for (int L = (int)levelLo.size() - 1; e == hipSuccess && L >= 0; --L) {
            const int cnt = levelHi[L] - levelLo[L];
            if (cnt <= 0) continue;
            hipLaunchKernelGGL(k_count, dim3((cnt + 255) / 256), dim3(256), 0, s, dRec, levelLo[L], levelHi[L]);
            e = hipGetLastError();
        }
        Record top{};
        if (e == hipSuccess) e = hipMemcpyAsync(&top, dRec, sizeof(top), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) {
            top.rank = 0;
            top.fid = 0;
            e = hipMemcpy(dRec, &top, sizeof(top), hipMemcpyHostToDevice);
        }
        for (size_t L = 0; e == hipSuccess && L < levelLo.size(); ++L) {
            const int cnt = levelHi[L] - levelLo[L];
            if (cnt <= 0) continue;
            hipLaunchKernelGGL(k_rank, dim3((cnt + 255) / 256), dim3(256), 0, s, dRec, levelLo[L], levelHi[L]);
            e = hipGetLastError();
        }
        const int total = 1 + 2 * top.ic;
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_emit_records, dim3((nRec + 255) / 256), dim3(256), 0, s, dRec, nRec, dOut);
            e = hipGetLastError();
        }
        if (e == hipSuccess && nSub > 0) {
            hipLaunchKernelGGL(k_emit_subtree, dim3(nSub), dim3(256), 0, s, dRec, dSub, nSub, dPool, dOut);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            out.nodes.resize(std::max(total, 1));
            out.used = total;
            e = hipMemcpyAsync(out.nodes.data(), dOut, (size_t)total * sizeof(RefNode), hipMemcpyDeviceToHost, s);
        }
        if (e == hipSuccess) e = hipMemcpyAsync(out.items.data(), dItems, (size_t)n * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(ctr, dCtr, sizeof(ctr), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) {
            out.degenerate = ctr[4] != 0;
            out.max_depth = ctr[5];
        }
    }
    for (void* p : {(void*)dBox, (void*)dMid, (void*)dItems, (void*)dTmp, (void*)dScan, (void*)dHole, (void*)dQ,
                    (void*)dCtr, (void*)dNext, (void*)dCur, (void*)dSub, (void*)dRec, (void*)dPool, (void*)dOut})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) {
        err = std::string("device BVH build: ") + hipGetErrorString(e);
        return false;
    }
    return true;
}

}  // namespace sr
