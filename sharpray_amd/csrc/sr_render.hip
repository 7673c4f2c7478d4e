// gfx950 wavefront path tracer behind the sharpray.h C ABI.
//
// Every pixel of a band owns one camera path at a time and the frame is rendered sample-major:
// for each sample, (bounces + 1) segments of three kernels run over compacted queues,
//   extend  - closest hit for every queued ray (two-level BVH traversal);
//   shade   - one path segment (scene.cs:641-737): fog scatter, surface hit or sky miss; the
//             direct-light sample becomes a deferred shadow ray; live paths are appended to the
//             next segment's queue;
//   shadow  - shadow transmittance and the deferred direct-light contribution;
// then `next` accumulates the sample, applies the convergence test and emits the pixel's next
// camera ray (renderer.cs:173-229). All rays of a launch sit at the same path depth and camera
// rays are queued in 8x8-tile order, which keeps waves coherent (mixing depths by regenerating
// paths as soon as they end measured slower on this workload). The per-pixel RNG state persists
// in HBM between samples, which keeps the reference's sequential xorwow stream per pixel
// (renderer.cs:156) and its sample-order sums. extend, shade and shadow run over a fixed grid with
// a wave-aligned grid-stride loop (one shared atomic work cursor measured slower).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sharpray.h"
#include "sr_device.h"
#include "sr_host.h"
#include "sr_trace.h"
#include "sr_adjoint.h"
#include "sr_cnn.h"

using namespace sr;

namespace {

constexpr int kBlock = kTravBlock;
#ifndef SR_TRACE_BLOCK
#define SR_TRACE_BLOCK 64
#endif
constexpr int kTraceBlock = SR_TRACE_BLOCK;  // threads per block of k_trace (one wave)
#ifndef SR_SHADE_BLOCK
#define SR_SHADE_BLOCK 256
#endif
constexpr int kShadeBlock = SR_SHADE_BLOCK;  // threads per block of k_shade (the hit / miss partition's span)
constexpr int kNextBlock = 256;
// Per-lane LDS columns of the mesh traversal stack (trav_lds_bytes(threads) per block).
extern __shared__ uint2 sr_lds[];
// This lane's column of its wave's LDS stack rows.
SR_DEV uint2* lds_column() { return sr_lds + (threadIdx.x >> 6) * (kLdsDepth * 64) + (threadIdx.x & 63); }

// ---- Frame state --------------------------------------------------------------------------------
// A path's state while a kernel works on it (registers, 64 B). In memory it is split into PathHot
// (48 B: what only the segments of the path read and write) and PathRad (16 B, a dense array: the
// radiance and flags, which the shadow rays update and the next-sample pass reads in slot order).
struct alignas(16) PathRec {
    vec3 energy;
    float scatterPdf;
    vec3 rad;
    uint32_t flags;  // bit0: scatterPdf valid, bit1: inside a medium, bit2: kPrimaryHit
    uint32_t rng[5];
    uint32_t seg;    // segment index of the queued ray (0 = camera ray)
    uint32_t pad[2];
};
struct alignas(16) PathHot {
    vec3 energy;
    float scatterPdf;
    uint32_t rng[5];
    uint32_t seg;
    uint32_t pad[2];
};
struct alignas(16) PathRad {
    vec3 rad;
    uint32_t flags;
};
// The rest of a path's state, touched only where it is used: the fog radiance (scenes with fog or
// a medium) and the interior medium's albedo (paths inside a transparent object).
struct alignas(16) PathCold {  // 32 B
    vec3 radFog;
    uint32_t pad0;
    vec3 medium;
    uint32_t pad1;
};
struct alignas(16) ShadowRec {  // 80 B
    vec3 o, d;
    vec3 skyRad, f;
    float w, pdf;
    vec3 energy;
    uint32_t path;  // bit31: add to radianceFog
};
struct alignas(16) PixSums {  // 48 B, changed by every sample
    vec3 radSum, radFogSum;
    float lumSum, lumSumSqr, lumSurf, lumSurfSqr;
    uint32_t sampleIndex;  // first sample of the batch in flight
    uint32_t active;
};
struct alignas(16) PixPrimary {  // 64 B, changed by samples whose camera ray hits a surface
    vec3 normalSum, normalFallback;
    uint32_t normalCount;
    vec2 uv;
    uint32_t hasUv;
    float depthSum;
    uint32_t depthCount;
    vec3 albedoSum;  // primary-hit Surface.Albedo, averaged over depthCount
    uint32_t pad;
};
struct alignas(16) PixAcc {  // 112 B per pixel: the sums of renderer.cs:160-203
    PixSums s;
    PixPrimary h;
};

struct ViewD {
    vec3 pos;
    quat rot;
    float nearPlane, tanFov, aspect;
};

// Render counters beyond the reference's SR_COUNTER_COUNT: the fused tail launch's share of
// BvhIntersectNode, BvhIntersectShape, SceneTrace and SceneOcclude.
// Then the reference-only share of BvhIntersectNode, BvhIntersectShape, MeshIntersect and
// SceneOcclude in the per-segment launches: traversal the reference does but a render without
// counters skips (side-culled shadow rays, see kCountOnly, and the nodes the reference's
// closest-hit walk of a shadow ray visits beyond those of the any-hit walk).
// Then three max-reduced words (atomicMax over unsigned keys, 0 = unset): SampleCountMin as the
// complement of the count, SampleCountMax, and SampleVarianceMax as a biased (long)(v * 1e4)
// (renderer.cs:224-226, counters.cs:195-219).
enum : int { kTailNode = SR_COUNTER_COUNT, kTailShape, kTailTrace, kTailOcclude, kRefNode, kRefShape, kRefMesh,
             kRefOcclude, kCountMinKey, kCountMaxKey, kVarMaxKey, kCtrWords };
// ShadowRec::path flags. kCountOnly marks a shadow ray whose light lies behind the shading or the
// geometric normal: its contribution is black, so the tracer skips it, but the reference traces
// it before its side tests (scene.cs:774-783) and counts that traversal. While counters are on,
// such rays are traced for their counts only.
constexpr uint32_t kToFog = 0x80000000u, kCountOnly = 0x40000000u, kPathMask = 0x3fffffffu;
constexpr int kCtrSpread = 16;  // copies of the counter words (a power of two)
// PathRec::flags bit: the sample's camera ray hit a surface, whose primary-hit record (Frame::prim)
// the next-sample pass folds into the pixel's sums.
constexpr uint32_t kPrimaryHit = 4u;
constexpr int kProgRing = 4;    // samples an adaptive render's driver runs ahead of its progress reads
constexpr uint32_t kMaxInFlight = 32;  // samples of a pixel in flight at most
constexpr uint32_t kTailAuto = 0xffffffffu;
// Device bytes per path slot (rays, hit, path record, two shadow records, two queue entries,
// primary-hit record) and the most the path slots of one render may take: samples in flight are
// capped to fit (a 1080p frame holds 32 samples per pixel in 24 GB).
constexpr size_t kPathSlotBytes = 3 * sizeof(float4) + sizeof(int32_t) + sizeof(PathHot) + sizeof(PathRad) +
                                  sizeof(PathCold) + 2 * sizeof(ShadowRec) + 2 * sizeof(uint32_t) + 2 * sizeof(float4) + sizeof(float);
static_assert(sizeof(PathRec) == 64 && sizeof(PathHot) == 48 && sizeof(PathRad) == 16 && sizeof(PathCold) == 32 &&
              sizeof(ShadowRec) == 80, "record sizes");
constexpr size_t kMaxPathBytes = (size_t)32 << 30;

// Queue counters (uint32 each), per slice. kOverflow collects the queues an append found full (bit
// 0: ray queues, bit 1: shadow queues); the append is then dropped and the host fails the render.
enum : int {
    kRayCount0 = 0, kRayCount1 = 1, kShadowCount0 = 2, kShadowCount1 = 3, kOverflow = 4, kCounterWords = 5
};

struct Frame {
    float4* rayO;       // origin per slot (xyz)
    float4* rayD;       // direction per slot (xyz)
    float4* hitA;       // t, u, v, object (int bits)
    int32_t* hitSlot;   // mesh item slot
    PathHot* paths;
    PathRad* prad;
    PathCold* cold;
    ShadowRec* shadows[2];  // shadow rays of even / odd segments
    PixAcc* acc;
    // Primary-hit record per path slot in three arrays (normal + depth, albedo + u, v), written by
    // the segment-0 shade and folded into the pixel's sums by the next-sample pass in sample order.
    float4* prim;
    float4* primB;
    float* primV;
    uint32_t* queue[2];
    uint32_t* counts;   // kCounterWords words
    unsigned long long* ctr;  // kCtrWords counters or null
    // Renderer.Tick's unit (renderer.cs:98-122): finished pixels per bs x bs block of the whole
    // image (bx blocks per row), and done[0], the count of blocks whose pixels are all final.
    uint32_t* blockPixels;
    uint32_t* blocksDone;
    uint32_t bs, bx;
    float* outRad;
    float* outRadFog;
    float* outNormal;
    float* outUv;
    float* outDepth;
    float* outAlbedo;
    uint32_t* outSamples;
    float* outVar;
    uint32_t W, H, y0, rows, tilesX, nSlots;
    // Samples in flight per pixel (a power of two): path slot j = q * batch + b runs sample (first
    // of the batch + b) of pixel slot q; nPaths = batch * nSlots is the capacity of every ray and
    // shadow queue.
    uint32_t batch, nPaths;
    // camSlots: with a batch in flight (every pixel of the band runs the same samples), k_start and
    // k_next write the camera-ray queue as every path slot in slot order, kNoPath where the slot
    // starts no sample, without a queue atomic; the segment-0 trace and shade skip kNoPath.
    uint32_t camSlots;
    // First sample of the batch (set for the segment-0 shade of a camSlots render): with the counter
    // generator a camera ray's random state is then (seed, pixel, sample0 + b, 2 draws) and is not
    // stored by k_start / k_next nor read back.
    uint32_t sample0;
    uint32_t fog;  // the scene has fog or a medium: paths carry fog radiance
    uint32_t minS, maxS, bounces, seed;
    float threshold, clamp;
    ViewD view;
};

// Slot q of a band maps to an 8x8 pixel tile, so a wave's 64 camera rays start coherent.
SR_DEV bool slot_pixel(const Frame& F, uint32_t q, uint32_t& x, uint32_t& y) {
    const uint32_t tile = q >> 6, in = q & 63u;
    x = (tile % F.tilesX) * 8u + (in & 7u);
    y = F.y0 + (tile / F.tilesX) * 8u + (in >> 3);
    return x < F.W && y < F.y0 + F.rows;
}

SR_DEV uint32_t lane_id() { return __lane_id(); }

// A hit record's last word (DevScene::packHit): -1 for a miss, else (object << 24 | item slot)
// (an analytic object's slot, unused, packs as 0).
SR_DEV int32_t pack_hit(const HitRec& h) {
    return h.obj < 0 ? -1 : (int32_t)(((uint32_t)h.obj << 24) | ((uint32_t)max(h.slot, 0) & 0xffffffu));
}
SR_DEV void unpack_hit(const Frame& F, const DevScene& S, uint32_t q, int32_t w, HitRec& h) {
    if (S.packHit) {
        h.obj = w < 0 ? -1 : (int32_t)((uint32_t)w >> 24);
        h.slot = w < 0 ? -1 : (int32_t)((uint32_t)w & 0xffffffu);
    } else {
        h.obj = w;
        h.slot = F.hitSlot[q];
    }
}

// A camera-ray queue entry of a slot that starts no sample (Frame::camSlots).
constexpr uint32_t kNoPath = 0xffffffffu;

// Queue appends are bounded by the queue's capacity (nPaths entries per queue and per shadow
// buffer): an append that finds the queue full is dropped and flagged in kOverflow, which the host
// turns into an error after the band instead of a write past the buffer.
SR_DEV bool queue_room(const Frame& F, uint32_t slot, uint32_t bit) {
    if (slot < F.nPaths) return true;
    atomicOr(&F.counts[kOverflow], bit);
    return false;
}
// A queue length read back for a launch, clamped to the capacity (only an overflowed queue's count
// exceeds it, and that render fails).
SR_DEV uint32_t queue_len(const Frame& F, int word) { return min(F.counts[word], F.nPaths); }


// Appends `pred` lanes of the whole block (NW waves) to a queue with one atomic per block; slots
// follow thread order. Every thread of the block calls it (it synchronises the block); `sh` is
// NW + 1 words of LDS.
template <int NW>
SR_DEV uint32_t block_append(uint32_t* count, bool pred, uint32_t* sh) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t wave = threadIdx.x >> 6;
    if (lane_id() == 0) sh[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = sh[w];
            sh[w] = total;
            total += c;
        }
        sh[NW] = total ? atomicAdd(count, total) : 0u;
    }
    __syncthreads();
    const uint32_t slot = sh[NW] + sh[wave] + (uint32_t)__popcll(mask & ((1ull << lane_id()) - 1ull));
    __syncthreads();  // sh serves the next call
    return slot;
}

// Appends the lanes whose key is in [0, NK) to a queue with one atomic per block, grouped by key:
// all of the block's key-0 items first, then key 1, ...; within a key in thread order. Every
// thread of the block calls it; `sh` is NW * NK + 1 words of LDS.
template <int NW, int NK>
SR_DEV uint32_t block_append_keyed(uint32_t* count, int key, uint32_t* sh) {
    const uint32_t wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane_id()) - 1ull;
    uint32_t rank = 0;
    for (int k = 0; k < NK; ++k) {
        const unsigned long long m = __ballot(key == k);
        if (lane_id() == 0) sh[k * NW + wave] = (uint32_t)__popcll(m);
        if (key == k) rank = (uint32_t)__popcll(m & below);
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // exclusive prefix over (key, wave)
        uint32_t total = 0;
        for (int i = 0; i < NW * NK; ++i) {
            const uint32_t c = sh[i];
            sh[i] = total;
            total += c;
        }
        sh[NW * NK] = total ? atomicAdd(count, total) : 0u;
    }
    __syncthreads();
    const uint32_t slot = (key >= 0 && key < NK) ? sh[NW * NK] + sh[key * NW + wave] + rank : 0u;
    __syncthreads();  // sh serves the next call
    return slot;
}

// Octant of a direction (sign bits of x, y, z): the key that groups a block's appended rays.
SR_DEV int dir_octant(vec3 d) { return (d.x < 0.0f ? 1 : 0) | (d.y < 0.0f ? 2 : 0) | (d.z < 0.0f ? 4 : 0); }

template <class R> SR_DEV R rng_load(const PathRec& p);
template <> SR_DEV XorwowRng rng_load<XorwowRng>(const PathRec& p) {
    XorwowRng r;
    r.s0 = p.rng[0]; r.s1 = p.rng[1]; r.s2 = p.rng[2]; r.s3 = p.rng[3]; r.s4 = p.rng[4];
    return r;
}
template <> SR_DEV CounterRng rng_load<CounterRng>(const PathRec& p) {
    CounterRng r;
    r.key0 = p.rng[0]; r.key1 = p.rng[1]; r.sample = p.rng[2]; r.draw = p.rng[3];
    return r;
}
SR_DEV void rng_store(PathRec& p, const XorwowRng& r) {
    p.rng[0] = r.s0; p.rng[1] = r.s1; p.rng[2] = r.s2; p.rng[3] = r.s3; p.rng[4] = r.s4;
}
SR_DEV void rng_store(PathRec& p, const CounterRng& r) {
    p.rng[0] = r.key0; p.rng[1] = r.key1; p.rng[2] = r.sample; p.rng[3] = r.draw;
}
// The RNG that draws sample `sample` of pixel (x, y): the pixel's persistent xorwow stream (new
// Rng(x, y) for sample 0, renderer.cs:156; the state the previous sample left in path slot j
// otherwise), or a fresh Philox counter for that (pixel, sample).
template <class R> SR_DEV R rng_for_sample(const Frame& F, size_t j, uint32_t x, uint32_t y, uint32_t sample);
template <> SR_DEV XorwowRng rng_for_sample<XorwowRng>(const Frame& F, size_t j, uint32_t x, uint32_t y, uint32_t sample) {
    if (sample == 0) return xorwow_seed(x, y);
    const uint32_t* w = F.paths[j].rng;
    XorwowRng r;
    r.s0 = w[0]; r.s1 = w[1]; r.s2 = w[2]; r.s3 = w[3]; r.s4 = w[4];
    return r;
}
template <> SR_DEV CounterRng rng_for_sample<CounterRng>(const Frame& F, size_t, uint32_t x, uint32_t y, uint32_t sample) {
    CounterRng r;
    r.key0 = F.seed;
    r.key1 = y * F.W + x;
    r.sample = sample;
    r.draw = 0;
    return r;
}
SR_DEV void rng_store_words(PathHot* p, const XorwowRng& r) {
    p->rng[0] = r.s0; p->rng[1] = r.s1; p->rng[2] = r.s2; p->rng[3] = r.s3; p->rng[4] = r.s4;
}
SR_DEV void rng_store_words(PathHot* p, const CounterRng& r) {
    *(uint4*)p->rng = make_uint4(r.key0, r.key1, r.sample, r.draw);
}

// The record of a path that ends in this segment: only what the next-sample pass reads (radiance
// and flags, 16 of its 64 bytes) and, for the xorwow stream, the random state the pixel's next
// sample continues from; a counter generator's next sample starts from its own counter.
SR_DEV void store_finished(const Frame& F, uint32_t q, const PathRec& p, const CounterRng&) {
    F.prad[q] = PathRad{p.rad, p.flags};
}
SR_DEV void store_finished(const Frame& F, uint32_t q, const PathRec& p, const XorwowRng& r) {
    F.prad[q] = PathRad{p.rad, p.flags};
    rng_store_words(F.paths + q, r);
}

// A path's state from its two records, and back.
SR_DEV PathRec load_path(const Frame& F, uint32_t j) {
    const PathHot h = F.paths[j];
    const PathRad r = F.prad[j];
    PathRec p;
    p.energy = h.energy;
    p.scatterPdf = h.scatterPdf;
    p.rad = r.rad;
    p.flags = r.flags;
    for (int k = 0; k < 5; ++k) p.rng[k] = h.rng[k];
    p.seg = h.seg;
    p.pad[0] = h.pad[0];
    p.pad[1] = h.pad[1];
    return p;
}
SR_DEV void store_path(const Frame& F, uint32_t j, const PathRec& p) {
    PathHot h;
    h.energy = p.energy;
    h.scatterPdf = p.scatterPdf;
    for (int k = 0; k < 5; ++k) h.rng[k] = p.rng[k];
    h.seg = p.seg;
    h.pad[0] = p.pad[0];
    h.pad[1] = p.pad[1];
    F.paths[j] = h;
    F.prad[j] = PathRad{p.rad, p.flags};
}

// Camera ray of one sample: jitter X then Y (renderer.cs:175), View.Ray (math.cs:1613-1636).
template <class R>
SR_DEV void camera_ray(const ViewD& v, uint32_t W, uint32_t H, uint32_t x, uint32_t y, R& rng, vec3& o, vec3& d) {
    const float jx = next_float(rng);
    const float jy = next_float(rng);
    const vec2 sp = mk2(((float)x + jx) / (float)W, ((float)y + jy) / (float)H);
    const float ndcX = sp.x * 2.0f - 1.0f;
    const float ndcY = -sp.y * 2.0f + 1.0f;
    float th, tv;
    if (v.aspect >= 1.0f) { th = v.tanFov; tv = th / v.aspect; }
    else { tv = v.tanFov; th = tv * v.aspect; }
    o = v.pos + rotate(v.rot, mk3(0.0f, 0.0f, v.nearPlane));
    d = rotate(v.rot, normalized(mk3(ndcX * th, ndcY * tv, 1.0f)));
}

// Starts sample `sample` of path slot j: its camera ray and the random state after the ray's two
// jitter draws. The shade of segment 0 starts the path record's other fields (start_path).
// Whether the random state after a camera ray is derived by the segment-0 shade (counter
// generator in a camSlots render) rather than stored.
template <class R> SR_DEV bool camera_rng_derived(const Frame& F);
template <> SR_DEV bool camera_rng_derived<XorwowRng>(const Frame&) { return false; }
template <> SR_DEV bool camera_rng_derived<CounterRng>(const Frame& F) { return F.camSlots != 0; }

template <class R>
SR_DEV void begin_sample(const Frame& F, size_t j, uint32_t x, uint32_t y, uint32_t sample) {
    R rng = rng_for_sample<R>(F, j, x, y, sample);
    vec3 o, d;
    camera_ray(F.view, F.W, F.H, x, y, rng, o, d);
    F.rayO[j] = make_float4(o.x, o.y, o.z, 0.0f);
    F.rayD[j] = make_float4(d.x, d.y, d.z, 0.0f);
    if (!camera_rng_derived<R>(F)) rng_store_words(F.paths + j, rng);
}

// The record of a path at its camera ray (Scene.Sample's initial state, scene.cs:632-640) with the
// random state begin_sample left: stored, or for a derived counter state the generator of
// (pixel, sample0 + b) after camera_ray's two draws.
template <class R>
SR_DEV PathRec start_path(const Frame& F, uint32_t j) {
    PathRec p;
    p.energy = splat3(1.0f);
    p.scatterPdf = 0.0f;
    p.rad = splat3(0.0f);
    p.flags = 0;
    if (camera_rng_derived<R>(F)) {
        const uint32_t q = j / F.batch, b = j - q * F.batch;
        uint32_t x, y;
        slot_pixel(F, q, x, y);
        p.rng[0] = F.seed;
        p.rng[1] = y * F.W + x;
        p.rng[2] = F.sample0 + b;
        p.rng[3] = 2;
        p.rng[4] = 0;
    } else {
        const uint32_t* w = F.paths[j].rng;
        for (int k = 0; k < 5; ++k) p.rng[k] = w[k];
    }
    p.seg = 0;
    p.pad[0] = p.pad[1] = 0;
    if (F.fog) F.cold[j].radFog = splat3(0.0f);  // the medium is written when a path enters one
    return p;
}

// The render counters are kept in kCtrSpread copies, picked by block index, so the many short
// blocks of a launch do not all queue their atomics on one address; reading sums the copies.
SR_DEV unsigned long long* ctr_at(const Frame& F, int idx) {
    return F.ctr + (size_t)(blockIdx.x & (kCtrSpread - 1)) * kCtrWords + idx;
}
// Sums a per-lane count over the wave and adds it to a render counter with one atomic; every lane
// of the wave calls it.
SR_DEV void bump_sum(const Frame& F, int idx, uint32_t v) {
    if (F.ctr == nullptr) return;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane_id() == 0 && v) atomicAdd(ctr_at(F, idx), (unsigned long long)v);
}
// Raises a max-reduced counter word to the wave's largest key (0 = nothing to add); every lane of
// the wave calls it.
SR_DEV void bump_max(const Frame& F, int idx, unsigned long long key) {
    if (F.ctr == nullptr) return;
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(key, o);
        key = other > key ? other : key;
    }
    if (lane_id() == 0 && key) atomicMax(ctr_at(F, idx), key);
}
// Counters.BumpMax(float) (counters.cs:214-219) keeps (long)(value * 1e4) in a signed word; this is
// that value as an unsigned key whose order matches the signed order (NaN adds nothing).
SR_DEV unsigned long long decimal_key(float v) {
    // NaN and values outside the long range convert to long.MinValue on x64 (the reduction's
    // initial value, i.e. no change).
    const float s = v * 1e4f;
    if (!(s == s) || s >= 9.2233720e18f || s < -9.2233720e18f) return 0ull;
    return (unsigned long long)(long long)s ^ 0x8000000000000000ull;
}
// Counts the lanes of the wave where pred holds with one atomic; every lane of the wave calls it.
SR_DEV void bump_wave(const Frame& F, int idx, bool pred) {
    if (F.ctr == nullptr) return;
    const unsigned long long mask = __ballot(pred);
    if (mask && (int)lane_id() == __ffsll(mask) - 1) atomicAdd(ctr_at(F, idx), (unsigned long long)__popcll(mask));
}

// ---- Kernels ------------------------------------------------------------------------------------
__global__ void k_reset(Frame F) {
    if (threadIdx.x < kCounterWords) F.counts[threadIdx.x] = 0;
}

// new Rng(x, y) per pixel (renderer.cs:156), zeroed sums, and the camera rays of the first batch
// (samples 0 .. batch - 1; one sample unless the samples are independent, see Frame::batch). One
// thread per path slot: slot j = q * batch + b is sample b of pixel slot q, so a pixel's samples sit
// next to each other in every buffer and in queue 0.
template <class R>
__global__ void __launch_bounds__(kBlock) k_start(Frame F) {
    __shared__ uint32_t sh[kBlock / 64 + 1];
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t q = j / F.batch, b = j - q * F.batch;
    const uint32_t nb = min(F.batch, F.maxS);
    bool live = false;
    if (q < F.nSlots) {
        uint32_t x, y;
        live = slot_pixel(F, q, x, y);
        if (b == 0) {
            PixAcc a;
            memset(&a, 0, sizeof(a));
            a.s.active = live ? 1u : 0u;
            F.acc[q] = a;
        }
        live = live && b < nb;
        if (live) begin_sample<R>(F, j, x, y, b);
    }
    bump_wave(F, 0, live && b == 0);
    bump_wave(F, 1, live);
    if (F.camSlots) {  // the queue in slot order, no atomic
        if (j < F.nPaths) F.queue[0][j] = live ? j : kNoPath;
        if (j == 0) F.counts[kRayCount0] = F.nPaths;
        return;
    }
    const uint32_t slot = block_append<kBlock / 64>(&F.counts[kRayCount0], live, sh);
    if (live && queue_room(F, slot, 1u)) F.queue[0][slot] = j;
}

// Counts of one shade pass, flushed into the render counters once per thread.
struct ShadeCount {
    uint32_t hit = 0, miss = 0, term = 0, fog = 0, esc = 0, occ = 0, occCulled = 0;
};

// One path segment (scene.cs:641-737) of path q at the hit h of `ray`: emission, the NEE shadow
// ray (returned in sr, not traced here), Russian roulette and the scatter. On return `cont` says
// whether the path continues with the ray (no, nd); p and rng are updated in place.
template <class R, bool FOG = true>
SR_DEV void shade_path(const Frame& F, const DevScene& S, uint32_t q, PathRec& p, R& rng, const HitRec& h,
                       const ray3& ray, bool& cont, bool& shadow, ShadowRec& sr, vec3& no, vec3& nd, ShadeCount& c) {
    cont = false;
    shadow = false;
    const uint32_t seg = p.seg;
    const bool primary = seg == 0;
    const bool last = seg == F.bounces;
    const bool hit = h.obj >= 0;
    float sd = 0.0f;
    no = ray.o;
    nd = ray.d;
    if (FOG && S.fog.enabled && medium_scatter(S.fog, ray, 0.0f, hit ? h.t : kInf, rng, sd)) {
        ++c.fog;
        p.energy = p.energy * S.fog.color;
        const vec3 pos = at(ray, sd);
        vec3 ld;
        float lpdf;
        SkyRad lrad;
        if (sky_light_rand(S, rng, ld, lpdf, nullptr, &lrad) && lpdf > 0.0f) {
            const float fogPdf = hg_pdf(dot3(ray.d, ld), S.fog.anisotropy);
            sr.o = pos; sr.d = ld;
            sr.skyRad = sky_light_radiance(S, lrad, ld);
            sr.f = splat3(fogPdf);
            sr.w = power_heuristic(lpdf, fogPdf);
            sr.pdf = lpdf;
            sr.energy = p.energy;
            sr.path = q | (primary ? kToFog : 0u);
            shadow = true;
        }
        bool term = false;
        if (seg >= 3) {
            const float survive = net_min(1.0f, max_comp(p.energy));
            if (next_float(rng) >= survive) { term = true; ++c.term; }
            else p.energy = p.energy / survive;
        }
        if (!term) {
            const vec3 dir = hg_scatter(ray.d, S.fog.anisotropy, rng);
            p.flags |= 1u;
            p.scatterPdf = hg_pdf(dot3(ray.d, dir), S.fog.anisotropy);
            no = pos;
            nd = dir;
            cont = !last;
        }
    } else if (hit) {
        ++c.hit;
        if (FOG && S.fog.enabled) ++c.esc;
        SurfaceD s;
        make_surface(S, ray, h, s);
        const float dist = h.t;
        const vec3 view = -ray.d;
        const float bias = net_max(1e-4f, dist * 1e-4f);
        const vec3 hitPos = at(ray, dist - bias);
        if (p.flags & 2u) p.energy = p.energy * beer(F.cold[q].medium, dist);
        p.rad = p.rad + s.radiance * p.energy;
        if (primary) {  // the pixel's normal, depth, albedo and uv sums take it in sample order
            F.prim[q] = make_float4(s.normal.x, s.normal.y, s.normal.z, dist);
            F.primB[q] = make_float4(s.albedo.x, s.albedo.y, s.albedo.z, s.uv.x);
            F.primV[q] = s.uv.y;
            p.flags |= kPrimaryHit;
        }
        // SampleSkyDirect (scene.cs:774-794): the sides are tested before the shadow ray,
        // which only skips rays whose contribution is black either way.
        vec3 ld;
        float lpdf;
        SkyRad lrad;
        if (sky_light_rand(S, rng, ld, lpdf, nullptr, &lrad) && lpdf > 0.0f) {
            const vec3 nn = dot3(s.normal, view) >= 0.0f ? s.normal : -s.normal;
            const vec3 ng = dot3(s.normalGeo, view) >= 0.0f ? s.normalGeo : -s.normalGeo;
            if (dot3(nn, ld) > 0.0f && dot3(ng, ld) > 0.0f) {
                sr.o = hitPos; sr.d = ld;
                sr.skyRad = sky_light_radiance(S, lrad, ld);
                sr.f = opaque_eval(s, view, ld);
                sr.w = power_heuristic(lpdf, opaque_pdf(s, view, ld, spec_prob(s, view)));
                sr.pdf = lpdf;
                sr.energy = p.energy;
                sr.path = q;
                shadow = true;
            } else if (F.ctr) {
                sr.o = hitPos; sr.d = ld;
                sr.path = q | kCountOnly;
                shadow = true;
            }
        }
        bool term = false;
        if (seg >= 3) {
            const float survive = net_min(1.0f, max_comp(p.energy));
            if (next_float(rng) >= survive) { term = true; ++c.term; }
            else p.energy = p.energy / survive;
        }
        if (!term) {
            const ScatterD sc = surface_scatter(s, ray.d, view, rng);
            p.energy = p.energy * sc.weight;
            p.energy = clamp_lum(p.energy, F.clamp);
            const bool enter = dot3(sc.dir, s.normalGeo) < 0.0f;
            if (sc.refracted) {
                if (enter) { p.flags |= 2u; F.cold[q].medium = s.albedo; }
                else p.flags &= ~2u;
            } else if (!enter) {
                p.flags &= ~2u;
            }
            if (sc.hasPdf) { p.flags |= 1u; p.scatterPdf = sc.pdf; }
            else p.flags &= ~1u;
            no = at(ray, dist) + sc.dir * bias;
            nd = sc.dir;
            cont = !last;
        }
    } else {
        ++c.miss;
        if (FOG && S.fog.enabled) ++c.esc;
        float mis = 1.0f;
        if (p.flags & 1u) {
            float lp = 0.0f;
            if (!sky_light_pdf(S, ray.d, lp)) lp = 0.0f;
            mis = power_heuristic(p.scatterPdf, lp);
        }
        p.rad = p.rad + sky_radiance(S, ray.d) * mis * p.energy;
    }
}

// One path segment for every queued path (scene.cs:641-737). Direct sky light becomes a deferred
// shadow ray whose contribution the next trace launch adds before the path's next segment, which
// keeps the reference's summation order.
// Waves per SIMD the fog-free shade is held to (at most 168 VGPRs); the instance with the
// participating-medium code runs at 2.
#ifndef SR_SHADE_MIN_WAVES
#define SR_SHADE_MIN_WAVES 3
#endif
// FOG = false compiles the participating-medium code out for scenes without fog or media (fewer
// registers: 3 waves per SIMD instead of 2).
// FIRST: the launch's rays are camera rays, whose path records start here (start_path); its own
// instance keeps the other segments' register count (a runtime flag cost the fog-free instance
// its third wave per SIMD).
template <class R, bool FOG, bool FIRST>
__global__ void __launch_bounds__(kShadeBlock, FOG ? 1 : SR_SHADE_MIN_WAVES) k_shade(Frame F, DevScene S, int cur) {
    const uint32_t n = queue_len(F, cur);
    const uint32_t* Q = F.queue[cur];
    ShadeCount c;
    const uint32_t stride = gridDim.x * blockDim.x;
    __shared__ uint32_t sQ[kShadeBlock];
    __shared__ uint32_t sHits[kShadeBlock / 64];
    __shared__ uint32_t sItems[kShadeBlock / 64];
    __shared__ uint32_t shApp[(kShadeBlock / 64) * 8 + 1];
    const uint32_t wave = threadIdx.x >> 6;
    // Grid-stride over the queue with a block-uniform base, so every lane of a wave runs the same
    // number of iterations (the wave-level appends need the whole wave).
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        // Hits first, then misses, within the block's chunk of the queue (each class in queue
        // order): most waves then run only one of the two shading paths (83 -> 67 ms of shade per
        // 1080p frame). Every path is shaded exactly once whatever lane takes it.
        // The chunk's paths: its queue entries, but for kNoPath entries (camera rays queued in slot
        // order, Frame::camSlots), which only a segment-0 queue holds.
        uint32_t items = 0;
        {
            const uint32_t qj = i < n ? Q[i] : kNoPath;
            const bool valid = qj != kNoPath;
            bool isHit = false;
            if (valid) isHit = __float_as_int(F.hitA[qj].w) >= 0;
            const unsigned long long hm = __ballot(isHit);
            const unsigned long long vm = FIRST ? __ballot(valid) : 0ull;
            if (lane_id() == 0) {
                sHits[wave] = (uint32_t)__popcll(hm);
                if (FIRST) sItems[wave] = (uint32_t)__popcll(vm);
            }
            __syncthreads();
            uint32_t hitsBefore = 0, hitsTotal = 0, itemsBefore = 0;
            for (uint32_t w = 0; w < kShadeBlock / 64; ++w) {
                const uint32_t hw = sHits[w];
                hitsTotal += hw;
                if (w < wave) hitsBefore += hw;
                if (FIRST) {
                    items += sItems[w];
                    if (w < wave) itemsBefore += sItems[w];
                }
            }
            if (!FIRST) items = min(n - base, (uint32_t)kShadeBlock);
            if (valid) {
                const unsigned long long below = (1ull << lane_id()) - 1ull;
                const uint32_t hitRank = hitsBefore + (uint32_t)__popcll(hm & below);
                const uint32_t itemRank = FIRST ? itemsBefore + (uint32_t)__popcll(vm & below) : threadIdx.x;
                sQ[isHit ? hitRank : hitsTotal + (itemRank - hitRank)] = qj;
            }
            __syncthreads();
        }
        bool cont = false, shadow = false;
        ShadowRec sr;
        uint32_t q = 0;
        vec3 nd = splat3(0.0f);
        if (threadIdx.x < items) {
            q = sQ[threadIdx.x];
            PathRec p = FIRST ? start_path<R>(F, q) : load_path(F, q);
            const float4 ho = F.hitA[q];
            const float4 ro = F.rayO[q], rd = F.rayD[q];
            HitRec h;
            h.t = ho.x; h.u = ho.y; h.v = ho.z;
            unpack_hit(F, S, q, __float_as_int(ho.w), h);
            R rng = rng_load<R>(p);
            const ray3 ray = mk_ray(mk3(ro.x, ro.y, ro.z), mk3(rd.x, rd.y, rd.z));
            vec3 no;
            shade_path<R, FOG>(F, S, q, p, rng, h, ray, cont, shadow, sr, no, nd, c);
            if (cont) {
                p.seg = p.seg + 1;
                F.rayO[q] = make_float4(no.x, no.y, no.z, 0.0f);
                F.rayD[q] = make_float4(nd.x, nd.y, nd.z, 0.0f);
            }
            if (cont) {
                rng_store(p, rng);
                store_path(F, q, p);
            } else {
                store_finished(F, q, p, rng);
            }
        }
        if (shadow) {
            if (sr.path & kCountOnly) ++c.occCulled;
            else ++c.occ;
        }
        // Block-aggregated appends: one atomic per block and queue (per-wave atomics on the two
        // counters cost 66 vs 54 ms of shade per 1080p frame). A block's rays come from a few
        // neighbouring tiles; grouping them by direction octant gives each traversal wave rays
        // that start close together and mostly share their child orderings.
        const uint32_t ns = block_append_keyed<kShadeBlock / 64, 8>(&F.counts[kShadowCount0 + cur],
                                                                    shadow ? dir_octant(sr.d) : -1, shApp);
        if (shadow && queue_room(F, ns, 2u)) F.shadows[cur][ns] = sr;
        const uint32_t nc = block_append_keyed<kShadeBlock / 64, 8>(&F.counts[cur ^ 1], cont ? dir_octant(nd) : -1, shApp);
        if (cont && queue_room(F, nc, 1u)) F.queue[cur ^ 1][nc] = q;
        __syncthreads();  // sQ and sHits serve the block's next chunk
    }
    if (F.ctr) {
        bump_sum(F, 3, c.hit); bump_sum(F, 4, c.miss); bump_sum(F, 5, c.term); bump_sum(F, 6, c.fog);
        bump_sum(F, 7, c.esc); bump_sum(F, 11, c.occ); bump_sum(F, kRefOcclude, c.occCulled);
    }
}

// ShadowTransmittance (scene.cs:576-600) of one shadow ray: the reference's closest-hit walk
// through transparent surfaces, or, in scenes without transparent objects, 0 or 1 from an any-hit
// traversal, which decides occlusion at the first hit the closest-hit walk would find. While
// counting (COUNT), the closest-hit walk runs in every scene so the counters follow the reference.
template <bool COUNT, bool TRANSP>
SR_DEV vec3 shadow_transmittance(const DevScene& S, const ShadowRec& sr, Tally& tl, uint32_t& extra) {
    ray3 ray = mk_ray(sr.o, sr.d);
    vec3 tr = splat3(1.0f);
    if (!TRANSP && !COUNT) {
        if (scene_any<COUNT>(S, ray, tl, lds_column())) tr = splat3(0.0f);
    } else {
        for (;;) {
            HitRec h;
            h.obj = -1;
            if (!scene_closest<COUNT>(S, ray, h, tl, lds_column())) break;
            const DevObject& ob = S.objects[h.obj];
            if (ob.transparency <= 0.0f) { tr = splat3(0.0f); break; }
            SurfaceD s;
            make_surface(S, ray, h, s);
            const float nDotV = fabsf(dot3(s.normalGeo, ray.d));
            const vec3 fr = fresnel3(nDotV, base_refl(ob.f0, s.albedo, ob.metallic));
            tr = tr * (s.albedo * ob.transparency * (splat3(1.0f) - fr));
            ray = mk_ray(at(ray, h.t + net_max(1e-4f, h.t * 1e-4f)), ray.d);
            ++extra;
            if (!(max_comp(tr) > 1e-4f)) break;
        }
    }
    return tr;
}

// The shadow ray's direct-light contribution ((((sky * T) * f) * w) / pdf) * energy
// (scene.cs:793, 809, 659-661, 699), added to the path's radiance or fog radiance.
SR_DEV void add_direct(const DevScene& S, const ShadowRec& sr, vec3 tr, vec3& rad, vec3& radFog) {
    if (!(max_comp(tr) > 0.0f)) return;
    if (S.fog.enabled) tr = tr * expf(-medium_optical_depth(S.fog, mk_ray(sr.o, sr.d), 0.0f, kInf));
    const vec3 c = sr.skyRad * tr * sr.f * sr.w / sr.pdf * sr.energy;
    if (sr.path & kToFog) radFog = radFog + c;
    else rad = rad + c;
}

template <bool COUNT, bool TRANSP>
SR_DEV void shadow_item(const Frame& F, const DevScene& S, const ShadowRec& sr, Tally& tl, uint32_t& extra) {
    const vec3 tr = shadow_transmittance<COUNT, TRANSP>(S, sr, tl, extra);
    if ((sr.path & kCountOnly) || !(max_comp(tr) > 0.0f)) return;
    const uint32_t j = sr.path & kPathMask;
    if (sr.path & kToFog) {
        vec3 rad, radFog = F.cold[j].radFog;
        add_direct(S, sr, tr, rad, radFog);
        F.cold[j].radFog = radFog;
    } else {
        vec3 rad = F.prad[j].rad, radFog;
        add_direct(S, sr, tr, rad, radFog);
        F.prad[j].rad = rad;
    }
}

// The remaining segments of every queued path in one launch: each lane runs its path's
// Scene.Sample loop (scene.cs:641-737) to the end, shading, tracing the shadow ray and adding its
// light at once (the order the wavefront launches keep), then tracing the next closest hit. Late
// segments hold few paths, and one launch per segment would mostly wait on each launch's slowest
// ray.
#ifndef SR_TAIL_MIN_WAVES
#define SR_TAIL_MIN_WAVES 1
#endif
template <class R, bool COUNT, bool TRANSP>
__global__ void __launch_bounds__(kBlock, SR_TAIL_MIN_WAVES) k_tail(Frame F, DevScene S, int cur, int first) {
    const uint32_t n = queue_len(F, cur);
    const uint32_t* Q = F.queue[cur];
    ShadeCount c;
    Tally tl = {0, 0, 0};
    uint32_t extra = 0, traces = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t q = Q[i];
        PathRec p = first ? start_path<R>(F, q) : load_path(F, q);
        vec3 radFog = F.fog ? F.cold[q].radFog : splat3(0.0f);
        R rng = rng_load<R>(p);
        const float4 ho = F.hitA[q];
        const float4 ro = F.rayO[q], rd = F.rayD[q];
        HitRec h;
        h.t = ho.x; h.u = ho.y; h.v = ho.z;
        unpack_hit(F, S, q, __float_as_int(ho.w), h);
        ray3 ray = mk_ray(mk3(ro.x, ro.y, ro.z), mk3(rd.x, rd.y, rd.z));
        for (;;) {
            bool cont, shadow;
            ShadowRec sr;
            vec3 no, nd;
            shade_path<R>(F, S, q, p, rng, h, ray, cont, shadow, sr, no, nd, c);
            if (shadow) {  // side-culled rays (counters only) count with this launch's traversal
                ++c.occ;
                const vec3 tr = shadow_transmittance<COUNT, TRANSP>(S, sr, tl, extra);
                if (!(sr.path & kCountOnly)) add_direct(S, sr, tr, p.rad, radFog);
            }
            if (!cont) break;
            p.seg = p.seg + 1;
            ray = mk_ray(no, nd);
            h.t = kInf; h.u = 0.0f; h.v = 0.0f; h.obj = -1; h.slot = -1;
            scene_closest<COUNT>(S, ray, h, tl, lds_column());
            ++traces;
        }
        store_finished(F, q, p, rng);
        if (F.fog) F.cold[q].radFog = radFog;
    }
    if (F.ctr) {
        bump_sum(F, 3, c.hit); bump_sum(F, 4, c.miss); bump_sum(F, 5, c.term); bump_sum(F, 6, c.fog);
        bump_sum(F, 7, c.esc); bump_sum(F, 11, c.occ + extra); bump_sum(F, 2, traces);
        // This launch's traversal work is kept apart (kTailNode..) so the trace kernel's own share
        // stays measurable; sr_counters_read adds it back.
        bump_sum(F, kTailTrace, traces); bump_sum(F, kTailOcclude, c.occ + extra);
        bump_sum(F, kTailNode, tl.box); bump_sum(F, kTailShape, tl.shape); bump_sum(F, 12, tl.mesh);
    }
}

// One traversal launch per segment: the closest hit of every ray queued for this segment, plus
// the shadow rays the previous segment's shade deferred. The previous segment's direct light thus
// lands in the path's radiance before this segment's shade adds anything (the reference's
// summation order), and the short tail launches of late segments share the GPU with the shadow
// rays instead of running alone. Each wave takes only one kind of item.
template <bool COUNT, bool TRANSP>
// Register budget of the trace kernel: 8 waves per SIMD, the most a SIMD holds (64 VGPRs, 24 spilled),
// on the 1080p / 1M-triangle frame 65.05 ms of trace against 65.75 with 7 waves (72 VGPRs) and
// 69.0 with 6.
#ifndef SR_TRACE_MIN_WAVES
#define SR_TRACE_MIN_WAVES 8
#endif
__global__ void __launch_bounds__(kTraceBlock, SR_TRACE_MIN_WAVES) k_trace(Frame F, DevScene S, int cur) {
    // The two queues this segment's shade fills start empty; nothing else touches them meanwhile.
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        F.counts[cur ^ 1] = 0;
        F.counts[kShadowCount0 + cur] = 0;
    }
    const uint32_t nE = queue_len(F, cur);
    const uint32_t nS = queue_len(F, kShadowCount0 + (cur ^ 1));
    const uint32_t nE64 = (nE + 63u) & ~63u;
    const uint32_t total = nE64 + nS;
    const uint32_t* Q = F.queue[cur];
    const ShadowRec* SR = F.shadows[cur ^ 1];
    if (COUNT && blockIdx.x == 0 && threadIdx.x == 0 && F.ctr) {
        atomicAdd(ctr_at(F, 2), (unsigned long long)nE);
        atomicAdd(ctr_at(F, 10), (unsigned long long)nE);
    }
    // tl: closest-hit rays; tlShadow: shadow rays as the reference walks them; tlAny: the any-hit
    // walks a render without counters does instead (counting, opaque scenes); tlRef: side-culled
    // shadow rays (counting only).
    Tally tl = {0, 0, 0}, tlShadow = {0, 0, 0}, tlAny = {0, 0, 0}, tlRef = {0, 0, 0};
    uint32_t extra = 0, extraRef = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < total; base += stride) {
        const uint32_t i = base + lane_id();
        if (base < nE64) {
            const uint32_t q = i < nE ? Q[i] : kNoPath;
            if (q != kNoPath) {
                const float4 o = F.rayO[q], d = F.rayD[q];
                const ray3 r = mk_ray(mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z));
                HitRec h;
                h.t = kInf; h.u = 0.0f; h.v = 0.0f; h.obj = -1; h.slot = -1;
                // Speculative mesh traversal outside counting renders: same hit, wider waves.
                scene_closest<COUNT, !COUNT>(S, r, h, tl, lds_column());
                if (S.packHit) {
                    F.hitA[q] = make_float4(h.t, h.u, h.v, __int_as_float(pack_hit(h)));
                } else {
                    F.hitA[q] = make_float4(h.t, h.u, h.v, __int_as_float(h.obj));
                    F.hitSlot[q] = h.slot;
                }
            }
        } else if (i - nE64 < nS) {
            const ShadowRec& rec = SR[i - nE64];
            if (rec.path & kCountOnly) {
                shadow_item<COUNT, TRANSP>(F, S, rec, tlRef, extraRef);
            } else {
                shadow_item<COUNT, TRANSP>(F, S, rec, tlShadow, extra);
                if (COUNT && !TRANSP) (void)scene_any<COUNT>(S, mk_ray(rec.o, rec.d), tlAny, lds_column());
            }
        }
    }
    if (COUNT && F.ctr) {
        // Reported totals follow the reference; the reference-only words hold what a render
        // without counters does not traverse (per ray the closest-hit walk visits at least the
        // any-hit walk's nodes).
        const Tally work = TRANSP ? tlShadow : tlAny;
        bump_sum(F, 13, tl.box + work.box);
        bump_sum(F, 14, tl.shape + work.shape);
        bump_sum(F, 12, tl.mesh + work.mesh);
        bump_sum(F, kRefNode, tlRef.box + tlShadow.box - work.box);
        bump_sum(F, kRefShape, tlRef.shape + tlShadow.shape - work.shape);
        bump_sum(F, kRefMesh, tlRef.mesh + tlShadow.mesh - work.mesh);
        bump_sum(F, kRefOcclude, extraRef);
    }
    bump_sum(F, 11, extra);
}

// Per texel of a texture sky: the direction toward its centre and its light pdf, then the sky
// radiance seen in that direction, so a light sample reads one 32-byte entry (sky_light_rand). S has
// no table while this runs.
__global__ void k_sky_table(DevScene S, float4* out) {
    const DevSky& k = S.sky;
    const uint32_t n = (uint32_t)k.w * (uint32_t)k.h;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = i % (uint32_t)k.w, y = i / (uint32_t)k.w;
    const vec3 d = sky_texel_dir(k, x, y);
    const float pdf = sky_texel_pdf(S, i);
    const vec3 r = sky_radiance(S, d);
    out[2 * (size_t)i] = make_float4(d.x, d.y, d.z, pdf);
    out[2 * (size_t)i + 1] = make_float4(r.x, r.y, r.z, 0.0f);
}

// After the fused tail launch: both ray queues are spent. k_next appends the next sample's camera
// rays to queue 0, which the tail launch read but did not empty.
__global__ void k_queue_reset(Frame F) {
    if (threadIdx.x == 0) {
        F.counts[kRayCount0] = 0;
        F.counts[kRayCount1] = 0;
    }
}

// The batch fold of k_next<R, true>: PixAcc viewed as 28 words, each updated by one lane per batch
// with the operation the sequential fold applies to it, from a column of the samples' LDS rows.
// Row columns: 0-2 radiance, 3-5 fog radiance, 6 luminance, 7 its square, 8 surface luminance,
// 9 its square, 10-12 primary normal, 13 depth, 14-16 albedo, 17-18 uv, 19 the constant 1.
constexpr uint32_t kFoldRow = 21;  // words per row (odd: a group's lanes reading a row hit distinct banks)
constexpr uint32_t kFoldAdd = 0, kFoldAddHit = 1, kFoldSetHit = 2, kFoldIncHit = 3, kFoldFirstHit = 4;
constexpr uint32_t kFoldJobs = 25;
#define SR_FOLD(word, src, kind) ((word) | ((src) << 8) | ((kind) << 16))
__constant__ uint32_t kFoldJob[kFoldJobs] = {
    SR_FOLD(0, 0, 0), SR_FOLD(1, 1, 0), SR_FOLD(2, 2, 0),          // radSum
    SR_FOLD(3, 3, 0), SR_FOLD(4, 4, 0), SR_FOLD(5, 5, 0),          // radFogSum
    SR_FOLD(6, 6, 0), SR_FOLD(7, 7, 0), SR_FOLD(8, 8, 0), SR_FOLD(9, 9, 0),  // lumSum, lumSumSqr, lumSurf, lumSurfSqr
    SR_FOLD(12, 10, 1), SR_FOLD(13, 11, 1), SR_FOLD(14, 12, 1),    // normalSum
    SR_FOLD(15, 10, 2), SR_FOLD(16, 11, 2), SR_FOLD(17, 12, 2),    // normalFallback (last hit)
    SR_FOLD(18, 0, 3),                                             // normalCount
    SR_FOLD(19, 17, 4), SR_FOLD(20, 18, 4), SR_FOLD(21, 19, 4),    // uv, hasUv (first hit)
    SR_FOLD(22, 13, 1), SR_FOLD(23, 0, 3),                         // depthSum, depthCount
    SR_FOLD(24, 14, 1), SR_FOLD(25, 15, 1), SR_FOLD(26, 16, 1),    // albedoSum
};
#undef SR_FOLD
static_assert(sizeof(PixAcc) == 28 * 4 && offsetof(PixAcc, s) == 0 && offsetof(PixAcc, h) == 12 * 4, "PixAcc words");
static_assert(offsetof(PixSums, lumSum) == 6 * 4 && offsetof(PixSums, lumSurfSqr) == 9 * 4, "PixSums words");
static_assert(offsetof(PixPrimary, normalFallback) == 3 * 4 && offsetof(PixPrimary, normalCount) == 6 * 4 &&
              offsetof(PixPrimary, uv) == 7 * 4 && offsetof(PixPrimary, hasUv) == 9 * 4 &&
              offsetof(PixPrimary, depthSum) == 10 * 4 && offsetof(PixPrimary, depthCount) == 11 * 4 &&
              offsetof(PixPrimary, albedoSum) == 12 * 4, "PixPrimary words");

// Renderer.RelativeStdErr (renderer.cs:232-237).
SR_DEV float rel_std_err(float sum, float sumSqr, uint32_t n) {
    const float mean = sum / (float)n;
    const float variance = net_max(0.0f, sumSqr / (float)n - mean * mean);
    return sqrtf(variance / (float)n) / net_max(mean, 1e-4f);
}

// After a batch's last segment, per active pixel and in sample order: accumulation, the primary
// hit's normal / depth / albedo / uv and the convergence test (renderer.cs:180-212); then either
// the pixel's next batch of camera rays, appended to queue 0 in slot order so the next camera rays
// stay tile-coherent, or its final AOVs (renderer.cs:214-229, 142-148).
// One thread per path slot (j = q * batch + b; batch a power of two, so a pixel's slots are a
// group of batch lanes of one wave): every thread loads its own path's radiance, flags and primary
// hit record (coalesced), the group's first lane folds them into the pixel's sums in sample order
// through lane shuffles, and every thread starts its own sample of the next batch.
// MULTI (more than one sample in flight): the fold runs on the group's lanes, one PixAcc word per
// lane (kFoldJobs), over the samples' values staged in LDS rows; the sequential fold below serves
// one sample in flight.
template <class R, bool MULTI>
__global__ void __launch_bounds__(kNextBlock) k_next(Frame F) {
    __shared__ uint32_t sh[kNextBlock / 64 + 1];
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t B = F.batch;
    const uint32_t q = j / B, b = j - q * B;
    const int lead = (int)(lane_id() - b);  // the pixel's first lane
    if (j == 0) {  // both shadow queues are spent; the next sample's first trace reads one of them
        F.counts[kShadowCount0] = 0;
        F.counts[kShadowCount1] = 0;
    }
    bool converged = false, maxed = false;
    uint32_t doneBlock = ~0u;  // the block of a pixel that finished in this pass
    unsigned long long countMinKey = 0, countMaxKey = 0, varKey = 0;
    PixAcc* A = F.acc + q;
    const bool pixel = q < F.nSlots;
    PixSums a = pixel ? A->s : PixSums{};  // the group's lanes read the same words
    const bool active = pixel && a.active;
    const uint32_t base = a.sampleIndex;
    const uint32_t nb = active ? min(B, F.maxS - base) : 0u;
    // This lane's path of the finished batch.
    float4 r0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), pr0 = r0, pr1 = r0, pr2 = r0;
    vec3 radFog = splat3(0.0f);
    if (b < nb) {
        r0 = *(const float4*)&F.prad[j];  // rad, flags
        if (F.fog) radFog = F.cold[j].radFog;
        if (__float_as_uint(r0.w) & kPrimaryHit) {
            pr0 = F.prim[j];
            const float4 b1 = F.primB[j];
            pr1 = make_float4(b1.x, b1.y, b1.z, 0.0f);
            pr2 = make_float4(b1.w, F.primV[j], 0.0f, 0.0f);
        }
    }
    PixPrimary h;
    bool hLoaded = false, done = false;
    uint32_t count = 0;
    if constexpr (MULTI) {
        // Every lane stages its sample's values in its LDS row; lane b of the group then folds PixAcc
        // words b, b + batch, ... over the group's rows in sample order, with the operation the
        // sequential fold applies to that word. With more than one sample in flight no sample but a
        // pixel's last can end it (no adaptive stop), so every sample of the batch is added.
        __shared__ float rows[kNextBlock * kFoldRow];
        const bool hitMine = b < nb && (__float_as_uint(r0.w) & kPrimaryHit) != 0;
        if (b < nb) {
            float* row = rows + threadIdx.x * kFoldRow;
            const vec3 rad = mk3(r0.x, r0.y, r0.z);
            const float lumSurf = luminance(rad);
            const float lum = lumSurf + luminance(radFog);
            row[0] = rad.x; row[1] = rad.y; row[2] = rad.z;
            row[3] = radFog.x; row[4] = radFog.y; row[5] = radFog.z;
            row[6] = lum; row[7] = lum * lum; row[8] = lumSurf; row[9] = lumSurf * lumSurf;
            row[10] = pr0.x; row[11] = pr0.y; row[12] = pr0.z; row[13] = pr0.w;
            row[14] = pr1.x; row[15] = pr1.y; row[16] = pr1.z;
            row[17] = pr2.x; row[18] = pr2.y; row[19] = __uint_as_float(1u);
        }
        const uint32_t groupHits = (uint32_t)((__ballot(hitMine) >> lead) & ((1ull << B) - 1ull));
        // hasUv before this batch (the word's own fold lane rewrites it).
        const bool hasUv0 = active && groupHits != 0 && A->h.hasUv != 0;
        __syncthreads();
        uint32_t* W = (uint32_t*)A;
        if (active) {
            for (uint32_t job = b; job < kFoldJobs; job += B) {
                const uint32_t d = kFoldJob[job];
                const uint32_t word = d & 0xffu, src = (d >> 8) & 0xffu, kind = d >> 16;
                if (kind != kFoldAdd && groupHits == 0) continue;  // primary-hit words change only at hits
                uint32_t acc = W[word];
                bool has = kind == kFoldFirstHit && hasUv0;
                const float* rk = rows + (threadIdx.x - b) * kFoldRow + src;
#pragma unroll 8
                for (uint32_t k = 0; k < nb; ++k, rk += kFoldRow) {
                    const bool hitK = ((groupHits >> k) & 1u) != 0;
                    const float v = *rk;
                    const uint32_t added = __float_as_uint(__uint_as_float(acc) + v);
                    const uint32_t nv = kind <= kFoldAddHit ? added : kind == kFoldIncHit ? acc + 1u : __float_as_uint(v);
                    const bool go = kind == kFoldAdd || (hitK && !(kind == kFoldFirstHit && has));
                    acc = go ? nv : acc;
                    has = has || (kind == kFoldFirstHit && hitK);
                }
                W[word] = acc;
            }
        }
        __syncthreads();  // the group's words are in the pixel's record
        done = active && base + nb >= F.maxS;
        if (active && b == 0 && done) {
            a = A->s;
            h = A->h;
            hLoaded = true;
            const uint32_t si = base + nb - 1u;
            count = si >= F.minS - 1u && rel_std_err(a.lumSum, a.lumSumSqr, si + 1) < F.threshold ? si + 1 : F.maxS + 1;
        }
    } else {
        // The first lane folds the group's samples in order (the others follow the shuffles).
        for (uint32_t k = 0; k < B; ++k) {
            const int src = lead + (int)k;
            const float4 s0 = make_float4(__shfl(r0.x, src), __shfl(r0.y, src), __shfl(r0.z, src), __shfl(r0.w, src));
            const vec3 sFog = mk3(__shfl(radFog.x, src), __shfl(radFog.y, src), __shfl(radFog.z, src));
            const bool hitK = (__float_as_uint(s0.w) & kPrimaryHit) != 0;
            float4 p0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), p1 = p0;
            float2 p2 = make_float2(0.0f, 0.0f);
            if (__ballot(hitK && b == 0 && k < nb) != 0) {
                p0 = make_float4(__shfl(pr0.x, src), __shfl(pr0.y, src), __shfl(pr0.z, src), __shfl(pr0.w, src));
                p1 = make_float4(__shfl(pr1.x, src), __shfl(pr1.y, src), __shfl(pr1.z, src), 0.0f);
                p2 = make_float2(__shfl(pr2.x, src), __shfl(pr2.y, src));
            }
            if (b != 0 || k >= nb || done) continue;
            const uint32_t si = base + k;
            if (hitK) {
                if (!hLoaded) { h = A->h; hLoaded = true; }
                const vec3 nrm = mk3(p0.x, p0.y, p0.z);
                h.normalSum = h.normalSum + nrm;
                h.normalFallback = nrm;
                h.normalCount += 1;
                h.depthSum += p0.w;
                h.depthCount += 1;
                h.albedoSum = h.albedoSum + mk3(p1.x, p1.y, p1.z);
                if (!h.hasUv) { h.hasUv = 1; h.uv = mk2(p2.x, p2.y); }
            }
            const vec3 rad = mk3(s0.x, s0.y, s0.z);
            a.radSum = a.radSum + rad;
            a.radFogSum = a.radFogSum + sFog;
            const float lum = luminance(rad) + luminance(sFog);
            a.lumSum = a.lumSum + lum;
            a.lumSumSqr = a.lumSumSqr + lum * lum;
            const float lumSurf = luminance(rad);
            a.lumSurf = a.lumSurf + lumSurf;
            a.lumSurfSqr = a.lumSurfSqr + lumSurf * lumSurf;
            if (si >= F.minS - 1u && rel_std_err(a.lumSum, a.lumSumSqr, si + 1) < F.threshold) {
                done = true;
                count = si + 1;
            } else if (si + 1 == F.maxS) {
                done = true;
                count = F.maxS + 1;  // the loop exits with sampleIndex == maxSamples (renderer.cs:173,215)
            }
        }
    }
    uint32_t x = 0, y = 0;
    if (pixel) slot_pixel(F, q, x, y);
    uint32_t again = 0;  // samples of the next batch this pixel starts
    if (active && b == 0) {
        if (!done) {
            const uint32_t next = base + nb;
            if (MULTI) {
                A->s.sampleIndex = next;  // the sums are in place
            } else {
                a.sampleIndex = next;
                A->s = a;
                if (hLoaded) A->h = h;
            }
            again = min(B, F.maxS - next);
        } else {
            if (!hLoaded) h = A->h;
            A->s.active = 0;
            const size_t o = (size_t)y * F.W + x;
            const float var = rel_std_err(a.lumSurf, a.lumSurfSqr, count);
            const vec3 r = a.radSum / (float)count;
            const vec3 rf = a.radFogSum / (float)count;
            const vec3 nrm = h.normalCount > 0 ? normalized_or(h.normalSum, h.normalFallback) : splat3(0.0f);
            const float depth = h.depthCount > 0 ? h.depthSum / (float)h.depthCount : kInf;
            F.outRad[o * 3] = r.x; F.outRad[o * 3 + 1] = r.y; F.outRad[o * 3 + 2] = r.z;
            F.outRadFog[o * 3] = rf.x; F.outRadFog[o * 3 + 1] = rf.y; F.outRadFog[o * 3 + 2] = rf.z;
            F.outNormal[o * 3] = nrm.x; F.outNormal[o * 3 + 1] = nrm.y; F.outNormal[o * 3 + 2] = nrm.z;
            F.outUv[o * 2] = h.hasUv ? h.uv.x : 0.0f;
            F.outUv[o * 2 + 1] = h.hasUv ? h.uv.y : 0.0f;
            F.outDepth[o] = depth;
            const vec3 alb = h.depthCount > 0 ? h.albedoSum / (float)h.depthCount : splat3(0.0f);
            F.outAlbedo[o * 3] = alb.x; F.outAlbedo[o * 3 + 1] = alb.y; F.outAlbedo[o * 3 + 2] = alb.z;
            F.outSamples[o] = count;
            F.outVar[o] = var;
            converged = count < F.maxS;
            maxed = !converged;
            countMinKey = ~(unsigned long long)count;
            countMaxKey = count;
            varKey = decimal_key(var);
            doneBlock = (y / F.bs) * F.bx + x / F.bs;
        }
    }
    // The first lane's decision reaches the group: every lane starts its own sample of the batch.
    again = (uint32_t)__shfl((int)again, lead);
    const uint32_t next = (uint32_t)__shfl((int)(base + nb), lead);
    const bool start = b < again;
    if (start) begin_sample<R>(F, j, x, y, next + b);
    // Tick's progress: the pixel's block (renderer.cs:125-134) is complete once all its pixels are
    // final. A wave's 8x8 tile lies in one block when the block size is a multiple of 8: then one
    // atomic per wave (per-pixel atomics on a few thousand addresses at a frame's last sample
    // serialise); otherwise one per pixel.
    // The waves of a workgroup whose finished pixels lie in one block pass (block, count) through
    // LDS and the workgroup's first thread adds each run of equal blocks with one atomic.
    __shared__ uint32_t waveBlk[kNextBlock / 64], waveAdd[kNextBlock / 64];
    const unsigned long long doneMask = __ballot(doneBlock != ~0u);
    const uint32_t wv = threadIdx.x / 64;
    uint32_t wBlk = ~0u, wAdd = 0;
    if (doneMask) {
        const int leader = __ffsll(doneMask) - 1;
        const uint32_t lb = __shfl(doneBlock, leader);
        const bool oneBlock = __ballot(doneBlock != ~0u && doneBlock != lb) == 0;
        if (oneBlock) {
            wBlk = lb;
            wAdd = (uint32_t)__popcll(doneMask);
        } else if (doneBlock != ~0u) {
            const uint32_t bxi = doneBlock % F.bx, byi = doneBlock / F.bx;
            const uint32_t bw = min(F.bs, F.W - bxi * F.bs), bh = min(F.bs, F.H - byi * F.bs);
            if (atomicAdd(&F.blockPixels[doneBlock], 1u) + 1u == bw * bh) atomicAdd(F.blocksDone, 1u);
        }
    }
    if (lane_id() == 0) {
        waveBlk[wv] = wBlk;
        waveAdd[wv] = wAdd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t runBlk = ~0u, runAdd = 0;
        for (uint32_t w = 0; w <= blockDim.x / 64; ++w) {
            const uint32_t blk = w < blockDim.x / 64 ? waveBlk[w] : ~0u;
            if (blk == runBlk && blk != ~0u) { runAdd += waveAdd[w]; continue; }
            if (runBlk != ~0u) {
                const uint32_t bxi = runBlk % F.bx, byi = runBlk / F.bx;
                const uint32_t bw = min(F.bs, F.W - bxi * F.bs), bh = min(F.bs, F.H - byi * F.bs);
                if (atomicAdd(&F.blockPixels[runBlk], runAdd) + runAdd == bw * bh) atomicAdd(F.blocksDone, 1u);
            }
            runBlk = blk;
            runAdd = blk != ~0u ? waveAdd[w] : 0u;
        }
    }
    bump_wave(F, 8, converged);
    bump_wave(F, 9, maxed);
    bump_max(F, kCountMinKey, countMinKey);
    bump_max(F, kCountMaxKey, countMaxKey);
    bump_max(F, kVarMaxKey, varKey);
    // One atomic per block: one per wave on the single queue counter serialised the launch
    // (26.1 -> 14.0 ms per 1080p frame). A pixel's samples of the batch sit together.
    bump_wave(F, 1, start);
    if (F.camSlots) {  // the queue in slot order, no atomic
        if (j < F.nPaths) F.queue[0][j] = start ? j : kNoPath;
        if (j == 0) F.counts[kRayCount0] = F.nPaths;
        return;
    }
    const uint32_t slot = block_append<kNextBlock / 64>(&F.counts[kRayCount0], start, sh);
    if (start && queue_room(F, slot, 1u)) F.queue[0][slot] = j;
}

// Primary-hit diagnostic: sample 0 of every pixel of a band.
// Adjoint backward pass: one lane per (pixel, sample) path, replayed with the forward render's
// counter-RNG stream (key (seed, y * W + x), counter (sample, draw)). Material gradients gather in
// LDS per block; sky-texel gradients go straight to global memory.
struct AdjointArgs {
    ViewD view;
    uint32_t W, H, spp, bounces, seed, nObj;
    float clampLum;
    const float* dImage;  // W * H * 3
    float* gMat;          // nObj * 4
    float* gEnv;          // envCopies copies of sky texels * 3, envStride floats apart
    size_t envStride;
    uint32_t envCopies;
    float* gTotal;        // one float: the through-total part of the sky-pdf derivatives
};
// Copies of the sky-texel gradient: block b adds into copy b % envCopies, so lanes that hit the
// same bright texels contend on envCopies addresses instead of one (512 x 512 x 16 spp backward
// against a 256 x 128 sky: 22.2 ms with one copy, 14.6 / 12.2 / 10.9 / 10.9 ms with 8 / 64 / 128 /
// 256, 9.4 ms with the sky gradient left out); k_env_reduce sums the copies in copy order. Up to
// kMaxEnvCopies copies, at least kMinEnvCopies, within kEnvCopyBytes of device memory.
constexpr uint32_t kMinEnvCopies = 8, kMaxEnvCopies = 128;
constexpr size_t kEnvCopyBytes = (size_t)1 << 30;
__global__ void k_env_reduce(const float* copies, size_t stride, size_t n, uint32_t ncopies, float* out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = 0.0f;
    for (uint32_t c = 0; c < ncopies; ++c) v += copies[(size_t)c * stride + i];
    out[i] = v;
}
// Register budget of the path-replay kernel: 4 waves per SIMD (128 VGPRs, some spilled) against
// 2 unconstrained (175 VGPRs): backward 25.8 -> 21.8-23.3 ms at 512 x 512 x 16 spp (3, 5, 6, 8 waves:
// 23.3, 22.9, 23.6, 22.0 ms).
#ifndef SR_ADJOINT_MIN_WAVES
#define SR_ADJOINT_MIN_WAVES 4
#endif
__global__ void __launch_bounds__(kBlock, SR_ADJOINT_MIN_WAVES) k_adjoint(AdjointArgs A, DevScene S) {
    __shared__ float gl[kAdjLdsObjects * 4];
    __shared__ float glTotal;
    const uint32_t nl = (A.nObj < (uint32_t)kAdjLdsObjects ? A.nObj : (uint32_t)kAdjLdsObjects) * 4;
    for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) gl[i] = 0.0f;
    if (threadIdx.x == 0) glTotal = 0.0f;
    __syncthreads();
    MatGrad gm{gl, A.gMat, A.nObj, 0.0f};
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < (uint64_t)A.W * A.H * A.spp) {
        const uint32_t smp = (uint32_t)(id % A.spp);
        const uint32_t pix = (uint32_t)(id / A.spp);
        const uint32_t x = pix % A.W, y = pix / A.W;
        const vec3 dI = mk3(A.dImage[pix * 3], A.dImage[pix * 3 + 1], A.dImage[pix * 3 + 2]);
        if (dI.x != 0.0f || dI.y != 0.0f || dI.z != 0.0f) {
            const vec3 delta = dI / (float)(A.spp + 1);  // Radiance = sum / (spp + 1) (renderer.cs:173,215)
            CounterRngT<false> rng;
            rng.key0 = A.seed;
            rng.key1 = y * A.W + x;
            rng.sample = smp;
            rng.draw = 0;
            vec3 o, d;
            camera_ray(A.view, A.W, A.H, x, y, rng, o, d);
            adjoint_path(S, mk_ray(o, d), rng, A.bounces, A.clampLum, delta, gm,
                         A.gEnv + (size_t)(blockIdx.x % A.envCopies) * A.envStride, lds_column());
        }
    }
    if (gm.gTotal != 0.0f) atomicAdd(&glTotal, gm.gTotal);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x)
        if (gl[i] != 0.0f) unsafeAtomicAdd(&A.gMat[i], gl[i]);
    if (threadIdx.x == 0 && glTotal != 0.0f) unsafeAtomicAdd(A.gTotal, glTotal);
}

__global__ void k_primary(Frame F, DevScene S, int32_t* obj, int32_t* tri, float* dist) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= F.nSlots) return;
    uint32_t x, y;
    if (!slot_pixel(F, q, x, y)) return;
    XorwowRng rng = xorwow_seed(x, y);
    vec3 o, d;
    camera_ray(F.view, F.W, F.H, x, y, rng, o, d);
    HitRec h;
    h.t = kInf; h.obj = -1; h.slot = -1;
    Tally tl = {0, 0, 0};
    const size_t idx = (size_t)y * F.W + x;
    if (scene_closest<false>(S, mk_ray(o, d), h, tl, lds_column())) {
        const DevObject& ob = S.objects[h.obj];
        obj[idx] = h.obj;
        tri[idx] = ob.kind == 0 ? S.slotTri[ob.triBase + h.slot] : -1;
        dist[idx] = h.t;
    } else {
        obj[idx] = -1; tri[idx] = -1; dist[idx] = kInf;
    }
}

// Compositor.Filter (compose.cs:101-204): joint bilateral filter over the circular window.
// Device counters of the filter kernels (compose.cs:101-245): sums, then max keys (decimal_key).
enum : int { kFltSample, kFltEarlyOut, kFltRejectMissing, kFltRejectNormal, kFltRejectDepth, kFltRejectLum,
             kFltWeightMax, kFltMaxLum, kFltMaxLumBoost, kFilterWords };
// Adds a per-lane count to a filter counter with one atomic per wave; every lane calls it.
SR_DEV void flt_sum(unsigned long long* fc, int idx, uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane_id() == 0 && v) atomicAdd(fc + idx, (unsigned long long)v);
}
SR_DEV void flt_max(unsigned long long* fc, int idx, unsigned long long key) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(key, o);
        key = other > key ? other : key;
    }
    if (lane_id() == 0 && key) atomicMax(fc + idx, key);
}

struct FilterParams {
    float strength, strengthMax, lumBoost, lumLimitInv, normalLimitInv, depthLimitInv, radiusInv;
    float albedoLimitInv;  // ALB only: exp(-|albedo delta|^2 * albedoLimitInv) on surface pixels
    int32_t kr;
};
// Compositor.Filter (compose.cs:101-204), one pixel per lane. ALB adds an albedo guide term (no
// reference counterpart) after the depth term.
// One pixel of Compositor.Filter. n (COUNT): FilterSample, DenoiseEarlyOut, DenoiseRejectMissing,
// -Normal, -Depth, -Lum; keys: DenoiseWeightMax, DenoiseMaxLum, DenoiseMaxLumBoost.
template <bool ALB, bool COUNT>
SR_DEV void filter_pixel(const FilterParams& fp, const float* rad, const float* nrm, const float* dep,
                         const float* var, const float* alb, int32_t W, int32_t H, float* out, int32_t cx,
                         int32_t cy, uint32_t* n, unsigned long long* keys) {
    const int32_t ci = cy * W + cx;
    const vec3 cr = mk3(rad[ci * 3], rad[ci * 3 + 1], rad[ci * 3 + 2]);
    const vec3 cn = mk3(nrm[ci * 3], nrm[ci * 3 + 1], nrm[ci * 3 + 2]);
    const float cd = dep[ci];
    const bool hasCN = len_sqr(cn) > 0.0f;
    const bool hasCD = !isinf(cd);
    const vec3 ca = ALB ? mk3(alb[ci * 3], alb[ci * 3 + 1], alb[ci * 3 + 2]) : splat3(0.0f);
    const float lum = luminance(cr);
    const float lb = 1.0f + lum * fp.lumBoost;
    const float dw = net_min(var[ci] * lb * fp.strength, fp.strengthMax);
    if (dw < 1e-4f) {
        out[ci * 3] = cr.x; out[ci * 3 + 1] = cr.y; out[ci * 3 + 2] = cr.z;
        if (COUNT) n[1] += 1;
        return;
    }
    if (COUNT) {
        keys[0] = decimal_key(dw);
        keys[1] = decimal_key(lum);
        keys[2] = decimal_key(lb);
    }
    float wSum = 1.0f;
    vec3 rSum = cr;
    const int32_t kr = fp.kr;
    for (int32_t ky = -kr; ky <= kr; ++ky) {
        const int32_t ky2 = ky * ky;
        const int32_t ny = cy + ky;
        for (int32_t kx = -kr; kx <= kr; ++kx) {
            if (kx == 0 && ky == 0) continue;
            const float kd2 = (float)(kx * kx + ky2);
            if (kd2 > (float)(kr * kr)) continue;
            const int32_t nx = cx + kx;
            if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
            const int32_t ni = ny * W + nx;
            const vec3 nr = mk3(rad[ni * 3], rad[ni * 3 + 1], rad[ni * 3 + 2]);
            const vec3 nn = mk3(nrm[ni * 3], nrm[ni * 3 + 1], nrm[ni * 3 + 2]);
            const float nd = dep[ni];
            const bool hasNN = len_sqr(nn) > 0.0f;
            const bool hasND = !isinf(nd);
            if (hasCN != hasNN || hasCD != hasND) {
                if (COUNT) n[2] += 1;
                continue;
            }
            float w = dw * expf(-kd2 * fp.radiusInv);
            if (hasCN && w > 0.0f) {
                w *= expf(-len_sqr(cn - nn) * fp.normalLimitInv);
                if (COUNT && w < 1e-4f) n[3] += 1;
            }
            if (hasCD && w > 0.0f) {
                const float dd = cd - nd;
                w *= expf(-(dd * dd) * fp.depthLimitInv);
                if (COUNT && w < 1e-4f) n[4] += 1;
            }
            if (ALB && hasCN && w > 0.0f) {
                const vec3 na = mk3(alb[ni * 3], alb[ni * 3 + 1], alb[ni * 3 + 2]);
                w *= expf(-len_sqr(ca - na) * fp.albedoLimitInv);
            }
            if (w > 0.0f) {
                const float ld = net_max(0.0f, luminance(nr) - lum);
                w *= expf(-(ld * ld) * fp.lumLimitInv);
                if (COUNT && w < 1e-4f) n[5] += 1;
            }
            wSum += w;
            rSum = rSum + nr * w;
            if (COUNT) n[0] += 1;
        }
    }
    const vec3 r = rSum / wSum;
    out[ci * 3] = r.x; out[ci * 3 + 1] = r.y; out[ci * 3 + 2] = r.z;
}

// COUNT bumps the reference's filter counters (fc, kFilterWords words) the way compose.cs:130-199
// does.
template <bool ALB, bool COUNT>
__global__ void __launch_bounds__(256) k_filter(FilterParams fp, const float* rad, const float* nrm, const float* dep,
                                                 const float* var, const float* alb, int32_t W, int32_t H,
                                                 float* out, unsigned long long* fc) {
    const int32_t cx = blockIdx.x * 16 + (threadIdx.x & 15);
    const int32_t cy = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (COUNT) {
        uint32_t n[6] = {0, 0, 0, 0, 0, 0};
        unsigned long long keys[3] = {0, 0, 0};
        if (cx < W && cy < H) filter_pixel<ALB, true>(fp, rad, nrm, dep, var, alb, W, H, out, cx, cy, n, keys);
        for (int k = 0; k < 6; ++k) flt_sum(fc, kFltSample + k, n[k]);
        for (int k = 0; k < 3; ++k) flt_max(fc, kFltWeightMax + k, keys[k]);
    } else if (cx < W && cy < H) {
        filter_pixel<ALB, false>(fp, rad, nrm, dep, var, alb, W, H, out, cx, cy, nullptr, nullptr);
    }
}

// The per-pixel sample mean of a render: Radiance * Samples / min(Samples, maxSamples). A pixel that
// ran to the sample cap holds the sum of M samples divided by M + 1 (renderer.cs:205-217); one that
// converged early holds the true mean. No reference counterpart (denoiser input).
__global__ void k_sample_mean(const float* rad, const uint32_t* samples, uint32_t maxS, int32_t n, float* out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = samples[i];
    const float f = c > maxS ? (float)c / (float)maxS : 1.0f;
    out[i * 3] = rad[i * 3] * f;
    out[i * 3 + 1] = rad[i * 3 + 1] * f;
    out[i * 3 + 2] = rad[i * 3 + 2] * f;
}

// Compositor.FilterFog (compose.cs:206-245): circular Gaussian of the fog radiance, centre weight 1,
// neighbours strength * exp(-d^2 / (2 r^2)), same loop order as the reference.
SR_DEV uint32_t filter_fog_pixel(float radiusInv, float strength, int32_t kr, const float* fog, int32_t W, int32_t H,
                                 float* out, int32_t cx, int32_t cy) {
    const int32_t ci = cy * W + cx;
    uint32_t taps = 0;  // neighbours weighted: FilterSample (compose.cs:240)
    float wSum = 1.0f;
    vec3 rSum = mk3(fog[ci * 3], fog[ci * 3 + 1], fog[ci * 3 + 2]);
    for (int32_t ky = -kr; ky <= kr; ++ky) {
        const int32_t ky2 = ky * ky;
        const int32_t ny = cy + ky;
        for (int32_t kx = -kr; kx <= kr; ++kx) {
            if (kx == 0 && ky == 0) continue;
            const float kd2 = (float)(kx * kx + ky2);
            if (kd2 > (float)(kr * kr)) continue;
            const int32_t nx = cx + kx;
            if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
            const int32_t ni = ny * W + nx;
            const float w = strength * expf(-kd2 * radiusInv);
            wSum += w;
            rSum = rSum + mk3(fog[ni * 3], fog[ni * 3 + 1], fog[ni * 3 + 2]) * w;
            ++taps;
        }
    }
    const vec3 r = rSum / wSum;
    out[ci * 3] = r.x; out[ci * 3 + 1] = r.y; out[ci * 3 + 2] = r.z;
    return taps;
}
// fc (optional, counting renders): the FilterSample filter counter.
__global__ void __launch_bounds__(256) k_filter_fog(float radiusInv, float strength, int32_t kr, const float* fog,
                                                     int32_t W, int32_t H, float* out, unsigned long long* fc) {
    const int32_t cx = blockIdx.x * 16 + (threadIdx.x & 15);
    const int32_t cy = blockIdx.y * 16 + (threadIdx.x >> 4);
    uint32_t taps = 0;
    if (cx < W && cy < H) taps = filter_fog_pixel(radiusInv, strength, kr, fog, W, H, out, cx, cy);
    if (fc) flt_sum(fc, kFltSample, taps);
}

// Color.ToPixel (math.cs:127-136): Clamp01 (Math.Clamp leaves NaN), sRGB power approximation, MathF.Max
// against 0, byte conversion (NaN converts to 0, as through .NET's int32 route).
__device__ __forceinline__ uint32_t srgb_byte(float c) {
    c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
    const float e = net_max(1.055f * powf(c, 0.416666667f) - 0.055f, 0.0f);
    const float v = e * 255.0f + 0.5f;
    return v == v ? (uint32_t)(int32_t)v & 0xFFu : 0u;
}

// Compositor.Tonemap (compose.cs:247-266) of main + fog, then ToPixel: 3 bytes (R, G, B) per pixel.
// tonemapper 0 = Reinhard, 1 = LinearSmooth.
__global__ void __launch_bounds__(256) k_tonemap(const float* mainRad, const float* fog, int32_t n, int32_t tonemapper,
                                                  float exposure, uint8_t* out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const vec3 rad = mk3(mainRad[i * 3], mainRad[i * 3 + 1], mainRad[i * 3 + 2]) +
                     mk3(fog[i * 3], fog[i * 3 + 1], fog[i * 3 + 2]);
    const vec3 x = rad * exposure;
    vec3 t;
    if (tonemapper == 0) {
        t = x / (x + splat3(1.0f));
    } else {
        const float a = 1.8f, b = 1.4f, c = 0.5f, d = 1.5f;
        t = (x * (a * x + splat3(b))) / (x * (a * x + splat3(c)) + splat3(d));
    }
    out[i * 3] = (uint8_t)srgb_byte(t.x);
    out[i * 3 + 1] = (uint8_t)srgb_byte(t.y);
    out[i * 3 + 2] = (uint8_t)srgb_byte(t.z);
}

}  // namespace

// Scratch of sr_compose_device_noalloc: the filtered radiance and the filtered fog radiance (two
// float RGB images) of up to kNoAllocMaxPixels pixels, in a device array of its own so the entry
// point allocates nothing. Host code uses the address hipGetSymbolAddress returns per device (the
// array's name is only a host-side handle); every call on one GPU shares that GPU's array.
constexpr uint32_t kNoAllocMaxPixels = 2u << 20;  // 2,097,152 pixels (1920 x 1080 and 2048 x 1024 fit)
static __device__ __align__(256) float scratch[2 * 3 * (size_t)kNoAllocMaxPixels];

// =================================================================================================
// Host side
// =================================================================================================
namespace {

thread_local std::string g_err;

// Per device: the address of `scratch` (hipGetSymbolAddress), resolved once with that device current.
std::mutex g_scratchMu;
std::vector<void*> g_scratchPtr;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace

int sr::sr_set_error(int code, const char* msg) { return fail(code, msg ? msg : ""); }

namespace {

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            (void)hipGetLastError(); /* a later launch check must not see this error */          \
            return fail(SR_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));            \
        }                                                                                    \
    } while (0)

struct HostTexture {
    std::vector<float> rgb;
    uint32_t w, h;
    int32_t filter;
    float tilx, tily;
};
struct HostMesh {
    std::vector<float> tris;  // 36 floats per Triangle
    BvhResult bvh;
    uint32_t bins;
};
struct HostObject {
    sr_transform trans;
    sr_material mat;
    int32_t kind, mesh;
    float shape[6];
};

template <class T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
};

}  // namespace

struct sr_ctx {
    int device = 0;
    int bvhBuilder = SR_BVH_BUILDER_HOST;
    // Segment from which a render finishes every path in one fused launch (k_tail); beyond the
    // bounce count it never does. SHARPRAY_TAIL_SEGMENT overrides the default, kTailAuto
    // (tail_segment: 1080p frame with 3 / 4 / 5 / 6: one sample in flight 240 / 261 / 278 / - ms,
    // 16 in flight 124.6 / 119.0 / 120.7 / 121.3 ms; with grid fog, whose paths scatter and live
    // longer, 4 in flight 29.2 / 23.6 / 22.5 / 22.1 ms and one in flight 34.2 / 30.1 / 29.6 / 30.1 ms).
    uint32_t tailSegment = kTailAuto;
    hipStream_t stream = nullptr;
    std::vector<HostTexture> textures;
    std::vector<HostMesh> meshes;
    std::vector<HostObject> objects;
    int skyKind = 0;
    int32_t skyTex = -1;
    float skyMult = 1.0f, skyYaw = 0.0f;
    sr_sky_procedural skyProc{};
    bool hasFog = false;
    sr_fog fog{};
    uint32_t fogKind = 0;  // 0 height fog, SR_MEDIUM_HOMOGENEOUS, SR_MEDIUM_GRID
    sr_fog_volume volume{};
    std::vector<uint16_t> fogGrid;
    bool built = false;
    BvhResult sceneBvh;
    std::vector<int32_t> meshNodeBase, meshTriBase;
    DevScene dev{};
    // device scene arrays
    DevBuf<float4> dNodes;
    DevBuf<float> dTris, dAttrs, dTexels, dCdf;
    DevBuf<float4> dAlias;
    DevBuf<float4> dSkyTab;  // texture sky: two float4 per texel (k_sky_table)
    DevBuf<uint16_t> dFogGrid;
    // Adjoint: the last forward render's view and parameters, and the backward pass's buffers.
    bool adjReady = false;
    sr_view adjView{};
    sr_render_params adjParams{};
    DevBuf<float> dAdjImage, dAdjGrad, dAdjEnv;
    sr::CnnNet* cnn = nullptr;  // kernel-predicting denoiser (weights loaded by sr_cnn_load)
    struct CnnExample {  // training pair kept in device memory
        DevBuf<float> rad, alb, nrm, dep, target;
        uint32_t w = 0, h = 0;
    };
    std::vector<CnnExample*> cnnExamples;
    DevBuf<float> dCnnLoss;
    DevBuf<int32_t> dSlotTri, dSceneItems;
    DevBuf<DevObject> dObjects;
    DevBuf<DevTexture> dTextures;
    // frame
    DevBuf<float4> dRayO, dRayD, dHitA;
    DevBuf<int32_t> dHitSlot;
    DevBuf<PathHot> dPaths;
    DevBuf<PathRad> dPrad;
    DevBuf<PathCold> dCold;
    DevBuf<float4> dPrim, dPrimB;  // one each per path slot
    DevBuf<float> dPrimV;
    DevBuf<ShadowRec> dShadows;
    DevBuf<PixAcc> dAcc;
    DevBuf<uint32_t> dQueue0, dQueue1, dCounts;
    DevBuf<unsigned long long> dCtr;
    DevBuf<float> dRad, dRadFog, dNormal, dUv, dDepth, dVar, dAlbedo;
    DevBuf<uint32_t> dSamples;
    // Progress snapshots of adaptive renders: after each sample, every slice copies its queue-0
    // count and the finished-block count into pinned memory and records an event; the driver reads
    // a snapshot kProgRing samples later (no stall), stops enqueueing a band once no pixel of it is
    // active and publishes the finished blocks to Tick.
    DevBuf<uint32_t> dBlockPixels, dProgress;
    uint32_t* progHost = nullptr;
    size_t progHostWords = 0;
    std::vector<hipEvent_t> progEv;
    bool rendering = false;
    bool countersOn = false;
    sr_render_params params{};
    ViewD view{};
    // Background driver (the analogue of the reference's worker threads, renderer.cs:88-122):
    // enqueues the iterations band by band; Tick waits on blocksDone, the count of bs x bs blocks
    // whose pixels are all final (renderer.cs:98-122).
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;
    std::vector<uint32_t> bandEnd;  // last row (exclusive) per band
    uint32_t blocksTotal = 0;
    uint32_t blocksDone = 0;     // published by the driver
    uint32_t ticksConsumed = 0;  // one Tick return per finished block, like the reference's semaphore
    bool workerDone = true;
    int workerStatus = SR_OK;
    std::string workerError;
    uint32_t iterations = 0;
    hipEvent_t evStart = nullptr, evEnd = nullptr;
    int gridCUs = 256;
    // Concurrent row slices of every band: each slice has its own queues, path records and
    // stream, so one slice's launch tails and fused tail run beside the others' work (one 1080p
    // frame: 252 ms as one slice, 237 ms as 2). The default keeps one slice, so
    // every kernel launch has the GPU to itself and its timing measures that kernel alone.
    // SHARPRAY_SLICES or sr_set_slices override it.
    uint32_t slices = 1;
    // Samples of a pixel in flight at once when samples are independent (counter RNG policy, a
    // fixed sample count, counters off): every launch then holds that many samples of every pixel,
    // so the fixed cost of each launch's slowest waves is paid once per batch instead of once per
    // sample. Results are identical for any value. SHARPRAY_IN_FLIGHT or
    // sr_set_samples_in_flight override it.
    uint32_t inFlight = kMaxInFlight;
    uint32_t batch = 1;                     // samples in flight of the current render
    uint32_t sliceCount = 1;                // slices of the current render
    uint32_t sliceRows = 0, sliceSlots = 0;  // rows (multiple of 8) and slots per slice
    std::vector<hipStream_t> sliceStreams;   // slices 1.. (slice 0 runs on `stream`)
    // Kernel timing per slice: one event after every launch; kernel i spans events i-1..i of its
    // slice.
    struct ProfLane {
        std::vector<hipEvent_t> ev;
        std::vector<int> kind;  // 0 extend, 1 shade, 2 shadow, 3 other
        size_t used = 0;
    };
    bool profileOn = false;
    std::vector<ProfLane> prof;
    // Counters.Type book (counters.cs:24-79) over the context's work; guarded by mu (the render
    // driver folds into it from its thread).
    int64_t book[SR_COUNTER_TYPE_COUNT];
    // Filter counters on the device (kFilterWords words), folded into the book at sync points.
    DevBuf<unsigned long long> dFilterCtr;
    bool filterCtrPending = false;
    std::chrono::steady_clock::time_point renderStart;

    void join() {
        if (worker.joinable()) worker.join();
    }

    ~sr_ctx() {
        join();
        for (ProfLane& l : prof)
            for (auto e : l.ev) (void)hipEventDestroy(e);
        for (hipStream_t st : sliceStreams) (void)hipStreamDestroy(st);
        if (progHost) (void)hipHostFree(progHost);
        for (auto e : progEv) (void)hipEventDestroy(e);
        dBlockPixels.release(); dProgress.release(); dFilterCtr.release();
        if (evStart) (void)hipEventDestroy(evStart);
        if (evEnd) (void)hipEventDestroy(evEnd);
        dNodes.release(); dTris.release(); dAttrs.release(); dTexels.release(); dCdf.release(); dAlias.release(); dSkyTab.release(); dFogGrid.release();
        dSlotTri.release(); dSceneItems.release(); dObjects.release(); dTextures.release();
        dAdjImage.release(); dAdjGrad.release(); dAdjEnv.release();
        delete cnn;
        cnn = nullptr;
        for (CnnExample* e : cnnExamples) {
            e->rad.release(); e->alb.release(); e->nrm.release(); e->dep.release(); e->target.release();
            delete e;
        }
        cnnExamples.clear();
        dCnnLoss.release();
        dRayO.release(); dRayD.release(); dHitA.release(); dHitSlot.release();
        dPaths.release(); dPrad.release(); dCold.release(); dPrim.release(); dPrimB.release(); dPrimV.release(); dShadows.release(); dAcc.release();
        dQueue0.release(); dQueue1.release(); dCounts.release(); dCtr.release();
        dRad.release(); dRadFog.release(); dNormal.release(); dUv.release(); dDepth.release(); dVar.release();
        dAlbedo.release();
        dSamples.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <class T>
int upload(DevBuf<T>& b, const T* src, size_t n, hipStream_t s) {
    b.release();
    if (n == 0) return SR_OK;
    HIP_TRY(hipMalloc((void**)&b.ptr, n * sizeof(T)));
    b.n = n;
    HIP_TRY(hipMemcpyAsync(b.ptr, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    return SR_OK;
}
template <class T>
int ensure(DevBuf<T>& b, size_t n) {
    if (b.n >= n && b.ptr) return SR_OK;
    b.release();
    HIP_TRY(hipMalloc((void**)&b.ptr, std::max<size_t>(n, 1) * sizeof(T)));
    b.n = n;
    return SR_OK;
}

// Tabulates a texture sky's light samples (k_sky_table) once its texels are on the device and S
// points at them. SHARPRAY_SKY_TABLE=0, or no memory for it, leaves the table out: sky_light_rand
// then evaluates the same values per sample.
int build_sky_table(sr_ctx* ctx, DevScene& S, hipStream_t s) {
    S.skyTab = nullptr;
    const char* env = getenv("SHARPRAY_SKY_TABLE");
    if (S.sky.kind != 1 || S.sky.total <= 0.0f || (env && strcmp(env, "0") == 0)) {
        ctx->dSkyTab.release();
        return SR_OK;
    }
    const size_t n = (size_t)S.sky.w * (size_t)S.sky.h;
    if (ensure(ctx->dSkyTab, 2 * n) != SR_OK) {
        ctx->dSkyTab.release();
        return SR_OK;
    }
    k_sky_table<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(S, ctx->dSkyTab.ptr);
    HIP_TRY(hipGetLastError());
    S.skyTab = ctx->dSkyTab.ptr;
    return SR_OK;
}

// ---- Counters.Type book ---------------------------------------------------------------------------
// Reductions of counters.cs:228-236: max for SampleVarianceMax, SampleCountMax and the Denoise*Max*
// entries, min for SampleCountMin, sum for the rest; initial values counters.cs:221-226.
int book_reduction(int k) {  // 0 sum, 1 min, 2 max
    switch (k) {
        case SR_CT_SAMPLE_COUNT_MIN: return 1;
        case SR_CT_SAMPLE_VARIANCE_MAX:
        case SR_CT_SAMPLE_COUNT_MAX:
        case SR_CT_DENOISE_WEIGHT_MAX:
        case SR_CT_DENOISE_MAX_LUM:
        case SR_CT_DENOISE_MAX_LUM_BOOST: return 2;
        default: return 0;
    }
}
void book_reset(sr_ctx* c) {
    for (int k = 0; k < SR_COUNTER_TYPE_COUNT; ++k) {
        const int r = book_reduction(k);
        c->book[k] = r == 1 ? INT64_MAX : (r == 2 ? INT64_MIN : 0);
    }
}
// Counters.Bump / BumpMin / BumpMax by the entry's reduction; the caller holds c->mu.
void book_bump(sr_ctx* c, int k, int64_t v) {
    const int r = book_reduction(k);
    if (r == 1) c->book[k] = std::min(c->book[k], v);
    else if (r == 2) c->book[k] = std::max(c->book[k], v);
    else c->book[k] += v;
}
void book_bump_locked(sr_ctx* c, int k, int64_t v) {
    std::lock_guard<std::mutex> lk(c->mu);
    book_bump(c, k, v);
}
int64_t micros_since(std::chrono::steady_clock::time_point t0) {
    return (int64_t)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
}
// An unsigned max key of the device counters (biased (long)(v * 1e4)) back to the signed value;
// 0 = never bumped.
bool decimal_from_key(unsigned long long key, int64_t& v) {
    if (!key) return false;
    v = (int64_t)(key ^ 0x8000000000000000ull);
    return true;
}

}  // namespace

void sr::sr_count_time(sr_ctx* ctx, int entry, int64_t micros) {
    if (ctx && ctx->countersOn) book_bump_locked(ctx, entry, micros);
}

namespace {

quat mkq(const float* r) { quat q; q.x = r[0]; q.y = r[1]; q.z = r[2]; q.w = r[3]; return q; }

// Quat.AngleAxis (math.cs:621-626).
quat angle_axis(float angle, vec3 axis) {
    const vec3 v = axis * sinf(angle * 0.5f);
    quat q; q.x = v.x; q.y = v.y; q.z = v.z; q.w = cosf(angle * 0.5f);
    return q;
}
// Quat.Look = FromMat4(Mat4.RotateLook(forward, upRef)) (math.cs:631-659, 742-758).
quat quat_look(vec3 fwdIn, vec3 upRef) {
    vec3 c0 = mk3(1, 0, 0), c1 = mk3(0, 1, 0), c2 = mk3(0, 0, 1);
    if (!(len_sqr(fwdIn) <= 1e-6f) && !(len_sqr(upRef) <= 1e-6f)) {
        const vec3 f = normalized(fwdIn);
        vec3 rt = cross3(upRef, f);
        rt = len_sqr(rt) > 1e-6f ? normalized(rt) : mk3(1, 0, 0);
        c0 = rt; c1 = cross3(f, rt); c2 = f;
    }
    quat q;
    const float tr = c0.x + c1.y + c2.z;
    if (tr > 1e-6f) {
        const float s = sqrtf(tr + 1.0f) * 2.0f;
        q.x = (c1.z - c2.y) / s; q.y = (c2.x - c0.z) / s; q.z = (c0.y - c1.x) / s; q.w = s * 0.25f;
    } else if (c0.x > c1.y && c0.x > c2.z) {
        const float s = sqrtf(1.0f + c0.x - c1.y - c2.z) * 2.0f;
        q.x = s * 0.25f; q.y = (c1.x + c0.y) / s; q.z = (c2.x + c0.z) / s; q.w = (c1.z - c2.y) / s;
    } else if (c1.y > c2.z) {
        const float s = sqrtf(1.0f + c1.y - c0.x - c2.z) * 2.0f;
        q.x = (c1.x + c0.y) / s; q.y = s * 0.25f; q.z = (c2.y + c1.z) / s; q.w = (c2.x - c0.z) / s;
    } else {
        const float s = sqrtf(1.0f + c2.z - c0.x - c1.y) * 2.0f;
        q.x = (c2.x + c0.z) / s; q.y = (c2.y + c1.z) / s; q.z = s * 0.25f; q.w = (c0.y - c1.x) / s;
    }
    return q;
}

// Object world bounds: AABox.Transform then Box.Bounds (scene.cs:221-222; math.cs:974, 1101-1114).
box3 object_world_bounds(const box3& lb, const sr_transform& t) {
    const quat rq = mkq(t.rot);
    const vec3 sc = mk3(t.scale[0], t.scale[1], t.scale[2]);
    const vec3 center = mk3(t.pos[0], t.pos[1], t.pos[2]) + rotate(rq, box_mid(lb) * sc);
    const vec3 size = box_ext(lb) * sc;
    box3 local;
    local.mn = center - size * 0.5f;
    local.mx = center + size * 0.5f;
    const vec3 c = box_mid(local);
    const vec3 h = box_ext(local) * 0.5f;
    const vec3 ax = rotate(rq, mk3(h.x, 0.0f, 0.0f));
    const vec3 ay = rotate(rq, mk3(0.0f, h.y, 0.0f));
    const vec3 az = rotate(rq, mk3(0.0f, 0.0f, h.z));
    const vec3 ext = mk3(fabsf(ax.x) + fabsf(ay.x) + fabsf(az.x), fabsf(ax.y) + fabsf(ay.y) + fabsf(az.y),
                         fabsf(ax.z) + fabsf(ay.z) + fabsf(az.z));
    box3 r;
    r.mn = c - (ext * 2.0f) * 0.5f;
    r.mx = c + (ext * 2.0f) * 0.5f;
    return r;
}

// MathF.Pow through libm (kept out of line so the compiler cannot fold pow(x, 2) into x * x).
__attribute__((noinline)) float libm_powf(float x, float y) {
    volatile float vy = y;
    return powf(x, vy);
}

// Appends a BVH to the node pool: one pad node, then the nodes, starting at an even index.
int append_bvh(std::vector<float4>& pool, const BvhResult& b, int32_t& base) {
    if (pool.size() % 4 != 0) pool.resize(pool.size() + 2, make_float4(0, 0, 0, 0));
    base = (int32_t)(pool.size() / 2);
    pool.resize(pool.size() + 2, make_float4(0, 0, 0, 0));
    for (int i = 0; i < b.used; ++i) {
        const RefNode& nd = b.nodes[i];
        if ((uint32_t)nd.child > kMaxChild || nd.count > 255 || nd.count < 0)
            return fail(SR_ERR_LIMIT, "BVH node exceeds the device layout (child < 2^24, leaf size < 256)");
        // The traversal's stack word (count << 24 | child) in place of the child index.
        const uint32_t word = ((uint32_t)nd.count << 24) | (uint32_t)nd.child;
        const int32_t cn = nd.count;
        pool.push_back(make_float4(nd.bounds.mn.x, nd.bounds.mn.y, nd.bounds.mn.z, nd.bounds.mx.x));
        pool.push_back(make_float4(nd.bounds.mx.y, nd.bounds.mx.z, __builtin_bit_cast(float, word), __builtin_bit_cast(float, cn)));
    }
    return SR_OK;
}

int grid_for(const sr_ctx* c, uint32_t maxItems) {
    const uint32_t want = (maxItems + kBlock - 1) / kBlock;
    const uint32_t cap = (uint32_t)c->gridCUs * 8u;
    return (int)std::max(1u, std::min(want, cap));
}

}  // namespace

namespace {
// Sky tables (scene.cs:389-404 / 341-352 / 290-298): the texture sky's row/column CDFs, alias
// table and pdf scale, or the procedural sky's constants.
void build_sky(sr_ctx* ctx, DevSky& K, std::vector<float>& cdf, std::vector<float4>& aliasTab) {
    K.kind = ctx->skyKind;
    if (ctx->skyKind == 1) {
        const HostTexture& t = ctx->textures[ctx->skyTex];
        const int W = (int)t.w, H = (int)t.h;
        K.tex = ctx->skyTex;
        K.multiplier = ctx->skyMult;
        K.rot = angle_axis(ctx->skyYaw, up3());
        K.rotInv = conj(K.rot);
        K.w = W; K.h = H;
        K.cdfOff = 0;
        cdf.assign((size_t)(H + 1) + (size_t)H * (W + 1), 0.0f);
        const size_t n = (size_t)W * H;
        std::vector<float> rowW(H), colW(W), allW(n);
        for (int y = 0; y < H; ++y) {
            const float s = sinf(kPi * ((float)y + 0.5f) / (float)H);
            for (int x = 0; x < W; ++x) {
                const float* c = t.rgb.data() + ((size_t)y * W + x) * 3;
                colW[x] = luminance(mk3(c[0], c[1], c[2])) * s;
                allW[(size_t)y * W + x] = colW[x];
            }
            rowW[y] = build_cdf1(colW.data(), W, cdf.data() + (H + 1) + (size_t)y * (W + 1));
        }
        K.total = build_cdf1(rowW.data(), H, cdf.data());
        K.pdfScale = (float)(W * H) / (K.total * 2.0f * kPi * kPi);
        // Bucket bounds of every CDF (cdf_sample_guided), kept only when no table decreases (weights
        // are non-negative unless a texture holds negative values). SHARPRAY_CDF_GUIDE=0 leaves
        // them out (the full searches pick the same texels).
        const char* guideEnv = getenv("SHARPRAY_CDF_GUIDE");
        bool monotone = !(guideEnv && strcmp(guideEnv, "0") == 0);
        for (size_t i = 1; i < cdf.size() && monotone; ++i) {
            const bool rowStart = i > (size_t)H && (i - (size_t)(H + 1)) % (size_t)(W + 1) == 0;
            if (i != (size_t)(H + 1) && !rowStart && !(cdf[i] >= cdf[i - 1])) monotone = false;
        }
        K.guideOff = -1;
        if (monotone) {
            auto first_false = [](const float* c, uint32_t n, float x) {
                uint32_t lo = 0, hi = n;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (c[mid + 1] <= x) lo = mid + 1;
                    else hi = mid;
                }
                return lo;
            };
            auto put_records = [&](const float* c, uint32_t n, uint32_t m, std::vector<float>& out) {
                uint32_t lo = first_false(c, n, 0.0f);
                for (uint32_t j = 0; j < m; ++j) {
                    const uint32_t hi = j + 1 < m ? first_false(c, n, (float)(j + 1) / (float)m) : n;
                    out.push_back(__builtin_bit_cast(float, lo));
                    out.push_back(__builtin_bit_cast(float, hi));
                    out.push_back(lo + 1 <= n ? c[lo + 1] : 0.0f);
                    out.push_back(lo + 2 <= n ? c[lo + 2] : 0.0f);
                    lo = hi;
                }
            };
            while (cdf.size() % 4) cdf.push_back(0.0f);  // float4 records
            std::vector<float> guide;
            guide.reserve(4 * ((size_t)kGuideMarg + (size_t)H * kGuideCond));
            put_records(cdf.data(), (uint32_t)H, kGuideMarg, guide);
            for (int y = 0; y < H; ++y)
                put_records(cdf.data() + (H + 1) + (size_t)y * (W + 1), (uint32_t)W, kGuideCond, guide);
            K.guideOff = (int32_t)cdf.size();
            cdf.insert(cdf.end(), guide.begin(), guide.end());
        }
        // Alias table over the same weights; each entry carries the pdf LightDirRand returns for
        // either outcome (luminance * pdfScale, scene.cs:414).
        std::vector<float> prob(n);
        std::vector<int32_t> alias(n);
        build_alias(allW.data(), n, prob.data(), alias.data());
        aliasTab.resize(n);
        auto texel_pdf = [&](size_t i) {
            const float* c = t.rgb.data() + i * 3;
            return luminance(mk3(c[0], c[1], c[2])) * K.pdfScale;
        };
        for (size_t i = 0; i < n; ++i) {
            float4 e;
            e.x = prob[i];
            memcpy(&e.y, &alias[i], sizeof(float));
            e.z = texel_pdf(i);
            e.w = texel_pdf((size_t)alias[i]);
            aliasTab[i] = e;
        }
    } else {
        const sr_sky_procedural& p = ctx->skyProc;
        K.top = mk3(p.radiance_top[0], p.radiance_top[1], p.radiance_top[2]);
        K.middle = mk3(p.radiance_middle[0], p.radiance_middle[1], p.radiance_middle[2]);
        K.bottom = mk3(p.radiance_bottom[0], p.radiance_bottom[1], p.radiance_bottom[2]);
        K.sunDir = mk3(p.sun_dir[0], p.sun_dir[1], p.sun_dir[2]);
        K.sunRad = mk3(p.sun_radiance[0], p.sun_radiance[1], p.sun_radiance[2]);
        K.sunCos = cosf(p.sun_angle);
        K.sunRot = quat_look(K.sunDir, up3());
        K.sunPdfDenom = kPi * (1.0f - K.sunCos);
    }
}

// Device object records (transform, material, BVH bases; Object ctor scene.cs:214-223).
void build_objects(sr_ctx* ctx, std::vector<DevObject>& dobj, int32_t& anyTransparent) {
    dobj.assign(ctx->objects.size(), DevObject{});
    anyTransparent = 0;
    for (size_t i = 0; i < ctx->objects.size(); ++i) {
        const HostObject& o = ctx->objects[i];
        DevObject& d = dobj[i];
        memset(&d, 0, sizeof(d));
        d.rot = mkq(o.trans.rot);
        d.pos = mk3(o.trans.pos[0], o.trans.pos[1], o.trans.pos[2]);
        d.scale = mk3(o.trans.scale[0], o.trans.scale[1], o.trans.scale[2]);
        d.kind = o.kind;
        if (o.kind == 0) {
            d.nodeBase = ctx->meshes[o.mesh].bvh.used > 0 ? ctx->meshNodeBase[o.mesh] : -1;
            d.triBase = ctx->meshTriBase[o.mesh];
        }
        d.texColor = o.mat.color_texture; d.texRough = o.mat.roughness_texture;
        d.texMetal = o.mat.metallic_texture; d.texNormal = o.mat.normal_texture;
        d.albedo = mk3(o.mat.albedo[0], o.mat.albedo[1], o.mat.albedo[2]);
        d.roughness = o.mat.roughness; d.metallic = o.mat.metallic;
        d.transparency = o.mat.transparency; d.ior = o.mat.ior;
        d.radiance = mk3(o.mat.radiance[0], o.mat.radiance[1], o.mat.radiance[2]);
        d.f0 = libm_powf((o.mat.ior - 1.0f) / (o.mat.ior + 1.0f), 2.0f);
        memcpy(d.shape, o.shape, sizeof(d.shape));
        if (o.mat.transparency > 0.0f) anyTransparent = 1;
        // Objects with bit-identical transforms share the object-space ray of a traversal.
        d.xform = (int32_t)i;
        for (size_t k = 0; k < i; ++k)
            if (memcmp(&dobj[k].rot, &d.rot, sizeof(quat)) == 0 && memcmp(&dobj[k].pos, &d.pos, sizeof(vec3)) == 0 &&
                memcmp(&dobj[k].scale, &d.scale, sizeof(vec3)) == 0) {
                d.xform = dobj[k].xform;
                break;
            }
    }
}

}  // namespace

extern "C" {

const char* sr_last_error(void) { return g_err.c_str(); }

int sr_create(int32_t device, sr_ctx** out) {
    if (!out) return fail(SR_ERR_INVALID, "out is null");
    *out = nullptr;
    if (device == -1) {  // host-only context: scene preparation and BVH export, no rendering
        sr_ctx* c = new sr_ctx();
        c->device = -1;
        book_reset(c);
        *out = c;
        return SR_OK;
    }
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(SR_ERR_INVALID, "no such HIP device");
    HIP_TRY(hipSetDevice(device));
    {
        // The allocation-free compose's scratch: resolve this device's copy of the array once.
        std::lock_guard<std::mutex> lk(g_scratchMu);
        if (g_scratchPtr.size() < (size_t)n) g_scratchPtr.resize(n, nullptr);
        if (!g_scratchPtr[device]) {
            void* ptr = nullptr;
            HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            if (!ptr) return fail(SR_ERR_HIP, "scratch: no device address");
            g_scratchPtr[device] = ptr;
        }
    }
    sr_ctx* c = new sr_ctx();
    c->device = device;
    book_reset(c);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail(SR_ERR_HIP, hipGetErrorString(e)); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->gridCUs = prop.multiProcessorCount;
    if (const char* t = getenv("SHARPRAY_TAIL_SEGMENT")) c->tailSegment = (uint32_t)strtoul(t, nullptr, 10);
    if (const char* t = getenv("SHARPRAY_SLICES")) c->slices = std::max(1u, (uint32_t)strtoul(t, nullptr, 10));
    if (const char* t = getenv("SHARPRAY_IN_FLIGHT"))
        c->inFlight = std::min(kMaxInFlight, std::max(1u, (uint32_t)strtoul(t, nullptr, 10)));
    *out = c;
    return SR_OK;
}

int sr_destroy(sr_ctx* ctx) {
    if (!ctx) return SR_OK;
    ctx->join();
    if (ctx->device >= 0) {
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    }
    delete ctx;
    return SR_OK;
}

int sr_texture_create(sr_ctx* ctx, const float* rgb, uint32_t w, uint32_t h, int32_t filter, float tx, float ty,
                      int32_t* out_id) {
    if (!ctx || !rgb || !out_id || w == 0 || h == 0) return fail(SR_ERR_INVALID, "texture: bad arguments");
    if (filter != 0 && filter != 1) return fail(SR_ERR_INVALID, "texture: filter must be 0 or 1");
    HostTexture t;
    t.rgb.assign(rgb, rgb + (size_t)w * h * 3);
    t.w = w; t.h = h; t.filter = filter; t.tilx = tx; t.tily = ty;
    ctx->textures.push_back(std::move(t));
    *out_id = (int32_t)ctx->textures.size() - 1;
    return SR_OK;
}

int sr_mesh_create(sr_ctx* ctx, const float* tris, uint32_t count, uint32_t bins, int32_t* out_id) {
    if (!ctx || !out_id || (count && !tris)) return fail(SR_ERR_INVALID, "mesh: bad arguments");
    if (bins < 2) return fail(SR_ERR_INVALID, "mesh: split_bins must be >= 2");
    if (ctx->built) return fail(SR_ERR_STATE, "Scene already built");
    if (count > kMaxChild) return fail(SR_ERR_LIMIT, "mesh: more than 2^24-1 triangles");
    HostMesh m;
    m.tris.assign(tris, tris + (size_t)count * 36);
    m.bins = bins;
    const auto t0 = std::chrono::steady_clock::now();  // TimeMeshBvhBuild (mesh.cs:38-42)
    if (ctx->bvhBuilder == SR_BVH_BUILDER_DEVICE && count > 0) {
        if (bins > 64) return fail(SR_ERR_LIMIT, "mesh: the device BVH builder takes at most 64 bins");
        HIP_TRY(hipSetDevice(ctx->device));
        std::vector<float> lean((size_t)count * 9);
        for (uint32_t i = 0; i < count; ++i) memcpy(&lean[(size_t)i * 9], tris + (size_t)i * 36, 36);
        float* dLean = nullptr;
        HIP_TRY(hipMalloc((void**)&dLean, lean.size() * 4));
        hipError_t e = hipMemcpy(dLean, lean.data(), lean.size() * 4, hipMemcpyHostToDevice);
        std::string err;
        const bool ok = e == hipSuccess && device_build_bvh(dLean, (int)count, (int)bins, 1.0f, 0.5f, m.bvh,
                                                            ctx->stream, err);
        (void)hipFree(dLean);
        if (e != hipSuccess) return fail(SR_ERR_HIP, hipGetErrorString(e));
        if (!ok) return fail(SR_ERR_HIP, err);
    } else {
        std::vector<box3> bb(count);
        for (uint32_t i = 0; i < count; ++i) bb[i] = lean_bounds(tris + (size_t)i * 36);
        m.bvh = build_bvh(bb.data(), (int)count, (int)bins, 1.0f, 0.5f);
    }
    if (m.bvh.degenerate) return fail(SR_ERR_BUILD, "mesh BVH: a split produced an empty child (bvh.cs:448)");
    if (ctx->countersOn) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        book_bump(ctx, SR_CT_TIME_MESH_BVH_BUILD, micros_since(t0));
        book_bump(ctx, SR_CT_MESH_TRIANGLE, (int64_t)count);
    }
    ctx->meshes.push_back(std::move(m));
    *out_id = (int32_t)ctx->meshes.size() - 1;
    return SR_OK;
}

int sr_scene_set_sky_texture(sr_ctx* ctx, int32_t tex, float mult, float yaw) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    if (ctx->built) return fail(SR_ERR_STATE, "Scene already built");
    if (tex < 0 || tex >= (int32_t)ctx->textures.size()) return fail(SR_ERR_INVALID, "sky: bad texture id");
    ctx->skyKind = 1; ctx->skyTex = tex; ctx->skyMult = mult; ctx->skyYaw = yaw;
    return SR_OK;
}

int sr_scene_set_sky_procedural(sr_ctx* ctx, const sr_sky_procedural* sky) {
    if (!ctx || !sky) return fail(SR_ERR_INVALID, "sky: null");
    if (ctx->built) return fail(SR_ERR_STATE, "Scene already built");
    ctx->skyKind = 2;
    ctx->skyProc = *sky;
    return SR_OK;
}

int sr_scene_set_fog(sr_ctx* ctx, const sr_fog* fog) {
    if (!ctx || !fog) return fail(SR_ERR_INVALID, "fog: null");
    if (ctx->built) return fail(SR_ERR_STATE, "Scene already built");
    ctx->hasFog = true;
    ctx->fog = *fog;
    ctx->fogKind = 0;
    return SR_OK;
}

int sr_scene_set_fog_volume(sr_ctx* ctx, const sr_fog_volume* v) {
    if (!ctx || !v) return fail(SR_ERR_INVALID, "fog volume: null");
    if (ctx->built) return fail(SR_ERR_STATE, "Scene already built");
    if (v->kind != SR_MEDIUM_HOMOGENEOUS && v->kind != SR_MEDIUM_GRID) return fail(SR_ERR_INVALID, "fog volume: bad kind");
    if (!(v->sigma >= 0.0f) || !(v->anisotropy > -1.0f && v->anisotropy < 1.0f))
        return fail(SR_ERR_INVALID, "fog volume: need sigma >= 0 and -1 < anisotropy < 1");
    for (int k = 0; k < 3; ++k)
        if (!(v->bounds_min[k] < v->bounds_max[k])) return fail(SR_ERR_INVALID, "fog volume: empty box");
    std::vector<uint16_t> grid;
    if (v->kind == SR_MEDIUM_GRID) {
        if (!v->density || v->res[0] == 0 || v->res[1] == 0 || v->res[2] == 0) return fail(SR_ERR_INVALID, "fog volume: grid required");
        if ((uint64_t)v->res[0] * v->res[1] * v->res[2] > (1ull << 30)) return fail(SR_ERR_LIMIT, "fog volume: grid too large");
        const size_t n = (size_t)v->res[0] * v->res[1] * v->res[2];
        grid.assign(v->density, v->density + n);
        for (uint16_t h : grid) {  // negative, infinite or NaN densities are rejected
            if ((h & 0x7c00u) == 0x7c00u) return fail(SR_ERR_INVALID, "fog volume: non-finite density");
            if ((h & 0x8000u) && (h & 0x7fffu)) return fail(SR_ERR_INVALID, "fog volume: negative density");
        }
    }
    ctx->hasFog = true;
    ctx->fogKind = v->kind;
    ctx->volume = *v;
    ctx->volume.density = nullptr;
    ctx->fogGrid.swap(grid);
    return SR_OK;
}

int sr_scene_add_object(sr_ctx* ctx, const sr_transform* trans, const sr_material* mat, int32_t kind, int32_t mesh,
                        const float* params, int32_t* out_id) {
    if (!ctx || !trans || !mat) return fail(SR_ERR_INVALID, "object: null argument");
    if (ctx->built) return fail(SR_ERR_STATE, "Scene already built");
    if (kind < 0 || kind > 2) return fail(SR_ERR_INVALID, "object: bad shape kind");
    if (kind == 0 && (mesh < 0 || mesh >= (int32_t)ctx->meshes.size())) return fail(SR_ERR_INVALID, "object: bad mesh id");
    if (kind != 0 && !params) return fail(SR_ERR_INVALID, "object: shape params required");
    const int32_t texs[4] = {mat->color_texture, mat->roughness_texture, mat->metallic_texture, mat->normal_texture};
    for (int32_t t : texs)
        if (t < -1 || t >= (int32_t)ctx->textures.size()) return fail(SR_ERR_INVALID, "object: bad texture id");
    HostObject o;
    o.trans = *trans;
    o.mat = *mat;
    o.kind = kind;
    o.mesh = mesh;
    memset(o.shape, 0, sizeof(o.shape));
    if (kind == 1) memcpy(o.shape, params, 4 * sizeof(float));
    if (kind == 2) memcpy(o.shape, params, 6 * sizeof(float));
    ctx->objects.push_back(o);
    if (out_id) *out_id = (int32_t)ctx->objects.size() - 1;
    return SR_OK;
}

int sr_scene_build(sr_ctx* ctx) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    if (ctx->built) return SR_OK;
    if (ctx->skyKind == 0) return fail(SR_ERR_STATE, "Scene sky is required");
    if (ctx->device >= 0) HIP_TRY(hipSetDevice(ctx->device));
    // Scene BVH over object world bounds (scene.cs:528-533).
    std::vector<box3> ob(ctx->objects.size());
    for (size_t i = 0; i < ctx->objects.size(); ++i) {
        const HostObject& o = ctx->objects[i];
        box3 lb;
        if (o.kind == 0) {
            const HostMesh& m = ctx->meshes[o.mesh];
            lb = m.bvh.used > 0 ? m.bvh.nodes[0].bounds : box_empty();
        } else if (o.kind == 1) {
            const vec3 c = mk3(o.shape[0], o.shape[1], o.shape[2]);
            lb.mn = c - splat3(o.shape[3]);
            lb.mx = c + splat3(o.shape[3]);
        } else {
            lb.mn = mk3(o.shape[0], o.shape[1], o.shape[2]);
            lb.mx = mk3(o.shape[3], o.shape[4], o.shape[5]);
        }
        ob[i] = object_world_bounds(lb, o.trans);
    }
    const auto t0 = std::chrono::steady_clock::now();
    ctx->sceneBvh = build_bvh(ob.data(), (int)ob.size(), 32, 1.0f, 10.0f);
    if (ctx->sceneBvh.degenerate) return fail(SR_ERR_BUILD, "scene BVH: a split produced an empty child");
    if (ctx->countersOn) {  // scene.cs:528-540
        const int64_t us = micros_since(t0);
        const BvhStatsHost st = bvh_stats(ctx->sceneBvh, 1.0f, 10.0f);
        std::lock_guard<std::mutex> lk(ctx->mu);
        book_bump(ctx, SR_CT_TIME_SCENE_BVH_BUILD, us);
        book_bump(ctx, SR_CT_SCENE_OBJECT, (int64_t)ctx->objects.size());
        book_bump(ctx, SR_CT_SCENE_BVH_NODES, st.nodeCount);
        book_bump(ctx, SR_CT_SCENE_BVH_DEPTH, st.depthMax);
    }

    // Node pool, triangles in item order, attributes in triangle order.
    std::vector<float4> nodes;
    std::vector<float> tris, attrs;
    std::vector<int32_t> slotTri;
    ctx->meshNodeBase.assign(ctx->meshes.size(), 0);
    ctx->meshTriBase.assign(ctx->meshes.size(), 0);
    for (size_t m = 0; m < ctx->meshes.size(); ++m) {
        const HostMesh& hm = ctx->meshes[m];
        int rc = append_bvh(nodes, hm.bvh, ctx->meshNodeBase[m]);
        if (rc) return rc;
        const int32_t base = (int32_t)slotTri.size();
        ctx->meshTriBase[m] = base;
        const size_t n = hm.bvh.items.size();
        // Triangles and their attributes both in item order: a hit's slot addresses both.
        for (size_t s = 0; s < n; ++s) {
            const int32_t t = hm.bvh.items[s];
            slotTri.push_back(t);
            const float* src = hm.tris.data() + (size_t)t * 36;
            tris.insert(tris.end(), src, src + 9);
            attrs.insert(attrs.end(), src + 9, src + 36);
        }
    }
    int32_t sceneBase = 0;
    int rc = append_bvh(nodes, ctx->sceneBvh, sceneBase);
    if (rc) return rc;

    std::vector<float> texels, cdf;
    std::vector<float4> aliasTab;
    std::vector<DevTexture> dtex;
    for (const HostTexture& t : ctx->textures) {
        DevTexture d;
        memset(&d, 0, sizeof(d));
        d.offset = (int32_t)(texels.size() / 3);
        d.w = (int32_t)t.w; d.h = (int32_t)t.h; d.filter = t.filter; d.tilx = t.tilx; d.tily = t.tily;
        texels.insert(texels.end(), t.rgb.begin(), t.rgb.end());
        dtex.push_back(d);
    }

    DevScene& S = ctx->dev;
    memset(&S, 0, sizeof(S));
    S.sceneBase = sceneBase;
    S.sceneNodes = ctx->sceneBvh.used;
    // Sky (scene.cs:389-404 / 341-352 / 290-298).
    build_sky(ctx, S.sky, cdf, aliasTab);
    S.fog.enabled = ctx->hasFog ? 1 : 0;
    S.fog.kind = (int32_t)ctx->fogKind;
    if (ctx->fogKind == 0) {
        S.fog.density = ctx->fog.density;
        S.fog.color = mk3(ctx->fog.color[0], ctx->fog.color[1], ctx->fog.color[2]);
        S.fog.anisotropy = ctx->fog.anisotropy;
        S.fog.heightFalloff = ctx->fog.height_falloff;
    } else {
        const sr_fog_volume& v = ctx->volume;
        S.fog.density = v.sigma;
        S.fog.color = mk3(v.color[0], v.color[1], v.color[2]);
        S.fog.anisotropy = v.anisotropy;
        S.fog.bmin = mk3(v.bounds_min[0], v.bounds_min[1], v.bounds_min[2]);
        S.fog.bmax = mk3(v.bounds_max[0], v.bounds_max[1], v.bounds_max[2]);
        if (ctx->fogKind == SR_MEDIUM_GRID) {
            S.fog.nx = (int32_t)v.res[0]; S.fog.ny = (int32_t)v.res[1]; S.fog.nz = (int32_t)v.res[2];
            S.fog.cell = mk3((S.fog.bmax.x - S.fog.bmin.x) / (float)S.fog.nx, (S.fog.bmax.y - S.fog.bmin.y) / (float)S.fog.ny,
                             (S.fog.bmax.z - S.fog.bmin.z) / (float)S.fog.nz);
            S.fog.cellInv = mk3(1.0f / S.fog.cell.x, 1.0f / S.fog.cell.y, 1.0f / S.fog.cell.z);
            float mx = 0.0f;
            for (uint16_t h : ctx->fogGrid) mx = std::max(mx, half_to_float_host(h));
            S.fog.majorant = v.sigma * mx;
        }
    }

    std::vector<DevObject> dobj;
    build_objects(ctx, dobj, S.anyTransparent);
    S.packHit = ctx->objects.size() < 128 ? 1 : 0;
    for (const HostMesh& hm : ctx->meshes)
        if (hm.bvh.items.size() >= 0xffffffu) S.packHit = 0;

    if (ctx->device < 0) {  // host-only context
        ctx->built = true;
        return SR_OK;
    }
    hipStream_t s = ctx->stream;
    if ((rc = upload(ctx->dNodes, nodes.data(), nodes.size(), s))) return rc;
    if ((rc = upload(ctx->dTris, tris.data(), tris.size(), s))) return rc;
    if ((rc = upload(ctx->dAttrs, attrs.data(), attrs.size(), s))) return rc;
    if ((rc = upload(ctx->dSlotTri, slotTri.data(), slotTri.size(), s))) return rc;
    if ((rc = upload(ctx->dSceneItems, ctx->sceneBvh.items.data(), ctx->sceneBvh.items.size(), s))) return rc;
    if ((rc = upload(ctx->dObjects, dobj.data(), dobj.size(), s))) return rc;
    if ((rc = upload(ctx->dTexels, texels.data(), texels.size(), s))) return rc;
    if ((rc = upload(ctx->dTextures, dtex.data(), dtex.size(), s))) return rc;
    if ((rc = upload(ctx->dCdf, cdf.data(), cdf.size(), s))) return rc;
    if ((rc = upload(ctx->dAlias, aliasTab.data(), aliasTab.size(), s))) return rc;
    if ((rc = upload(ctx->dFogGrid, ctx->fogGrid.data(), ctx->fogGrid.size(), s))) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    S.nodes = ctx->dNodes.ptr;
    S.tris = ctx->dTris.ptr;
    S.attrs = ctx->dAttrs.ptr;
    S.slotTri = ctx->dSlotTri.ptr;
    S.sceneItems = ctx->dSceneItems.ptr;
    S.objects = ctx->dObjects.ptr;
    S.texels = ctx->dTexels.ptr;
    S.textures = ctx->dTextures.ptr;
    S.cdf = ctx->dCdf.ptr;
    S.alias = ctx->dAlias.ptr;
    S.fog.grid = ctx->dFogGrid.ptr;
    if ((rc = build_sky_table(ctx, S, s))) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    ctx->built = true;
    return SR_OK;
}

}  // extern "C"

namespace {

ViewD make_view(const sr_view* v, uint32_t W, uint32_t H) {
    ViewD d;
    d.pos = mk3(v->pos[0], v->pos[1], v->pos[2]);
    d.rot = mkq(v->rot);
    d.nearPlane = v->near_plane;
    d.tanFov = tanf(v->fov * 0.5f);  // MathF.Tan(Fov * 0.5f) (math.cs:1623)
    d.aspect = (float)W / (float)H;  // renderer.cs:62
    return d;
}

// The frame state of rows [y0, y0 + rows) run as slice `slice` (its own part of every buffer).
// Segment from which the fused tail launch finishes every path (SHARPRAY_TAIL_SEGMENT /
// sr_set_tail_segment, else 3 with one or a few samples in flight and 4 with 8 or more; in scenes
// with fog or a medium 5 with one or two in flight and 6 with more).
uint32_t tail_segment(const sr_ctx* c, uint32_t batch) {
    if (c->tailSegment != kTailAuto) return c->tailSegment;
    if (c->dev.fog.enabled) return batch >= 4 ? 6u : 5u;
    return batch >= 8 ? 4u : 3u;
}

Frame frame_for(sr_ctx* c, uint32_t y0, uint32_t rows, uint32_t slice = 0) {
    Frame F;
    memset(&F, 0, sizeof(F));
    const size_t o = (size_t)slice * c->sliceSlots;             // pixel slots
    const size_t pathSlots = (size_t)c->sliceSlots * c->batch;  // path slots per slice
    const size_t op = (size_t)slice * pathSlots;
    F.rayO = c->dRayO.ptr + op;
    F.rayD = c->dRayD.ptr + op;
    F.hitA = c->dHitA.ptr + op;
    F.hitSlot = c->dHitSlot.ptr + op;
    F.paths = c->dPaths.ptr + op;
    F.prad = c->dPrad.ptr + op;
    F.cold = c->dCold.ptr + op;
    F.prim = c->dPrim.ptr + op;
    F.primB = c->dPrimB.ptr + op;
    F.primV = c->dPrimV.ptr + op;
    F.shadows[0] = c->dShadows.ptr + 2 * op;
    F.shadows[1] = c->dShadows.ptr + 2 * op + pathSlots;
    F.acc = c->dAcc.ptr + o;
    F.queue[0] = c->dQueue0.ptr + op;
    F.queue[1] = c->dQueue1.ptr + op;
    F.counts = c->dCounts.ptr + (size_t)slice * kCounterWords;
    F.ctr = c->countersOn ? c->dCtr.ptr : nullptr;
    F.outRad = c->dRad.ptr;
    F.outRadFog = c->dRadFog.ptr;
    F.outNormal = c->dNormal.ptr;
    F.outUv = c->dUv.ptr;
    F.outDepth = c->dDepth.ptr;
    F.outAlbedo = c->dAlbedo.ptr;
    F.outSamples = c->dSamples.ptr;
    F.outVar = c->dVar.ptr;
    F.W = c->params.width;
    F.H = c->params.height;
    F.y0 = y0;
    F.rows = rows;
    F.tilesX = (F.W + 7) / 8;
    F.nSlots = F.tilesX * ((rows + 7) / 8) * 64;
    F.batch = c->batch;
    F.nPaths = F.nSlots * F.batch;
    // A batch's camera rays can be queued in slot order when every pixel runs the batch's samples
    // (more than one in flight) and a segment-0 trace and shade consume them (not the fused tail).
    F.camSlots = c->batch > 1 && tail_segment(c, c->batch) > 0 ? 1u : 0u;
    F.sample0 = 0;
    F.fog = c->dev.fog.enabled ? 1u : 0u;
    F.minS = c->params.min_samples;
    F.maxS = c->params.max_samples;
    F.bounces = c->params.bounces;
    F.seed = c->params.rng_seed;
    F.threshold = c->params.variance_threshold;
    F.clamp = c->params.indirect_clamp;
    F.view = c->view;
    F.blockPixels = c->dBlockPixels.ptr;
    F.blocksDone = c->dProgress.ptr;
    F.bs = std::max(1u, c->params.block_size);
    F.bx = (F.W + F.bs - 1) / F.bs;
    return F;
}

// Frame buffers for `slices` slices of `sliceSlots` pixel slots each, with c->batch path slots per
// pixel slot.
int alloc_frame(sr_ctx* c, uint32_t sliceSlots, uint32_t slices) {
    const size_t px = (size_t)c->params.width * c->params.height;
    const size_t maxSlots = (size_t)sliceSlots * slices;
    const size_t maxPaths = maxSlots * c->batch;
    c->sliceSlots = sliceSlots;
    int rc;
    if ((rc = ensure(c->dRayO, maxPaths)) || (rc = ensure(c->dRayD, maxPaths)) || (rc = ensure(c->dHitA, maxPaths)) ||
        (rc = ensure(c->dHitSlot, maxPaths)) || (rc = ensure(c->dPaths, maxPaths)) || (rc = ensure(c->dPrad, maxPaths)) || (rc = ensure(c->dShadows, maxPaths * 2)) ||
        (rc = ensure(c->dPrim, maxPaths)) || (rc = ensure(c->dPrimB, maxPaths)) || (rc = ensure(c->dPrimV, maxPaths)) || (rc = ensure(c->dCold, maxPaths)) ||
        (rc = ensure(c->dAcc, maxSlots)) || (rc = ensure(c->dQueue0, maxPaths)) || (rc = ensure(c->dQueue1, maxPaths)) ||
        (rc = ensure(c->dCounts, (size_t)kCounterWords * slices)) ||
        (rc = ensure(c->dCtr, (size_t)kCtrWords * kCtrSpread)) || (rc = ensure(c->dRad, px * 3)) || (rc = ensure(c->dRadFog, px * 3)) ||
        (rc = ensure(c->dNormal, px * 3)) || (rc = ensure(c->dUv, px * 2)) || (rc = ensure(c->dDepth, px)) || (rc = ensure(c->dAlbedo, px * 3)) ||
        (rc = ensure(c->dVar, px)) || (rc = ensure(c->dSamples, px)) ||
        (rc = ensure(c->dBlockPixels, c->blocksTotal)) || (rc = ensure(c->dProgress, 1)))
        return rc;
    // Pinned progress snapshots: kProgRing samples x slices x (queue-0 count, finished blocks).
    const size_t words = (size_t)kProgRing * slices * 2;
    if (c->progHostWords < words) {
        if (c->progHost) HIP_TRY(hipHostFree(c->progHost));
        c->progHost = nullptr;
        c->progHostWords = 0;
        HIP_TRY(hipHostMalloc((void**)&c->progHost, words * sizeof(uint32_t), hipHostMallocDefault));
        c->progHostWords = words;
    }
    while (c->progEv.size() < (size_t)kProgRing * slices) {
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->progEv.push_back(e);
    }
    return SR_OK;
}


hipStream_t slice_stream(const sr_ctx* c, uint32_t slice) { return slice == 0 ? c->stream : c->sliceStreams[slice - 1]; }

void mark(sr_ctx* c, uint32_t slice, int kind) {
    if (!c->profileOn) return;
    sr_ctx::ProfLane& l = c->prof[slice];
    if (l.used == l.ev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        l.ev.push_back(e);
        l.kind.push_back(kind);
    }
    l.kind[l.used] = kind;
    (void)hipEventRecord(l.ev[l.used++], slice_stream(c, slice));
}

// One slice of a band: its frame state, stream and launch shapes. Traversal and shade launches
// get one block per queue chunk of the largest possible queue (blocks past the live count exit at
// once): the dispatcher then hands chunks to whichever CU frees up, where a grid-stride loop over
// a resident-sized grid fixes each wave's share in advance and the launch waits for the unluckiest
// wave (trace 207 -> 188 ms, shade 85.5 -> 82.8 ms per 1080p frame; one-wave trace blocks,
// 188 -> 179 ms, free a CU slot per finished wave). A trace launch holds at most nPaths rays
// (padded to whole waves) plus nPaths shadow rays; a shade launch at most nPaths rays.
struct SliceRun {
    Frame F;
    uint32_t slice;
    hipStream_t s;
    int slotBlocks, gTrace, gShade, gTail;
};

SliceRun slice_run(sr_ctx* c, const Frame& F, uint32_t slice) {
    SliceRun r;
    r.F = F;
    r.slice = slice;
    r.s = slice_stream(c, slice);
    r.slotBlocks = (int)((F.nPaths + kBlock - 1) / kBlock);
    r.gTrace = (int)((2u * F.nPaths + 64u + kTraceBlock - 1) / kTraceBlock);
    r.gShade = (int)((F.nPaths + kShadeBlock - 1) / kShadeBlock);
    r.gTail = c->gridCUs * 8;
    return r;
}

// Sample 0's camera rays of a slice.
template <class R>
void enqueue_start(sr_ctx* c, const SliceRun& r) {
    mark(c, r.slice, 3);
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, r.s, r.F);
    hipLaunchKernelGGL(k_start<R>, dim3(r.slotBlocks), dim3(kBlock), 0, r.s, r.F);
    mark(c, r.slice, 3);
}

// One sample of a slice: (bounces + 1) segments of extend, shade and shadow (the segments from the
// tail segment on in one fused launch), then the accumulate / next-camera-ray pass. Queues that run
// empty (converged pixels, terminated paths) make the remaining launches exit at once.
template <class R, bool COUNT>
void enqueue_sample(sr_ctx* c, const SliceRun& r, uint32_t smp) {
    const Frame& F = r.F;
    Frame F0 = F;  // the segment-0 shade's view: the batch's first sample
    F0.sample0 = smp;
    const hipStream_t s = r.s;
    auto trace = c->dev.anyTransparent ? k_trace<COUNT, true> : k_trace<COUNT, false>;
    auto tail = c->dev.anyTransparent ? k_tail<R, COUNT, true> : k_tail<R, COUNT, false>;
    auto shade = c->dev.fog.enabled ? k_shade<R, true, false> : k_shade<R, false, false>;
    auto shadeFirst = c->dev.fog.enabled ? k_shade<R, true, true> : k_shade<R, false, true>;
    int cur = 0;
    bool fused = false;
    const uint32_t tailSeg = tail_segment(c, F.batch);
    for (uint32_t seg = 0; seg <= F.bounces; ++seg) {
        hipLaunchKernelGGL(trace, dim3(r.gTrace), dim3(kTraceBlock), trav_lds_bytes(kTraceBlock), s, F, c->dev, cur);
        mark(c, r.slice, 0);
        if (seg == tailSeg) {  // the rest of every path in one launch
            hipLaunchKernelGGL(tail, dim3(r.gTail), dim3(kBlock), trav_lds_bytes(kBlock), s, F, c->dev, cur, (int)(seg == 0));
            mark(c, r.slice, 2);
            hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, s, F);
            fused = true;
            break;
        }
        hipLaunchKernelGGL(seg == 0 ? shadeFirst : shade, dim3(r.gShade), dim3(kShadeBlock), 0, s, seg == 0 ? F0 : F,
                           c->dev, cur);
        mark(c, r.slice, 1);
        cur ^= 1;
    }
    if (!fused) {  // the last segment's shadow rays (its ray queue is empty)
        hipLaunchKernelGGL(trace, dim3(r.gTrace), dim3(kTraceBlock), trav_lds_bytes(kTraceBlock), s, F, c->dev, cur);
        mark(c, r.slice, 2);
    }
    auto nextPass = F.batch > 1 ? k_next<R, true> : k_next<R, false>;
    hipLaunchKernelGGL(nextPass, dim3((F.nPaths + kNextBlock - 1) / kNextBlock), dim3(kNextBlock), 0, s, F);
    mark(c, r.slice, 3);
}

int fold_render_counters(sr_ctx* c, std::string& msg);

// Publishes the finished-block count for Tick.
void publish_blocks(sr_ctx* c, uint32_t done) {
    std::lock_guard<std::mutex> lk(c->mu);
    if (done > c->blocksDone) {
        c->blocksDone = std::min(done, c->blocksTotal);
        c->cv.notify_all();
    }
}

// Every band as concurrent row slices, one stream each. The host enqueues sample by sample across
// the slices, so every stream has work queued from the start; the band is done when all are. In
// adaptive renders (renderer.cs:205-211) each sample ends with a progress snapshot per slice; the
// driver reads the snapshot kProgRing samples behind the last enqueued one and stops enqueueing
// the band once no pixel of it is active (the reference's workers stop per pixel; the GPU launches
// after that point would only find empty queues).
template <class R, bool COUNT>
int run_frame(sr_ctx* c) {
    const bool adaptive = c->params.min_samples < c->params.max_samples && c->params.variance_threshold > 0.0f;
    for (size_t b = 0; b < c->bandEnd.size(); ++b) {
        const uint32_t y0 = b == 0 ? 0 : c->bandEnd[b - 1];
        std::vector<SliceRun> runs;
        for (uint32_t j = 0; j < c->sliceCount; ++j) {
            const uint32_t sy = y0 + j * c->sliceRows;
            if (sy >= c->bandEnd[b]) break;
            runs.push_back(slice_run(c, frame_for(c, sy, std::min(c->sliceRows, c->bandEnd[b] - sy), j), j));
        }
        const size_t nr = runs.size();
        for (const SliceRun& r : runs) enqueue_start<R>(c, r);
        std::vector<uint32_t> pending;  // ring slots of snapshots not read yet, oldest first
        bool bandIdle = false;
        // Reads the oldest snapshot: its finished blocks go to Tick, and an all-empty camera-ray
        // queue means every pixel of the band is final.
        auto consume = [&]() -> int {
            const uint32_t slot = pending.front();
            pending.erase(pending.begin());
            uint32_t live = 0, done = 0;
            for (size_t j = 0; j < nr; ++j) {
                HIP_TRY(hipEventSynchronize(c->progEv[slot * nr + j]));
                live += c->progHost[(slot * nr + j) * 2];
                done = std::max(done, c->progHost[(slot * nr + j) * 2 + 1]);
            }
            publish_blocks(c, done);
            if (live == 0) bandIdle = true;
            return SR_OK;
        };
        for (uint32_t smp = 0; smp < c->params.max_samples && !bandIdle; smp += c->batch) {
            for (const SliceRun& r : runs) enqueue_sample<R, COUNT>(c, r, smp);
            c->iterations++;
            if (!adaptive) continue;
            if (pending.size() == (size_t)kProgRing) {
                int rc = consume();
                if (rc) return rc;
            }
            const uint32_t slot = smp % kProgRing;
            for (size_t j = 0; j < nr; ++j) {
                uint32_t* hp = c->progHost + (slot * nr + j) * 2;
                HIP_TRY(hipMemcpyAsync(hp, runs[j].F.counts + kRayCount0, 4, hipMemcpyDeviceToHost, runs[j].s));
                HIP_TRY(hipMemcpyAsync(hp + 1, c->dProgress.ptr, 4, hipMemcpyDeviceToHost, runs[j].s));
                HIP_TRY(hipEventRecord(c->progEv[slot * nr + j], runs[j].s));
            }
            pending.push_back(slot);
        }
        HIP_TRY(hipGetLastError());
        for (const SliceRun& r : runs) HIP_TRY(hipStreamSynchronize(r.s));
        uint32_t overflow = 0;
        for (const SliceRun& r : runs) {
            uint32_t ov = 0;
            HIP_TRY(hipMemcpy(&ov, r.F.counts + kOverflow, 4, hipMemcpyDeviceToHost));
            overflow |= ov;
        }
        if (overflow)
            return fail(SR_ERR_LIMIT, std::string("render: a ") + ((overflow & 1u) ? "ray" : "shadow-ray") +
                                          " queue reached its capacity; the render was stopped");
        uint32_t done = 0;
        HIP_TRY(hipMemcpy(&done, c->dProgress.ptr, 4, hipMemcpyDeviceToHost));
        publish_blocks(c, done);
    }
    return SR_OK;
}

void drive(sr_ctx* c) {
    int rc = SR_OK;
    std::string msg;
    if (hipSetDevice(c->device) != hipSuccess || hipEventRecord(c->evStart, c->stream) != hipSuccess) {
        rc = SR_ERR_HIP;
        msg = "render worker: could not start";
    } else {
        for (uint32_t j = 1; j < c->sliceCount && rc == SR_OK; ++j)
            if (hipStreamWaitEvent(slice_stream(c, j), c->evStart, 0) != hipSuccess) {
                rc = SR_ERR_HIP;
                msg = "render worker: could not order the slice streams";
            }
        const bool counter = c->params.rng_policy == SR_RNG_COUNTER;
        if (rc == SR_OK) {
            if (counter) rc = c->countersOn ? run_frame<CounterRng, true>(c) : run_frame<CounterRng, false>(c);
            else rc = c->countersOn ? run_frame<XorwowRng, true>(c) : run_frame<XorwowRng, false>(c);
            if (rc) msg = g_err;
        }
        hipError_t e = hipEventRecord(c->evEnd, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(c->evEnd);
        if (!rc && e != hipSuccess) { rc = SR_ERR_HIP; msg = hipGetErrorString(e); }
        if (!rc && c->blocksDone != c->blocksTotal) {
            rc = SR_ERR_STATE;
            msg = "render: " + std::to_string(c->blocksDone) + " of " + std::to_string(c->blocksTotal) +
                  " blocks reported finished";
        }
        if (!rc && c->countersOn) rc = fold_render_counters(c, msg);
    }
    std::lock_guard<std::mutex> lk(c->mu);
    c->workerStatus = rc;
    c->workerError = msg;
    c->workerDone = true;
    c->blocksDone = c->blocksTotal;
    c->cv.notify_all();
}

// Waits for the background render; reports its error on the calling thread.
int wait_render(sr_ctx* c) {
    c->join();
    if (c->workerStatus != SR_OK) return fail(c->workerStatus, "render: " + c->workerError);
    return SR_OK;
}

}  // namespace

extern "C" {

int sr_render_begin(sr_ctx* ctx, const sr_view* view, const sr_render_params* p) {
    if (!ctx || !view || !p) return fail(SR_ERR_INVALID, "render: null argument");
    if (ctx->device < 0) return fail(SR_ERR_STATE, "host-only context (device -1) cannot render");
    if (p->width == 0 || p->height == 0 || p->block_size == 0) return fail(SR_ERR_INVALID, "render: zero size");
    if (p->min_samples == 0 || p->min_samples > p->max_samples) return fail(SR_ERR_INVALID, "render: need 0 < minSamples <= maxSamples");
    if (p->rng_policy > 1) return fail(SR_ERR_INVALID, "render: unknown rng policy");
    if (p->light_sampler > 1) return fail(SR_ERR_INVALID, "render: unknown light sampler");
    if (p->width > 32768 || p->height > 32768) return fail(SR_ERR_LIMIT, "render: image larger than 32768");
    ctx->join();  // a previous render finishes first
    // From here on a failure leaves no render: read and tick report SR_ERR_STATE instead of using
    // the previous render's buffers with this call's sizes.
    ctx->rendering = false;
    HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->built) {
        int rc = sr_scene_build(ctx);
        if (rc) return rc;
    }
    ctx->params = *p;
    ctx->dev.sky.sampler = (int32_t)p->light_sampler;
    if (ctx->params.tick_bands == 0) ctx->params.tick_bands = 1;
    const uint32_t bands = std::min(ctx->params.tick_bands, (p->height + 7) / 8);
    const uint32_t rowsPer = ((p->height + bands - 1) / bands + 7) / 8 * 8;
    const uint32_t tilesX = (p->width + 7) / 8;
    // Each band runs as up to `slices` concurrent slices of whole 8-row tile rows.
    const uint32_t tileRows = rowsPer / 8;
    ctx->sliceCount = std::max(1u, std::min(ctx->slices, tileRows));
    ctx->sliceRows = (tileRows + ctx->sliceCount - 1) / ctx->sliceCount * 8;
    ctx->sliceCount = (rowsPer + ctx->sliceRows - 1) / ctx->sliceRows;
    while (ctx->sliceStreams.size() + 1 < ctx->sliceCount) {
        hipStream_t st;
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        ctx->sliceStreams.push_back(st);
    }
    ctx->blocksTotal = ((p->width + p->block_size - 1) / p->block_size) * ((p->height + p->block_size - 1) / p->block_size);
    // Several samples of a pixel in flight need independent samples: the counter RNG (the xorwow
    // stream of a pixel runs through its samples in order), no adaptive stop (it decides after every
    // sample whether the next one runs) and no counters (they follow the reference's work).
    const bool adaptive = p->min_samples < p->max_samples && p->variance_threshold > 0.0f;
    const size_t pixelSlots = (size_t)tilesX * (ctx->sliceRows / 8) * 64 * ctx->sliceCount;
    const uint32_t fit = (uint32_t)std::max<size_t>(1, kMaxPathBytes / (pixelSlots * kPathSlotBytes));
    uint32_t batch = (p->rng_policy == SR_RNG_COUNTER && !adaptive && !ctx->countersOn)
                         ? std::min(std::min(ctx->inFlight, p->max_samples), fit) : 1u;
    while (batch & (batch - 1)) batch &= batch - 1;  // a power of two: a pixel's slots share a wave
    ctx->batch = batch;
    int rc = alloc_frame(ctx, tilesX * (ctx->sliceRows / 8) * 64, ctx->sliceCount);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(ctx->dBlockPixels.ptr, 0, (size_t)ctx->blocksTotal * 4, ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->dProgress.ptr, 0, 4, ctx->stream));
    if (!ctx->evStart) HIP_TRY(hipEventCreate(&ctx->evStart));
    if (!ctx->evEnd) HIP_TRY(hipEventCreate(&ctx->evEnd));
    if (ctx->countersOn)
        HIP_TRY(hipMemsetAsync(ctx->dCtr.ptr, 0, (size_t)kCtrWords * kCtrSpread * sizeof(unsigned long long), ctx->stream));
    ctx->prof.resize(std::max<size_t>(ctx->prof.size(), ctx->sliceCount));
    for (sr_ctx::ProfLane& l : ctx->prof) l.used = 0;
    ctx->iterations = 0;
    ctx->view = make_view(view, p->width, p->height);
    ctx->bandEnd.clear();
    for (uint32_t y0 = 0; y0 < p->height; y0 += rowsPer) ctx->bandEnd.push_back(std::min(p->height, y0 + rowsPer));
    ctx->blocksDone = 0;
    ctx->ticksConsumed = 0;
    ctx->renderStart = std::chrono::steady_clock::now();
    ctx->workerDone = false;
    ctx->workerStatus = SR_OK;
    ctx->workerError.clear();
    ctx->rendering = true;
    ctx->worker = std::thread(drive, ctx);
    return SR_OK;
}

int sr_render_tick(sr_ctx* ctx, uint32_t* step, uint32_t* total) {
    if (!ctx || !step || !total) return fail(SR_ERR_INVALID, "tick: null argument");
    if (!ctx->rendering) return fail(SR_ERR_STATE, "no render in progress");
    std::unique_lock<std::mutex> lk(ctx->mu);
    *total = ctx->blocksTotal;
    if (ctx->blocksDone >= ctx->blocksTotal && ctx->workerDone) {
        if (ctx->workerStatus != SR_OK) return fail(ctx->workerStatus, "render: " + ctx->workerError);
        *step = ctx->blocksTotal;  // renderer.cs:100-101
        return SR_OK;
    }
    ctx->cv.wait(lk, [&] { return ctx->workerDone || ctx->ticksConsumed < ctx->blocksDone; });
    if (ctx->workerDone && ctx->workerStatus != SR_OK) return fail(ctx->workerStatus, "render: " + ctx->workerError);
    if (ctx->ticksConsumed < ctx->blocksDone) ctx->ticksConsumed++;
    *step = ctx->blocksDone;
    return SR_OK;
}

int sr_render_read(sr_ctx* ctx, float* rad, float* radFog, float* nrm, float* uv, float* depth, uint32_t* samples,
                   float* var) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    if (!ctx->rendering) return fail(SR_ERR_STATE, "no render in progress");
    int rc = wait_render(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t px = (size_t)ctx->params.width * ctx->params.height;
    if (rad) HIP_TRY(hipMemcpy(rad, ctx->dRad.ptr, px * 12, hipMemcpyDeviceToHost));
    if (radFog) HIP_TRY(hipMemcpy(radFog, ctx->dRadFog.ptr, px * 12, hipMemcpyDeviceToHost));
    if (nrm) HIP_TRY(hipMemcpy(nrm, ctx->dNormal.ptr, px * 12, hipMemcpyDeviceToHost));
    if (uv) HIP_TRY(hipMemcpy(uv, ctx->dUv.ptr, px * 8, hipMemcpyDeviceToHost));
    if (depth) HIP_TRY(hipMemcpy(depth, ctx->dDepth.ptr, px * 4, hipMemcpyDeviceToHost));
    if (samples) HIP_TRY(hipMemcpy(samples, ctx->dSamples.ptr, px * 4, hipMemcpyDeviceToHost));
    if (var) HIP_TRY(hipMemcpy(var, ctx->dVar.ptr, px * 4, hipMemcpyDeviceToHost));
    return SR_OK;
}

int sr_render_read_albedo(sr_ctx* ctx, float* albedo) {
    if (!ctx || !albedo) return fail(SR_ERR_INVALID, "read_albedo: null argument");
    if (!ctx->rendering) return fail(SR_ERR_STATE, "no render in progress");
    int rc = wait_render(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t px = (size_t)ctx->params.width * ctx->params.height;
    HIP_TRY(hipMemcpy(albedo, ctx->dAlbedo.ptr, px * 12, hipMemcpyDeviceToHost));
    return SR_OK;
}

int sr_render_device_aovs(sr_ctx* ctx, void** rad, void** radFog, void** nrm, void** uv, void** depth, void** samples,
                          void** var) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    if (!ctx->rendering) return fail(SR_ERR_STATE, "no render in progress");
    int rc = wait_render(ctx);
    if (rc) return rc;
    if (rad) *rad = ctx->dRad.ptr;
    if (radFog) *radFog = ctx->dRadFog.ptr;
    if (nrm) *nrm = ctx->dNormal.ptr;
    if (uv) *uv = ctx->dUv.ptr;
    if (depth) *depth = ctx->dDepth.ptr;
    if (samples) *samples = ctx->dSamples.ptr;
    if (var) *var = ctx->dVar.ptr;
    return SR_OK;
}

int sr_render_gpu_seconds(sr_ctx* ctx, double* out) {
    if (!ctx || !out) return fail(SR_ERR_INVALID, "null argument");
    if (!ctx->rendering) return fail(SR_ERR_STATE, "no render in progress");
    int rc = wait_render(ctx);
    if (rc) return rc;
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->evStart, ctx->evEnd));
    *out = ms * 1e-3;
    return SR_OK;
}

int sr_profile_enable(sr_ctx* ctx, int32_t enable) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    ctx->profileOn = enable != 0;
    return SR_OK;
}

int sr_profile_read(sr_ctx* ctx, double* ms, uint32_t* launches) {
    if (!ctx || !ms || !launches) return fail(SR_ERR_INVALID, "null argument");
    if (!ctx->rendering || !ctx->profileOn) return fail(SR_ERR_STATE, "kernel timing was not enabled");
    int rc = wait_render(ctx);
    if (rc) return rc;
    for (int k = 0; k < 4; ++k) { ms[k] = 0.0; launches[k] = 0; }
    for (uint32_t j = 0; j < ctx->sliceCount; ++j) {
        const sr_ctx::ProfLane& l = ctx->prof[j];
        for (size_t i = 1; i < l.used; ++i) {
            float t = 0.0f;
            HIP_TRY(hipEventElapsedTime(&t, l.ev[i - 1], l.ev[i]));
            const int k = l.kind[i];
            ms[k] += t;
            launches[k] += 1;
        }
    }
    return SR_OK;
}

int sr_profile_busy(sr_ctx* ctx, double* ms) {
    if (!ctx || !ms) return fail(SR_ERR_INVALID, "null argument");
    if (!ctx->rendering || !ctx->profileOn) return fail(SR_ERR_STATE, "kernel timing was not enabled");
    int rc = wait_render(ctx);
    if (rc) return rc;
    // Each launch spans [t(i-1), t(i)] on its slice's clock (times from the render's start event);
    // per kind, the union of those spans over all slices.
    std::vector<std::pair<float, float>> iv[4];
    for (uint32_t j = 0; j < ctx->sliceCount; ++j) {
        const sr_ctx::ProfLane& l = ctx->prof[j];
        float prev = 0.0f;
        for (size_t i = 0; i < l.used; ++i) {
            float t = 0.0f;
            HIP_TRY(hipEventElapsedTime(&t, ctx->evStart, l.ev[i]));
            if (i > 0) iv[l.kind[i]].push_back({prev, t});
            prev = t;
        }
    }
    for (int k = 0; k < 4; ++k) {
        std::sort(iv[k].begin(), iv[k].end());
        double total = 0.0;
        float a = 0.0f, b = -1.0f;
        for (const auto& x : iv[k]) {
            if (x.first > b) {
                if (b > a) total += b - a;
                a = x.first;
                b = x.second;
            } else if (x.second > b) {
                b = x.second;
            }
        }
        if (b > a) total += b - a;
        ms[k] = total;
    }
    return SR_OK;
}

int sr_set_slices(sr_ctx* ctx, uint32_t slices) {
    if (!ctx || slices == 0) return fail(SR_ERR_INVALID, "set_slices: need a context and at least one slice");
    ctx->slices = slices;
    return SR_OK;
}

int sr_set_samples_in_flight(sr_ctx* ctx, uint32_t samples) {
    if (!ctx || samples == 0 || samples > kMaxInFlight)
        return fail(SR_ERR_INVALID, "set_samples_in_flight: need a context and 1 to " + std::to_string(kMaxInFlight) + " samples");
    ctx->inFlight = samples;
    return SR_OK;
}

int sr_counters_enable(sr_ctx* ctx, int32_t enable) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    ctx->countersOn = enable != 0;
    return SR_OK;
}

namespace {
// The counter words of the last render, summed over their kCtrSpread copies.
int read_counters(sr_ctx* ctx, unsigned long long* v) {
    std::vector<unsigned long long> all((size_t)kCtrWords * kCtrSpread);
    HIP_TRY(hipMemcpy(all.data(), ctx->dCtr.ptr, all.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int i = 0; i < kCtrWords; ++i) {
        const bool isMax = i == kCountMinKey || i == kCountMaxKey || i == kVarMaxKey;
        v[i] = 0;
        for (int k = 0; k < kCtrSpread; ++k) {
            const unsigned long long x = all[(size_t)k * kCtrWords + i];
            v[i] = isMax ? std::max(v[i], x) : v[i] + x;
        }
    }
    return SR_OK;
}

// The reference-taxonomy render counters (SR_COUNTER_COUNT words) from the device words: the fused
// tail launch's traversal and the reference-only share are added back.
void render_counters(const unsigned long long* v, int64_t* out) {
    for (int i = 0; i < SR_COUNTER_COUNT; ++i) out[i] = (int64_t)v[i];
    out[10] += (int64_t)v[kTailTrace];
    out[13] += (int64_t)v[kTailNode];
    out[14] += (int64_t)v[kTailShape];
    out[11] += (int64_t)v[kRefOcclude];
    out[12] += (int64_t)v[kRefMesh];
    out[13] += (int64_t)v[kRefNode];
    out[14] += (int64_t)v[kRefShape];
}

// After a counting render (driver thread): the render's counters into the book.
int fold_render_counters(sr_ctx* c, std::string& msg) {
    unsigned long long v[kCtrWords];
    if (read_counters(c, v) != SR_OK) {
        msg = g_err;
        return SR_ERR_HIP;
    }
    int64_t r[SR_COUNTER_COUNT];
    render_counters(v, r);
    // Position of each SR_COUNTER_COUNT word in Counters.Type.
    static const int kMap[SR_COUNTER_COUNT] = {
        SR_CT_PIXEL, SR_CT_SAMPLE, SR_CT_SAMPLE_BOUNCE, SR_CT_SAMPLE_HIT, SR_CT_SAMPLE_MISS, SR_CT_SAMPLE_TERMINATE,
        SR_CT_SAMPLE_FOG_SCATTER, SR_CT_SAMPLE_FOG_ESCAPE, SR_CT_SAMPLE_PIXEL_CONVERGED, SR_CT_SAMPLE_PIXEL_MAXED,
        SR_CT_SCENE_TRACE, SR_CT_SCENE_OCCLUDE, SR_CT_MESH_INTERSECT, SR_CT_BVH_INTERSECT_NODE,
        SR_CT_BVH_INTERSECT_SHAPE};
    std::lock_guard<std::mutex> lk(c->mu);
    for (int i = 0; i < SR_COUNTER_COUNT; ++i) book_bump(c, kMap[i], r[i]);
    book_bump(c, SR_CT_WORKER, (int64_t)c->sliceCount);
    book_bump(c, SR_CT_BLOCK, (int64_t)c->blocksTotal);
    if (v[kCountMinKey]) book_bump(c, SR_CT_SAMPLE_COUNT_MIN, (int64_t)(~v[kCountMinKey]));
    if (v[kCountMaxKey]) book_bump(c, SR_CT_SAMPLE_COUNT_MAX, (int64_t)v[kCountMaxKey]);
    int64_t var;
    if (decimal_from_key(v[kVarMaxKey], var)) book_bump(c, SR_CT_SAMPLE_VARIANCE_MAX, var);
    book_bump(c, SR_CT_TIME_RENDER, micros_since(c->renderStart));
    return SR_OK;
}

// The filter kernels' device counters into the book (waits for the context's stream).
int fold_filter_counters(sr_ctx* c) {
    if (!c->filterCtrPending) return SR_OK;
    HIP_TRY(hipStreamSynchronize(c->stream));
    unsigned long long v[kFilterWords];
    HIP_TRY(hipMemcpy(v, c->dFilterCtr.ptr, sizeof(v), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemset(c->dFilterCtr.ptr, 0, sizeof(v)));
    c->filterCtrPending = false;
    std::lock_guard<std::mutex> lk(c->mu);
    book_bump(c, SR_CT_FILTER_SAMPLE, (int64_t)v[kFltSample]);
    book_bump(c, SR_CT_DENOISE_EARLY_OUT, (int64_t)v[kFltEarlyOut]);
    book_bump(c, SR_CT_DENOISE_REJECT_MISSING, (int64_t)v[kFltRejectMissing]);
    book_bump(c, SR_CT_DENOISE_REJECT_NORMAL, (int64_t)v[kFltRejectNormal]);
    book_bump(c, SR_CT_DENOISE_REJECT_DEPTH, (int64_t)v[kFltRejectDepth]);
    book_bump(c, SR_CT_DENOISE_REJECT_LUM, (int64_t)v[kFltRejectLum]);
    int64_t x;
    if (decimal_from_key(v[kFltWeightMax], x)) book_bump(c, SR_CT_DENOISE_WEIGHT_MAX, x);
    if (decimal_from_key(v[kFltMaxLum], x)) book_bump(c, SR_CT_DENOISE_MAX_LUM, x);
    if (decimal_from_key(v[kFltMaxLumBoost], x)) book_bump(c, SR_CT_DENOISE_MAX_LUM_BOOST, x);
    return SR_OK;
}
}  // namespace

int sr_counters_read(sr_ctx* ctx, int64_t* out) {
    if (!ctx || !out) return fail(SR_ERR_INVALID, "null argument");
    if (!ctx->rendering || !ctx->countersOn) return fail(SR_ERR_STATE, "counters were not collected");
    int rc = wait_render(ctx);
    if (rc) return rc;
    unsigned long long v[kCtrWords];
    int rc2 = read_counters(ctx, v);
    if (rc2) return rc2;
    render_counters(v, out);
    return SR_OK;
}

int sr_counters_read_all(sr_ctx* ctx, int64_t* out, uint32_t count) {
    if (!ctx || !out) return fail(SR_ERR_INVALID, "null argument");
    if (count > SR_COUNTER_TYPE_COUNT) return fail(SR_ERR_INVALID, "counters: count above SR_COUNTER_TYPE_COUNT");
    if (ctx->rendering) {
        int rc = wait_render(ctx);
        if (rc) return rc;
    }
    if (ctx->device >= 0) {
        HIP_TRY(hipSetDevice(ctx->device));
        int rc = fold_filter_counters(ctx);
        if (rc) return rc;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (uint32_t i = 0; i < count; ++i) out[i] = ctx->book[i];
    return SR_OK;
}

int sr_counters_reset(sr_ctx* ctx) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    if (ctx->rendering) {
        int rc = wait_render(ctx);
        if (rc) return rc;
    }
    if (ctx->device >= 0) {
        HIP_TRY(hipSetDevice(ctx->device));
        int rc = fold_filter_counters(ctx);
        if (rc) return rc;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    book_reset(ctx);
    return SR_OK;
}

int sr_bvh_stats_get(sr_ctx* ctx, int32_t mesh, sr_bvh_stats* out) {
    if (!ctx || !out) return fail(SR_ERR_INVALID, "bvh stats: null argument");
    const BvhResult* b;
    float ci;
    if (mesh < 0) {
        if (!ctx->built) return fail(SR_ERR_STATE, "Scene not built");
        b = &ctx->sceneBvh;
        ci = 10.0f;  // scene.cs:530
    } else {
        if ((size_t)mesh >= ctx->meshes.size()) return fail(SR_ERR_INVALID, "bvh stats: no such mesh");
        b = &ctx->meshes[mesh].bvh;
        ci = 0.5f;  // mesh.cs:40
    }
    const BvhStatsHost st = bvh_stats(*b, 1.0f, ci);
    out->node_count = st.nodeCount;
    out->depth_max = st.depthMax;
    out->leaf_count = st.leafCount;
    out->leaf_size_min = st.leafMin;
    out->leaf_size_max = st.leafMax;
    // bvh.cs:90-91
    out->leaf_size_avg = st.leafCount > 0 ? (float)st.shapeCount / (float)st.leafCount : 0.0f;
    out->leaf_depth_avg = st.shapeCount > 0 ? (float)st.depthWeighted / (float)st.shapeCount : 0.0f;
    out->sah_cost = st.sah;
    return SR_OK;
}

int sr_counters_read_reference_only(sr_ctx* ctx, int64_t* out) {
    if (!ctx || !out) return fail(SR_ERR_INVALID, "null argument");
    if (!ctx->rendering || !ctx->countersOn) return fail(SR_ERR_STATE, "counters were not collected");
    int rc = wait_render(ctx);
    if (rc) return rc;
    unsigned long long v[kCtrWords];
    int rc2 = read_counters(ctx, v);
    if (rc2) return rc2;
    out[0] = (int64_t)v[kRefNode];
    out[1] = (int64_t)v[kRefShape];
    out[2] = (int64_t)v[kRefMesh];
    out[3] = (int64_t)v[kRefOcclude];
    return SR_OK;
}

int sr_counters_read_fused(sr_ctx* ctx, int64_t* out) {
    if (!ctx || !out) return fail(SR_ERR_INVALID, "null argument");
    if (!ctx->rendering || !ctx->countersOn) return fail(SR_ERR_STATE, "counters were not collected");
    int rc = wait_render(ctx);
    if (rc) return rc;
    unsigned long long v[kCtrWords];
    int rc2 = read_counters(ctx, v);
    if (rc2) return rc2;
    out[0] = (int64_t)v[kTailNode];
    out[1] = (int64_t)v[kTailShape];
    out[2] = (int64_t)v[kTailTrace];
    out[3] = (int64_t)v[kTailOcclude];
    return SR_OK;
}

int sr_set_tail_segment(sr_ctx* ctx, uint32_t segment) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    ctx->tailSegment = segment;
    return SR_OK;
}

int sr_trace_primary(sr_ctx* ctx, const sr_view* view, uint32_t w, uint32_t h, int32_t* obj, int32_t* tri, float* dist) {
    if (!ctx || !view || !obj || !tri || !dist || w == 0 || h == 0) return fail(SR_ERR_INVALID, "trace_primary: bad arguments");
    if (ctx->device < 0) return fail(SR_ERR_STATE, "host-only context (device -1) cannot render");
    HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->built) {
        int rc = sr_scene_build(ctx);
        if (rc) return rc;
    }
    if (ctx->rendering) {
        int rc = wait_render(ctx);
        if (rc) return rc;
    }
    const size_t px = (size_t)w * h;
    int32_t *dObj = nullptr, *dTri = nullptr;
    float* dDist = nullptr;
    HIP_TRY(hipMalloc((void**)&dObj, px * 4));
    HIP_TRY(hipMalloc((void**)&dTri, px * 4));
    HIP_TRY(hipMalloc((void**)&dDist, px * 4));
    sr_render_params saved = ctx->params;
    ctx->params.width = w;
    ctx->params.height = h;
    Frame F = frame_for(ctx, 0, h);
    ctx->params = saved;
    F.view = make_view(view, w, h);
    hipLaunchKernelGGL(k_primary, dim3((F.nSlots + kBlock - 1) / kBlock), dim3(kBlock), trav_lds_bytes(kBlock), ctx->stream, F, ctx->dev, dObj,
                       dTri, dDist);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(obj, dObj, px * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(tri, dTri, px * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(dist, dDist, px * 4, hipMemcpyDeviceToHost);
    (void)hipFree(dObj);
    (void)hipFree(dTri);
    (void)hipFree(dDist);
    if (e != hipSuccess) return fail(SR_ERR_HIP, hipGetErrorString(e));
    return SR_OK;
}

int sr_bvh_export(sr_ctx* ctx, int32_t mesh, float* nodes, int32_t* items, uint32_t* node_count, uint32_t* item_count) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    const BvhResult* b = nullptr;
    if (mesh < 0) {
        if (!ctx->built) return fail(SR_ERR_STATE, "Scene not built");
        b = &ctx->sceneBvh;
    } else {
        if (mesh >= (int32_t)ctx->meshes.size()) return fail(SR_ERR_INVALID, "bad mesh id");
        b = &ctx->meshes[mesh].bvh;
    }
    if (node_count) *node_count = (uint32_t)b->used;
    if (item_count) *item_count = (uint32_t)b->items.size();
    if (nodes) {
        for (int i = 0; i < b->used; ++i) {
            const RefNode& nd = b->nodes[i];
            float* o = nodes + (size_t)i * 8;
            o[0] = nd.bounds.mn.x; o[1] = nd.bounds.mn.y; o[2] = nd.bounds.mn.z;
            o[3] = nd.bounds.mx.x; o[4] = nd.bounds.mx.y; o[5] = nd.bounds.mx.z;
            memcpy(o + 6, &nd.child, 4);
            memcpy(o + 7, &nd.count, 4);
        }
    }
    if (items) memcpy(items, b->items.data(), b->items.size() * 4);
    return SR_OK;
}

namespace {
// The device filter counters of a counting context (allocated and zeroed on first use; null
// otherwise). Kernels that bump them leave the book to be folded at the next sync point.
hipError_t filter_counters(sr_ctx* ctx, unsigned long long*& fc) {
    fc = nullptr;
    if (!ctx->countersOn) return hipSuccess;
    if (!ctx->dFilterCtr.ptr) {
        hipError_t e = hipMalloc((void**)&ctx->dFilterCtr.ptr, kFilterWords * sizeof(unsigned long long));
        if (e != hipSuccess) return e;
        ctx->dFilterCtr.n = kFilterWords;
        e = hipMemset(ctx->dFilterCtr.ptr, 0, kFilterWords * sizeof(unsigned long long));
        if (e != hipSuccess) return e;
    }
    fc = ctx->dFilterCtr.ptr;
    ctx->filterCtrPending = true;
    return hipSuccess;
}

int launch_filter(sr_ctx* ctx, const sr_denoise_params* p, const float* rad, const float* nrm, const float* dep,
                  const float* var, const float* alb, float albedoLimit, uint32_t W, uint32_t H, float* out,
                  bool countable = true) {
    if (!ctx || !p || !rad || !nrm || !dep || !var || !out || W == 0 || H == 0)
        return fail(SR_ERR_INVALID, "denoise: bad arguments");
    if (alb && !(albedoLimit > 0.0f)) return fail(SR_ERR_INVALID, "denoise: albedo limit must be > 0");
    if (ctx->device < 0) return fail(SR_ERR_STATE, "host-only context (device -1) cannot render");
    HIP_TRY(hipSetDevice(ctx->device));
    FilterParams fp;
    // Compositor ctor (compose.cs:42-44) and Filter's per-image constants (compose.cs:113-115).
    fp.strength = p->strength;
    fp.strengthMax = p->strength_max;
    fp.lumBoost = p->luminance_boost;
    fp.lumLimitInv = 1.0f / (p->luminance_limit * p->luminance_limit * 2.0f);
    fp.normalLimitInv = 1.0f / (p->normal_limit * p->normal_limit * 2.0f);
    fp.depthLimitInv = 1.0f / (p->depth_limit * p->depth_limit * 2.0f);
    const float radiusPixels = p->radius * sqrtf((float)(W * H));
    fp.radiusInv = 1.0f / (radiusPixels * radiusPixels * 2.0f);
    fp.kr = (int32_t)ceilf(radiusPixels * 3.0f);
    fp.albedoLimitInv = alb ? 1.0f / (albedoLimit * albedoLimit * 2.0f) : 0.0f;
    const dim3 grid((W + 15) / 16, (H + 15) / 16);
    unsigned long long* fc = nullptr;
    if (countable) HIP_TRY(filter_counters(ctx, fc));
    auto kern = alb ? (fc ? k_filter<true, true> : k_filter<true, false>)
                    : (fc ? k_filter<false, true> : k_filter<false, false>);
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, ctx->stream, fp, rad, nrm, dep, var, alb, (int32_t)W, (int32_t)H, out,
                       fc);
    HIP_TRY(hipGetLastError());
    return SR_OK;
}

// Filters the last render's AOVs into host memory.
int filter_to_host(sr_ctx* ctx, const sr_denoise_params* p, bool guided, float albedoLimit, float* out_rgb,
                   bool sampleMean = false) {
    if (!ctx || !p || !out_rgb) return fail(SR_ERR_INVALID, "denoise: null argument");
    if (!ctx->rendering) return fail(SR_ERR_STATE, "no render to denoise");
    int rc0 = wait_render(ctx);
    if (rc0) return rc0;
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t W = ctx->params.width, H = ctx->params.height;
    float* dOut = nullptr;
    HIP_TRY(hipMalloc((void**)&dOut, (size_t)W * H * (sampleMean ? 24 : 12)));
    const float* rad = ctx->dRad.ptr;
    if (sampleMean) {
        float* dMean = dOut + (size_t)W * H * 3;
        hipLaunchKernelGGL(k_sample_mean, dim3((W * H + 255) / 256), dim3(256), 0, ctx->stream, ctx->dRad.ptr,
                           ctx->dSamples.ptr, ctx->params.max_samples, (int32_t)(W * H), dMean);
        rad = dMean;
    }
    int rc = launch_filter(ctx, p, rad, ctx->dNormal.ptr, ctx->dDepth.ptr, ctx->dVar.ptr,
                           guided ? ctx->dAlbedo.ptr : nullptr, albedoLimit, W, H, dOut);
    hipError_t e = rc == SR_OK ? hipStreamSynchronize(ctx->stream) : hipSuccess;
    if (rc == SR_OK && e == hipSuccess) e = hipMemcpy(out_rgb, dOut, (size_t)W * H * 12, hipMemcpyDeviceToHost);
    (void)hipFree(dOut);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SR_ERR_HIP, hipGetErrorString(e));
    return SR_OK;
}
}  // namespace

int sr_denoise_bilateral_device(sr_ctx* ctx, const sr_denoise_params* p, const float* rad, const float* nrm,
                                const float* dep, const float* var, uint32_t W, uint32_t H, float* out) {
    return launch_filter(ctx, p, rad, nrm, dep, var, nullptr, 0.0f, W, H, out);
}

int sr_denoise_bilateral(sr_ctx* ctx, const sr_denoise_params* p, float* out_rgb) {
    return filter_to_host(ctx, p, false, 0.0f, out_rgb);
}

int sr_denoise_bilateral_guided(sr_ctx* ctx, const sr_denoise_params* p, float albedo_limit, float* out_rgb) {
    return filter_to_host(ctx, p, true, albedo_limit, out_rgb);
}

int sr_denoise_bilateral_mean(sr_ctx* ctx, const sr_denoise_params* p, float* out_rgb) {
    return filter_to_host(ctx, p, false, 0.0f, out_rgb, true);
}

// ---- Compose tail ----------------------------------------------------------------------------------
namespace {
int launch_filter_fog(sr_ctx* ctx, float radius, float strength, const float* fog, uint32_t W, uint32_t H,
                      float* out, bool countable = true) {
    if (!(radius >= 0.0f)) return fail(SR_ERR_INVALID, "fog filter: radius must be >= 0");
    // FilterFog's per-image constants (compose.cs:211-213).
    const float radiusPixels = radius * sqrtf((float)(W * H));
    const float radiusInv = 1.0f / (radiusPixels * radiusPixels * 2.0f);
    const float krf = ceilf(radiusPixels * 3.0f);
    if (krf > 4096.0f) return fail(SR_ERR_LIMIT, "fog filter: kernel radius above 4096 pixels");
    const dim3 grid((W + 15) / 16, (H + 15) / 16);
    unsigned long long* fc = nullptr;
    if (countable) HIP_TRY(filter_counters(ctx, fc));
    hipLaunchKernelGGL(k_filter_fog, grid, dim3(256), 0, ctx->stream, radiusInv, strength, (int32_t)krf, fog,
                       (int32_t)W, (int32_t)H, out, fc);
    HIP_TRY(hipGetLastError());
    return SR_OK;
}

// The last render's AOVs, finished, on a device context.
int compose_ready(sr_ctx* ctx) {
    if (!ctx->rendering) return fail(SR_ERR_STATE, "no render to compose");
    int rc = wait_render(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    return SR_OK;
}

int tonemap_to_host(sr_ctx* ctx, const float* dMain, const float* dFog, int32_t tonemapper, float exposure,
                    uint8_t* out_rgb) {
    if (tonemapper != SR_TONEMAP_REINHARD && tonemapper != SR_TONEMAP_LINEAR_SMOOTH)
        return fail(SR_ERR_INVALID, "unknown tonemapper");
    const uint32_t n = ctx->params.width * ctx->params.height;
    uint8_t* dOut = nullptr;
    HIP_TRY(hipMalloc((void**)&dOut, (size_t)n * 3));
    hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, dMain, dFog, (int32_t)n,
                       tonemapper, exposure, dOut);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out_rgb, dOut, (size_t)n * 3, hipMemcpyDeviceToHost);
    (void)hipFree(dOut);
    if (e != hipSuccess) return fail(SR_ERR_HIP, hipGetErrorString(e));
    return SR_OK;
}
}  // namespace

int sr_filter_fog(sr_ctx* ctx, float fog_radius, float fog_strength, float* out_rgb) {
    if (!ctx || !out_rgb) return fail(SR_ERR_INVALID, "fog filter: null argument");
    int rc = compose_ready(ctx);
    if (rc) return rc;
    const uint32_t W = ctx->params.width, H = ctx->params.height;
    float* dOut = nullptr;
    HIP_TRY(hipMalloc((void**)&dOut, (size_t)W * H * 12));
    rc = launch_filter_fog(ctx, fog_radius, fog_strength, ctx->dRadFog.ptr, W, H, dOut);
    hipError_t e = rc == SR_OK ? hipStreamSynchronize(ctx->stream) : hipSuccess;
    if (rc == SR_OK && e == hipSuccess) e = hipMemcpy(out_rgb, dOut, (size_t)W * H * 12, hipMemcpyDeviceToHost);
    (void)hipFree(dOut);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SR_ERR_HIP, hipGetErrorString(e));
    return SR_OK;
}

int sr_preview(sr_ctx* ctx, int32_t tonemapper, float exposure, uint8_t* out_rgb) {
    if (!ctx || !out_rgb) return fail(SR_ERR_INVALID, "preview: null argument");
    int rc = compose_ready(ctx);
    if (rc) return rc;
    return tonemap_to_host(ctx, ctx->dRad.ptr, ctx->dRadFog.ptr, tonemapper, exposure, out_rgb);
}

int sr_compose(sr_ctx* ctx, const sr_compose_params* p, uint8_t* out_rgb) {
    if (!ctx || !p || !out_rgb) return fail(SR_ERR_INVALID, "compose: null argument");
    int rc = compose_ready(ctx);
    if (rc) return rc;
    const auto t0 = std::chrono::steady_clock::now();  // TimeCompose (program.cs:279-283)
    const uint32_t W = ctx->params.width, H = ctx->params.height;
    float* dTmp = nullptr;
    HIP_TRY(hipMalloc((void**)&dTmp, (size_t)W * H * 24));
    float* dMain = dTmp;
    float* dFog = dTmp + (size_t)W * H * 3;
    rc = launch_filter(ctx, &p->denoise, ctx->dRad.ptr, ctx->dNormal.ptr, ctx->dDepth.ptr, ctx->dVar.ptr, nullptr,
                       0.0f, W, H, dMain);
    if (rc == SR_OK) rc = launch_filter_fog(ctx, p->fog_radius, p->fog_strength, ctx->dRadFog.ptr, W, H, dFog);
    if (rc == SR_OK) rc = tonemap_to_host(ctx, dMain, dFog, p->tonemapper, p->exposure, out_rgb);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(dTmp);
    if (rc == SR_OK && ctx->countersOn) book_bump_locked(ctx, SR_CT_TIME_COMPOSE, micros_since(t0));
    return rc;
}

int sr_compose_device_noalloc(sr_ctx* ctx, const sr_compose_params* p, const float* d_radiance,
                              const float* d_radiance_fog, const float* d_normals, const float* d_depth,
                              const float* d_variance, uint32_t width, uint32_t height, uint8_t* d_out_rgb) {
    if (!ctx || !p || !d_radiance || !d_radiance_fog || !d_normals || !d_depth || !d_variance || !d_out_rgb ||
        width == 0 || height == 0)
        return fail(SR_ERR_INVALID, "compose_device_noalloc: bad arguments");
    if (ctx->device < 0) return fail(SR_ERR_STATE, "host-only context (device -1) cannot render");
    if ((uint64_t)width * height > kNoAllocMaxPixels)
        return fail(SR_ERR_LIMIT, "compose_device_noalloc: images above 2,097,152 pixels exceed the scratch");
    if (p->tonemapper != SR_TONEMAP_REINHARD && p->tonemapper != SR_TONEMAP_LINEAR_SMOOTH)
        return fail(SR_ERR_INVALID, "unknown tonemapper");
    HIP_TRY(hipSetDevice(ctx->device));
    float* sc;
    {
        std::lock_guard<std::mutex> lk(g_scratchMu);
        sc = (size_t)ctx->device < g_scratchPtr.size() ? (float*)g_scratchPtr[ctx->device] : nullptr;
    }
    if (!sc) return fail(SR_ERR_STATE, "compose_device_noalloc: scratch not resolved for this device");
    float* dMain = sc;                              // filtered radiance, w * h * 3
    float* dFog = sc + 3 * (size_t)kNoAllocMaxPixels;  // filtered fog radiance, w * h * 3
    // Counting stays off here: the filter counters' buffer would be an allocation.
    int rc = launch_filter(ctx, &p->denoise, d_radiance, d_normals, d_depth, d_variance, nullptr, 0.0f, width, height,
                           dMain, false);
    if (rc) return rc;
    rc = launch_filter_fog(ctx, p->fog_radius, p->fog_strength, d_radiance_fog, width, height, dFog, false);
    if (rc) return rc;
    const uint32_t n = width * height;
    hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, dMain, dFog, (int32_t)n,
                       p->tonemapper, p->exposure, d_out_rgb);
    HIP_TRY(hipGetLastError());
    return SR_OK;
}

int sr_image_encode_bmp(const uint8_t* rgb, uint32_t width, uint32_t height, uint8_t* out, uint64_t* size) {
    if (!rgb || !size || width == 0 || height == 0) return fail(SR_ERR_INVALID, "bmp: bad arguments");
    if (width > (1u << 20) || height > (1u << 20)) return fail(SR_ERR_LIMIT, "bmp: image above 2^20 pixels a side");
    // Image.WriteBmp (image.cs:147-191): 14-byte file header, 40-byte DIB header, 24-bit BGR rows
    // padded to 4 bytes, top-down (negative height).
    const uint64_t stride = ((uint64_t)width * 3 + 3) & ~3ull;
    const uint64_t pixelBytes = stride * height;
    const uint64_t total = 54 + pixelBytes;
    if (total > 0xFFFFFFFFull) return fail(SR_ERR_LIMIT, "bmp: file above 4 GiB");
    *size = total;
    if (!out) return SR_OK;
    auto put32 = [](uint8_t* p, uint32_t v) { p[0] = v & 255; p[1] = (v >> 8) & 255; p[2] = (v >> 16) & 255; p[3] = v >> 24; };
    memset(out, 0, 54);
    out[0] = 'B';
    out[1] = 'M';
    put32(out + 2, (uint32_t)total);
    put32(out + 10, 54);
    put32(out + 14, 40);
    put32(out + 18, width);
    put32(out + 22, (uint32_t)(-(int32_t)height));
    out[26] = 1;   // colour planes
    out[28] = 24;  // bits per pixel
    put32(out + 34, (uint32_t)pixelBytes);
    for (uint32_t y = 0; y < height; ++y) {
        uint8_t* row = out + 54 + stride * y;
        const uint8_t* src = rgb + (size_t)y * width * 3;
        for (uint32_t x = 0; x < width; ++x) {
            row[x * 3] = src[x * 3 + 2];
            row[x * 3 + 1] = src[x * 3 + 1];
            row[x * 3 + 2] = src[x * 3];
        }
        memset(row + (size_t)width * 3, 0, stride - (uint64_t)width * 3);
    }
    return SR_OK;
}

int sr_image_write_bmp(const char* path, const uint8_t* rgb, uint32_t width, uint32_t height) {
    if (!path) return fail(SR_ERR_INVALID, "bmp: null path");
    uint64_t size = 0;
    int rc = sr_image_encode_bmp(rgb, width, height, nullptr, &size);
    if (rc) return rc;
    std::vector<uint8_t> buf(size);
    rc = sr_image_encode_bmp(rgb, width, height, buf.data(), &size);
    if (rc) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) return fail(SR_ERR_INVALID, std::string("bmp: cannot open ") + path);
    const bool ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (fclose(f) != 0 || !ok) return fail(SR_ERR_INVALID, std::string("bmp: write failed: ") + path);
    return SR_OK;
}

// ---- Adjoint -------------------------------------------------------------------------------------
int sr_adjoint_param_count(sr_ctx* ctx, uint32_t* count) {
    if (!ctx || !count) return fail(SR_ERR_INVALID, "adjoint: null argument");
    size_t n = ctx->objects.size() * 4;
    if (ctx->skyKind == 1) n += (size_t)ctx->textures[ctx->skyTex].w * ctx->textures[ctx->skyTex].h * 3;
    if (n > 0xffffffffu) return fail(SR_ERR_LIMIT, "adjoint: too many parameters");
    *count = (uint32_t)n;
    return SR_OK;
}

int sr_adjoint_params_get(sr_ctx* ctx, float* params) {
    if (!ctx || !params) return fail(SR_ERR_INVALID, "adjoint: null argument");
    size_t k = 0;
    for (const HostObject& o : ctx->objects) {
        params[k++] = o.mat.albedo[0]; params[k++] = o.mat.albedo[1]; params[k++] = o.mat.albedo[2];
        params[k++] = o.mat.roughness;
    }
    if (ctx->skyKind == 1) {
        const std::vector<float>& rgb = ctx->textures[ctx->skyTex].rgb;
        memcpy(params + k, rgb.data(), rgb.size() * sizeof(float));
    }
    return SR_OK;
}

int sr_adjoint_params_set(sr_ctx* ctx, const float* params) {
    if (!ctx || !params) return fail(SR_ERR_INVALID, "adjoint: null argument");
    ctx->join();
    uint32_t n = 0;
    int rc = sr_adjoint_param_count(ctx, &n);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; ++i)
        if (!std::isfinite(params[i])) return fail(SR_ERR_INVALID, "adjoint: non-finite parameter");
    size_t k = 0;
    for (HostObject& o : ctx->objects) {
        o.mat.albedo[0] = params[k]; o.mat.albedo[1] = params[k + 1]; o.mat.albedo[2] = params[k + 2];
        o.mat.roughness = params[k + 3];
        k += 4;
    }
    if (ctx->skyKind == 1) {
        std::vector<float>& rgb = ctx->textures[ctx->skyTex].rgb;
        memcpy(rgb.data(), params + k, rgb.size() * sizeof(float));
    }
    ctx->adjReady = false;
    if (!ctx->built) return SR_OK;
    // Refresh what depends on the parameters: object records, the sky texture's texels and the sky
    // sampling tables (the BVHs do not change).
    std::vector<DevObject> dobj;
    build_objects(ctx, dobj, ctx->dev.anyTransparent);
    std::vector<float> cdf;
    std::vector<float4> aliasTab;
    const int32_t sampler = ctx->dev.sky.sampler;
    build_sky(ctx, ctx->dev.sky, cdf, aliasTab);
    ctx->dev.sky.sampler = sampler;
    if (ctx->device < 0) return SR_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    if ((rc = upload(ctx->dObjects, dobj.data(), dobj.size(), s))) return rc;
    if (ctx->skyKind == 1) {
        size_t off = 0;
        for (int32_t t = 0; t < ctx->skyTex; ++t) off += ctx->textures[t].rgb.size();
        const std::vector<float>& rgb = ctx->textures[ctx->skyTex].rgb;
        HIP_TRY(hipMemcpyAsync(ctx->dTexels.ptr + off, rgb.data(), rgb.size() * sizeof(float), hipMemcpyHostToDevice, s));
        if ((rc = upload(ctx->dCdf, cdf.data(), cdf.size(), s))) return rc;
        if ((rc = upload(ctx->dAlias, aliasTab.data(), aliasTab.size(), s))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(s));
    ctx->dev.objects = ctx->dObjects.ptr;
    ctx->dev.cdf = ctx->dCdf.ptr;
    ctx->dev.alias = ctx->dAlias.ptr;
    if ((rc = build_sky_table(ctx, ctx->dev, s))) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    return SR_OK;
}

int sr_adjoint_forward(sr_ctx* ctx, const sr_view* view, const sr_render_params* p, float* image) {
    if (!ctx || !view || !p) return fail(SR_ERR_INVALID, "adjoint forward: null argument");
    if (p->rng_policy != SR_RNG_COUNTER) return fail(SR_ERR_INVALID, "adjoint forward: needs the counter RNG policy");
    if (p->min_samples != p->max_samples || p->variance_threshold != 0.0f)
        return fail(SR_ERR_INVALID, "adjoint forward: needs a fixed sample count (min = max, threshold 0)");
    if (p->bounces + 1 > (uint32_t)kAdjMaxVertices) return fail(SR_ERR_LIMIT, "adjoint forward: at most 16 bounces");
    if (ctx->hasFog) return fail(SR_ERR_STATE, "adjoint: scenes with fog are not differentiated");
    for (const HostObject& o : ctx->objects)
        if (o.mat.transparency > 0.0f) return fail(SR_ERR_STATE, "adjoint: transparent objects are not differentiated");
    int rc = sr_render_begin(ctx, view, p);
    if (rc) return rc;
    if ((rc = wait_render(ctx))) return rc;
    if (image) {
        HIP_TRY(hipMemcpy(image, ctx->dRad.ptr, (size_t)p->width * p->height * 12, hipMemcpyDeviceToHost));
    }
    ctx->adjView = *view;
    ctx->adjParams = *p;
    ctx->adjReady = true;
    return SR_OK;
}

int sr_adjoint_backward(sr_ctx* ctx, const float* d_image, float* grad) {
    if (!ctx || !d_image || !grad) return fail(SR_ERR_INVALID, "adjoint backward: null argument");
    if (!ctx->adjReady) return fail(SR_ERR_STATE, "adjoint backward: no forward render with the current parameters");
    ctx->join();
    HIP_TRY(hipSetDevice(ctx->device));
    const sr_render_params& p = ctx->adjParams;
    uint32_t n = 0;
    int rc = sr_adjoint_param_count(ctx, &n);
    if (rc) return rc;
    const size_t px = (size_t)p.width * p.height;
    const size_t envN = (size_t)n - ctx->objects.size() * 4;  // sky texels * 3 (0 without a texture sky)
    const size_t envStride = (envN + 63) / 64 * 64;
    const uint32_t envCopies = (uint32_t)std::min<size_t>(
        kMaxEnvCopies, std::max<size_t>(kMinEnvCopies, kEnvCopyBytes / std::max<size_t>(1, envStride * sizeof(float))));
    if ((rc = ensure(ctx->dAdjImage, px * 3)) || (rc = ensure(ctx->dAdjGrad, (size_t)n + 1)) ||
        (rc = ensure(ctx->dAdjEnv, std::max<size_t>(1, envStride * envCopies))))
        return rc;
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(ctx->dAdjImage.ptr, d_image, px * 12, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(ctx->dAdjGrad.ptr, 0, ((size_t)n + 1) * sizeof(float), s));
    if (envN) HIP_TRY(hipMemsetAsync(ctx->dAdjEnv.ptr, 0, envStride * envCopies * sizeof(float), s));
    AdjointArgs A;
    A.view = make_view(&ctx->adjView, p.width, p.height);
    A.W = p.width; A.H = p.height; A.spp = p.max_samples; A.bounces = p.bounces; A.seed = p.rng_seed;
    A.nObj = (uint32_t)ctx->objects.size();
    A.clampLum = p.indirect_clamp;
    A.dImage = ctx->dAdjImage.ptr;
    A.gMat = ctx->dAdjGrad.ptr;
    A.gEnv = ctx->dAdjEnv.ptr;
    A.envStride = envStride;
    A.envCopies = envCopies;
    A.gTotal = ctx->dAdjGrad.ptr + n;
    const uint64_t total = (uint64_t)px * p.max_samples;
    const uint64_t blocks = (total + kBlock - 1) / kBlock;
    if (blocks > 0x7fffffffu) return fail(SR_ERR_LIMIT, "adjoint backward: too many paths");
    // The replay keeps its whole traversal stack in private memory (LDS rows measured 2.7 ms slower
    // on the 512^2 / 16 spp backward pass).
    hipLaunchKernelGGL(k_adjoint, dim3((uint32_t)blocks), dim3(kBlock), 0, s, A, ctx->dev);
    if (envN)
        hipLaunchKernelGGL(k_env_reduce, dim3((uint32_t)((envN + 255) / 256)), dim3(256), 0, s, ctx->dAdjEnv.ptr, envStride,
                           envN, envCopies, ctx->dAdjGrad.ptr + ctx->objects.size() * 4);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(grad, ctx->dAdjGrad.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    float gTotal = 0.0f;
    HIP_TRY(hipMemcpyAsync(&gTotal, ctx->dAdjGrad.ptr + n, sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // The sky pdf's dependence on the CDF total reaches every texel: d total / d texel(x, y, c) =
    // luminance weight c * sin(pi (y + 0.5) / H).
    if (ctx->skyKind == 1 && gTotal != 0.0f) {
        const HostTexture& t = ctx->textures[ctx->skyTex];
        float* g = grad + ctx->objects.size() * 4;
        for (uint32_t y = 0; y < t.h; ++y) {
            const float sy = sinf(kPi * ((float)y + 0.5f) / (float)t.h);
            for (uint32_t x = 0; x < t.w; ++x) {
                const size_t i = ((size_t)y * t.w + x) * 3;
                g[i] += gTotal * 0.2126f * sy;
                g[i + 1] += gTotal * 0.7152f * sy;
                g[i + 2] += gTotal * 0.0722f * sy;
            }
        }
    }
    return SR_OK;
}

// ---- Kernel-predicting denoiser --------------------------------------------------------------------
int sr_cnn_weight_count(uint64_t* count) {
    if (!count) return fail(SR_ERR_INVALID, "cnn: null argument");
    *count = sr::cnn_weight_count();
    return SR_OK;
}

int sr_cnn_load(sr_ctx* ctx, const float* weights, uint64_t count) {
    if (!ctx || !weights) return fail(SR_ERR_INVALID, "cnn: null argument");
    if (count != sr::cnn_weight_count()) return fail(SR_ERR_INVALID, "cnn: weight count mismatch");
    if (ctx->device < 0) return fail(SR_ERR_STATE, "host-only context (device -1) cannot denoise");
    for (uint64_t i = 0; i < count; ++i)
        if (!std::isfinite(weights[i])) return fail(SR_ERR_INVALID, "cnn: non-finite weight");
    ctx->join();
    HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->cnn) ctx->cnn = new sr::CnnNet();
    if (!ctx->cnn->load(weights, ctx->stream)) return fail(SR_ERR_HIP, "cnn: weight upload failed");
    return SR_OK;
}

int sr_cnn_denoise_device(sr_ctx* ctx, const float* rad, const float* alb, const float* nrm, const float* dep,
                          uint32_t W, uint32_t H, float* out) {
    if (!ctx || !rad || !alb || !nrm || !dep || !out || W == 0 || H == 0) return fail(SR_ERR_INVALID, "cnn: bad arguments");
    if (!ctx->cnn || !ctx->cnn->loaded()) return fail(SR_ERR_STATE, "cnn: no weights loaded");
    if (W > 32768 || H > 32768) return fail(SR_ERR_LIMIT, "cnn: image larger than 32768");
    HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->cnn->run(rad, alb, nrm, dep, W, H, out, ctx->stream)) return fail(SR_ERR_HIP, "cnn: launch failed");
    return SR_OK;
}

int sr_cnn_denoise(sr_ctx* ctx, float* out_rgb) {
    if (!ctx || !out_rgb) return fail(SR_ERR_INVALID, "cnn: null argument");
    if (!ctx->rendering) return fail(SR_ERR_STATE, "no render to denoise");
    int rc = wait_render(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t W = ctx->params.width, H = ctx->params.height;
    float* dOut = nullptr;
    HIP_TRY(hipMalloc((void**)&dOut, (size_t)W * H * 24));
    float* dMean = dOut + (size_t)W * H * 3;
    hipLaunchKernelGGL(k_sample_mean, dim3((W * H + 255) / 256), dim3(256), 0, ctx->stream, ctx->dRad.ptr,
                       ctx->dSamples.ptr, ctx->params.max_samples, (int32_t)(W * H), dMean);
    rc = sr_cnn_denoise_device(ctx, dMean, ctx->dAlbedo.ptr, ctx->dNormal.ptr, ctx->dDepth.ptr, W, H, dOut);
    hipError_t e = rc == SR_OK ? hipStreamSynchronize(ctx->stream) : hipSuccess;
    if (rc == SR_OK && e == hipSuccess) e = hipMemcpy(out_rgb, dOut, (size_t)W * H * 12, hipMemcpyDeviceToHost);
    (void)hipFree(dOut);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SR_ERR_HIP, hipGetErrorString(e));
    return SR_OK;
}

int sr_cnn_layer_read(sr_ctx* ctx, int32_t layer, float* out, uint32_t* channels, uint32_t* width, uint32_t* height) {
    if (!ctx || !channels || !width || !height) return fail(SR_ERR_INVALID, "cnn: null argument");
    if (!ctx->cnn) return fail(SR_ERR_STATE, "cnn: no weights loaded");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<float> v;
    if (!ctx->cnn->read_layer(layer, v, *channels, *width, *height)) return fail(SR_ERR_STATE, "cnn: no such layer or no run yet");
    if (out) memcpy(out, v.data(), v.size() * sizeof(float));
    return SR_OK;
}

int sr_cnn_benchmark(sr_ctx* ctx, uint32_t iterations, double* ms_per_run) {
    if (!ctx || !ms_per_run || iterations == 0) return fail(SR_ERR_INVALID, "cnn benchmark: bad arguments");
    if (!ctx->rendering) return fail(SR_ERR_STATE, "no render to denoise");
    int rc = wait_render(ctx);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t W = ctx->params.width, H = ctx->params.height;
    float* dOut = nullptr;
    HIP_TRY(hipMalloc((void**)&dOut, (size_t)W * H * 24));
    float* dMean = dOut + (size_t)W * H * 3;
    const dim3 gm((W * H + 255) / 256);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // One untimed run sizes the activations; the timed runs (sample mean + network) are back to back.
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sample_mean, gm, dim3(256), 0, ctx->stream, ctx->dRad.ptr, ctx->dSamples.ptr,
                           ctx->params.max_samples, (int32_t)(W * H), dMean);
        rc = sr_cnn_denoise_device(ctx, dMean, ctx->dAlbedo.ptr, ctx->dNormal.ptr, ctx->dDepth.ptr, W, H, dOut);
    }
    if (e == hipSuccess && rc == SR_OK) e = hipEventRecord(e0, ctx->stream);
    for (uint32_t i = 0; e == hipSuccess && rc == SR_OK && i < iterations; ++i) {
        hipLaunchKernelGGL(k_sample_mean, gm, dim3(256), 0, ctx->stream, ctx->dRad.ptr, ctx->dSamples.ptr,
                           ctx->params.max_samples, (int32_t)(W * H), dMean);
        rc = sr_cnn_denoise_device(ctx, dMean, ctx->dAlbedo.ptr, ctx->dNormal.ptr, ctx->dDepth.ptr, W, H, dOut);
    }
    if (e == hipSuccess && rc == SR_OK) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess && rc == SR_OK) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess && rc == SR_OK) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(dOut);
    if (rc) return rc;
    if (e != hipSuccess) return fail(SR_ERR_HIP, hipGetErrorString(e));
    *ms_per_run = (double)ms / iterations;
    return SR_OK;
}

int sr_cnn_train_begin(sr_ctx* ctx) {
    if (!ctx) return fail(SR_ERR_INVALID, "cnn: null argument");
    if (!ctx->cnn || !ctx->cnn->loaded()) return fail(SR_ERR_STATE, "cnn: load initial weights first");
    HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->cnn->train_init(ctx->stream)) return fail(SR_ERR_HIP, "cnn: training state allocation failed");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SR_OK;
}

int sr_cnn_train_add(sr_ctx* ctx, const float* rad, const float* alb, const float* nrm, const float* dep,
                     const float* target, uint32_t W, uint32_t H) {
    if (!ctx || !rad || !alb || !nrm || !dep || !target || W == 0 || H == 0) return fail(SR_ERR_INVALID, "cnn: bad arguments");
    if (ctx->device < 0) return fail(SR_ERR_STATE, "host-only context (device -1)");
    if (W > 8192 || H > 8192) return fail(SR_ERR_LIMIT, "cnn: training images up to 8192 x 8192");
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n = (size_t)W * H;
    sr_ctx::CnnExample* e = new sr_ctx::CnnExample();
    int rc;
    hipStream_t s = ctx->stream;
    if ((rc = upload(e->rad, rad, n * 3, s)) || (rc = upload(e->alb, alb, n * 3, s)) || (rc = upload(e->nrm, nrm, n * 3, s)) ||
        (rc = upload(e->dep, dep, n, s)) || (rc = upload(e->target, target, n * 3, s))) {
        e->rad.release(); e->alb.release(); e->nrm.release(); e->dep.release(); e->target.release();
        delete e;
        return rc;
    }
    HIP_TRY(hipStreamSynchronize(s));
    e->w = W;
    e->h = H;
    ctx->cnnExamples.push_back(e);
    return SR_OK;
}

int sr_cnn_train_run(sr_ctx* ctx, uint32_t steps, float learning_rate, uint32_t seed, float* mean_loss) {
    if (!ctx || !mean_loss || steps == 0) return fail(SR_ERR_INVALID, "cnn: bad arguments");
    if (!ctx->cnn || !ctx->cnn->training()) return fail(SR_ERR_STATE, "cnn: sr_cnn_train_begin first");
    if (ctx->cnnExamples.empty()) return fail(SR_ERR_STATE, "cnn: no training examples");
    if (!(learning_rate > 0.0f)) return fail(SR_ERR_INVALID, "cnn: learning rate must be > 0");
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = ensure(ctx->dCnnLoss, 1);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(ctx->dCnnLoss.ptr, 0, sizeof(float), ctx->stream));
    // Examples in a seeded shuffled order, reshuffled every pass.
    std::vector<uint32_t> order(ctx->cnnExamples.size());
    uint64_t st = ((uint64_t)seed << 32) ^ 0x9e3779b97f4a7c15ull;
    size_t pos = order.size();
    for (uint32_t k = 0; k < steps; ++k) {
        if (pos == order.size()) {
            for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
            for (size_t i = order.size(); i > 1; --i) {
                st = st * 6364136223846793005ull + 1442695040888963407ull;
                std::swap(order[i - 1], order[(size_t)((st >> 33) % i)]);
            }
            pos = 0;
        }
        const sr_ctx::CnnExample* e = ctx->cnnExamples[order[pos++]];
        if (!ctx->cnn->train_step(e->rad.ptr, e->alb.ptr, e->nrm.ptr, e->dep.ptr, e->target.ptr, e->w, e->h, learning_rate,
                                  ctx->dCnnLoss.ptr, ctx->stream))
            return fail(SR_ERR_HIP, "cnn: training step failed");
    }
    float loss = 0.0f;
    HIP_TRY(hipMemcpyAsync(&loss, ctx->dCnnLoss.ptr, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *mean_loss = loss / (float)steps;
    return SR_OK;
}

int sr_cnn_train_gradient(sr_ctx* ctx, uint32_t example, float* gradient, float* loss) {
    if (!ctx || !gradient || !loss) return fail(SR_ERR_INVALID, "cnn: null argument");
    if (!ctx->cnn || !ctx->cnn->training()) return fail(SR_ERR_STATE, "cnn: sr_cnn_train_begin first");
    if (example >= ctx->cnnExamples.size()) return fail(SR_ERR_INVALID, "cnn: no such example");
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = ensure(ctx->dCnnLoss, 1);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(ctx->dCnnLoss.ptr, 0, sizeof(float), ctx->stream));
    const sr_ctx::CnnExample* e = ctx->cnnExamples[example];
    if (!ctx->cnn->train_step(e->rad.ptr, e->alb.ptr, e->nrm.ptr, e->dep.ptr, e->target.ptr, e->w, e->h, 1.0f,
                              ctx->dCnnLoss.ptr, ctx->stream, false))
        return fail(SR_ERR_HIP, "cnn: gradient pass failed");
    std::vector<float> blob;
    if (!ctx->cnn->export_grads(blob, ctx->stream)) return fail(SR_ERR_HIP, "cnn: gradient export failed");
    memcpy(gradient, blob.data(), blob.size() * sizeof(float));
    HIP_TRY(hipMemcpy(loss, ctx->dCnnLoss.ptr, sizeof(float), hipMemcpyDeviceToHost));
    return SR_OK;
}

int sr_cnn_train_export(sr_ctx* ctx, float* weights) {
    if (!ctx || !weights) return fail(SR_ERR_INVALID, "cnn: null argument");
    if (!ctx->cnn || !ctx->cnn->training()) return fail(SR_ERR_STATE, "cnn: not training");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<float> blob;
    if (!ctx->cnn->export_weights(blob)) return fail(SR_ERR_HIP, "cnn: export failed");
    memcpy(weights, blob.data(), blob.size() * sizeof(float));
    return SR_OK;
}

int sr_cnn_flops(uint32_t width, uint32_t height, double* flops) {
    if (!flops) return fail(SR_ERR_INVALID, "cnn: null argument");
    *flops = sr::CnnNet::flops(width, height);
    return SR_OK;
}

int sr_set_bvh_builder(sr_ctx* ctx, int32_t builder) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    if (builder != SR_BVH_BUILDER_HOST && builder != SR_BVH_BUILDER_DEVICE)
        return fail(SR_ERR_INVALID, "unknown BVH builder");
    if (builder == SR_BVH_BUILDER_DEVICE && ctx->device < 0)
        return fail(SR_ERR_STATE, "host-only context (device -1) has no device builder");
    ctx->bvhBuilder = builder;
    return SR_OK;
}

int sr_scene_counts(sr_ctx* ctx, uint32_t* textures, uint32_t* meshes, uint32_t* objects) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    if (textures) *textures = (uint32_t)ctx->textures.size();
    if (meshes) *meshes = (uint32_t)ctx->meshes.size();
    if (objects) *objects = (uint32_t)ctx->objects.size();
    return SR_OK;
}

int sr_texture_read(sr_ctx* ctx, int32_t id, float* rgb, uint32_t* width, uint32_t* height) {
    if (!ctx || !width || !height) return fail(SR_ERR_INVALID, "texture read: null argument");
    if (id < 0 || id >= (int32_t)ctx->textures.size()) return fail(SR_ERR_INVALID, "texture read: bad id");
    const HostTexture& t = ctx->textures[id];
    *width = t.w;
    *height = t.h;
    if (rgb) memcpy(rgb, t.rgb.data(), t.rgb.size() * 4);
    return SR_OK;
}

int sr_mesh_read(sr_ctx* ctx, int32_t id, float* triangles, uint32_t* count) {
    if (!ctx || !count) return fail(SR_ERR_INVALID, "mesh read: null argument");
    if (id < 0 || id >= (int32_t)ctx->meshes.size()) return fail(SR_ERR_INVALID, "mesh read: bad id");
    const HostMesh& m = ctx->meshes[id];
    *count = (uint32_t)(m.tris.size() / 36);
    if (triangles) memcpy(triangles, m.tris.data(), m.tris.size() * 4);
    return SR_OK;
}

int sr_scene_object_read(sr_ctx* ctx, int32_t id, sr_transform* trans, sr_material* material, int32_t* shape_kind,
                         int32_t* mesh) {
    if (!ctx) return fail(SR_ERR_INVALID, "ctx is null");
    if (id < 0 || id >= (int32_t)ctx->objects.size()) return fail(SR_ERR_INVALID, "object read: bad id");
    const HostObject& o = ctx->objects[id];
    if (trans) *trans = o.trans;
    if (material) *material = o.mat;
    if (shape_kind) *shape_kind = o.kind;
    if (mesh) *mesh = o.mesh;
    return SR_OK;
}

int sr_triangles_build(const float* pos, const float* nrm, const float* uv, uint32_t n, int32_t smooth, float* out) {
    if ((n && (!pos || !out))) return fail(SR_ERR_INVALID, "triangles: null argument");
    build_triangles(pos, nrm, uv, n, smooth != 0, out);
    return SR_OK;
}

}  // extern "C"
