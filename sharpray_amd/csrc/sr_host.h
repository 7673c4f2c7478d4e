// Host-side scene preparation for the gfx950 tracer: binned-SAH BVH construction in the
// reference's node order, Triangle construction, and tabulated CDFs for environment sampling.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "sr_math.h"

struct sr_ctx;

namespace sr {

// One BVH node exactly as bvh.cs:26-31 lays it out (AABox 24 B, Child, ShapeCount): 32 bytes.
struct RefNode {
    box3 bounds;
    int32_t child;
    int32_t count;
};
static_assert(sizeof(RefNode) == 32, "RefNode must match the reference's 32-byte node");

struct BvhResult {
    std::vector<RefNode> nodes;  // nodes[0..used)
    std::vector<int32_t> items;  // item permutation (bvh.cs:53)
    int used = 0;
    bool degenerate = false;     // a split left one side empty
    int max_depth = 0;
};

// Builds the BVH of bvh.cs:58-75/285-476 over `n` primitive bounds: same split candidates, costs,
// partition and depth-first pair allocation, so nodes and items match the reference bit for bit.
BvhResult build_bvh(const box3* bounds, int n, int bins, float cost_traverse, float cost_intersect);

// Bvh.GetStats (bvh.cs:79-125) of a built BVH: the reference's recursion and merge order, so the
// float SAH cost and averages round as the reference's do.
struct BvhStatsHost {
    int32_t nodeCount = 0, depthMax = 0, leafCount = 0, leafMin = 0, leafMax = 0, shapeCount = 0, depthWeighted = 0;
    float sah = 0.0f;
};
BvhStatsHost bvh_stats(const BvhResult& b, float cost_traverse, float cost_intersect);

// Triangle ctors (math.cs:1503-1531) and Mesh.ComputeSmoothNormals (mesh.cs:99-123).
// has_nrm (optional, with nrm): per triangle, 0 = the face-normal ctor despite nrm (obj.cs:430-441).
void build_triangles(const float* pos, const float* nrm, const float* uv, uint32_t n, bool smooth, float* out36,
                     const uint8_t* has_nrm = nullptr);

// The same build on the current HIP device from n TriangleLean records (9 floats each, device
// memory), with 2..64 bins: identical nodes and items (defined in sr_bvh.hip). Returns false with
// err set on a HIP error.
bool device_build_bvh(const float* dLean, int n, int bins, float cost_traverse, float cost_intersect, BvhResult& out,
                      hipStream_t stream, std::string& err);

// Records an error message for sr_last_error and returns code (defined with the C ABI).
int sr_set_error(int code, const char* msg);
// Counters.Bump(entry, micros) on a context that collects counters (defined with the C ABI).
void sr_count_time(::sr_ctx* ctx, int entry, int64_t micros);

// Bounds of a TriangleLean given as (A, B-A, C-A) (math.cs:1353-1359).
box3 lean_bounds(const float* t9);

// Cdf1 (cdf.cs:7-57): prefix sums normalised by the total; returns the total weight.
float build_cdf1(const float* w, int n, float* cdf_out /* n+1 */);
// Walker alias table of n weights (Vose's construction in double precision): prob[i] is the
// acceptance threshold of column i, alias[i] the column taken otherwise. All-zero weights leave
// prob = 1, alias = i.
void build_alias(const float* w, size_t n, float* prob, int32_t* alias);
// IEEE half-precision bits to float (exact).
float half_to_float_host(uint16_t h);

}  // namespace sr
