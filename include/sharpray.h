/*
 * sharpray.h - C ABI of the MI355X (gfx950) path tracer.
 *
 * This is the drop-in boundary for SharpRay's per-pixel path-tracing loop. The reference keeps
 * that loop behind in-process C# classes (Scene, Renderer, Compositor); a C# host binds the
 * functions below through P/Invoke instead. Each declaration names the reference member it
 * replaces (file:line in BastianBlokland/sharpray, src/).
 *
 * Conventions
 *  - Every function returns an int32 status: SR_OK (0) or a negative SR_ERR_* code; the text of
 *    the last error on the calling thread is available from sr_last_error().
 *  - Host arrays use the reference's blittable layouts, so pinned C# arrays pass straight through:
 *    Color/Vec3 = 3 x f32, Vec2 = 2 x f32, Vec4 = 4 x f32, Triangle = 36 x f32 (math.cs:1491-1500).
 *  - The caller owns every host buffer; the library owns all device memory.
 *  - One context per thread; calls on a context are serialised by the caller. GPU work runs
 *    asynchronously on the context's HIP stream and sr_render_tick waits on it.
 */
#ifndef SHARPRAY_H
#define SHARPRAY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SR_OK 0
#define SR_ERR_INVALID -1 /* bad argument (null pointer, zero size, out-of-range id) */
#define SR_ERR_STATE -2   /* call out of order (scene already built / not built, no render) */
#define SR_ERR_HIP -3     /* a HIP runtime call failed */
#define SR_ERR_LIMIT -4   /* input exceeds a documented device-layout limit */
#define SR_ERR_BUILD -5   /* BVH build hit a split with an empty side (bvh.cs:448) */

typedef struct sr_ctx sr_ctx;

/* Transform (math.cs:773-786): position, unit quaternion (x, y, z, w), non-zero scale. */
typedef struct sr_transform {
    float pos[3];
    float rot[4];
    float scale[3];
} sr_transform;

/* Material (scene.cs:161-171). Texture ids come from sr_texture_create; -1 = none. */
typedef struct sr_material {
    float albedo[3];
    float roughness;
    float metallic;
    float transparency;
    float ior;
    float radiance[3];
    int32_t color_texture;
    int32_t roughness_texture;
    int32_t metallic_texture;
    int32_t normal_texture;
} sr_material;

/* SkyProcedural (scene.cs:341-352). */
typedef struct sr_sky_procedural {
    float radiance_top[3];
    float radiance_middle[3];
    float radiance_bottom[3];
    float sun_dir[3];
    float sun_radiance[3];
    float sun_angle;
} sr_sky_procedural;

/* Fog (scene.cs:438-443): exponential height fog. */
typedef struct sr_fog {
    float density;
    float color[3];
    float anisotropy;
    float height_falloff;
} sr_fog;

/* A bounded participating medium (no reference counterpart; it takes the place of Fog in
 * Scene.Sample, scene.cs:641-672, with the same color / anisotropy meaning). kind
 * SR_MEDIUM_HOMOGENEOUS: extinction sigma inside the box bounds_min..bounds_max; kind
 * SR_MEDIUM_GRID: extinction sigma * value of the voxel containing the point, from a
 * res[0] x res[1] x res[2] grid of IEEE half-precision values (x fastest) spanning the box.
 * Free flight is sampled by delta tracking against sigma * max(grid); shadow rays take the
 * exact voxel-by-voxel transmittance. The grid is copied at the call. */
#define SR_MEDIUM_HOMOGENEOUS 1
#define SR_MEDIUM_GRID 2
typedef struct sr_fog_volume {
    uint32_t kind;
    float sigma;
    float color[3];
    float anisotropy;
    float bounds_min[3];
    float bounds_max[3];
    uint32_t res[3];
    const uint16_t* density;
} sr_fog_volume;

/* View (math.cs:1582-1611): transform without scale, horizontal fov (radians), near plane. */
typedef struct sr_view {
    float pos[3];
    float rot[4];
    float fov;
    float near_plane;
} sr_view;

/* Renderer ctor arguments (renderer.cs:40-51). block_size is the unit of sr_render_tick's progress;
 * the GPU renders sample-major over whole bands of rows and counts a block as completed once all
 * its pixels are final.
 * rng_policy: SR_RNG_REFERENCE reproduces the reference's per-pixel stateful xorwow stream
 * (renderer.cs:156); SR_RNG_COUNTER draws from a counter-based hash of (seed, pixel, sample,
 * draw index) so one path's divergence cannot shift another sample's draws (no reference
 * counterpart). tick_bands: number of row bands the frame is rendered in, one after the other
 * (>= 1); each band's blocks complete when the band does. light_sampler picks how a texture sky's light directions are drawn:
 * SR_LIGHT_CDF walks the reference's row/column CDFs (Cdf2, cdf.cs:62-98); SR_LIGHT_ALIAS draws
 * the same discrete distribution from a Walker alias table with one 16-byte load (no reference
 * counterpart; same two uniforms, same pdf). */
#define SR_RNG_REFERENCE 0
#define SR_RNG_COUNTER 1
#define SR_LIGHT_CDF 0
#define SR_LIGHT_ALIAS 1
typedef struct sr_render_params {
    uint32_t width;
    uint32_t height;
    uint32_t block_size;
    uint32_t min_samples;
    uint32_t max_samples;
    float variance_threshold;
    uint32_t bounces;
    float indirect_clamp;
    uint32_t rng_policy;
    uint32_t rng_seed;
    uint32_t tick_bands;
    uint32_t light_sampler;
} sr_render_params;

/* Compositor ctor denoise arguments (compose.cs:22-34). */
typedef struct sr_denoise_params {
    float radius;
    float strength;
    float strength_max;
    float luminance_boost;
    float luminance_limit;
    float normal_limit;
    float depth_limit;
} sr_denoise_params;

/* ---- Context ------------------------------------------------------------------------------- */
/* device = -1 creates a host-only context: scene preparation and BVH export work, rendering
 * calls return SR_ERR_STATE. */
int sr_create(int32_t device, sr_ctx** out);
int sr_destroy(sr_ctx* ctx);
const char* sr_last_error(void);

/* ---- Scene assets -------------------------------------------------------------------------- */
/* Texture (texture.cs:17-30): linear RGB float texels, row-major w*h*3. filter 0 = bilinear,
 * 1 = point (texture.cs:4); tiling (texture.cs:10). */
int sr_texture_create(sr_ctx* ctx, const float* rgb, uint32_t width, uint32_t height, int32_t filter,
                      float tiling_x, float tiling_y, int32_t* out_id);
/* Mesh ctor (mesh.cs:18-46): triangles in the C# Triangle layout (36 floats each). Builds the
 * binned-SAH BVH with sahCostIntersect 0.5 (mesh.cs:41) and `split_bins` bins (reference: 8). */
int sr_mesh_create(sr_ctx* ctx, const float* triangles, uint32_t count, uint32_t split_bins, int32_t* out_id);

/* Which builder sr_mesh_create uses for the mesh BVH: the host builder (default) or the device
 * builder, which produces the identical node array and item permutation (bvh.cs:285-476) on the
 * context's GPU (2..64 bins). No counterpart in the reference. */
#define SR_BVH_BUILDER_HOST 0
#define SR_BVH_BUILDER_DEVICE 1
int sr_set_bvh_builder(sr_ctx* ctx, int32_t builder);

/* ---- Scene (scene.cs:507-602) -------------------------------------------------------------- */
/* Scene.SetSky(new SkyTexture(texture, multiplier, rotYaw)) (scene.cs:543-551, 389-404). */
int sr_scene_set_sky_texture(sr_ctx* ctx, int32_t texture, float multiplier, float rot_yaw);
/* Scene.SetSky(new SkyProcedural(...)) (scene.cs:341-352). */
int sr_scene_set_sky_procedural(sr_ctx* ctx, const sr_sky_procedural* sky);
/* Scene.SetFog (scene.cs:553-561). */
int sr_scene_set_fog(sr_ctx* ctx, const sr_fog* fog);
/* Scene medium from sr_fog_volume (replaces a height fog set before; no reference counterpart). */
int sr_scene_set_fog_volume(sr_ctx* ctx, const sr_fog_volume* volume);
/* Scene.AddObject(new Object(name, trans, material, shape)) (scene.cs:563-571, 214-223).
 * shape_kind 0 = mesh (mesh id), 1 = Sphere (shape_params = center xyz, radius; math.cs:1171),
 * 2 = AABox (shape_params = min xyz, max xyz; math.cs:936). */
#define SR_SHAPE_MESH 0
#define SR_SHAPE_SPHERE 1
#define SR_SHAPE_BOX 2
int sr_scene_add_object(sr_ctx* ctx, const sr_transform* trans, const sr_material* material, int32_t shape_kind,
                        int32_t mesh, const float* shape_params, int32_t* out_id);
/* Scene.Build (scene.cs:517-541): scene BVH (32 bins, sahCostIntersect 10) and device upload.
 * Idempotent like the reference. */
int sr_scene_build(sr_ctx* ctx);

/* ---- Renderer (renderer.cs:40-237) --------------------------------------------------------- */
/* Renderer ctor (renderer.cs:40-96): builds the scene if needed, allocates the seven AOVs and
 * starts rendering asynchronously. */
int sr_render_begin(sr_ctx* ctx, const sr_view* view, const sr_render_params* params);
/* Renderer.Tick (renderer.cs:98-105): returns (completed blocks, total blocks) in the ctor's
 * block_size x block_size blocks; a block counts as completed once every one of its pixels has its
 * final AOVs. Like the reference's semaphore (one release per completed block) each call consumes
 * one completion: it returns at once while completions are unconsumed and otherwise blocks until
 * the next block completes; (total, total) without blocking once all are done. The GPU renders
 * sample-major over whole bands (tick_bands), so blocks complete in bursts: at the end of each
 * band with fixed sample counts, and as pixels converge with adaptive sampling. */
int sr_render_tick(sr_ctx* ctx, uint32_t* step, uint32_t* total);
/* Copies the AOVs of renderer.cs:14-20 to host arrays (any may be null): Radiance and
 * RadianceFog (w*h*3), Normals (w*h*3, zero when no hit), Uv (w*h*2), Depth (w*h, +inf when no
 * hit), Samples (w*h), Variance (w*h). Waits for the render to finish. */
int sr_render_read(sr_ctx* ctx, float* radiance, float* radiance_fog, float* normals, float* uv, float* depth,
                   uint32_t* samples, float* variance);
/* Device addresses of the same seven AOVs (valid until the next sr_render_begin / sr_destroy),
 * for callers that keep the image resident in HBM. No counterpart in the reference. */
int sr_render_device_aovs(sr_ctx* ctx, void** radiance, void** radiance_fog, void** normals, void** uv,
                          void** depth, void** samples, void** variance);
/* Albedo AOV (w*h*3): mean of the primary surface's Surface.Albedo over the samples whose camera
 * ray hit a surface (the samples Depth averages over), 0 where none did. A denoiser guide with no
 * counterpart in the reference (its Fragment, scene.cs:154-159, has no albedo). */
int sr_render_read_albedo(sr_ctx* ctx, float* albedo);

/* ---- Compositor (compose.cs:101-204) ------------------------------------------------------- */
/* Compositor.Filter over the last render's AOVs (joint bilateral, variance-driven). Writes w*h*3
 * floats to out_rgb (host). */
int sr_denoise_bilateral(sr_ctx* ctx, const sr_denoise_params* params, float* out_rgb);
/* The same filter with the albedo AOV as a fourth guide: on pixels with a surface, a neighbour's
 * weight is also multiplied by exp(-|albedo_c - albedo_n|^2 / (2 albedo_limit^2)), applied after
 * the depth term (albedo_limit > 0). No counterpart in the reference. */
int sr_denoise_bilateral_guided(sr_ctx* ctx, const sr_denoise_params* params, float albedo_limit, float* out_rgb);
/* The same filter on the per-pixel sample mean Radiance * Samples / min(Samples, max_samples)
 * instead of the Radiance AOV (the reference divides capped pixels by M + 1, renderer.cs:205-217).
 * No counterpart in the reference. */
int sr_denoise_bilateral_mean(sr_ctx* ctx, const sr_denoise_params* params, float* out_rgb);
/* Same filter on caller-provided device arrays (radiance/normals: w*h*3, depth/variance: w*h). */
int sr_denoise_bilateral_device(sr_ctx* ctx, const sr_denoise_params* params, const float* d_radiance,
                                const float* d_normals, const float* d_depth, const float* d_variance,
                                uint32_t width, uint32_t height, float* d_out_rgb);

/* Tonemapper (compose.cs:5). */
#define SR_TONEMAP_REINHARD 0
#define SR_TONEMAP_LINEAR_SMOOTH 1

/* Compositor ctor arguments (compose.cs:22-34). */
typedef struct sr_compose_params {
    sr_denoise_params denoise;
    int32_t tonemapper;
    float exposure;
    float fog_radius;   /* denoiseFogRadius: Gaussian radius as a fraction of sqrt(w*h) */
    float fog_strength; /* denoiseFogStrength: neighbour weight scale */
} sr_compose_params;

/* Compositor.Compose (compose.cs:74-99) of the last render without the overlay: Filter of the
 * radiance + FilterFog of the fog radiance, Tonemap, Color.ToPixel. Writes w*h*3 bytes (R, G, B). */
int sr_compose(sr_ctx* ctx, const sr_compose_params* params, uint8_t* out_rgb);
/* No counterpart in the reference: sr_compose for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Same filters,
 * tonemap and bytes as sr_compose, on caller device arrays (radiance, fog radiance, normals: w*h*3;
 * depth, variance: w*h) into caller device memory (d_out_rgb: w*h*3 bytes), asynchronous on the
 * context's stream. It calls no allocating routine: its two intermediate images live in a static
 * device array of the library (48 MiB per GPU, held until the process ends), so w*h is limited to
 * 2,097,152 pixels (SR_ERR_LIMIT above). Calls that overlap on one GPU share that array and race;
 * the filter counters (sr_counters_read_all) are not collected by this call. */
int sr_compose_device_noalloc(sr_ctx* ctx, const sr_compose_params* p, const float* d_radiance,
                              const float* d_radiance_fog, const float* d_normals, const float* d_depth,
                              const float* d_variance, uint32_t width, uint32_t height, uint8_t* d_out_rgb);
/* Compositor.Preview (compose.cs:57-65) without the overlay: Tonemap(radiance + fog radiance). */
int sr_preview(sr_ctx* ctx, int32_t tonemapper, float exposure, uint8_t* out_rgb);
/* Compositor.FilterFog (compose.cs:206-245) of the last render's fog radiance: w*h*3 floats. */
int sr_filter_fog(sr_ctx* ctx, float fog_radius, float fog_strength, float* out_rgb);
/* Image.WriteBmp (image.cs:147-191): 24-bit top-down BMP of w*h*3 RGB bytes. encode writes the file
 * bytes to out (null: only *size is set); write_bmp stores them at path (Image.Save). Host only. */
int sr_image_encode_bmp(const uint8_t* rgb, uint32_t width, uint32_t height, uint8_t* out, uint64_t* size);
int sr_image_write_bmp(const char* path, const uint8_t* rgb, uint32_t width, uint32_t height);

/* ---- Diagnostics --------------------------------------------------------------------------- */
/* Sample-0 primary hit of every pixel (renderer.cs:156,175; math.cs:1613-1636; scene.cs:602-625):
 * object index, triangle index (-1 for analytic shapes / misses) and world distance (+inf). */
int sr_trace_primary(sr_ctx* ctx, const sr_view* view, uint32_t width, uint32_t height, int32_t* out_object,
                     int32_t* out_triangle, float* out_dist);
/* The reference-order BVH of a mesh (mesh >= 0) or of the scene (mesh = -1): node_count nodes of
 * 8 words (min xyz, max xyz as f32, child and shapeCount as i32; bvh.cs:26-31) and the item
 * permutation (bvh.cs:53). Pass null arrays to query node_count / item_count only. */
int sr_bvh_export(sr_ctx* ctx, int32_t mesh, float* nodes, int32_t* items, uint32_t* node_count,
                  uint32_t* item_count);
/* Bvh.GetStats (bvh.cs:79-125, the BvhStats record bvh.cs:5-14) of a mesh's BVH (mesh >= 0) or the
 * scene's (mesh = -1, after sr_scene_build), as Mesh.Describe (mesh.cs:87-96) and the scene dump
 * (program.cs:139-144) print them. All zero for an empty BVH. */
typedef struct sr_bvh_stats {
    int32_t node_count;
    int32_t depth_max;
    int32_t leaf_count;
    int32_t leaf_size_min;
    int32_t leaf_size_max;
    float leaf_size_avg;
    float leaf_depth_avg;
    float sah_cost;
} sr_bvh_stats;
int sr_bvh_stats_get(sr_ctx* ctx, int32_t mesh, sr_bvh_stats* out);
/* Render-path counters in Counters.Type order subset (counters.cs:24-79):
 * [0] Pixel [1] Sample [2] SampleBounce [3] SampleHit [4] SampleMiss [5] SampleTerminate
 * [6] SampleFogScatter [7] SampleFogEscape [8] SamplePixelConverged [9] SamplePixelMaxed
 * [10] SceneTrace [11] SceneOcclude [12] MeshIntersect [13] BvhIntersectNode
 * [14] BvhIntersectShape. Collected only when enabled with sr_counters_enable. */
#define SR_COUNTER_COUNT 15
int sr_counters_enable(sr_ctx* ctx, int32_t enable);
int sr_counters_read(sr_ctx* ctx, int64_t* out);
/* The whole Counters.Type taxonomy (counters.cs:24-79), in the reference's order, accumulated over
 * the context's work since creation or the last sr_counters_reset, the way the reference's one
 * Counters object accumulates over a program run. Collected while sr_counters_enable is on:
 *  - mesh creation: MeshTriangle, TimeMeshBvhBuild (mesh.cs:38-45); scene build: SceneObject,
 *    SceneBvhNodes, SceneBvhDepth, TimeSceneBvhBuild (scene.cs:528-540); sr_obj_load: TimeObjLoad
 *    (obj.cs:147);
 *  - every finished render: Worker (one per concurrent row slice, the GPU analogue of a worker
 *    thread), Block, the render-path counters of sr_counters_read, SampleVarianceMax,
 *    SampleCountMin, SampleCountMax (renderer.cs:223-226) and TimeRender (sr_render_begin to the
 *    last AOV written; program.cs:177 times the same span around Tick);
 *  - every bilateral / fog filter: FilterSample, DenoiseEarlyOut, DenoiseWeightMax, DenoiseMaxLum,
 *    DenoiseMaxLumBoost, DenoiseReject* (compose.cs:101-245); sr_compose: TimeCompose.
 * Sum counters start at 0; the max-reduced ones (SampleVarianceMax, SampleCountMax, Denoise*Max*)
 * start at INT64_MIN and SampleCountMin at INT64_MAX, the reference's "unset" values
 * (counters.cs:221-226); float ones hold (long)(value * 1e4) (counters.cs:214-219). Entries the
 * library has no part in (Overlay*, Rt*, BvhOverlap*, MeshIntersectAny, TimeSetup, TimeTotal) stay
 * unset for the host's own Counters to fill. Waits for a running render. count: entries to copy
 * (at most SR_COUNTER_TYPE_COUNT). */
enum sr_counter_type {
    SR_CT_WORKER, SR_CT_BLOCK, SR_CT_PIXEL, SR_CT_SAMPLE, SR_CT_SAMPLE_BOUNCE, SR_CT_SAMPLE_HIT, SR_CT_SAMPLE_MISS,
    SR_CT_SAMPLE_TERMINATE, SR_CT_SAMPLE_FOG_SCATTER, SR_CT_SAMPLE_FOG_ESCAPE, SR_CT_SAMPLE_VARIANCE_MAX,
    SR_CT_SAMPLE_PIXEL_CONVERGED, SR_CT_SAMPLE_PIXEL_MAXED, SR_CT_SAMPLE_COUNT_MIN, SR_CT_SAMPLE_COUNT_MAX,
    SR_CT_SCENE_OBJECT, SR_CT_SCENE_BVH_NODES, SR_CT_SCENE_BVH_DEPTH, SR_CT_SCENE_TRACE, SR_CT_SCENE_OCCLUDE,
    SR_CT_MESH_TRIANGLE, SR_CT_MESH_INTERSECT, SR_CT_MESH_INTERSECT_ANY, SR_CT_BVH_INTERSECT_NODE,
    SR_CT_BVH_INTERSECT_SHAPE, SR_CT_BVH_OVERLAP_NODE, SR_CT_BVH_OVERLAP_SHAPE, SR_CT_FILTER_SAMPLE,
    SR_CT_DENOISE_EARLY_OUT, SR_CT_DENOISE_WEIGHT_MAX, SR_CT_DENOISE_MAX_LUM, SR_CT_DENOISE_MAX_LUM_BOOST,
    SR_CT_DENOISE_REJECT_MISSING, SR_CT_DENOISE_REJECT_NORMAL, SR_CT_DENOISE_REJECT_DEPTH, SR_CT_DENOISE_REJECT_LUM,
    SR_CT_OVERLAY_LINE, SR_CT_OVERLAY_TEXT, SR_CT_RT_PEAK_WORKING_SET, SR_CT_RT_GC_ALLOCATED_BYTES,
    SR_CT_RT_GC_CURRENT_BYTES, SR_CT_RT_GC_GEN0, SR_CT_RT_GC_GEN1, SR_CT_RT_GC_GEN2, SR_CT_TIME_SETUP,
    SR_CT_TIME_OBJ_LOAD, SR_CT_TIME_MESH_BVH_BUILD, SR_CT_TIME_SCENE_BVH_BUILD, SR_CT_TIME_RENDER,
    SR_CT_TIME_COMPOSE, SR_CT_TIME_TOTAL, SR_COUNTER_TYPE_COUNT
};
int sr_counters_read_all(sr_ctx* ctx, int64_t* out, uint32_t count);
int sr_counters_reset(sr_ctx* ctx);
/* The share of BvhIntersectNode, BvhIntersectShape, SceneTrace and SceneOcclude (out[0..3]) done by
 * the fused launch that finishes paths from the tail segment on (sr_set_tail_segment); the rest
 * was done by the per-segment traversal launches. No counterpart in the reference. */
int sr_counters_read_fused(sr_ctx* ctx, int64_t* out);
/* The reference-only share of BvhIntersectNode, BvhIntersectShape, MeshIntersect and SceneOcclude
 * (out[0..3]) in the per-segment traversal launches: traversal the reference does that a render
 * without counters skips. The reference traces the direct-light shadow ray before testing whether
 * the light lies behind the shading or the geometric normal (scene.cs:774-783), and walks every
 * shadow ray with its closest-hit traversal (scene.cs:576-600), where an any-hit walk decides
 * occlusion at the first hit. While counters are enabled the tracer traces those rays and walks
 * shadow rays the reference's way, so sr_counters_read follows the reference's taxonomy; the work
 * of a render without counters is the totals minus this share. No counterpart in the reference. */
int sr_counters_read_reference_only(sr_ctx* ctx, int64_t* out);
/* Later renders trace segments 0..segment-1 as separate wavefront launches and finish every path
 * from `segment` on in one fused launch (by default 3, or 4 with 8 or more samples in flight, and
 * in scenes with fog or a medium 5, or 6 with 4 or more samples in flight; a
 * value above the bounce count disables it; results are identical either way). The environment variable SHARPRAY_TAIL_SEGMENT sets the
 * default of new contexts. No counterpart in the reference. */
int sr_set_tail_segment(sr_ctx* ctx, uint32_t segment);
/* Later renders run every band as `slices` concurrent row slices (whole 8-row tile rows), each with
 * its own queues, path records and HIP stream, so the launch tails of one slice overlap the work of
 * the others (default 1, one wavefront per band; 2 measured 6 % faster on the 1080p frame; results
 * are identical either way). The
 * environment variable SHARPRAY_SLICES sets the default of new contexts. No counterpart in the
 * reference (its worker threads claim 32x32 blocks, renderer.cs:122-151). */
int sr_set_slices(sr_ctx* ctx, uint32_t slices);
/* Later renders keep up to `samples` samples of every pixel in flight at once (1 to 32, default 32;
 * the largest power of two not above it, the sample count and a 32 GiB budget is used) when the
 * samples are independent: counter RNG policy, a fixed sample count (min = max or threshold 0) and
 * counters off; other renders run one sample at a time. Every launch then holds that many samples
 * of every pixel, so the fixed cost of a launch's slowest waves is paid once per batch of samples. Each pixel still folds its samples into its sums in sample order, so results
 * are identical for any value. The environment variable SHARPRAY_IN_FLIGHT sets the default of new
 * contexts. No counterpart in the reference (its workers run a pixel's samples one after another,
 * renderer.cs:173-212). */
int sr_set_samples_in_flight(sr_ctx* ctx, uint32_t samples);
/* Kernel timing of the next renders: with timing enabled, an event is recorded on the slice's
 * stream after every launch; sr_profile_read sums the spans per kernel kind of the last render
 * over all slices (concurrent slices overlap, so sums can exceed the frame time):
 * ms[4] / launches[4] for [0] trace (closest hit of a segment's rays plus the previous segment's
 * shadow rays), [1] shade, [2] trace of a sample's last shadow rays, [3] raygen + accumulate.
 * No counterpart in the reference (its TimeScope covers whole phases, counters.cs:4-20). */
int sr_profile_enable(sr_ctx* ctx, int32_t enable);
int sr_profile_read(sr_ctx* ctx, double* ms, uint32_t* launches);
/* ms[4]: per kernel kind of the last render, the time during which at least one launch of that
 * kind was running on any slice (the union of the launch spans). No counterpart in the reference. */
int sr_profile_busy(sr_ctx* ctx, double* ms);
/* Seconds of GPU time of the last render (HIP events around the whole frame). */
int sr_render_gpu_seconds(sr_ctx* ctx, double* out_seconds);

/* ---- Adjoint rendering (no counterpart in the reference) -------------------------------------
 * Differentiable parameters, in order: per object (in sr_scene_add_object order) albedo r, g, b
 * and roughness (4 floats), then, for a texture sky, the sky texture's texels (w*h*3, row-major).
 * The backward pass returns d/dθ of sum(d_image * Radiance) for the last forward render, replaying
 * every path with its counter-RNG stream; it differentiates the BSDF values, the sky radiance, the
 * MIS weights and the density of the reference's BSDF sampling (whose ClampToGeoSide mirror makes
 * it differ from OpaquePdf). Opaque scenes without fog; fixed sample counts. */
int sr_adjoint_param_count(sr_ctx* ctx, uint32_t* count);
int sr_adjoint_params_get(sr_ctx* ctx, float* params);
/* Sets the parameters (materials and sky texels; the sky's sampling tables are rebuilt). */
int sr_adjoint_params_set(sr_ctx* ctx, const float* params);
/* Renders like sr_render_begin + sr_render_read's Radiance (w*h*3, may be null) and keeps the view
 * and parameters for the backward pass. Requires SR_RNG_COUNTER, min = max samples, threshold 0,
 * at most 16 bounces. */
int sr_adjoint_forward(sr_ctx* ctx, const sr_view* view, const sr_render_params* params, float* image);
/* grad (param_count floats) = d sum(d_image * Radiance) / d params; d_image is w*h*3 (host). */
int sr_adjoint_backward(sr_ctx* ctx, const float* d_image, float* grad);

/* ---- Kernel-predicting denoiser (no counterpart in the reference) ----------------------------
 * A 5-level U-Net (32/48/64/96/128 channels, two 3x3 conv + ReLU per level, 2x2 average pooling,
 * nearest upsampling with skip concatenation) on bf16 MFMA with fp32 accumulation. Inputs per
 * pixel: log(1 + radiance), albedo, normal, 1 / (1 + depth) (0 without a surface); the last conv
 * predicts 4 x 25 logits: four 5x5 kernels (softmax-normalised) applied in cascade to the noisy
 * radiance with their taps 1, 2, 4 and 8 pixels apart (edge-clamped). Weights: float blob, per layer [cout][cin][3][3] then bias[cout], layers in
 * the order of sharpray_amd.cnn_layers(); concatenated inputs list the upsampled channels first. */
int sr_cnn_weight_count(uint64_t* count);
int sr_cnn_load(sr_ctx* ctx, const float* weights, uint64_t count);
/* Denoises the last render into out_rgb (host, w*h*3). The network sees the per-pixel sample mean,
 * Radiance * Samples / min(Samples, max_samples): pixels that ran to the sample cap hold M samples
 * divided by M + 1 (renderer.cs:205-217), a bias of 1/(M+1) that no kernel of convex weights could
 * undo. Albedo, normal and depth AOVs are the guides. */
int sr_cnn_denoise(sr_ctx* ctx, float* out_rgb);
/* Same on caller device arrays (rad: the per-pixel sample mean, alb/nrm: w*h*3, dep: w*h), out
 * device w*h*3; asynchronous on the context's stream. */
int sr_cnn_denoise_device(sr_ctx* ctx, const float* d_radiance, const float* d_albedo, const float* d_normals,
                          const float* d_depth, uint32_t width, uint32_t height, float* d_out);
/* Activations of a layer after the last denoise (layer -1: the input features) as floats,
 * [(y * width + x) * channels + c]; out may be null to query the sizes. Layer 1000 + L reads
 * layer L's output gradient (after its ReLU) from the last training step. */
int sr_cnn_layer_read(sr_ctx* ctx, int32_t layer, float* out, uint32_t* channels, uint32_t* width, uint32_t* height);
/* Milliseconds per denoise of the last render, averaged over `iterations` back-to-back device runs
 * (HIP events on the context's stream). */
int sr_cnn_benchmark(sr_ctx* ctx, uint32_t iterations, double* ms_per_run);
/* Training with hand-written kernels: Adam on fp32 master weights starting from the loaded ones;
 * loss mean (out - target)^2 / (target^2 + 0.01) of the denoised radiance against a reference.
 * sr_cnn_train_add keeps one example (noisy AOVs + target radiance, host arrays) in device
 * memory; sr_cnn_train_run takes `steps` steps over the examples in a seeded shuffled order and
 * returns their mean loss; sr_cnn_train_export writes the weights (sr_cnn_weight_count floats);
 * the context's denoiser uses the trained weights meanwhile. */
int sr_cnn_train_begin(sr_ctx* ctx);
int sr_cnn_train_add(sr_ctx* ctx, const float* radiance, const float* albedo, const float* normals, const float* depth,
                     const float* target, uint32_t width, uint32_t height);
int sr_cnn_train_run(sr_ctx* ctx, uint32_t steps, float learning_rate, uint32_t seed, float* mean_loss);
int sr_cnn_train_export(sr_ctx* ctx, float* weights);
/* Loss and gradient (blob layout) of one stored example, without an update. */
int sr_cnn_train_gradient(sr_ctx* ctx, uint32_t example, float* gradient, float* loss);
/* Floating-point operations of one denoise at width x height (2 per multiply-add). */
int sr_cnn_flops(uint32_t width, uint32_t height, double* flops);

/* ---- Asset ingest (host only; the reference's ObjLoader, ImageHdr and Image loaders) --------- */
/* ObjLoader.Load (obj.cs:138-258) with ObjConfig (obj.cs:127-134): parses an OBJ file and the MTL
 * files it names, and adds one mesh object per o / g / usemtl block to the scene (textures of
 * map_Kd / map_Pm / map_Pr / map_bump via Image.Load, TGA only). transform null = identity;
 * material null = Material(White, roughness 1) for objects without an MTL material; smooth_normals
 * != 0 computes area-weighted vertex normals when the file has no vn. Errors (SR_ERR_INVALID) carry
 * the reference's exception texts; objects added before an error stay in the scene. */
int sr_obj_load(sr_ctx* ctx, const char* path, const sr_transform* transform, const sr_material* material,
                int32_t smooth_normals, uint32_t* objects_added);
/* ImageHdr.Load (image.cs:307-382) + Texture.FromHdr (texture.cs:154-163): Radiance RGBE .hdr
 * (flat or RLE scanlines, -Y h +X w) to linear float RGB w*h*3, luminance clamped to
 * max_luminance. rgb null: only width/height. */
int sr_hdr_load(const char* path, float max_luminance, float* rgb, uint32_t* width, uint32_t* height);
int sr_hdr_decode(const uint8_t* data, uint64_t size, float max_luminance, float* rgb, uint32_t* width,
                  uint32_t* height);
/* Image.LoadTga (image.cs:217-266): uncompressed 24-bit / 8-bit grey TGA to top-down RGB bytes. */
int sr_tga_decode(const uint8_t* data, uint64_t size, uint8_t* rgb, uint32_t* width, uint32_t* height);
/* What the scene holds (for hosts and tests): counts, a texture's texels, a mesh's Triangle
 * records (36 floats each, as given to sr_mesh_create) and an object's transform, material and
 * shape. Arrays may be null to query sizes. */
int sr_scene_counts(sr_ctx* ctx, uint32_t* textures, uint32_t* meshes, uint32_t* objects);
int sr_texture_read(sr_ctx* ctx, int32_t id, float* rgb, uint32_t* width, uint32_t* height);
int sr_mesh_read(sr_ctx* ctx, int32_t id, float* triangles, uint32_t* count);
int sr_scene_object_read(sr_ctx* ctx, int32_t id, sr_transform* trans, sr_material* material, int32_t* shape_kind,
                         int32_t* mesh);

/* ---- Host data preparation (the C# host computes these in Triangle's constructors) ---------- */
/* Triangle ctors (math.cs:1503-1531) and Mesh.ComputeSmoothNormals (mesh.cs:99-123): pos n*9,
 * normals n*9 or null, uv n*6 or null (null = the OBJ defaults (0,0),(1,0),(0,1); obj.cs:417-427).
 * smooth != 0 with null normals recomputes area-weighted vertex normals. out: n*36. */
int sr_triangles_build(const float* pos, const float* normals, const float* uv, uint32_t count, int32_t smooth,
                       float* out_triangles);

#ifdef __cplusplus
}
#endif

#endif /* SHARPRAY_H */
