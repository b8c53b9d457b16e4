/*
 * spk_pt.h — C ABI of the MI355X path-tracer hot path.
 *
 * A C++ host that keeps Sparkle's scene, material, camera and renderer semantics swaps its
 * per-pixel sample loop for these calls. Nothing C++ crosses the boundary: plain structs,
 * pointers and sizes. Every call returns SPK_OK (0) or a negative SPK_ERR_* code; none aborts
 * or throws; spk_last_error() gives the message of the last failing call on a context.
 * A context is used by one host thread at a time; separate contexts are independent.
 *
 * Reference interfaces replaced (paths relative to the Sparkle repository):
 *   spk_scene_upload  SceneRenderProxy::UpdateBVH / TLAS::Build / BLAS::Build
 *                     (libraries/source/renderer/proxy/SceneRenderProxy.cpp:269-297, 31-69;
 *                      libraries/source/renderer/proxy/MeshRenderProxy.cpp:31-65, 310-315)
 *   spk_trace_closest SceneRenderProxy::Intersect<false> (SceneRenderProxy.h:119,
 *                     SceneRenderProxy.cpp:299-309) down to MeshRenderProxy::IntersectTriangle
 *                     (MeshRenderProxy.cpp:128-194)
 *   spk_set_camera    CameraRenderProxy posture + focus plane (CameraRenderProxy.h:18-57,
 *                     CameraRenderProxy.cpp:113-136)
 *   spk_render_frame  CPURenderer::BasePass -> RenderPixel -> SamplePixel
 *                     (libraries/include/renderer/renderer/CPURenderer.h:37,
 *                      libraries/source/renderer/renderer/CPURenderer.cpp:240-450)
 *   spk_accumulate    CPURenderer::DenoisePass, temporal part (CPURenderer.cpp:506-552)
 *   spk_tonemap       CPURenderer::ToneMappingPass (CPURenderer.cpp:554-576)
 * Materials mirror MaterialRenderProxy::MaterialRenderData (MaterialRenderProxy.h:39-58);
 * the camera mirrors CameraRenderProxy::UniformBufferData (CameraRenderProxy.h:18-28).
 *
 * Device pointers: arguments named *_dev are HIP device pointers on the context's device.
 * Everything else is host memory, copied by the call.
 */
#ifndef SPK_PT_H
#define SPK_PT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPK_OK 0
#define SPK_ERR_INVALID_ARGUMENT (-1)
#define SPK_ERR_HIP (-2)
#define SPK_ERR_OUT_OF_MEMORY (-3)
#define SPK_ERR_NO_SCENE (-4)
#define SPK_ERR_INVALID_SCENE (-5)
#define SPK_ERR_STACK_DEPTH (-6)
#define SPK_ERR_INVALID_RAY (-7) /* some rays were NaN/inf/zero-length; they were reported as misses */
#define SPK_ERR_NO_DEVICE (-8)
#define SPK_ERR_UNSUPPORTED (-9)

#define SPK_MISS 0xFFFFFFFFu

typedef struct spk_context spk_context;

/* 32-byte ray record: origin, tmin, direction (need not be normalised), tmax. A hit needs
   tmin < t and t <= tmax. */
typedef struct spk_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} spk_ray;

/* 8-byte hit record: global triangle id (SPK_MISS when nothing is hit) and distance t. */
typedef struct spk_hit {
    uint32_t prim;
    float t;
} spk_hit;

/* One triangle mesh. Global triangle ids number the triangles of every instance in the order
   the instances are given: id = (triangles of earlier instances) + face index. */
typedef struct spk_mesh_desc {
    const float *positions; /* 3 floats per vertex */
    uint32_t num_vertices;
    const uint32_t *indices; /* 3 per triangle */
    uint32_t num_triangles;
    const float *normals;  /* 3 per vertex, or NULL */
    const float *tangents; /* 4 per vertex (xyz, handedness), or NULL -> (1,0,0,1) */
    const float *uvs;      /* 2 per vertex, or NULL */
} spk_mesh_desc;

/* Instance of a mesh with a material. world_from_local is row-major 3x4. Instances are baked
   into world space at upload (one flattened tree), or kept as TLAS leaves over object-space BLASes
   with spk_build_options.two_level. */
typedef struct spk_instance_desc {
    uint32_t mesh;
    uint32_t material;
    float world_from_local[12];
} spk_instance_desc;

#define SPK_MATERIAL_PBR 0u
#define SPK_MATERIAL_DIELECTRIC 1u

/* Field set of MaterialRenderData (MaterialRenderProxy.h:39-58) and MaterialResource's textures
   (io/Material.h:9-23). A texture field holds 0 for none (so a zeroed struct has no textures) or
   k + 1 for entry k of the context's texture table (spk_textures_upload); then the getters of
   MaterialRenderProxy.h:117-162 apply: base = tex(uv).rgb * base_color, metallic = tex(uv).b *
   metallic, roughness = tex(uv).g * roughness, emissive = tex(uv).rgb * emissive, and the normal
   map perturbs the shading normal in the tangent frame (MeshRenderProxy.cpp:350-356). */
typedef struct spk_material_desc {
    uint32_t type; /* SPK_MATERIAL_* */
    float base_color[3];
    float emissive[3];
    float metallic;
    float roughness;
    float eta;
    uint32_t base_color_texture;
    uint32_t metallic_roughness_texture;
    uint32_t emissive_texture;
    uint32_t normal_texture;
} spk_material_desc;

#define SPK_TEXTURE_RGBA8_UNORM 0u /* 4 bytes per texel, value / 255 */
#define SPK_TEXTURE_RGBA8_SRGB 1u  /* as UNORM, then the sRGB decode of Image::Sample */
#define SPK_TEXTURE_RGBA32F 2u     /* 16 bytes per texel */

/* One 2D image (io/Image.h): row-major texels, sampled bilinearly with wrapping (Image.h:112-152). */
typedef struct spk_texture_desc {
    uint32_t width;
    uint32_t height;
    uint32_t format; /* SPK_TEXTURE_* */
    uint32_t reserved;
    const void *pixels;
} spk_texture_desc;

typedef struct spk_build_options {
    uint32_t bvh_width;     /* 8 (default) or 4 */
    uint32_t max_leaf_tris; /* triangles per leaf slot, 1..3 (default 3) */
    uint32_t num_threads;   /* host build threads, 0 = all */
    uint32_t two_level;     /* 1: one BLAS per mesh in object space under a TLAS over the instances
                               (SceneRenderProxy's TLAS / MeshRenderProxy's BLAS), so that
                               spk_scene_update_instances moves instances without rebuilding any
                               BLAS; 0: one tree over the flattened triangles (default, fastest) */
} spk_build_options;

typedef struct spk_scene_stats {
    uint64_t num_triangles;
    uint64_t num_nodes;
    uint32_t bvh_width;
    uint32_t depth;           /* deepest node level (root = 1) */
    uint64_t node_bytes;      /* bytes per node record */
    uint64_t triangle_bytes;  /* bytes per triangle record */
    uint64_t device_bytes;    /* scene bytes resident on the device */
    uint64_t degenerate_tris; /* zero-area triangles (never hit) */
    double sah_cost;          /* SAH cost of the wide tree (node traversal = 1, triangle = 1) */
    double build_seconds;
} spk_scene_stats;

/* Camera exactly as SetupViewRay consumes it (CPURenderer.cpp:220-238). */
typedef struct spk_camera_desc {
    float position[3];
    float right[3];
    float up[3];
    float front[3];
    float lower_left[3];
    float max_u[3];
    float max_v[3];
    float lens_radius;
    float far_plane;
} spk_camera_desc;

/* RenderConfig::DebugMode (RenderConfig.h:33-47); every mode 0-11 is served. Modes 2, 5 and 6
   rewrite the colour after the loop and take neither SPK_FRAME_NEE nor SPK_FRAME_CACHE. */
#define SPK_DEBUG_COLOR 0u
#define SPK_DEBUG_DEBUG 1u
#define SPK_DEBUG_RAY_DEPTH 2u
#define SPK_DEBUG_NORMAL 3u
#define SPK_DEBUG_RAY_DIRECTION 4u
#define SPK_DEBUG_INDIRECT_LIGHTING 5u
#define SPK_DEBUG_DIRECTIONAL_LIGHTING 6u
#define SPK_DEBUG_METALLIC 7u
#define SPK_DEBUG_ROUGHNESS 8u
#define SPK_DEBUG_ALBEDO 9u
#define SPK_DEBUG_EMISSIVE 10u
#define SPK_DEBUG_DEPTH 11u

#define SPK_FRAME_SEED64 1u /* compose the pixel seed in 64 bits (no uint32 wrap) */
#define SPK_FRAME_NEE 2u    /* next-event estimation on emissive triangles (not in the reference) */
#define SPK_FRAME_CACHE 4u  /* terminate paths into the learned radiance cache and train it (needs
                               spk_cache_create; not in the reference) */
#define SPK_FRAME_NO_SORT 0x80000000u /* shade in queue order instead of material order */

typedef struct spk_frame_desc {
    uint32_t width;
    uint32_t height;
    uint32_t frame_seed; /* dispatched sample count, CPURenderer.cpp:432 */
    uint32_t max_bounce;
    uint32_t debug_mode;
    uint32_t sky_enabled;
    float sky_color[3];
    uint32_t flags; /* SPK_FRAME_* */
    float output_limit; /* per-sample clamp; 0 means the reference's 6 (CameraRenderProxy.h:16) */
} spk_frame_desc;

typedef struct spk_frame_stats {
    uint64_t rays_traced; /* extend rays (one per path segment) */
    uint64_t shadow_rays;
    double milliseconds;  /* device time of the frame */
} spk_frame_stats;

/* ---- context ------------------------------------------------------------------------------ */
int spk_create(int device, spk_context **out);
int spk_destroy(spk_context *ctx);
const char *spk_last_error(const spk_context *ctx);
/* Run all work on this HIP stream (hipStream_t); NULL restores the context's own stream. */
int spk_set_stream(spk_context *ctx, void *hip_stream);
void *spk_get_stream(spk_context *ctx);
int spk_synchronize(spk_context *ctx);
int spk_device_alloc(spk_context *ctx, size_t bytes, void **out_dev);
int spk_device_free(spk_context *ctx, void *dev);
int spk_copy_to_device(spk_context *ctx, void *dst_dev, const void *src, size_t bytes);
int spk_copy_to_host(spk_context *ctx, void *dst, const void *src_dev, size_t bytes);

/* ---- scene -------------------------------------------------------------------------------- */
int spk_scene_upload(spk_context *ctx, const spk_mesh_desc *meshes, uint32_t num_meshes,
                     const spk_instance_desc *instances, uint32_t num_instances,
                     const spk_material_desc *materials, uint32_t num_materials,
                     const spk_build_options *options);
int spk_scene_stats_get(spk_context *ctx, spk_scene_stats *out);
/* Moves the instances of a two-level scene (spk_build_options.two_level): world_from_local holds
   num_instances row-major 3x4 transforms in upload order (num_instances = the uploaded count).
   Replaces SceneRenderProxy::UpdateBVH's TLAS rebuild on primitive moves (SceneRenderProxy.cpp:
   269-297): the TLAS is rebuilt over the moved world boxes, every BLAS is kept, and the world-space
   shading data of the instances whose transform changed is recomputed. Hits after the update are
   those of a fresh upload of the moved scene. SPK_ERR_UNSUPPORTED for a flattened scene. */
int spk_scene_update_instances(spk_context *ctx, const float *world_from_local, uint32_t num_instances);
/* Analytic sphere (SpherePrimitive / SphereRenderProxy.cpp:5-48, SphereRenderProxy.h:22-36):
   hit t = -half_b -+ sqrt(half_b^2 - c) with half_b = d.(o - center), c = |o - center|^2 - r^2
   (near root from outside, far root from inside; ray directions are taken as unit length, as in
   the reference), normal (p - center)/|p - center|, tangent the normal's major axis, uv 0. */
typedef struct spk_sphere_desc {
    float center[3];
    float radius;
    uint32_t material;
    uint32_t reserved[3];
} spk_sphere_desc;

/* Replaces the uploaded scene's spheres (none after spk_scene_upload). Sphere k has the global
   primitive id num_triangles + k in hit records. */
int spk_scene_set_spheres(spk_context *ctx, const spk_sphere_desc *spheres, uint32_t num_spheres);

/* Replaces the context's texture table (shared by the scene and batch uploads that follow; a
   material naming a texture index beyond the table is rejected at upload). Call before
   spk_scene_upload / spk_batch_upload; changing it invalidates an uploaded scene. */
int spk_textures_upload(spk_context *ctx, const spk_texture_desc *textures, uint32_t num_textures);

/* Host-only check of the BVH builder (no device needed): flattens and builds exactly as
   spk_scene_upload does, verifies the traversal invariants (every triangle referenced once and
   inside each dequantised ancestor box, indices in range, depth as reported) and fills stats.
   Returns SPK_OK or SPK_ERR_INVALID_SCENE; errmsg (may be NULL) receives the reason. */
int spk_build_check(const spk_mesh_desc *meshes, uint32_t num_meshes, const spk_instance_desc *instances,
                    uint32_t num_instances, const spk_build_options *options, spk_scene_stats *stats,
                    char *errmsg, size_t errmsg_size);

/* ---- closest hit -------------------------------------------------------------------------- */
/* Closest hit of n rays (device arrays). Blocks until done. Exact-t ties go to the smaller id. */
int spk_trace_closest(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, spk_hit *hits_dev);
/* Same, enqueued on the context stream without waiting; invalid-ray counts are reported by the
   next blocking call. */
int spk_trace_closest_async(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, spk_hit *hits_dev);
/* Same with rays and hits in host memory (pageable is fine): the batch goes through pinned staging
   buffers in chunks, with host copies, uploads, traversal and downloads overlapped. Blocks until
   done. The hits equal those of spk_trace_closest over the same rays. */
int spk_trace_closest_host(spk_context *ctx, const spk_ray *rays, uint64_t n, spk_hit *hits);
/* No counterpart in the reference. For callers that cannot allocate device memory during a call
   (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
   closest hit as spk_trace_closest in the context's trace configuration, with the coherence sort's
   scratch taken from a fixed device array of the library (44.3 MiB per GPU, held until the
   process ends) instead of context buffers grown on demand; the call allocates nothing. At most
   SPK_NOALLOC_MAX_RAYS rays per call (more: SPK_ERR_INVALID_ARGUMENT). Every such call on one GPU
   shares that array: calls that overlap on one GPU (other streams, other contexts) race. Blocks
   until done; invalid rays are reported as by spk_trace_closest. */
#define SPK_NOALLOC_MAX_RAYS 2097152ull
int spk_trace_closest_noalloc(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, spk_hit *hits_dev);
/* Traversal statistics for n rays: per ray node visits and triangle tests (device arrays of
   uint32, either may be NULL). */
int spk_trace_count(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, uint32_t *nodes_dev,
                    uint32_t *tris_dev);
/* Kernel timing, a diagnostic with no counterpart in the reference (its rays are traced on the CPU):
   records HIP events on the context stream around the closest-hit kernel of each of the next
   `capacity` spk_trace_closest(_async) calls, leaving the coherence sort of a ray_sort configuration
   outside, so that the traversal kernel can be priced against its memory traffic on its own.
   capacity 0 switches it off; a new capacity forgets earlier calls. */
int spk_trace_kernel_timing(spk_context *ctx, uint32_t capacity);
/* Waits for the timed calls and writes each one's kernel time in milliseconds, in call order
   (up to max); count receives the number of timed calls. */
int spk_trace_kernel_times(spk_context *ctx, float *ms, uint32_t max, uint32_t *count);

/* Traversal kernel configuration: BVH width (4 or 8), LDS stack entries per lane, threads per
   block. The stack must hold the tree depth (spk_scene_stats.depth). */
typedef struct spk_trace_config {
    uint32_t bvh_width;
    uint32_t stack;
    uint32_t block;
    uint32_t refill; /* 0: one ray per thread; k > 0: persistent lanes, a wave takes new rays once k of
                        its 64 lanes are idle */
    uint32_t ray_sort; /* 1 (one ray per thread only): trace the rays in the order of a counting sort
                          by direction bin (cube face x 4 x 4) and origin cell (8^3 of the scene box),
                          so a wave's rays start close and point alike; hits keep the input order.
                          Calls of 2^32 rays or more trace in input order. */
} spk_trace_config;

#define SPK_TUNE_WIDTH 1u /* also try the other BVH width (rebuilds the tree on the host) */

int spk_get_trace_config(spk_context *ctx, spk_trace_config *out);
/* Selects a compiled variant; a different bvh_width rebuilds and re-uploads the BVH. */
int spk_set_trace_config(spk_context *ctx, const spk_trace_config *config);
/* Lists the compiled variants valid for the current scene depth (up to max_out). */
int spk_list_trace_configs(spk_context *ctx, spk_trace_config *out, uint32_t max_out, uint32_t *count);
/* Autotuner: times every valid variant on the given rays (device array) and keeps the fastest.
   best_ms receives the fastest time per launch. */
int spk_autotune(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, uint32_t flags, spk_trace_config *chosen,
                 double *best_ms);

/* ---- path tracing ------------------------------------------------------------------------- */
int spk_set_camera(spk_context *ctx, const spk_camera_desc *camera);
/* Sky cube map (SkyRenderProxy::Evaluate with a sky map, SkyRenderProxy.cpp:50-56, Image2DCube::Sample,
   io/Image.cpp:298-357): texture index + 1 of the faces +X, -X, +Y, -Y, +Z, -Z; all zero (or NULL)
   for the gradient sky. Missed rays see it when spk_frame_desc.sky_enabled is set. A new texture
   table clears it. */
int spk_set_sky_cubemap(spk_context *ctx, const uint32_t faces[6]);
/* Host-only cook of a textured sky light (SkyLightCookJob::Execute, SkyLight.cpp:83-135; no device
   needed): samples the equirectangular map (width x height RGBA floats, row-major, as Image::Sample
   does, at CartesianToEquirectangular of each cube texel's direction) into six faces of
   face_size^2 RGBA floats (+X, -X, +Y, -Y, +Z, -Z, rows of face_size texels), each value rounded to
   FP16 like the reference's RGBAFloat16 cube (alpha 1). The reference cooks 1024^2 faces. Upload the
   faces as RGBA32F textures and pass their indices to spk_set_sky_cubemap. */
int spk_sky_cook(const float *equirect_rgba, uint32_t width, uint32_t height, uint32_t face_size,
                 float *faces_rgba);
/* One BasePass: one path per pixel; rgba = clamped colour and valid flag (+1/-1); normal =
   first-hit world normal (may be NULL). Device arrays of width*height*4 / *3 floats. */
int spk_render_frame(spk_context *ctx, const spk_frame_desc *frame, float *rgba_dev, float *normal_dev,
                     spk_frame_stats *stats);
/* accum = rgba + (accum - rgba) * cumulated / (cumulated + spp), on the device. */
int spk_accumulate(spk_context *ctx, const float *rgba_dev, float *accum_dev, uint64_t num_pixels,
                   uint32_t cumulated, uint32_t spp);
/* CPURenderer::DenoisePass spatial part (CPURenderer.cpp:452-525): every invalid pixel (w <= 0) of
   rgba_dev takes the colour of a valid, non-sky neighbour with a normal within dot >= 0.99 (the
   reference's direction order; where several neighbours qualify, the last in row-major order, as
   the reference's loop on one thread). rgba_dev and normal_dev (3 floats per pixel) are
   spk_render_frame's outputs; out_dev receives the filled frame (not in place). */
int spk_spatial_fill(spk_context *ctx, const float *rgba_dev, const float *normal_dev, float *out_dev, uint32_t width,
                     uint32_t height);
/* ACES film curve with exposure, vertically flipped, to RGBA float (CPURenderer.cpp:554-576). */
int spk_tonemap(spk_context *ctx, const float *accum_dev, float *out_dev, uint32_t width, uint32_t height,
                float exposure);

/* ---- batches of independent scenes in one launch ------------------------------------------ */
/* Each scene keeps its own BVH (rebased into one device array), materials, camera and image
   size; one frame renders every scene's pixels with the same persistent kernels, each path finding
   its scene's BVH root. Pixel seeds are per scene (as if the scene were rendered alone), so every
   scene's image is bit-identical to spk_render_frame on that scene alone. NEE and the cache are
   single-scene features. */
typedef struct spk_batch_scene {
    const spk_mesh_desc *meshes;
    uint32_t num_meshes;
    const spk_instance_desc *instances;
    uint32_t num_instances;
    const spk_material_desc *materials;
    uint32_t num_materials;
    spk_camera_desc camera;
    uint32_t width;
    uint32_t height;
} spk_batch_scene;

int spk_batch_upload(spk_context *ctx, const spk_batch_scene *scenes, uint32_t num_scenes,
                     const spk_build_options *options);
/* frame->width/height are ignored; rgba_dev holds the scenes' images back to back, scene s
   starting at pixel offsets[s] (spk_batch_offsets). */
int spk_batch_render(spk_context *ctx, const spk_frame_desc *frame, float *rgba_dev, spk_frame_stats *stats);
/* num_scenes + 1 pixel offsets (the last is the total pixel count). */
int spk_batch_offsets(spk_context *ctx, uint64_t *offsets, uint32_t max_count);
/* Launch schedule of the path tracer's persistent kernels. Images never depend on it. */
typedef struct spk_schedule {
    uint32_t blocks_per_cu; /* persistent grid: 128-thread blocks per CU, 1..32 (default 8) */
    uint32_t material_sort; /* 1: shade each bounce in material order (counting sort; default 1) */
    uint32_t refill;        /* extend after the camera rays (which run one path per thread): a wave
                               takes new paths once this many of its 64 lanes are idle, 1..64
                               (default 16); 0: one path (and one NEE shadow ray) per thread over a
                               full grid, like the camera rays (the per-path tail then refills at 8) */
    uint32_t tail_bounce;   /* segments from this index on run one path per lane to its end (closest
                               hit, shading and shadow ray in one persistent kernel) instead of one
                               bounce-synchronous launch per segment; 0 = none (default), <= 64 */
} spk_schedule;

/* batch = 0: the schedule of spk_render_frame; batch != 0: of spk_batch_render. */
int spk_get_schedule(spk_context *ctx, int batch, spk_schedule *out);
int spk_set_schedule(spk_context *ctx, int batch, const spk_schedule *schedule);
/* Schedule autotuners: time blocks_per_cu {2, 4, 8, 16} x material_sort {0, 1} x refill {0, 16, 64} on
   the median of five frames (after a capture frame per candidate), then tail_bounce {1, 2, 3, 4, 5} (below max_bounce)
   on the three fastest of those; keep and return the fastest. */
int spk_render_autotune(spk_context *ctx, const spk_frame_desc *frame, float *rgba_dev, spk_schedule *chosen,
                        double *best_ms);
int spk_batch_autotune(spk_context *ctx, const spk_frame_desc *frame, float *rgba_dev, spk_schedule *chosen,
                       double *best_ms);

/* ---- learned radiance cache (no counterpart in the reference) ---------------------------- */
/* A multiresolution hash-grid encoding of the position (levels x 2 features, 2^log2_table FP16
   entries per level) and SPK_CACHE_ATTRS surface attributes per query (shading normal xyz, albedo
   rgb, roughness, and 1.0 for the first layer's bias; zeros pad the input to 64) feeding a bias-free
   ReLU MLP 64-64-64-64-64-3 on BF16 matrix cores with FP32 accumulation, trained online with Adam
   on a relative L2 loss. Positions are normalised by the scene box. In the path tracer the network
   predicts outgoing radiance divided by the vertex albedo (floored at SPK_CACHE_ALBEDO_FLOOR), which
   the frame multiplies back. */
#define SPK_CACHE_ATTRS 8
#define SPK_CACHE_ALBEDO_FLOOR 0.02f
typedef struct spk_cache_config {
    uint32_t levels;           /* 16 */
    uint32_t features;         /* 2 */
    uint32_t log2_table;       /* 19 */
    uint32_t base_resolution;  /* 16 */
    float per_level_scale;     /* growth between levels, e.g. 1.3819 */
    float learning_rate;       /* 1e-2 */
    uint32_t batch;            /* training samples per frame (65536) */
    uint32_t terminate_bounce; /* paths terminate into the cache at this hit index (2) */
    uint32_t seed;             /* initialisation seed */
    float train_fraction;      /* fraction of paths that are training paths, 0 = batch / (2 pixels) */
    uint32_t train_suffix;     /* a training path ends in the cache from hit index terminate_bounce +
                                  train_suffix on (self-training: its targets include the network's
                                  prediction there), except one in 16 traced to its end; 0: every
                                  training path is traced to its end (2) */
} spk_cache_config;

int spk_cache_create(spk_context *ctx, const spk_cache_config *config);
/* Back to the initial parameters of the configured seed (Adam state and step count cleared),
   keeping every buffer, so a captured frame graph stays valid. */
int spk_cache_reset(spk_context *ctx);
int spk_cache_destroy(spk_context *ctx);
/* n points (device, 3 floats each; world space when world != 0, else in [0,1]^3) and their
   attributes (device, SPK_CACHE_ATTRS floats each; NULL: all zero) -> n x 3 outputs. */
int spk_cache_infer(spk_context *ctx, const float *points_dev, const float *attrs_dev, uint32_t n, int world,
                    float *out_dev);
/* One training step (forward, backward, Adam) on n samples; loss (host, may be NULL) receives the
   batch loss before the update. */
int spk_cache_train(spk_context *ctx, const float *points_dev, const float *attrs_dev, const float *targets_dev,
                    uint32_t n, int world, float *loss);
/* Forward + backward without an update: loss, weight gradient (host, cache_num_weights floats) and
   table gradient (host, levels*2^log2_table*2 floats; may be NULL). */
int spk_cache_gradients(spk_context *ctx, const float *points_dev, const float *attrs_dev, const float *targets_dev,
                        uint32_t n, int world, float *loss, float *dweights, float *dtable);
/* FP32 master parameters to / from host arrays (either pointer may be NULL). */
int spk_cache_get_params(spk_context *ctx, float *weights, float *table);
int spk_cache_set_params(spk_context *ctx, const float *weights, const float *table);
/* Number of MLP weights (14,528 for the default network); training steps taken; last frame's loss. */
int spk_cache_info(spk_context *ctx, uint64_t *num_weights, uint64_t *num_table_params, uint32_t *steps,
                   float *last_loss);

#ifdef __cplusplus
}
#endif

#endif /* SPK_PT_H */
