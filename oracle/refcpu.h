/*
 * refcpu — CPU restatement of Sparkle's `cpu` pipeline hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker for the HIP
 * implementation in sparkle_amd/. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * What it restates (paths relative to the Sparkle repository):
 *   - Xoshiro128+ per-pixel reseeding and RandomUnit/UnitDisk
 *       libraries/include/core/math/Sampler.h:16-53
 *   - camera ray generation          renderer/renderer/CPURenderer.cpp:220-238, 398-417
 *   - the SamplePixel bounce loop    renderer/renderer/CPURenderer.cpp:240-396
 *   - BasePass fan-out and seeding   renderer/renderer/CPURenderer.cpp:419-450
 *   - temporal accumulation          renderer/renderer/CPURenderer.cpp:506-552
 *   - two-level BVH, no tmax culling renderer/proxy/SceneRenderProxy.cpp:24-130,
 *                                    renderer/proxy/MeshRenderProxy.cpp:24-126
 *   - Moller-Trumbore + world_t      renderer/proxy/MeshRenderProxy.cpp:128-194
 *   - hit attributes                 renderer/proxy/MeshRenderProxy.cpp:340-361, io/Mesh.h:86-134
 *   - BxDFs                          scene/material/BxDF.h:32-141,
 *                                    renderer/proxy/PbrMaterialRenderProxy.h:15-51,
 *                                    renderer/proxy/DialetricMaterialRenderProxy.h:16-38
 *   - sky miss shading               renderer/proxy/SkyRenderProxy.cpp:50-64
 * plus checkers the reference does not have: a watertight FP32 triangle test
 * (the GPU's bit-exact twin), an FP64 closest-hit oracle with edge/tie/grazing
 * flags, brute force, and an FP64 hash-grid + MLP with Adam for the learned
 * radiance cache.
 *
 * Parity pins (tests/test_oracle_*.py): known-answer triangle tests, FP64 vs
 * brute force, xoshiro128+ against an independent numpy implementation and
 * the published first outputs, closed-form camera/BxDF statistics, furnace
 * and emitter scenes, finite-difference gradients for the MLP.
 * Parity unpinned (no reference binary can be built here): the exact
 * SplitMix64 -> 4x32-bit seed expansion of Xoshiro-cpp, madmann91/bvh tree
 * shape and visit order, Eigen's reduction order, libm transcendental bits.
 *
 * Every entry point is a plain C function over plain pointers (ctypes).
 */
#ifndef SPARKLE_ORACLE_REFCPU_H
#define SPARKLE_ORACLE_REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orcl_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} orcl_ray;

typedef struct orcl_hit {
    uint32_t prim; /* global triangle id, 0xFFFFFFFF = miss */
    float t;
} orcl_hit;

typedef struct orcl_mesh {
    const float *positions; /* 3 floats per vertex */
    uint32_t num_vertices;
    const uint32_t *indices; /* 3 per triangle */
    uint32_t num_triangles;
    const float *normals;  /* 3 per vertex, may be NULL */
    const float *tangents; /* 4 per vertex, may be NULL */
    const float *uvs;      /* 2 per vertex, may be NULL */
} orcl_mesh;

typedef struct orcl_instance {
    uint32_t mesh;
    uint32_t material;
    float world_from_local[12]; /* row-major 3x4 */
    float local_from_world[12]; /* row-major 3x4 */
} orcl_instance;

enum { ORCL_MAT_PBR = 0, ORCL_MAT_DIELECTRIC = 1 };

typedef struct orcl_material {
    uint32_t type;
    float base_color[3];
    float emissive[3];
    float metallic;
    float roughness;
    float eta;
    int32_t base_color_texture; /* texture index, -1 = none */
    int32_t metallic_roughness_texture;
    int32_t emissive_texture;
    int32_t normal_texture;
} orcl_material;

/* RGBA texels, row-major: 8-bit channels (format 0 UNORM, 1 sRGB) or float (format 2). */
typedef struct orcl_texture {
    uint32_t width;
    uint32_t height;
    uint32_t format;
    uint32_t reserved;
    const void *pixels;
} orcl_texture;

typedef struct orcl_camera {
    float position[3];
    float right[3];
    float up[3];
    float front[3];
    float lower_left[3];
    float max_u[3];
    float max_v[3];
    float lens_radius;
    float far_plane;
} orcl_camera;

typedef struct orcl_frame {
    uint32_t width;
    uint32_t height;
    uint32_t frame_seed;
    uint32_t max_bounce;
    uint32_t debug_mode;
    uint32_t sky_enabled;
    float sky_color[3];
    uint32_t seed_hi; /* high 32 bits of a 64-bit pixel seed (0 = reference uint32 wrap) */
    uint32_t nee;      /* next-event estimation as the device's SPK_FRAME_NEE (no reference counterpart) */
    float output_limit; /* per-sample clamp; 0 = the reference's 6 */
    uint32_t rr_off;    /* checker switch, no reference counterpart: 1 keeps every path past bounce 3
                           (the roulette draw is still consumed, its outcome ignored, so the two
                           estimators share every other random number); 0 = the reference's roulette */
} orcl_frame;

enum {
    ORCL_TRACE_REF_MT = 0,   /* reference semantics: MT, |det|<1e-6 cutoff, no tmax culling */
    ORCL_TRACE_WT = 1,       /* watertight FP32 (GPU twin), order-independent ties */
    ORCL_TRACE_WT_BRUTE = 2, /* watertight FP32 over every triangle */
    ORCL_TRACE_FP64 = 3,     /* FP64 oracle with flags */
    ORCL_TRACE_WT_TWO = 4    /* watertight FP32 twin of a two-level scene: every instance's triangles
                                in object space, ray through the float inverse transform */
};

enum {
    ORCL_FLAG_EDGE = 1,  /* FP64 hit within eps_edge (barycentric) of an edge */
    ORCL_FLAG_TIE = 2,   /* two hits within eps_tie relative t */
    ORCL_FLAG_GRAZE = 4  /* |cos(normal, dir)| below eps_graze */
};

typedef struct orcl_scene orcl_scene;

/* Builds one BLAS per mesh and a TLAS over instances (SAH, binned). */
orcl_scene *orcl_scene_create(const orcl_mesh *meshes, uint32_t num_meshes, const orcl_instance *instances,
                              uint32_t num_instances, const orcl_material *materials, uint32_t num_materials);
void orcl_scene_destroy(orcl_scene *scene);
/* Analytic spheres (SphereRenderProxy.cpp:5-48): closed-form hit with unit directions assumed;
   sphere k has the global id num_triangles + k. Ignored by the FP64 trace mode. */
typedef struct orcl_sphere {
    float center[3];
    float radius;
    uint32_t material;
    uint32_t reserved[3];
} orcl_sphere;
void orcl_scene_set_spheres(orcl_scene *scene, const orcl_sphere *spheres, uint32_t num_spheres);
/* Sky cube map faces (+X, -X, +Y, -Y, +Z, -Z) as texture indices; faces[0] < 0 = gradient sky. */
void orcl_scene_set_sky_faces(orcl_scene *scene, const int32_t *faces);
/* Texture table the materials' texture indices refer to (copied). */
void orcl_scene_set_textures(orcl_scene *scene, const orcl_texture *textures, uint32_t num_textures);
/* Image::Sample of texture `index` at (u, v): rgb into out[3]. */
void orcl_texture_sample(const orcl_scene *scene, int index, float u, float v, float *out);
uint64_t orcl_scene_num_triangles(const orcl_scene *scene);

/* Closest hit for each ray. flags may be NULL (only written in FP64 mode). */
int orcl_trace(const orcl_scene *scene, const orcl_ray *rays, uint64_t num_rays, int mode, int num_threads,
               orcl_hit *hits, uint32_t *flags, double *t64);
void orcl_set_fp64_eps(double eps_edge, double eps_tie, double eps_graze);

/* Single-triangle tests (known-answer tests). Return 1 on hit. */
int orcl_tri_mt(const float *org, const float *dir, const float *v0, const float *v1, const float *v2, float *t,
                float *u, float *v);
int orcl_tri_wt(const float *org, const float *dir, float tmin, const float *v0, const float *v1, const float *v2,
                float *t, float *u, float *v);
float orcl_refine_t(const float *org, const float *dir, const float *v0, const float *v1, const float *v2);

/* RNG */
void orcl_xoshiro_stream(uint64_t seed, uint32_t *out, uint32_t n);
void orcl_random_unit_stream(uint64_t seed, float *out, uint32_t n);
/* n samples of one BxDF lobe in the local frame (normal +z) from one xoshiro stream: kind 0
   Lambertian, 1 specular, 2 dielectric, 3 VNDF microfacet normal. thr / valid may be NULL. */
/* Textured sky light cook (SkyLightCookJob::Execute, SkyLight.cpp:83-135): six face_size^2 RGBA faces
   (+X, -X, +Y, -Y, +Z, -Z) sampled from an equirectangular RGBA float map, values rounded to FP16. */
void orcl_sky_cook(const float *equirect, uint32_t w, uint32_t h, uint32_t face_size, float *faces);
/* float -> binary16 (round to nearest even) -> float. */
float orcl_half_round(float x);
/* Advances a raw xoshiro128+ state by n steps with the same transition as the renderer's generator
   (Sampler.h:16-41); writes the n outputs to out when out is not null. */
void orcl_xoshiro_advance(uint32_t state[4], uint64_t n, uint32_t *out);
void orcl_bxdf_sample(int kind, const float *wo, const float *base, float roughness, float metallic, float eta,
                      uint64_t seed, uint32_t n, float *wi, float *thr, int32_t *valid);
/* SampleSurface's shading frame (u, v, w) from a normal and a tangent. */
void orcl_local_axes(const float *normal, const float *tangent, float *u, float *v, float *w);

/* One frame of BasePass: rgba = (color clamped to 6, valid flag), normal = first-hit normal. */
int orcl_render_frame(const orcl_scene *scene, const orcl_camera *camera, const orcl_frame *frame, int num_threads,
                      float *rgba, float *normal);
/* Temporal accumulation as in CPURenderer::DenoisePass (spatial fill off). */
void orcl_accumulate(const float *rgba, float *accum, uint64_t num_pixels, uint32_t cumulated, uint32_t spp);
void orcl_spatial_fill(const float *rgba, const float *normal, float *out, uint32_t w, uint32_t h);
/* Ray-segment statistics of the last orcl_render_frame call (extend rays traced). */
uint64_t orcl_last_render_rays(void);

/* Camera set-up from physical attributes (CameraComponent.cpp:120-128, CameraRenderProxy.cpp:24-27, 113-136). */
void orcl_camera_setup(const float position[3], const float right[3], const float front[3], const float up[3],
                       float vertical_fov, float focus_distance, float aperture_radius, float aspect,
                       orcl_camera *camera);

/* ---- learned radiance cache oracle (FP64) -------------------------------- */
typedef struct orcl_cache_config {
    uint32_t levels;          /* 16 */
    uint32_t features;        /* 2 */
    uint32_t log2_table;      /* 19 */
    uint32_t base_resolution; /* 16 */
    float per_level_scale;    /* growth factor between levels */
    uint32_t hidden;          /* 64 */
    uint32_t hidden_layers;   /* 4 */
    uint32_t outputs;         /* 3 */
    uint32_t inputs;          /* MLP input width: the encoding, then per-query attributes and zero
                                 padding (0: the encoding alone, levels * features) */
} orcl_cache_config;

/* Encode points (in [0,1]^3) -> features [n][levels*features] from an FP32 table. */
void orcl_hashgrid_encode(const orcl_cache_config *cfg, const float *table, const float *points, uint32_t n,
                          double *features);
/* Hash index of a grid corner at a level (Instant-NGP primes). */
uint32_t orcl_hashgrid_index(const orcl_cache_config *cfg, uint32_t level, uint32_t x, uint32_t y, uint32_t z);
/* MLP forward/backward in FP64. weights: per layer W[out][in] row-major, no biases.
   loss = mean over samples and outputs of relative L2: (y - target)^2 / (sg(y)^2 + 0.01).
   Returns loss; writes dW (same layout as weights) and dfeatures [n][in]. */
double orcl_mlp_loss_grad(const orcl_cache_config *cfg, const double *weights, const double *features,
                          const double *targets, uint32_t n, double *outputs, double *dweights, double *dfeatures);
/* Adam step (Kingma & Ba 2015) in FP64 with bias correction at step t (1-based). */
void orcl_adam(double *param, const double *grad, double *m, double *v, uint64_t count, double lr, double beta1,
               double beta2, double eps, uint32_t step);

#ifdef __cplusplus
}
#endif

#endif
