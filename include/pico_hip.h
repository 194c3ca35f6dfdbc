/*
 * pico_hip.h — C ABI of the MI355X (gfx950) rendering hot path of Pico.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, never
 * throws, and returns an int status (PICO_OK == 0, negative on error; the
 * text of the last error of a context is returned by pico_hip_last_error).
 *
 * Which reference interface each entry point replaces (reference files are
 * cited relative to the Pico source tree):
 *
 *   pico_hip_upload_scene   Scene::Scene ctor + UpperLevelBVH::build
 *                           (Source/Core/Scene.cpp:36-84, Source/Core/UpperLevelBVH.cpp:27-57,
 *                           Source/Core/LowerLevelMeshBVH.cpp:11-79)
 *   pico_hip_render         Scene::Scene::render_scene_to_memory
 *                           (Source/Core/Scene.hpp:68, Source/Core/Scene.cpp:143-235)
 *   pico_hip_gbuffer        Scene::generate_denoiser_inputs (Source/Core/Scene.cpp:248-294)
 *   pico_hip_atrous         Util::atrous_denoise (Source/Util/Denoisers.hpp:9,
 *                           Source/Util/Denoisers.cpp:107-126)
 *   pico_hip_tonemap        Util::reinhard_tone_mapping (Source/Util/ToneMappers.cpp:12-44)
 *   pico_hip_closest_hit    UpperLevelBVH::get_closest_intersection
 *                           (Source/Core/UpperLevelBVH.cpp:12-20) for a batch of rays
 *   pico_hip_upload_scene_ex  the same scene set-up with the BVH builds on the GPU
 *                           (Source/Core/BVH.cpp:73-197, Source/Core/LowerLevelMeshBVH.cpp:48-77)
 *   pico_hip_set_instance_transforms  UpperLevelBVH::add_lower_level_bvh + build, per frame
 *                           (Source/Core/UpperLevelBVH.cpp:22-57)
 *   pico_scene_load_json    Scene::Scene(ThreadPool&, path) and its section handlers
 *                           (Source/Core/Scene.cpp:36-84, 297-720)
 *   pico_scene_load_obj     Scene::Scene(ThreadPool&, dir, aiScene*, options) with parse_node
 *                           (Source/Core/Scene.cpp:86-141, 722-971) for an OBJ model file
 *   pico_scene_set_sun / _set_sky_image  the -SunDirection/-SunColour/-Skybox options of
 *                           Scene::Scene(ThreadPool&, path, aiScene, options) (Source/Core/Scene.cpp:86-123)
 *   pico_write_hdr          the file half of Scene::render_scene_to_file (Source/Core/Scene.cpp:236-246)
 *
 * Buffers passed as `float*` RGB/vec3 images use the reference layout: packed
 * 3 x f32 per pixel (glm::vec3, 12 bytes, no padding), pixel (x, y) at y*W+x.
 * Entry points whose name ends in `_d` take device pointers (memory obtained
 * from pico_hip_malloc); all others take host pointers and stage internally.
 */
#ifndef PICO_HIP_H
#define PICO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PICO_OK              0
#define PICO_ERR_ARG        -1   /* invalid argument                        */
#define PICO_ERR_HIP        -2   /* HIP runtime error                       */
#define PICO_ERR_STATE      -3   /* call out of order (e.g. no scene)       */
#define PICO_ERR_CANCELLED  -4   /* cancel flag observed                    */
#define PICO_ERR_LIMIT      -5   /* size beyond a documented limit          */
#define PICO_ERR_NOMEM      -6   /* host or device allocation failed        */

/* ---- scene description ------------------------------------------------- */
/* BSDF kinds, Source/Render/BSRDF.hpp:33-40 and the "BSRDF" keys of
 * Source/Core/Scene.cpp:367-417. */
#define PICO_BSDF_DIFFUSE     0u
#define PICO_BSDF_SPECULAR    1u
#define PICO_BSDF_DIELECTRIC  2u
#define PICO_BSDF_LIGHT       3u
#define PICO_BSDF_DELTA       4u
/* Transmissive BSDFs (BSRDF_Type::kBTDF): Transparent_BTDF and Fresnel_BTDF
 * (Source/Render/BSRDF.cpp:240-421) with the ray's index-of-refraction stack
 * (Source/Core/AABB.cpp:54-79); their material needs index_of_refraction. */
#define PICO_BSDF_TRANSPARENT 5u
#define PICO_BSDF_FRESNEL     6u

/* Mesh shapes (pico_scene_desc.mesh_shape): a triangle mesh, or one of the
 * reference's built-in implicit shapes, LowerLevelSphereBVH(0.5) and
 * LowerLevelCube (Source/Core/LowerLevelImplicitShapesBVH.cpp:16-113,
 * registered as "Sphere" and "Cube" by Source/Core/Scene.cpp:49-56), which
 * have no vertices or triangles of their own. */
#define PICO_SHAPE_TRIANGLES 0u
#define PICO_SHAPE_SPHERE    1u   /* centre 0, radius 0.5 */
#define PICO_SHAPE_CUBE      2u   /* [-0.5, 0.5]^3 */

/* Microfacet distribution used by every specular lobe.  BECKMANN_ALL is the
 * reference's Beckmann_All_Microfacet_Distribution
 * (Source/Render/Distributions.cpp:31-78); GGX is a new Distribution subclass
 * (Walter et al. 2007) plugged into the same interface
 * (Source/Render/Distributions.hpp:10-21). */
#define PICO_MICROFACET_BECKMANN_ALL 0u
#define PICO_MICROFACET_GGX          1u

/* Core::EvaluatedMaterial (Source/Core/MaterialManager.hpp:17-46) of a
 * constant material, without the unused normal field, plus the index of
 * refraction of the ConstantTransparent* materials
 * (Source/Render/BasicMaterials.hpp:41-65); 0 for the others. */
typedef struct pico_material {
    float diffuse[3];
    float specular[3];
    float roughness;
    float emissive[3];
    float index_of_refraction;
} pico_material;

/* Sky image, Core::ImageCube (Source/Core/Image.cpp:361-466): looked up
 * with the nearest texel in the direction of every ray that leaves the scene
 * and for the G-buffer's sky pixels.  faces = 6: a cube map, faces +x, -x,
 * +y, -y, +z, -z one after the other (a JSON "Skybox" array); faces = 1: an
 * equirectangular image (a JSON "Skybox" HDRI file).  Texels are row-major,
 * the first row first; RGBA8 texels are divided by 255, alpha is unused.
 * At most 2^26 texels (all faces). */
#define PICO_TEXEL_RGBA8   0u   /* 4 x u8  */
#define PICO_TEXEL_RGBA32F 1u   /* 4 x f32 */
typedef struct pico_sky_image {
    uint32_t    width;
    uint32_t    height;
    uint32_t    faces;      /* 6 or 1 */
    uint32_t    format;     /* PICO_TEXEL_* */
    const void* texels;     /* faces x height x width texels */
} pico_sky_image;

/* Flattened scene.  Meshes are object-space triangle soups with shared
 * vertices; instances reference a mesh, a material and a BSDF kind with a 4x4
 * column-major transform (glm::mat4 layout).  Lights are instances whose BSDF
 * is PICO_BSDF_LIGHT, listed in light-selection order (Scene::m_lights,
 * Source/Core/Scene.cpp:407-411). */
typedef struct pico_scene_desc {
    uint32_t        n_meshes;
    const uint32_t* mesh_vertex_offset;    /* n_meshes + 1 */
    const uint32_t* mesh_triangle_offset;  /* n_meshes + 1 */
    const float*    positions;             /* 3 x f32 per vertex */
    const float*    normals;               /* 3 x f32 per vertex */
    const uint32_t* indices;               /* 3 x u32 per triangle, mesh-local */
    uint32_t        n_instances;
    const float*    transforms;            /* 16 x f32 per instance */
    const uint32_t* instance_mesh;
    const uint32_t* instance_material;
    const uint32_t* instance_bsdf;
    uint32_t        n_materials;
    const pico_material* materials;
    uint32_t        n_lights;
    const uint32_t* light_instances;
    float           sky_rgb[3];            /* constant sky (1x1 cube map) */
    uint32_t        microfacet;            /* PICO_MICROFACET_* */
    const uint32_t* mesh_shape;            /* n_meshes PICO_SHAPE_*, or NULL: all triangle meshes */
    /* The sun (Scene::Sun, Source/Core/Scene.hpp:52-58): with use_sun set,
     * next-event estimation picks among n_lights + 1 lights, the sun last,
     * and the sun lights a point when a ray from it along -sun_direction
     * meets nothing (Source/Render/Integrators.cpp:55-77).  sun_direction is
     * the direction the light travels, used as given. */
    uint32_t        use_sun;
    float           sun_direction[3];
    float           sun_colour[3];
    /* Sky image replacing the constant sky_rgb, or NULL. */
    const struct pico_sky_image* sky_image;
} pico_scene_desc;

/* Scene::Camera (Source/Core/Camera.hpp:14-35): pinhole with
 * tan(half-FOV) = 0.5, up = (0,1,0); the far plane is the miss distance of
 * G-buffer rays. */
typedef struct pico_camera {
    float position[3];
    float direction[3];
    float up[3];
    float aspect;
    float far_plane;
} pico_camera;

/* Scene::RenderParams (Source/Core/Scene.hpp:30-43) minus the buffers, plus
 * the frame seed of the per-pixel PCG32 streams and the tile size used by the
 * tile-local denoiser constants. */
typedef struct pico_render_params {
    uint32_t width;
    uint32_t height;
    uint32_t spp;            /* m_maxSamples                               */
    uint32_t max_depth;      /* m_maxRayDepth                              */
    uint32_t tile_w;         /* tile size (Util::Tiler)                    */
    uint32_t tile_h;
    uint32_t frame_seed;     /* seeds PCG32: state=mix(frame,sample), seq=pixel */
    uint32_t view_id;        /* batched views: folded into the seed         */
    float    max_variance;   /* m_maxVariance (early-out)                  */
    uint32_t flags;          /* PICO_RENDER_*                              */
} pico_render_params;

#define PICO_RENDER_RESUME  1u  /* accumulate onto the caller's buffers (sample counts > 0) */
#define PICO_RENDER_COUNT   2u  /* count BVH node visits and triangle tests (slower) */

typedef struct pico_stats {
    uint64_t camera_rays;
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t gbuffer_rays;
    uint64_t node_visits;    /* PICO_RENDER_COUNT only: extension + camera rays */
    uint64_t tri_tests;      /* PICO_RENDER_COUNT only */
    uint64_t instance_visits;/* PICO_RENDER_COUNT only */
    uint64_t shadow_node_visits;     /* PICO_RENDER_COUNT only: shadow rays */
    uint64_t shadow_tri_tests;       /* PICO_RENDER_COUNT only */
    uint64_t shadow_instance_visits; /* PICO_RENDER_COUNT only */
    uint64_t samples;        /* (pixel, sample) pairs accumulated */
    uint64_t accumulated_rays; /* rays of the accumulated samples only */
    double   render_ms;      /* device time of the render, HIP events; for each view
                                of a batched launch: launch start -> end of the chunk
                                holding the view's last sample */
    int32_t  cancelled;
    int32_t  denoise_fallback; /* pico_hip_render_views_denoise_d: 1 when the view's output is
                                  the a-trous fallback (see pico_hip_denoise_gbuffer_d) */
    double   done_ms;        /* batched launches: launch start -> the view's output is
                                final (after its denoiser when one is requested) */
    uint64_t spill_pushes;   /* PICO_RENDER_COUNT only: traversal-stack entries written past */
    uint64_t spill_pops;     /* the LDS part (device memory), and read back, all rays */
    uint64_t finish_rays;    /* rays of a chunk's last bounces, which one persistent launch traces path
                                by path once few paths are left (included in the ray counts above) */
    uint64_t finish_node_visits;     /* PICO_RENDER_COUNT only: of those rays (included above) */
    uint64_t finish_tri_tests;       /* PICO_RENDER_COUNT only */
    uint64_t finish_instance_visits; /* PICO_RENDER_COUNT only */
} pico_stats;

/* ---- context ----------------------------------------------------------- */
typedef struct pico_hip_ctx pico_hip_ctx;

int  pico_hip_create(int device, pico_hip_ctx** out);
/* Frees the context's device memory, its scene and every pico_cnn created on it. */
void pico_hip_destroy(pico_hip_ctx* ctx);
const char* pico_hip_last_error(const pico_hip_ctx* ctx);
int  pico_hip_synchronize(pico_hip_ctx* ctx);
int  pico_hip_device_name(pico_hip_ctx* ctx, char* buf, size_t n);

/* Device memory owned by the caller, used by the `_d` entry points. */
int  pico_hip_malloc(pico_hip_ctx* ctx, size_t bytes, void** dptr);
int  pico_hip_free(pico_hip_ctx* ctx, void* dptr);
int  pico_hip_memcpy_h2d(pico_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int  pico_hip_memcpy_d2h(pico_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int  pico_hip_memset_d(pico_hip_ctx* ctx, void* dst, int value, size_t bytes);

/* Per-kernel timing with HIP events on the context's stream.  When enabled,
 * every launch of every kernel is bracketed by events; pico_hip_kernel_time
 * returns the launch count and mean duration (ms) of the named kernel since
 * the last reset. */
int  pico_hip_enable_kernel_timing(pico_hip_ctx* ctx, int enable);
int  pico_hip_kernel_time(pico_hip_ctx* ctx, const char* kernel, uint64_t* launches, double* mean_ms);

/* Measurement aid with no counterpart in the reference: a streaming copy of
 * `bytes` (multiple of 16) with 16-byte loads and stores (kernel "copy"), used
 * to measure achievable HBM bandwidth and to calibrate byte counters. */
int  pico_hip_copy_d(pico_hip_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

/* ---- scene ------------------------------------------------------------- */
/* Builds the BLAS (one per mesh, 16-bin binned SAH) and the TLAS over
 * instances on the host, lays them out for the device and uploads them. */
int  pico_hip_upload_scene(pico_hip_ctx* ctx, const pico_scene_desc* scene);

/* Where pico_hip_upload_scene_ex builds the BVHs. */
#define PICO_BUILD_DEVICE   1u   /* BLAS and TLAS built by kernels instead of on the host */

typedef struct pico_build_stats {
    double   total_ms;      /* the whole upload: checks, tables, copies, BVH build */
    float    bvh_ms;        /* the BVH build: host wall time, or the device kernels event-timed */
    uint32_t binary_nodes;  /* binary BLAS nodes over every mesh */
    uint32_t blas_nodes4;   /* 4-wide BLAS nodes over every mesh */
    uint32_t blas_bound;    /* most traversal-stack entries any BLAS can need */
    uint32_t stack_bound;   /* the same for the whole scene (<= 64) */
    uint32_t levels;        /* deepest binary BLAS, in levels of internal nodes */
} pico_build_stats;

/* pico_hip_upload_scene with a choice of builder and build statistics
 * (stats may be NULL).  With PICO_BUILD_DEVICE every mesh BVH is built on
 * the GPU, all meshes at once, level by level: binned SAH with the host
 * builder's split rule (16 bins on each axis, surface-area cost, leaves of
 * at most 4 triangles), so both builders make the same binary tree except
 * where all centroids of a node coincide; then the same greedy collapse to
 * 4-wide nodes, and the TLAS as in pico_hip_set_instance_transforms.  The
 * host only checks the description and builds the material, instance and
 * light tables.  Replaces the serial host build of the reference's scene
 * set-up: BVHFactory::generate_BVH / SAH_parition_Scheme
 * (Source/Core/BVH.cpp:73-197), LowerLevelMeshBVH::build
 * (Source/Core/LowerLevelMeshBVH.cpp:48-77) and UpperLevelBVH::build
 * (Source/Core/UpperLevelBVH.cpp:27-57).  Errors as pico_hip_upload_scene;
 * on any error no scene is left uploaded. */
int  pico_hip_upload_scene_ex(pico_hip_ctx* ctx, const pico_scene_desc* scene, uint32_t flags,
                              pico_build_stats* stats);
/* Node/triangle/instance byte sizes and tree depths of the uploaded scene. */
int  pico_hip_scene_info(pico_hip_ctx* ctx, uint64_t* bytes_nodes, uint64_t* bytes_tris,
                         uint64_t* bytes_instances, uint32_t* tlas_depth, uint32_t* blas_depth);

/* Result of a device TLAS rebuild. */
typedef struct pico_rebuild_stats {
    float    device_ms;     /* the rebuild's kernels, event-timed on the context's stream */
    uint32_t binary_depth;  /* deepest internal node of the binary (Karras) hierarchy */
    uint32_t stack_bound;   /* traversal-stack entries the rebuilt scene can need (<= 64) */
    uint32_t wide_nodes;    /* 4-wide TLAS nodes written */
} pico_rebuild_stats;

/* New transforms for every instance of the uploaded scene (n_instances x 16
 * f32, column-major like glm::mat4, the layout of pico_scene_desc.transforms);
 * the instance inverses and the TLAS are rebuilt on the device (linear BVH
 * over Morton codes of the instance world boxes, collapsed to 4-wide nodes)
 * and the BLAS are kept.  Replaces the TLAS part of the reference's scene
 * build, UpperLevelBVH::add_lower_level_bvh / UpperLevelBVH::build
 * (Source/Core/UpperLevelBVH.cpp:22-57), which the reference runs once on the
 * host.  Closest hits do not depend on the tree, so a rebuild with the
 * upload's own transforms renders the same image.  A transform that is not
 * affine or is singular gives PICO_ERR_ARG and a tree too deep for the
 * traversal stack PICO_ERR_LIMIT; in both cases the scene is left unchanged.
 * `_d`: the transforms are in device memory. */
int  pico_hip_set_instance_transforms(pico_hip_ctx* ctx, const float* transforms, pico_rebuild_stats* stats);
int  pico_hip_set_instance_transforms_d(pico_hip_ctx* ctx, const float* d_transforms, pico_rebuild_stats* stats);

/* Closest hit of n world-space rays (origins/dirs: 3 x f32 each).  Outputs
 * t (world distance), instance, mesh-local primitive and barycentrics (u, v);
 * misses give t = +inf and instance = prim = 0xffffffff. */
int  pico_hip_closest_hit(pico_hip_ctx* ctx, uint32_t n, const float* origins, const float* dirs,
                          float* t, uint32_t* instance, uint32_t* prim, float* uv);

/* Primary-hit ids of every pixel for a fixed sub-pixel sample xi (the G-buffer
 * pass uses xi = (0,0), Source/Core/Scene.cpp:271). */
int  pico_hip_primary_ids(pico_hip_ctx* ctx, const pico_camera* cam, uint32_t width, uint32_t height,
                          float xi_x, float xi_y, uint32_t* instance, uint32_t* prim);

/* ---- render ------------------------------------------------------------ */
/* Path-traces the frame and accumulates into the reference's buffers:
 * pixels / variance 3 x f32 per pixel, sample_count u32 per pixel (running
 * mean/variance update and early-out of Source/Core/Scene.cpp:163-197).
 * cancel (may be NULL; may be written by another thread) is read before each
 * chunk of samples, after the previous chunk has completed on the device:
 * a non-zero value stops the render with PICO_ERR_CANCELLED, the pixels
 * accumulated so far kept (Scene.cpp:158 polls `should_quit` per pixel). */
int  pico_hip_render(pico_hip_ctx* ctx, const pico_camera* cam, const pico_render_params* params,
                     float* pixels, uint32_t* sample_count, float* variance,
                     const volatile int32_t* cancel, pico_stats* stats);
int  pico_hip_render_d(pico_hip_ctx* ctx, const pico_camera* cam, const pico_render_params* params,
                       float* d_pixels, uint32_t* d_sample_count, float* d_variance,
                       const volatile int32_t* cancel, pico_stats* stats);

/* Progressive display: pico_hip_render (host buffers) that, after every
 * chunk of samples, copies the accumulated pixels and sample counts to a
 * snapshot and calls `cb` on the calling thread with it (valid during the
 * call only); samples_done is the sample index the chunk reached.  A
 * non-zero return stops the render with PICO_ERR_CANCELLED, the samples so
 * far kept.  Replaces the reference's viewer reading frame_memory while the
 * render thread writes it (Source/main.cpp:119-132) with a consistent copy
 * between two chunks. */
typedef int (*pico_pass_callback)(void* user, const float* pixels, const uint32_t* sample_count, uint32_t width,
                                  uint32_t height, uint32_t samples_done);
int  pico_hip_render_progressive(pico_hip_ctx* ctx, const pico_camera* cam, const pico_render_params* params,
                                 float* pixels, uint32_t* sample_count, float* variance, pico_pass_callback cb,
                                 void* user, pico_stats* stats);

/* Radiance of every (pixel, sample) with sample index in
 * [sample_base, sample_base + params->spp), without accumulation: out is
 * spp x H x W x 3 f32 (sample-major).  Used for per-path parity checks. */
int  pico_hip_render_samples(pico_hip_ctx* ctx, const pico_camera* cam, const pico_render_params* params,
                             uint32_t sample_base, float* out);

/* FP64 accumulation of spp samples per pixel (no early-out) into
 * d_sum (3 x f64 per pixel, added to), for the converged-reference oracle. */
int  pico_hip_render_accum64_d(pico_hip_ctx* ctx, const pico_camera* cam, const pico_render_params* params,
                               uint32_t sample_base, double* d_sum, pico_stats* stats);

/* G-buffer of Source/Core/Scene.cpp:248-294: world normal, view-space
 * position (glm::lookAt view matrix) and diffuse albedo, 3 x f32 each. */
int  pico_hip_gbuffer(pico_hip_ctx* ctx, const pico_camera* cam, uint32_t width, uint32_t height,
                      float* normal, float* position, float* albedo);
int  pico_hip_gbuffer_d(pico_hip_ctx* ctx, const pico_camera* cam, uint32_t width, uint32_t height,
                        float* d_normal, float* d_position, float* d_albedo);

/* ---- denoise / post ---------------------------------------------------- */
/* A-trous wavelet filter of Source/Util/Denoisers.cpp:18-126: `levels`
 * passes at dilation 2^l, edge-stopping on normal, position and albedo with
 * tile-local range constants over tiles of tile_w x tile_h. */
int  pico_hip_atrous(pico_hip_ctx* ctx, const float* rgb, const float* normal, const float* position,
                     const float* albedo, uint32_t width, uint32_t height, uint32_t tile_w,
                     uint32_t tile_h, uint32_t levels, float* out_rgb);
/* Device variant; d_scratch must hold width*height*3 f32 (ping-pong). */
int  pico_hip_atrous_d(pico_hip_ctx* ctx, const float* d_rgb, const float* d_normal, const float* d_position,
                       const float* d_albedo, uint32_t width, uint32_t height, uint32_t tile_w,
                       uint32_t tile_h, uint32_t levels, float* d_scratch, float* d_out);

/* Extended Reinhard with a global luminance white point, in place. */
/* Labelled mode with no counterpart in the reference: edge-avoiding a-trous
 * (Dammertz et al., HPG 2010) with FIXED edge-stopping widths in place of the
 * reference's tile-local ones (Source/Util/Denoisers.cpp:45-63), B3-spline
 * taps, `levels` passes at dilation 2^l.  Widths: normal distance, depth
 * difference relative to the centre's |position|, albedo distance, and
 * colour distance (halved every level); 0 drops a term.  The pixel's own tap
 * always has weight > 0, so no tile can divide by zero. */
int  pico_hip_atrous_fixed_d(pico_hip_ctx* ctx, const float* d_rgb, const float* d_normal, const float* d_position,
                             const float* d_albedo, uint32_t width, uint32_t height, uint32_t levels,
                             float sigma_normal, float sigma_depth, float sigma_albedo, float sigma_color,
                             float* d_scratch, float* d_out);
int  pico_hip_tonemap(pico_hip_ctx* ctx, float* rgb, uint32_t width, uint32_t height);

/* relMSE = mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2) with the
 * reference r = d_ref_sum / ref_count (FP64), reduced in FP64 in a fixed
 * order (the same inputs give the same bits).  At most 2^28 pixels;
 * n_pixels = 0 gives 0.  Synchronises the context's stream. */
int  pico_hip_relmse_d(pico_hip_ctx* ctx, const float* d_img, const double* d_ref_sum, double ref_count,
                       uint32_t n_pixels, double* out);

/* Allocation-free variant of pico_hip_relmse_d.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  It
 * calls no allocator and does not synchronise: the mean is written to
 * d_out (one double in device memory) in stream order, and the per-block
 * parts of the reduction live in a fixed 8 KiB static device array, one copy
 * per device, held until the process ends.  Calls that overlap on one device
 * race on that array.  Same kernels and grid as pico_hip_relmse_d, so the
 * same bits; at most 2^28 pixels (PICO_ERR_LIMIT beyond); n_pixels = 0
 * writes 0. */
int  pico_hip_relmse_noalloc_d(pico_hip_ctx* ctx, const float* d_img, const double* d_ref_sum, double ref_count,
                               uint32_t n_pixels, double* d_out);

/* ---- learned denoiser (kernel-predicting U-Net) ------------------------ */
/* Input: 10 channels per pixel (radiance 3, albedo 3, normal 3, depth 1),
 * planar layout [10][H][W] f32.  Network: 3-level U-Net 32/64/128 with 3x3
 * convolutions (BF16 weights/activations on MFMA, FP32 accumulation) and a
 * 1x1 head predicting a 5x5 per-pixel kernel applied to the radiance. */
typedef struct pico_cnn pico_cnn;
int  pico_hip_cnn_create(pico_hip_ctx* ctx, uint32_t max_width, uint32_t max_height, uint32_t max_batch,
                         pico_cnn** out);
void pico_hip_cnn_destroy(pico_cnn* cnn);
uint32_t pico_hip_cnn_num_params(void);
int  pico_hip_cnn_init(pico_cnn* cnn, uint32_t seed);
int  pico_hip_cnn_set_params(pico_cnn* cnn, const float* params);
int  pico_hip_cnn_get_params(pico_cnn* cnn, float* params);
int  pico_hip_cnn_denoise(pico_cnn* cnn, const float* inputs, uint32_t width, uint32_t height, float* out_rgb);
int  pico_hip_cnn_denoise_d(pico_cnn* cnn, const float* d_inputs, uint32_t width, uint32_t height, float* d_out_rgb);
/* Network input from a G-buffer: radiance, albedo, normal (packed RGB per
 * pixel, as pico_hip_gbuffer writes them) and depth = |view-space position|,
 * written planar [10][H][W]. */
int  pico_hip_cnn_inputs_d(pico_hip_ctx* ctx, const float* d_rgb, const float* d_normal, const float* d_position,
                           const float* d_albedo, uint32_t width, uint32_t height, float* d_inputs);
/* The learned filter with the argument list of Util::atrous_denoise
 * (Source/Util/Denoisers.hpp:9: pixels, normals, positions, diffuse), so it
 * drops in where the a-trous filter is called (Source/Core/Scene.cpp:214-223). */
int  pico_hip_cnn_denoise_gbuffer(pico_cnn* cnn, const float* rgb, const float* normal, const float* position,
                                  const float* albedo, uint32_t width, uint32_t height, float* out_rgb);
int  pico_hip_cnn_denoise_gbuffer_d(pico_cnn* cnn, const float* d_rgb, const float* d_normal, const float* d_position,
                                    const float* d_albedo, uint32_t width, uint32_t height, float* d_out_rgb);
/* The denoise step of Scene::render_scene_to_memory (Source/Core/Scene.cpp:210-222)
 * with the learned filter in front and the reference's filter as its
 * fallback: the network's output, unless the network is NULL or refuses the
 * call (e.g. the image exceeds its size), or any value of its output is not
 * finite; then the whole image is Util::atrous_denoise's (reference mode, 5
 * levels, tiles of tile_w x tile_h, as pico_hip_atrous_d), the call site the
 * learned filter replaces.  The check and the fallback run on the device in
 * stream order (no synchronisation); d_fallback (one device u32) receives 1
 * when the output is the a-trous filter's, else 0.  d_scratch: W*H*3 f32.
 * The host variant stages the buffers and returns the flag in
 * *used_fallback (may be NULL). */
int  pico_hip_denoise_gbuffer_d(pico_hip_ctx* ctx, pico_cnn* cnn, const float* d_rgb, const float* d_normal,
                                const float* d_position, const float* d_albedo, uint32_t width, uint32_t height,
                                uint32_t tile_w, uint32_t tile_h, float* d_scratch, float* d_out, uint32_t* d_fallback);
int  pico_hip_denoise_gbuffer(pico_hip_ctx* ctx, pico_cnn* cnn, const float* rgb, const float* normal,
                              const float* position, const float* albedo, uint32_t width, uint32_t height,
                              uint32_t tile_w, uint32_t tile_h, float* out_rgb, int32_t* used_fallback);
/* Loss (mean relMSE of the filtered radiance vs target, FP64 reduction) and
 * FP32 gradients of every parameter for one batch (host buffers). */
int  pico_hip_cnn_loss_grad(pico_cnn* cnn, const float* inputs, const float* targets, uint32_t batch,
                            uint32_t width, uint32_t height, double* loss, float* grads);
/* Global gradient-norm clipping of the training step (default 1.0, 0 = off)
 * and the pre-clipping norm of the last step. */
int  pico_hip_cnn_set_clip(pico_cnn* cnn, float max_grad_norm);
int  pico_hip_cnn_last_grad_norm(pico_cnn* cnn, double* out);
/* One Adam step on a device-resident batch. */
int  pico_hip_cnn_train_step_d(pico_cnn* cnn, const float* d_inputs, const float* d_targets, uint32_t batch,
                               uint32_t width, uint32_t height, float lr, double* loss);

/* ---- batched uneven views ---------------------------------------------- */
typedef struct pico_view {
    pico_camera        camera;
    pico_render_params params;
    float*    d_pixels;         /* device, W*H*3 */
    uint32_t* d_sample_count;   /* device, W*H   */
    float*    d_variance;       /* device, W*H*3 */
} pico_view;

/* Launch shape of the wavefront, chosen by the autotuner:
 *   queue_paths      paths in flight per pass (samples per chunk = queue_paths / pixels); 0 = 128M
 *                    (216 B of path state each: 27 GB of device memory at 128M)
 *   tile_w, tile_h   order of the pixels in the active list, i.e. which pixels
 *                    neighbouring lanes trace: tiles of tile_w x tile_h in
 *                    row-major tile order, row-major inside a tile; 0 x 0 =
 *                    plain row-major image order.
 * The output never depends on the tuning: every (pixel, sample) has its own
 * random stream and samples are accumulated in sample order. */
typedef struct pico_tuning {
    uint32_t queue_paths;
    uint32_t tile_w;
    uint32_t tile_h;
    uint32_t flags;          /* PICO_TUNE_* */
} pico_tuning;
/* pico_hip_render_views_denoise_d runs each view's denoiser on a second
 * stream (overlapping the remaining passes) instead of between two chunks. */
#define PICO_TUNE_SIDE_STREAM 1u
/* Batched views run shortest first: ordered by pixels x samples, as
 * successive launches whose cumulative work doubles (1/16, 1/8, ... of the
 * batch), so cheap views finish early; per-view times count from the first. */
#define PICO_TUNE_SHORTEST_FIRST 2u
int  pico_hip_set_tuning(pico_hip_ctx* ctx, const pico_tuning* tuning);
int  pico_hip_get_tuning(pico_hip_ctx* ctx, pico_tuning* tuning);

/* Renders n views in one wavefront: every pass of the path tracer covers the
 * live paths of all views.  stats may be NULL or hold n entries (per-view
 * completion times in render_ms / done_ms).  queue_paths overrides the
 * tuning's (0 = use the tuning). */
int  pico_hip_render_views_d(pico_hip_ctx* ctx, const pico_view* views, uint32_t n, uint32_t queue_paths,
                             pico_stats* stats);

/* pico_hip_render_views_d for views in host memory (the reference's
 * layout per view): staged to the device, rendered in one launch, copied
 * back.  With PICO_RENDER_RESUME in a view's params its buffers are uploaded
 * first. */
typedef struct pico_view_host {
    pico_camera        camera;
    pico_render_params params;
    float*    pixels;           /* host, W*H*3 */
    uint32_t* sample_count;     /* host, W*H   */
    float*    variance;         /* host, W*H*3 */
} pico_view_host;
int  pico_hip_render_views(pico_hip_ctx* ctx, const pico_view_host* views, uint32_t n, uint32_t queue_paths,
                           pico_stats* stats);

/* The same launch, each view followed by the learned denoiser as soon as its
 * last chunk of samples is done: G-buffer + network between chunks or on a
 * second stream (PICO_TUNE_SIDE_STREAM), overlapping the remaining passes.
 * d_denoised[i] receives view i's filtered image (W*H*3 f32, device); views
 * must fit the network's size.  A view whose network output holds a value
 * that is not finite gets the reference's a-trous filter instead (tiles of
 * its params' tile size, 64 x 64 when 0), as pico_hip_denoise_gbuffer_d;
 * stats[i].denoise_fallback says which. */
int  pico_hip_render_views_denoise_d(pico_hip_ctx* ctx, const pico_view* views, uint32_t n, pico_cnn* cnn,
                                     float* const* d_denoised, pico_stats* stats);

/* Autotuner: times the batch at reduced samples per pixel (the largest view
 * at calib_spp, the others scaled alike) for a fixed set of launch shapes
 * (queue depth x pixel-tile order, the default first), keeps the fastest
 * (ties: the earlier) as the context's tuning and caches the choice per
 * (scene, view sizes, spp, depth), so a repeated call returns it without
 * measuring.  Renders into the views' buffers.  chosen_ms / default_ms (may
 * be NULL) are the calibration times of the choice and of the default. */
int  pico_hip_autotune_views(pico_hip_ctx* ctx, const pico_view* views, uint32_t n, uint32_t calib_spp,
                             pico_tuning* chosen, float* chosen_ms, float* default_ms);

/* ---- file output --------------------------------------------------------- */
/* The file half of Scene::render_scene_to_file (Source/Core/Scene.cpp:236-246):
 * writes 3 x f32 pixels (host memory) as a Radiance .hdr file (RGBE); with
 * `reverse` the pixel order is reversed end to end first, as the reference
 * does before stbi_write_hdr.  Host-only; needs no context. */
int  pico_write_hdr(const char* path, uint32_t width, uint32_t height, const float* rgb, int reverse);

/* ---- procedural scenes (inputs of the configurations) ------------------ */
typedef struct pico_scene pico_scene;
/* Cornell box: 5 walls x 2 tris, two boxes of 12 tris, one quad light,
 * Dielectric walls/boxes with albedo U[0.2,0.8], roughness U[0.1,0.9]. */
int  pico_scene_cornell(uint32_t seed, uint32_t microfacet, pico_scene** out);
/* The same Cornell box with its two boxes transmissive: a Fresnel glass
 * block (index 1.5, roughness 0.1) and a Transparent one (index 1.33,
 * roughness 0.2), both with 0.95 transmitted energy. */
int  pico_scene_cornell_glass(uint32_t seed, uint32_t microfacet, pico_scene** out);
/* A Cornell room of the built-in implicit shapes: a Dielectric sphere, a
 * metal (Specular) cube turned 30 degrees, a Fresnel glass sphere, and three
 * lights, the ceiling quad, an emitting sphere and a small emitting cube. */
int  pico_scene_shapes(uint32_t seed, uint32_t microfacet, pico_scene** out);
/* Instanced scene: n_mesh_types meshes (icospheres and noise-displaced tori
 * of min_tris..max_tris triangles, geometric), n_instances random affine
 * instances in a box of side `extent`, 3 quad lights and a constant sky. */
int  pico_scene_instanced(uint32_t seed, uint32_t n_mesh_types, uint32_t n_instances, uint32_t min_tris,
                          uint32_t max_tris, float extent, uint32_t microfacet, pico_scene** out);
/* Randomised scenes used to train the learned denoiser, by seed % 3: a
 * jittered Cornell box, a small instanced cluster, or a 100^3 instanced field
 * like the benchmark scene seen from a random outside viewpoint. */
int  pico_scene_training(uint32_t seed, uint32_t microfacet, pico_scene** out);
/* A scene file in the reference's JSON format (sections GLOBALS, MESH,
 * MATERIALS, INSTANCE, CAMERA), read as Scene::Scene(ThreadPool&, path) and
 * its section handlers read it (Source/Core/Scene.cpp:36-84, 297-720):
 * meshes from Wavefront .obj files (triangulated, flat normals where the file
 * has none, equal corners joined, the file's first mesh), Constant materials,
 * instances T * R(quaternion) * S with the Diffuse, Specular, Dielectric,
 * Light, Delta, Transparent and Fresnel BSRDFs (Beckmann_All distribution),
 * the built-in Sphere and Cube assets (PICO_SHAPE_SPHERE / _CUBE meshes),
 * named cameras, and from GLOBALS the sun (SunDirection, SunColour) and the
 * sky (Skybox: six PNG cube faces, or one PNG or Radiance .hdr
 * equirectangular image).  Entries of a section are taken in the sorted
 * order of their names.  Image materials are refused.  On error PICO_ERR_ARG and
 * the reason (file, line, entry) in `err` (err_len bytes, may be NULL). */
int  pico_scene_load_json(const char* path, pico_scene** out, char* err, size_t err_len);
/* A model file through the reference's import path, Scene::Scene(ThreadPool&,
 * dir, aiScene*, options) with parse_node and add_material
 * (Source/Core/Scene.cpp:86-141, 722-971), for Wavefront .obj files as assimp
 * imports them with the flags of Source/main.cpp:53-59: the file's meshes
 * (split at objects, groups and material changes), each an instance with
 * the identity transform, a Light when its material emits (Ke) and otherwise
 * a Dielectric with the Beckmann_All lobe; materials from the file's mtllib
 * libraries (Kd, Ks, Ke; roughness 0 as the reference derives it; assimp's
 * default diffuse 0.6 where none applies).  Black sky and no sun: set them
 * with pico_scene_set_sky_image / pico_scene_set_sun as the -Skybox, -Sun*
 * options do.  No cameras (the reference takes the camera from its options).
 * Errors as pico_scene_load_json. */
int  pico_scene_load_obj(const char* path, pico_scene** out, char* err, size_t err_len);
/* The named cameras of a loaded scene, in name order (none for procedural
 * scenes); pico_scene_default_camera gives the first one with its own aspect
 * replaced by the argument. */
int  pico_scene_camera_count(const pico_scene* scene, uint32_t* count);
int  pico_scene_camera_get(const pico_scene* scene, uint32_t index, char* name, size_t name_len, pico_camera* out);
int  pico_scene_desc_get(const pico_scene* scene, pico_scene_desc* out);
/* The sun and the sky image of a scene, as the command-line options
 * -SunDirection/-SunColour and -Skybox set them for a scene the reference
 * imports (Source/Core/Scene.cpp:86-123): direction = colour = NULL turns the
 * sun off; img = NULL restores the constant sky.  The texels are copied. */
int  pico_scene_set_sun(pico_scene* scene, const float* direction, const float* colour);
int  pico_scene_set_sky_image(pico_scene* scene, const pico_sky_image* img);
/* A camera that frames the scene as its generator intends. */
int  pico_scene_default_camera(const pico_scene* scene, float aspect, pico_camera* out);
void pico_scene_destroy(pico_scene* scene);

#ifdef __cplusplus
}
#endif

#endif /* PICO_HIP_H */
