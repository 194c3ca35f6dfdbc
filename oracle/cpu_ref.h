/*
 * cpu_ref.h — TEST INFRASTRUCTURE ONLY.
 *
 * C interface of the CPU restatement of Pico's rendering hot path (the
 * oracle).  Only tests, the smoke check and the CPU-baseline leg of the bench
 * may load it; the product library never links or calls it.
 */
#ifndef PICO_CPU_REF_H
#define PICO_CPU_REF_H

#include <stdint.h>
#include "../include/pico_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Random-number source of the integrator.
 * CPUREF_RNG_REFERENCE: the reference's own generators (xorshift/Hammersley
 *   for sample dimensions and light-table draws, std::mt19937_64 with
 *   std::uniform_real_distribution<float> for light pick and Russian
 *   roulette), one integrator per pixel seeded with a per-pixel seed that
 *   persists over the pixel's samples (Source/Core/Scene.cpp:171-176).
 * CPUREF_RNG_PCG: two PCG32 streams per (pixel, sample) in place of those two
 *   generators, consumed in the same order; identical to the device path. */
#define CPUREF_RNG_REFERENCE 0u
#define CPUREF_RNG_PCG       1u

typedef struct cpuref_scene cpuref_scene;

cpuref_scene* cpuref_scene_create(const pico_scene_desc* desc, uint32_t sah_buckets);
void     cpuref_scene_destroy(cpuref_scene* s);
uint32_t cpuref_scene_node_count(const cpuref_scene* s, uint32_t* tlas_nodes, uint32_t* max_depth);

void cpuref_closest_hit(const cpuref_scene* s, uint32_t n, const float* origins, const float* dirs,
                        float* t, uint32_t* inst, uint32_t* prim, float* uv, float* pos, float* nrm);
/* Brute force over every instance and triangle (no BVH), for pinning. */
void cpuref_closest_hit_brute(const cpuref_scene* s, uint32_t n, const float* origins, const float* dirs,
                              float* t, uint32_t* inst, uint32_t* prim);
void cpuref_primary_ids(const cpuref_scene* s, const pico_camera* cam, uint32_t w, uint32_t h,
                        float xi_x, float xi_y, uint32_t* inst, uint32_t* prim);
void cpuref_camera_rays(const pico_camera* cam, uint32_t w, uint32_t h, uint32_t n, const uint32_t* px,
                        const uint32_t* py, const float* xi, float* dirs);

/* Per-pixel seed of CPUREF_RNG_REFERENCE mode. */
uint32_t cpuref_pixel_seed(uint32_t frame_seed, uint32_t pixel);

int  cpuref_render(const cpuref_scene* s, const pico_camera* cam, const pico_render_params* p,
                   uint32_t rng_mode, uint32_t threads, float* pixels, uint32_t* counts, float* variance,
                   pico_stats* stats);
/* Per-sample radiance, spp x H x W x 3, sample indices sample_base.. (PCG
 * mode) or the first spp samples of each pixel's integrator (reference mode). */
int  cpuref_render_samples(const cpuref_scene* s, const pico_camera* cam, const pico_render_params* p,
                           uint32_t rng_mode, uint32_t threads, uint32_t sample_base, float* out,
                           uint32_t* rays_out);
void cpuref_gbuffer(const cpuref_scene* s, const pico_camera* cam, uint32_t w, uint32_t h,
                    float* normal, float* position, float* albedo);

void cpuref_atrous(const float* rgb, const float* normal, const float* position, const float* albedo,
                   uint32_t w, uint32_t h, uint32_t tile_w, uint32_t tile_h, uint32_t levels, float* out);
void cpuref_atrous_fixed(const float* rgb, const float* normal, const float* position, const float* albedo,
                         uint32_t w, uint32_t h, uint32_t levels, float sigma_normal, float sigma_depth,
                         float sigma_albedo, float sigma_color, float* out);
void cpuref_tonemap(float* rgb, uint32_t w, uint32_t h);
double cpuref_relmse(const float* img, const double* ref_sum, double ref_count, uint32_t n_pixels);

/* Building blocks (unit-level pinning). kind: 0 cosine hemisphere,
 * 1 Beckmann_All, 2 GGX. */
void cpuref_dist_sample(uint32_t kind, uint32_t n, const float* xi, const float* v, const float* r, float* h);
void cpuref_dist_pdf(uint32_t kind, uint32_t n, const float* wo, const float* h, const float* r, float* pdf);
void cpuref_tangent_frame(uint32_t n, const float* v, const float* nrm, float* rows9);
void cpuref_xorshift(uint32_t seed, uint32_t n, uint32_t* out);
void cpuref_hammersley(uint32_t n, const uint32_t* bits, float* out2);
void cpuref_pcg32(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out);
void cpuref_pcg32_advance(uint64_t initstate, uint64_t initseq, uint64_t delta, uint32_t n, uint32_t* out);
void cpuref_alias_build(uint32_t n, const float* weights, float* bucket_weight, uint32_t* bucket_alias);
uint32_t cpuref_alias_sample(uint32_t n, const float* bucket_weight, const uint32_t* bucket_alias,
                             uint32_t seed, uint32_t draws, uint32_t* out);
void cpuref_triangle_areas(uint32_t n_tris, const float* positions, const uint32_t* indices, float* areas);
/* Transparent / Fresnel samples from rays whose IOR stack is the camera's
 * 1.0 plus the material's index where inside[i]; depth_after[i] is the
 * stack size after the sample. */
/* The sky seen in n directions (ImageCube::sample4, Source/Core/Image.cpp:361-466);
 * img NULL: the constant sky_rgb. */
void cpuref_sky_sample(const pico_sky_image* img, const float* sky_rgb, uint32_t n, const float* dirs, float* out);
void cpuref_btdf_sample(uint32_t bsdf, uint32_t microfacet, const pico_material* m, uint32_t n, const float* v,
                        const float* nrm, const uint8_t* inside, uint32_t rng_mode, uint64_t seed, float* l, float* p,
                        float* energy, uint32_t* depth_after);
/* BSDF sample at a shading point with a fixed random sequence (PCG). */
void cpuref_bsdf_sample(uint32_t bsdf, uint32_t microfacet, const pico_material* m, uint32_t n,
                        const float* v, const float* nrm, uint32_t rng_mode, uint64_t seed, float* l, float* p,
                        float* energy);

#ifdef __cplusplus
}
#endif
#endif
