/*
 * ref_harness.cpp — TEST INFRASTRUCTURE ONLY.
 *
 * A thin C interface over Pico's own hot-path sources, compiled from where
 * they lie in the reference tree (oracle/Makefile builds oracle/_ref/).  It
 * composes the reference's classes exactly as Source/Core/Scene.cpp does
 * (mesh BLAS from an aiMesh, UpperLevelBVH, MaterialManager, BSRDFs, lights,
 * Monte_Carlo_Integrator, Util::Tiler over the reference ThreadPool,
 * Util::atrous_denoise, Util::reinhard_tone_mapping) so that the CPU
 * restatement can be checked against the reference itself, and so that the
 * reference CPU renderer can be timed as the CPU baseline.
 *
 * Differences from Scene.cpp, all in scene set-up or seeding, none in the
 * hot path:
 *  - meshes come from a flattened pico_scene_desc instead of assimp/JSON;
 *    vertices are de-indexed (3 per triangle) with the triangle id stored in
 *    the vertex colour so that the hit record names the primitive (the
 *    intersection and interpolation results are unchanged by this);
 *  - materials are constant EvaluatedMaterials;
 *  - per-pixel integrator seeds come from a fixed hash of (frame, pixel)
 *    instead of std::random_device, so runs are repeatable.
 */
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "Core/Camera.hpp"
#include "Core/Image.hpp"
#include "Core/LowerLevelImplicitShapesBVH.hpp"
#include "Core/LowerLevelMeshBVH.hpp"
#include "Core/MaterialManager.hpp"
#include "Core/RandUtils.hpp"
#include "Core/ThreadPool.hpp"
#include "Core/UpperLevelBVH.hpp"
#include "Core/vectorUtils.hpp"
#include "Render/BSRDF.hpp"
#include "Render/BasicMaterials.hpp"
#include "Render/Distributions.hpp"
#include "Render/Integrators.hpp"
#include "Util/AliasTable.hpp"
#include "Util/Denoisers.hpp"
#include "Util/Tiler.hpp"
#include "Util/ToneMappers.hpp"
#include "assimp/mesh.h"
#include "stb_image.h"

#include "../include/pico_hip.h"

namespace {

using namespace Core::Acceleration_Structures;

class FixedMaterial : public Core::Material {
public:
    explicit FixedMaterial(const Core::EvaluatedMaterial& m) : m_(m) {}
    Core::EvaluatedMaterial evaluate_material(const glm::vec2&) const override { return m_; }
    bool is_light() const override { return m_.emits_light(); }
    size_t get_residence_size() const override { return 0; }
    bool is_resident() const override { return true; }
    void make_resident(void*) override {}
    void make_nonresident() override {}
private:
    Core::EvaluatedMaterial m_;
};

// A constant material with an index of refraction: the transmissive BSDFs
// read it through Render::ConstantTransparentMaterial (BSRDF.hpp:193-199).
class FixedTransparentMaterial : public Render::ConstantTransparentMaterial {
public:
    FixedTransparentMaterial(const Core::EvaluatedMaterial& m, float ior) : ConstantTransparentMaterial(1.0f, ior), m_(m) {}
    Core::EvaluatedMaterial evaluate_material(const glm::vec2&) const override { return m_; }
    bool is_light() const override { return m_.emits_light(); }
private:
    Core::EvaluatedMaterial m_;
};

std::unique_ptr<Core::Material> make_material(const pico_material& pm) {
    Core::EvaluatedMaterial em{};
    em.diffuse = glm::vec3(pm.diffuse[0], pm.diffuse[1], pm.diffuse[2]);
    em.specular = glm::vec3(pm.specular[0], pm.specular[1], pm.specular[2]);
    em.roughness = pm.roughness;
    em.normal = glm::vec3(0.0f, 1.0f, 0.0f);
    em.emissive = glm::vec3(pm.emissive[0], pm.emissive[1], pm.emissive[2]);
    if (pm.index_of_refraction > 0.0f) return std::make_unique<FixedTransparentMaterial>(em, pm.index_of_refraction);
    return std::make_unique<FixedMaterial>(em);
}

uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
uint32_t pixel_seed(uint32_t frame, uint32_t pixel) {
    return (uint32_t)mix64(((uint64_t)frame << 32) ^ (uint64_t)pixel ^ 0x5851F42D4C957F2DULL);
}

}  // namespace

struct ref_scene {
    std::vector<std::unique_ptr<LowerLevelBVH>> blas;
    std::vector<uint32_t> inst_shape;  // PICO_SHAPE_* of each instance's mesh
    UpperLevelBVH tlas;
    Core::MaterialManager mm;
    std::vector<Scene::Light> lights;
    std::unique_ptr<Core::ImageCube> sky;
    Scene::Sun sun{};
    std::map<const Render::BSRDF*, uint32_t> inst_of;
};

// The sky as Scene::process_globals makes it (Source/Core/Scene.cpp:639-690):
// a cube of 6 RGBA8 faces, a single RGBA float image, or here also the
// constant sky as a 1x1 float image.  Image frees its data with free().
static std::unique_ptr<Core::ImageCube> make_sky(const pico_sky_image* img, const float* rgb) {
    if (!img) {
        float* px = static_cast<float*>(std::malloc(16));
        px[0] = rgb[0]; px[1] = rgb[1]; px[2] = rgb[2]; px[3] = 1.0f;
        return std::make_unique<Core::ImageCube>(reinterpret_cast<unsigned char*>(px), Core::ImageExtent{1, 1, 1},
                                                 Core::Format::kRGBA_F);
    }
    const size_t bytes = (size_t)img->width * img->height * img->faces * (img->format == PICO_TEXEL_RGBA32F ? 16 : 4);
    unsigned char* px = static_cast<unsigned char*>(std::malloc(bytes));
    std::memcpy(px, img->texels, bytes);
    return std::make_unique<Core::ImageCube>(px, Core::ImageExtent{img->width, img->height, img->faces},
                                             img->format == PICO_TEXEL_RGBA32F ? Core::Format::kRGBA_F
                                                                               : Core::Format::kRBGA_8UNorm);
}

static Scene::Camera make_camera(const pico_camera* c, uint32_t w, uint32_t h) {
    Scene::Camera cam(glm::vec3(c->position[0], c->position[1], c->position[2]),
                      glm::vec3(c->direction[0], c->direction[1], c->direction[2]), c->aspect, 0.1f, c->far_plane);
    cam.setUp(glm::vec3(c->up[0], c->up[1], c->up[2]));
    cam.set_resolution(glm::uvec2(w, h));
    return cam;
}

extern "C" {

ref_scene* ref_scene_create(const pico_scene_desc* d) {
    auto* s = new ref_scene();
    for (uint32_t mi = 0; mi < d->n_meshes; ++mi) {
        // the built-in shapes, as Scene::Scene registers them (Source/Core/Scene.cpp:49-56)
        const uint32_t shape = d->mesh_shape ? d->mesh_shape[mi] : PICO_SHAPE_TRIANGLES;
        if (shape == PICO_SHAPE_SPHERE) {
            s->blas.push_back(std::make_unique<LowerLevelSphereBVH>(0.5f));
            continue;
        }
        if (shape == PICO_SHAPE_CUBE) {
            s->blas.push_back(std::make_unique<LowerLevelCube>());
            continue;
        }
        const uint32_t v0 = d->mesh_vertex_offset[mi];
        const uint32_t t0 = d->mesh_triangle_offset[mi], t1 = d->mesh_triangle_offset[mi + 1];
        const uint32_t nt = t1 - t0;
        aiMesh m;
        m.mNumVertices = nt * 3;
        m.mNumFaces = nt;
        m.mVertices = new aiVector3D[nt * 3];
        m.mNormals = new aiVector3D[nt * 3];
        m.mColors[0] = new aiColor4D[nt * 3];
        m.mFaces = new aiFace[nt];
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t t = 0; t < nt; ++t) {
            m.mFaces[t].mNumIndices = 3;
            m.mFaces[t].mIndices = new unsigned int[3];
            for (int k = 0; k < 3; ++k) {
                const uint32_t v = v0 + d->indices[3 * (t0 + t) + k];
                const uint32_t o = 3 * t + k;
                m.mVertices[o] = aiVector3D(d->positions[3 * v], d->positions[3 * v + 1], d->positions[3 * v + 2]);
                m.mNormals[o] = aiVector3D(d->normals[3 * v], d->normals[3 * v + 1], d->normals[3 * v + 2]);
                m.mColors[0][o] = aiColor4D((float)t, 0.0f, 0.0f, 1.0f);
                m.mFaces[t].mIndices[k] = o;
                for (int c = 0; c < 3; ++c) {
                    const float p = d->positions[3 * v + c];
                    mn[c] = p < mn[c] ? p : mn[c];
                    mx[c] = mx[c] < p ? p : mx[c];
                }
            }
        }
        m.mAABB.mMin = aiVector3D(mn[0], mn[1], mn[2]);
        m.mAABB.mMax = aiVector3D(mx[0], mx[1], mx[2]);
        s->blas.push_back(std::make_unique<LowerLevelMeshBVH>(&m));
    }
    std::vector<Core::MaterialManager::MaterialID> mids;
    for (uint32_t i = 0; i < d->n_materials; ++i) {
        std::unique_ptr<Core::Material> mat = make_material(d->materials[i]);
        mids.push_back(s->mm.add_material(mat));
    }
    std::vector<char> is_light(d->n_instances, 0);
    for (uint32_t l = 0; l < d->n_lights; ++l) is_light[d->light_instances[l]] = 1;
    std::vector<glm::mat4> xf(d->n_instances);
    for (uint32_t i = 0; i < d->n_instances; ++i) {
        glm::mat4 M;
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) M[c][r] = d->transforms[16 * i + 4 * c + r];
        xf[i] = M;
        const auto mat = mids[d->instance_material[i]];
        std::unique_ptr<Render::BSRDF> b;
        // Source/Core/Scene.cpp:367-417
        switch (d->instance_bsdf[i]) {
            case PICO_BSDF_DIFFUSE: {
                std::unique_ptr<Render::Distribution> dist = std::make_unique<Render::Cos_Weighted_Hemisphere_Distribution>();
                b = std::make_unique<Render::Diffuse_BRDF>(dist, s->mm, mat);
                break;
            }
            case PICO_BSDF_SPECULAR: {
                std::unique_ptr<Render::Distribution> dist = std::make_unique<Render::Beckmann_All_Microfacet_Distribution>();
                b = std::make_unique<Render::Specular_BRDF>(dist, s->mm, mat);
                break;
            }
            case PICO_BSDF_DIELECTRIC: {
                std::unique_ptr<Render::Distribution> dd = std::make_unique<Render::Cos_Weighted_Hemisphere_Distribution>();
                std::unique_ptr<Render::Distribution> sd = std::make_unique<Render::Beckmann_All_Microfacet_Distribution>();
                b = std::make_unique<Render::Dielectric_BRDF>(dd, sd, s->mm, mat);
                break;
            }
            case PICO_BSDF_LIGHT:
                b = std::make_unique<Render::Light_BRDF>(s->mm, mat);
                break;
            case PICO_BSDF_TRANSPARENT: {
                std::unique_ptr<Render::Distribution> dist = std::make_unique<Render::Beckmann_All_Microfacet_Distribution>();
                b = std::make_unique<Render::Transparent_BTDF>(dist, s->mm, mat);
                break;
            }
            case PICO_BSDF_FRESNEL: {
                std::unique_ptr<Render::Distribution> sd = std::make_unique<Render::Beckmann_All_Microfacet_Distribution>();
                std::unique_ptr<Render::Distribution> td = std::make_unique<Render::Beckmann_All_Microfacet_Distribution>();
                b = std::make_unique<Render::Fresnel_BTDF>(sd, td, s->mm, mat);
                break;
            }
            default:
                b = std::make_unique<Render::Specular_Delta_BRDF>(s->mm, mat);
                break;
        }
        const Render::BSRDF* raw = b.get();
        s->inst_shape.push_back(d->mesh_shape ? d->mesh_shape[d->instance_mesh[i]] : PICO_SHAPE_TRIANGLES);
        s->tlas.add_lower_level_bvh(s->blas[d->instance_mesh[i]].get(), M, b);
        s->inst_of[raw] = i;
    }
    for (uint32_t l = 0; l < d->n_lights; ++l) {
        const uint32_t ii = d->light_instances[l];
        auto* g = s->blas[d->instance_mesh[ii]].get();
        g->generate_sampling_data();
        s->lights.push_back({xf[ii], glm::inverse(xf[ii]), g});
    }
    s->tlas.build();
    s->sky = make_sky(d->sky_image, d->sky_rgb);
    s->sun.m_sky_box = s->sky.get();
    s->sun.m_use_sun = d->use_sun != 0;
    s->sun.m_sun_direction = glm::vec3(d->sun_direction[0], d->sun_direction[1], d->sun_direction[2]);
    s->sun.m_sun_colour = glm::vec3(d->sun_colour[0], d->sun_colour[1], d->sun_colour[2]);
    return s;
}

/* ImageCube::sample4 of a sky image (or the 1x1 constant sky) for n
 * directions; rgb out (alpha dropped as the integrator's vec3 does). */
void ref_sky_sample(const pico_sky_image* img, const float* sky_rgb, uint32_t n, const float* dirs, float* out) {
    auto sky = make_sky(img, sky_rgb);
    for (uint32_t i = 0; i < n; ++i) {
        const glm::vec4 c = sky->sample4(glm::vec3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
}

void ref_scene_destroy(ref_scene* s) { delete s; }

/* The reference's image reader (its stb_image, compiled from the reference
 * tree) as Scene::process_globals calls it: stbi_load (want_float 0) or
 * stbi_loadf (want_float 1) with 4 channels.  Returns stb's buffer (free with
 * ref_stbi_free) or NULL. */
void* ref_stbi_load(const char* path, int want_float, int* w, int* h) {
    int comp = 0;
    if (want_float) return stbi_loadf(path, w, h, &comp, 4);
    return stbi_load(path, w, h, &comp, 4);
}
void ref_stbi_free(void* p) { stbi_image_free(p); }

static bool ref_hit(const ref_scene* s, Core::Ray& r, InterpolatedVertex& v, float& t, uint32_t& inst, uint32_t& prim) {
    if (!s->tlas.get_closest_intersection(r, &v)) return false;
    glm::vec4 dd = v.mPosition - r.mOrigin;
    t = glm::length(dd);
    inst = s->inst_of.at(v.m_bsrdf);
    prim = s->inst_shape[inst] ? 0u : (uint32_t)std::lround(v.mVertexColour.x);
    return true;
}

void ref_closest_hit(const ref_scene* s, uint32_t n, const float* o, const float* d, float* t, uint32_t* inst,
                     uint32_t* prim, float* pos, float* nrm) {
    for (uint32_t i = 0; i < n; ++i) {
        Core::Ray r{};
        r.mOrigin = glm::vec4(o[3 * i], o[3 * i + 1], o[3 * i + 2], 1.0f);
        r.mDirection = glm::vec3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        r.mLenght = 10000.0f;
        InterpolatedVertex v{};
        float tt = INFINITY; uint32_t ii = ~0u, pp = ~0u;
        bool ok = ref_hit(s, r, v, tt, ii, pp);
        t[i] = tt; inst[i] = ii; prim[i] = pp;
        if (pos) { pos[3 * i] = ok ? v.mPosition.x : 0; pos[3 * i + 1] = ok ? v.mPosition.y : 0; pos[3 * i + 2] = ok ? v.mPosition.z : 0; }
        if (nrm) { nrm[3 * i] = ok ? v.mNormal.x : 0; nrm[3 * i + 1] = ok ? v.mNormal.y : 0; nrm[3 * i + 2] = ok ? v.mNormal.z : 0; }
    }
}

void ref_camera_rays(const pico_camera* c, uint32_t w, uint32_t h, uint32_t n, const uint32_t* px, const uint32_t* py,
                     const float* xi, float* dirs) {
    Scene::Camera cam = make_camera(c, w, h);
    for (uint32_t i = 0; i < n; ++i) {
        Core::Ray r = cam.generate_ray(glm::vec2(xi[2 * i], xi[2 * i + 1]), glm::uvec2(px[i], py[i]));
        dirs[3 * i] = r.mDirection.x; dirs[3 * i + 1] = r.mDirection.y; dirs[3 * i + 2] = r.mDirection.z;
    }
}

void ref_primary_ids(const ref_scene* s, const pico_camera* c, uint32_t w, uint32_t h, float xx, float xy,
                     uint32_t* inst, uint32_t* prim) {
    Scene::Camera cam = make_camera(c, w, h);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            Core::Ray r = cam.generate_ray(glm::vec2(xx, xy), glm::uvec2(x, y));
            InterpolatedVertex v{};
            float t; uint32_t ii = ~0u, pp = ~0u;
            ref_hit(s, r, v, t, ii, pp);
            inst[y * w + x] = ii; prim[y * w + x] = pp;
        }
}

/* Per-sample radiance from the reference integrator: one
 * Monte_Carlo_Integrator per pixel (seeded as in Source/Core/Scene.cpp:171),
 * spp calls of integrate_ray.  out: spp x H x W x 3. */
void ref_render_samples(const ref_scene* s, const pico_camera* c, const pico_render_params* p, float* out) {
    Scene::Camera cam = make_camera(c, p->width, p->height);
    const uint32_t W = p->width, H = p->height, N = W * H;
    ref_scene* ms = const_cast<ref_scene*>(s);
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            const uint32_t flat = y * W + x;
            Render::Monte_Carlo_Integrator integ(ms->tlas, ms->mm, ms->lights, ms->sun, pixel_seed(p->frame_seed, flat));
            for (uint32_t i = 0; i < p->spp; ++i) {
                glm::vec3 r = integ.integrate_ray(cam, glm::uvec2(x, y), p->max_depth);
                const size_t o = ((size_t)i * N + flat) * 3;
                out[o] = r.x; out[o + 1] = r.y; out[o + 2] = r.z;
            }
        }
}

/* The reference render loop (tile lambda of Source/Core/Scene.cpp:145-201
 * without the per-tile pixel shuffle) over Util::Tiler and the reference
 * ThreadPool with `threads` workers.  Returns the number of closest-hit
 * queries is not observable from the reference, so the caller counts rays
 * with the restatement on the same seeds. */
int ref_render(const ref_scene* s, const pico_camera* c, const pico_render_params* p, uint32_t threads, float* pixels,
               uint32_t* counts, float* variance) {
    Scene::Camera cam = make_camera(c, p->width, p->height);
    ref_scene* ms = const_cast<ref_scene*>(s);
    if (!(p->flags & PICO_RENDER_RESUME)) {
        std::memset(pixels, 0, sizeof(float) * 3 * p->width * p->height);
        std::memset(counts, 0, sizeof(uint32_t) * p->width * p->height);
        std::memset(variance, 0, sizeof(float) * 3 * p->width * p->height);
    }
    glm::vec3* P = reinterpret_cast<glm::vec3*>(pixels);
    glm::vec3* V = reinterpret_cast<glm::vec3*>(variance);
    auto tile_fn = [&](const glm::uvec2 start, const glm::uvec2& ts, const glm::uvec2& res, const uint32_t,
                       const Scene::Camera& camera) -> bool {
        for (uint32_t ty = 0; ty < ts.y; ++ty)
            for (uint32_t tx = 0; tx < ts.x; ++tx) {
                const uint32_t flat = (start.y + ty) * res.x + start.x + tx;
                if (counts[flat] >= p->spp) continue;
                const glm::uvec2 loc(flat % res.x, flat / res.x);
                Render::Monte_Carlo_Integrator integ(ms->tlas, ms->mm, ms->lights, ms->sun, pixel_seed(p->frame_seed, flat));
                for (uint32_t i = 0; i < p->spp; ++i) {
                    uint32_t prev = counts[flat];
                    glm::vec3 r = integ.integrate_ray(camera, loc, p->max_depth);
                    if (prev < 1) {
                        counts[flat] = 1;
                        V[flat] = glm::vec4(0, 0, 0, 0);
                    } else {
                        const glm::vec3& pp = P[flat];
                        const glm::vec3& pv = V[flat];
                        const glm::vec3 np = pp + ((r - pp) * (1.0f / static_cast<float>(prev)));
                        V[flat] = ((pv * (float)prev) + ((r - pp) * (r - np))) / (float)(prev + 1);
                        counts[flat] = prev + 1;
                        r = np;
                    }
                    P[flat] = r;
                    if (prev > 4 && glm::all(glm::lessThanEqual(V[flat], glm::vec3(p->max_variance)))) break;
                }
            }
        return false;
    };
    ThreadPool pool(threads ? threads : std::thread::hardware_concurrency());
    Core::Rand::xorshift_random rng(p->frame_seed);
    Util::Tiler tiler(pool, rng, glm::uvec2(p->width, p->height),
                      glm::uvec2(p->tile_w ? p->tile_w : 64, p->tile_h ? p->tile_h : 64));
    tiler.execute_over_surface(tile_fn, cam);
    return 0;
}

void ref_gbuffer(const ref_scene* s, const pico_camera* c, uint32_t w, uint32_t h, float* normal, float* position,
                 float* albedo) {
    Scene::Camera cam = make_camera(c, w, h);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const uint32_t flat = y * w + x;
            Core::Ray ray = cam.generate_ray(glm::vec2(0.0f, 0.0f), glm::uvec2(x, y));
            InterpolatedVertex frag{};
            glm::vec3 n, p, a;
            if (s->tlas.get_closest_intersection(ray, &frag)) {
                n = frag.mNormal;
                p = cam.getViewMatrix() * frag.mPosition;
                a = s->mm.evaluate_material(frag.m_bsrdf->get_material_id(), frag.mUV).diffuse;
            } else {
                n = -ray.mDirection;
                p = ray.mDirection * ray.mLenght;
                a = s->sun.m_sky_box->sample4(ray.mDirection);
            }
            std::memcpy(normal + 3 * flat, &n, 12);
            std::memcpy(position + 3 * flat, &p, 12);
            std::memcpy(albedo + 3 * flat, &a, 12);
        }
}

void ref_atrous(const float* rgb, const float* normal, const float* position, const float* albedo, uint32_t w, uint32_t h,
                uint32_t tw, uint32_t th, float* out) {
    ThreadPool pool(4);
    Core::Rand::xorshift_random rng(1);
    Util::Tiler tiler(pool, rng, glm::uvec2(w, h), glm::uvec2(tw, th));
    glm::vec3* r = Util::atrous_denoise(reinterpret_cast<const glm::vec3*>(rgb), reinterpret_cast<const glm::vec3*>(normal),
                                        reinterpret_cast<const glm::vec3*>(position),
                                        reinterpret_cast<const glm::vec3*>(albedo), tiler);
    std::memcpy(out, r, sizeof(float) * 3 * w * h);
    delete[] r;
}

void ref_tonemap(float* rgb, uint32_t w, uint32_t h) {
    ThreadPool pool(4);
    Core::Rand::xorshift_random rng(1);
    Util::Tiler tiler(pool, rng, glm::uvec2(w, h), glm::uvec2(64, 64));
    Util::reinhard_tone_mapping(reinterpret_cast<glm::vec3*>(rgb), glm::uvec2(w, h), tiler);
}

void ref_dist_sample(uint32_t kind, uint32_t n, const float* xi, const float* v, const float* r, float* h) {
    Render::Cos_Weighted_Hemisphere_Distribution cd;
    Render::Beckmann_All_Microfacet_Distribution bd;
    Render::Distribution* d = kind == 0 ? (Render::Distribution*)&cd : (Render::Distribution*)&bd;
    for (uint32_t i = 0; i < n; ++i) {
        glm::vec3 H = d->sample(glm::vec2(xi[2 * i], xi[2 * i + 1]), glm::vec3(v[3 * i], v[3 * i + 1], v[3 * i + 2]), r[i]);
        h[3 * i] = H.x; h[3 * i + 1] = H.y; h[3 * i + 2] = H.z;
    }
}

void ref_dist_pdf(uint32_t kind, uint32_t n, const float* wo, const float* h, const float* r, float* pdf) {
    Render::Cos_Weighted_Hemisphere_Distribution cd;
    Render::Beckmann_All_Microfacet_Distribution bd;
    Render::Distribution* d = kind == 0 ? (Render::Distribution*)&cd : (Render::Distribution*)&bd;
    for (uint32_t i = 0; i < n; ++i)
        pdf[i] = d->pdf(glm::vec3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), glm::vec3(h[3 * i], h[3 * i + 1], h[3 * i + 2]), r[i]);
}

void ref_tangent_frame(uint32_t n, const float* v, const float* nrm, float* rows9) {
    for (uint32_t i = 0; i < n; ++i) {
        glm::mat3 m = Core::TangentSpace::construct_world_to_tangent_transform(glm::vec3(v[3 * i], v[3 * i + 1], v[3 * i + 2]),
                                                                                glm::vec3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]));
        for (int row = 0; row < 3; ++row)
            for (int col = 0; col < 3; ++col) rows9[9 * i + 3 * row + col] = m[col][row];
    }
}

void ref_xorshift(uint32_t seed, uint32_t n, uint32_t* out) {
    Core::Rand::xorshift_random x(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = x.next();
}

void ref_hammersley(uint32_t n, const uint32_t* bits, float* out2) {
    for (uint32_t i = 0; i < n; ++i) {
        glm::vec2 v = Core::Rand::hammersley(bits[i], std::numeric_limits<uint32_t>::max());
        out2[2 * i] = v.x; out2[2 * i + 1] = v.y;
    }
}

void ref_alias_sample(uint32_t n, const float* weights, uint32_t seed, uint32_t draws, uint32_t* out) {
    Util::AliasTable t(std::vector<float>(weights, weights + n));
    Core::Rand::xorshift_random x(seed);
    float pdf;
    for (uint32_t i = 0; i < draws; ++i) out[i] = t.sample(x, pdf);
}

/* BSDF sample with the reference generator (Hammersley_Generator(seed)).
 * v is the direction towards the viewer (the incoming ray is -v). */
void ref_bsdf_sample(uint32_t bsdf, const pico_material* pm, uint32_t n, const float* v, const float* nrm, uint32_t seed,
                     float* l, float* p, float* energy) {
    Core::MaterialManager mm;
    Core::EvaluatedMaterial em{};
    em.diffuse = glm::vec3(pm->diffuse[0], pm->diffuse[1], pm->diffuse[2]);
    em.specular = glm::vec3(pm->specular[0], pm->specular[1], pm->specular[2]);
    em.roughness = pm->roughness;
    em.emissive = glm::vec3(pm->emissive[0], pm->emissive[1], pm->emissive[2]);
    std::unique_ptr<Core::Material> mat = std::make_unique<FixedMaterial>(em);
    auto id = mm.add_material(mat);
    std::unique_ptr<Render::BSRDF> b;
    std::unique_ptr<Render::Distribution> d1 = std::make_unique<Render::Cos_Weighted_Hemisphere_Distribution>();
    std::unique_ptr<Render::Distribution> d2 = std::make_unique<Render::Beckmann_All_Microfacet_Distribution>();
    switch (bsdf) {
        case PICO_BSDF_DIFFUSE: b = std::make_unique<Render::Diffuse_BRDF>(d1, mm, id); break;
        case PICO_BSDF_SPECULAR: b = std::make_unique<Render::Specular_BRDF>(d2, mm, id); break;
        case PICO_BSDF_DIELECTRIC: b = std::make_unique<Render::Dielectric_BRDF>(d1, d2, mm, id); break;
        case PICO_BSDF_LIGHT: b = std::make_unique<Render::Light_BRDF>(mm, id); break;
        default: b = std::make_unique<Render::Specular_Delta_BRDF>(mm, id); break;
    }
    Core::Rand::Hammersley_Generator g(seed);
    for (uint32_t i = 0; i < n; ++i) {
        InterpolatedVertex frag{};
        frag.mNormal = glm::vec3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        Core::Ray ray{};
        ray.mDirection = -glm::vec3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        Render::Sample smp = b->sample(g, frag, ray);
        l[3 * i] = smp.L.x; l[3 * i + 1] = smp.L.y; l[3 * i + 2] = smp.L.z;
        p[i] = smp.P;
        energy[3 * i] = smp.energy.x; energy[3 * i + 1] = smp.energy.y; energy[3 * i + 2] = smp.energy.z;
    }
}


/* Transparent / Fresnel BTDF samples with the reference generator: each from
 * a ray whose IOR stack is the camera's 1.0 (Camera.cpp:23), plus the
 * material's index when inside[i]; inside_after[i] is Ray::inside_geometry
 * after the sample. */
void ref_btdf_sample(uint32_t bsdf, const pico_material* pm, uint32_t n, const float* v, const float* nrm,
                     const uint8_t* inside, uint32_t seed, float* l, float* p, float* energy, uint8_t* inside_after) {
    Core::MaterialManager mm;
    std::unique_ptr<Core::Material> mat = make_material(*pm);
    auto id = mm.add_material(mat);
    std::unique_ptr<Render::BSRDF> b;
    std::unique_ptr<Render::Distribution> d1 = std::make_unique<Render::Beckmann_All_Microfacet_Distribution>();
    std::unique_ptr<Render::Distribution> d2 = std::make_unique<Render::Beckmann_All_Microfacet_Distribution>();
    if (bsdf == PICO_BSDF_FRESNEL) b = std::make_unique<Render::Fresnel_BTDF>(d1, d2, mm, id);
    else b = std::make_unique<Render::Transparent_BTDF>(d1, mm, id);
    Core::Rand::Hammersley_Generator g(seed);
    for (uint32_t i = 0; i < n; ++i) {
        InterpolatedVertex frag{};
        frag.mNormal = glm::vec3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
        Core::Ray ray{};
        ray.push_index_of_refraction(1.0f);
        if (inside[i]) ray.push_index_of_refraction(pm->index_of_refraction);
        ray.mDirection = -glm::vec3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        Render::Sample smp = b->sample(g, frag, ray);
        l[3 * i] = smp.L.x; l[3 * i + 1] = smp.L.y; l[3 * i + 2] = smp.L.z;
        p[i] = smp.P;
        energy[3 * i] = smp.energy.x; energy[3 * i + 1] = smp.energy.y; energy[3 * i + 2] = smp.energy.z;
        inside_after[i] = ray.inside_geometry() ? 1 : 0;
    }
}

// Radiance .hdr through the reference's own writer and reader (the file half
// of Scene::render_scene_to_file, Source/Core/Scene.cpp:236-246).
int stbi_write_hdr(char const* filename, int w, int h, int comp, const float* data);
float* stbi_loadf(char const* filename, int* x, int* y, int* comp, int req_comp);
void stbi_image_free(void* retval_from_stbi_load);

int ref_write_hdr(const char* path, int w, int h, const float* rgb) { return stbi_write_hdr(path, w, h, 3, rgb); }

int ref_load_hdr(const char* path, int w, int h, float* out) {
    int x = 0, y = 0, comp = 0;
    float* d = stbi_loadf(path, &x, &y, &comp, 3);
    if (!d) return -1;
    const int ok = (x == w && y == h) ? 0 : -2;
    if (!ok) std::memcpy(out, d, sizeof(float) * 3 * (size_t)w * h);
    stbi_image_free(d);
    return ok;
}
}  // extern "C"
