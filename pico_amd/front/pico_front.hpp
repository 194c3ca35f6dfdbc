// C++ front of the MI355X path with the call shapes of Pico's own renderer:
//
//   Scene::Scene::render_scene_to_memory(const Camera&, const RenderParams&, const bool* should_quit)
//       (Source/Core/Scene.hpp:68, Source/Core/Scene.cpp:143-235)
//   Scene::Scene::render_scene_to_file(const Camera&, RenderParams&, const char*)
//       (Source/Core/Scene.cpp:236-246)
//   glm::vec3* Util::atrous_denoise(pixels, normals, position, diffuse, Tiler&)
//       (Source/Util/Denoisers.hpp:9; returns a new[] buffer the caller delete[]s)
//   Scene::Scene(ThreadPool&, const std::filesystem::path& sceneFile) and
//   Scene::get_camera(name) (Source/Core/Scene.hpp:64-79): SceneFile
//
// Header-only; it calls the C ABI of libpico_hip.so (include/pico_hip.h) and
// throws nothing: failures are reported through the int status of each call,
// with the text in last_error().  Vec3 is any 12-byte {x, y, z} float type
// (glm::vec3 in Pico).
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <thread>

#include "../../include/pico_hip.h"

namespace pico_front {

// Scene::RenderParams (Source/Core/Scene.hpp:30-43), same members and meaning.
template <class Vec3>
struct RenderParams {
    uint32_t m_maxRayDepth;
    uint32_t m_maxSamples;
    float m_maxVariance;
    uint32_t m_Height;
    uint32_t m_Width;
    bool m_tonemap;
    bool m_denoise;
    Vec3* m_Pixels;
    uint32_t* m_SampleCount;
    Vec3* m_variance;
};

// A scene file in the reference's JSON format, loaded as Scene::Scene(ThreadPool&,
// path) loads it, with its named cameras as Scene::get_camera gives them.
class SceneFile {
public:
    // as Pico's main() chooses (Source/main.cpp:45-60): a .json scene file,
    // or any other file as a model through the import path
    explicit SceneFile(const std::string& path) {
        const bool json = path.size() >= 5 && path.compare(path.size() - 5, 5, ".json") == 0;
        status_ = json ? pico_scene_load_json(path.c_str(), &scene_, error_, sizeof error_)
                       : pico_scene_load_obj(path.c_str(), &scene_, error_, sizeof error_);
        if (status_ != PICO_OK) return;
        status_ = pico_scene_desc_get(scene_, &desc_);
        uint32_t n = 0;
        if (status_ == PICO_OK) status_ = pico_scene_camera_count(scene_, &n);
        for (uint32_t i = 0; status_ == PICO_OK && i < n; ++i) {
            char name[256];
            pico_camera cam{};
            status_ = pico_scene_camera_get(scene_, i, name, sizeof name, &cam);
            if (status_ == PICO_OK) cameras_[name] = cam;
        }
    }
    ~SceneFile() { pico_scene_destroy(scene_); }
    SceneFile(const SceneFile&) = delete;
    SceneFile& operator=(const SceneFile&) = delete;

    int status() const { return status_; }
    const char* error() const { return error_; }
    const pico_scene_desc& desc() const { return desc_; }
    // nullptr when the file has no camera of that name, as the reference
    const pico_camera* get_camera(const std::string& name) const {
        auto it = cameras_.find(name);
        return it == cameras_.end() ? nullptr : &it->second;
    }

private:
    pico_scene* scene_ = nullptr;
    pico_scene_desc desc_{};
    std::map<std::string, pico_camera> cameras_;
    int status_ = PICO_ERR_STATE;
    char error_[512] = {0};
};

class Renderer {
public:
    explicit Renderer(int device = 0) { status_ = pico_hip_create(device, &ctx_); }
    ~Renderer() { pico_hip_destroy(ctx_); }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    int status() const { return status_; }
    pico_hip_ctx* context() const { return ctx_; }
    std::string last_error() const { return ctx_ ? pico_hip_last_error(ctx_) : "no device context"; }

    // Scene::Scene's BVH build (Source/Core/Scene.cpp:36-141): a flattened scene.
    int upload(const pico_scene_desc& scene) { return ctx_ ? pico_hip_upload_scene(ctx_, &scene) : status_; }
    int upload(const SceneFile& file) { return file.status() != PICO_OK ? file.status() : upload(file.desc()); }

    // Scene::render_scene_to_memory: render (resuming from the pixels'
    // sample counts like the reference does), then, as the reference orders
    // them, a-trous on the G-buffer and Reinhard tone mapping, in place.
    template <class Vec3>
    int render_scene_to_memory(const pico_camera& camera, const RenderParams<Vec3>& params, const bool* should_quit,
                               uint32_t frame_seed = 0) {
        static_assert(sizeof(Vec3) == 12, "packed 3 x f32 per pixel");
        if (!ctx_) return status_;
        pico_render_params p{};
        p.width = params.m_Width;
        p.height = params.m_Height;
        p.spp = params.m_maxSamples;
        p.max_depth = params.m_maxRayDepth;
        p.tile_w = p.tile_h = 64;  // Scene.cpp:206
        p.frame_seed = frame_seed;
        p.max_variance = params.m_maxVariance;
        p.flags = PICO_RENDER_RESUME;
        // the C ABI reads an int32 flag; mirror the caller's bool into it
        // while the render runs
        volatile int32_t cancel = should_quit && *should_quit ? 1 : 0;
        std::atomic<bool> done{false};
        std::thread watch;
        if (should_quit)
            watch = std::thread([&] {
                while (!done.load(std::memory_order_acquire)) {
                    if (*reinterpret_cast<const volatile bool*>(should_quit)) cancel = 1;
                    std::this_thread::sleep_for(std::chrono::milliseconds(1));
                }
            });
        int rc = pico_hip_render(ctx_, &camera, &p, reinterpret_cast<float*>(params.m_Pixels), params.m_SampleCount,
                                 reinterpret_cast<float*>(params.m_variance), &cancel, nullptr);
        done.store(true, std::memory_order_release);
        if (watch.joinable()) watch.join();
        // a cancelled render still runs the denoiser and the tone mapping on
        // the partial frame, as the reference does: its tiles return, then
        // Scene.cpp:205-234 runs unconditionally
        const int render_rc = rc;
        if (rc && rc != PICO_ERR_CANCELLED) return rc;
        const size_t n = (size_t)params.m_Width * params.m_Height;
        if (params.m_denoise) {
            Vec3* nrm = new Vec3[n];
            Vec3* pos = new Vec3[n];
            Vec3* alb = new Vec3[n];
            rc = pico_hip_gbuffer(ctx_, &camera, params.m_Width, params.m_Height, reinterpret_cast<float*>(nrm),
                                  reinterpret_cast<float*>(pos), reinterpret_cast<float*>(alb));
            if (!rc)
                rc = pico_hip_atrous(ctx_, reinterpret_cast<const float*>(params.m_Pixels), reinterpret_cast<float*>(nrm),
                                     reinterpret_cast<float*>(pos), reinterpret_cast<float*>(alb), params.m_Width,
                                     params.m_Height, 64, 64, 5, reinterpret_cast<float*>(params.m_Pixels));
            delete[] nrm;
            delete[] pos;
            delete[] alb;
            if (rc) return rc;
        }
        if (params.m_tonemap)
            rc = pico_hip_tonemap(ctx_, reinterpret_cast<float*>(params.m_Pixels), params.m_Width, params.m_Height);
        return rc ? rc : render_rc;
    }

    // Scene::render_scene_to_file (Source/Core/Scene.cpp:236-246): render,
    // reverse the pixels end to end in the caller's buffer (as the reference
    // does), write them as a Radiance .hdr file.
    template <class Vec3>
    int render_scene_to_file(const pico_camera& camera, RenderParams<Vec3>& params, const char* path,
                             uint32_t frame_seed = 0) {
        const bool quit = false;
        int rc = render_scene_to_memory(camera, params, &quit, frame_seed);
        if (rc) return rc;
        std::reverse(params.m_Pixels, params.m_Pixels + (size_t)params.m_Width * params.m_Height);
        return pico_write_hdr(path, params.m_Width, params.m_Height, reinterpret_cast<const float*>(params.m_Pixels), 0);
    }

private:
    pico_hip_ctx* ctx_ = nullptr;
    int status_ = PICO_OK;
};

// Util::atrous_denoise: returns a new[] buffer of width * height pixels (the
// caller delete[]s it), or nullptr on failure.  tile is the Tiler's tile size.
template <class Vec3>
Vec3* atrous_denoise(pico_hip_ctx* ctx, const Vec3* pixels, const Vec3* normals, const Vec3* position, const Vec3* diffuse,
                     uint32_t width, uint32_t height, uint32_t tile = 64) {
    static_assert(sizeof(Vec3) == 12, "packed 3 x f32 per pixel");
    Vec3* out = new Vec3[(size_t)width * height];
    const int rc = pico_hip_atrous(ctx, reinterpret_cast<const float*>(pixels), reinterpret_cast<const float*>(normals),
                                   reinterpret_cast<const float*>(position), reinterpret_cast<const float*>(diffuse), width,
                                   height, tile, tile, 5, reinterpret_cast<float*>(out));
    if (rc) {
        delete[] out;
        return nullptr;
    }
    return out;
}

// The learned filter in the same shape (no counterpart in the reference),
// with Util::atrous_denoise as its fallback: when cnn is NULL, refuses the
// image or gives any value that is not finite, the returned buffer is the
// a-trous filter's (tiles of `tile`), and *used_fallback (may be NULL) says
// so.  Returns a new[] buffer, or nullptr on failure.
template <class Vec3>
Vec3* cnn_denoise(pico_hip_ctx* ctx, pico_cnn* cnn, const Vec3* pixels, const Vec3* normals, const Vec3* position,
                  const Vec3* diffuse, uint32_t width, uint32_t height, uint32_t tile = 64,
                  int32_t* used_fallback = nullptr) {
    static_assert(sizeof(Vec3) == 12, "packed 3 x f32 per pixel");
    Vec3* out = new Vec3[(size_t)width * height];
    const int rc = pico_hip_denoise_gbuffer(ctx, cnn, reinterpret_cast<const float*>(pixels),
                                            reinterpret_cast<const float*>(normals),
                                            reinterpret_cast<const float*>(position),
                                            reinterpret_cast<const float*>(diffuse), width, height, tile, tile,
                                            reinterpret_cast<float*>(out), used_fallback);
    if (rc) {
        delete[] out;
        return nullptr;
    }
    return out;
}

}  // namespace pico_front
