// Library context: device, stream, uploaded scene, work buffers, error text
// and per-kernel event timing.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/pico_hip.h"
#include "bvh_build.h"
#include "trace.h"

struct KernelTimes {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    uint64_t launches = 0;
    double total_ms = 0.0;
};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Default path budget of a render pass (paths in flight; 184 B of path state
// each, 23.6 GB at 128M).  Samples per chunk = budget / live pixels, so a 1080p
// 64-spp frame runs as 2 chunks (6 + 58 samples) instead of 9 at 16M: the
// last bounces, whose few long rays set each launch's duration, are paid
// twice instead of nine times (1680 against 1150 Mrays/s).
constexpr uint32_t kDefaultQueuePaths = 128u << 20;

// Shape of a render chunk (live pixels, first sample, samples, depth): the
// key of the hand-over point learned from earlier renders.
struct FinishKey {
    uint64_t live;
    uint32_t off, spp, depth;
    bool operator<(const FinishKey& o) const {
        if (live != o.live) return live < o.live;
        if (off != o.off) return off < o.off;
        if (spp != o.spp) return spp < o.spp;
        return depth < o.depth;
    }
};

struct TuneEntry {
    pico_tuning tuning;
    float chosen_ms, default_ms;
};

struct pico_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;   // every launch goes here
    hipStream_t side = nullptr;     // second stream of batched render + denoise (created on first use)
    std::string err;
    int num_cus = 256;
    // scene
    bool has_scene = false;
    uint64_t scene_serial = 0;      // bumped by every upload (keys the tuning cache)
    pico::HostScene host;
    pico::DevScene dev{};
    std::vector<void*> scene_allocs;
    uint32_t stack_need = 0;
    // kept from the upload for the device TLAS rebuild: the node array has
    // room for the largest TLAS a rebuild can emit after tlas_base
    const uint32_t* d_instance_mesh = nullptr;
    const float* d_mesh_box = nullptr;
    const uint32_t* d_mesh_vrange = nullptr;  // per mesh (first, end) vertex in dev.pos; empty for built-in shapes
    uint32_t max_mesh_vertices = 0;           // the largest of those ranges
    uint32_t tlas_base = 0;
    uint32_t blas_bound = 0;
    // reported by pico_hip_scene_info (either builder)
    uint64_t n_nodes4 = 0, n_tris = 0;
    uint32_t tlas_depth = 0, blas_depth = 0;
    // named scratch buffers (grown on demand, freed at destroy)
    std::map<std::string, DevBuf> bufs;
    // launch shape of the wavefront and the autotuner's cache
    pico_tuning tuning{kDefaultQueuePaths, 0, 0, 0};  // queue depth, tile, flags
    std::map<uint64_t, TuneEntry> tune_cache;
    // per chunk shape, the first bounce an earlier render left to its
    // hand-over kernel at the latest: a render launches the bounces before it
    // only
    std::map<FinishKey, uint32_t> finish_hints;
    // timing
    bool timing = false;
    std::map<std::string, KernelTimes> timers;
    // pinned host staging of the progressive render's per-pass snapshots
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // this device's copy of the static scratch of the allocation-free entry
    // point (resolved at create with the device current)
    void* static_scratch = nullptr;
    // learned-denoiser instances created on this context (destroyed with it)
    std::set<pico_cnn*> cnns;
};

namespace pico {

inline int fail(pico_hip_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define PICO_HIP(ctx, expr)                                                                        \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::pico::fail((ctx), PICO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Points the context's launches at another stream for a scope (the second
// stream of the batched render + denoise); restores the main one after.
struct StreamSwap {
    pico_hip_ctx* c;
    hipStream_t saved;
    StreamSwap(pico_hip_ctx* ctx, hipStream_t s) : c(ctx), saved(ctx->stream) { c->stream = s; }
    ~StreamSwap() { c->stream = saved; }
};

// Learned-denoiser queries used by the batched render (defined with the network).
bool cnn_belongs_to(const pico_cnn* n, const pico_hip_ctx* c);
int cnn_check_size(pico_cnn* n, uint32_t w, uint32_t h);

// Resolves the context's device address of the allocation-free entry point's
// static scratch (post.hip).
int resolve_static_scratch(pico_hip_ctx* c);

// Reference-mode a-trous (post.hip); every kernel is skipped unless
// gate == NULL or *gate != 0 on the device.
int atrous_gated(pico_hip_ctx* c, const float* rgb, const float* nrm, const float* pos, const float* alb, uint32_t w,
                 uint32_t h, uint32_t tw, uint32_t th, uint32_t levels, float* scratch_buf, float* out,
                 const uint32_t* gate);

// Returns a device buffer of at least `bytes` under `name` (contents undefined).
void* scratch(pico_hip_ctx* c, const std::string& name, size_t bytes);

// Device build of the whole scene, BLAS and TLAS (blas_build.hip); the
// previous scene is already freed.
int upload_scene_device(pico_hip_ctx* c, const pico_scene_desc& d, pico_build_stats* stats);

// Sky image and sun of a scene being uploaded (api.hip).
int upload_sky_sun(pico_hip_ctx* c, HostScene& hs, DevScene& s);

// Device TLAS rebuild over the context's n instances (tlas.hip).
int tlas_rebuild(pico_hip_ctx* c, uint32_t n, const float* d_transforms, pico_rebuild_stats* stats);

// Bracket a launch with events when timing is enabled; checks the launch.
struct LaunchScope {
    pico_hip_ctx* c;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    LaunchScope(pico_hip_ctx* ctx, const char* n) : c(ctx), name(n) {
        if (c->timing) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, c->stream);
        }
    }
    int done() {
        hipError_t e = hipGetLastError();
        if (c->timing) {
            hipEventRecord(b, c->stream);
            c->timers[name].pending.push_back({a, b});
        }
        if (e != hipSuccess) return fail(c, PICO_ERR_HIP, std::string("launch of ") + name + ": " + hipGetErrorString(e));
        return PICO_OK;
    }
};

#define PICO_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                               \
    do {                                                                                        \
        ::pico::LaunchScope ls_(ctx, name);                                                     \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem, (ctx)->stream, __VA_ARGS__); \
        int rc_ = ls_.done();                                                                   \
        if (rc_) return rc_;                                                                    \
    } while (0)

}  // namespace pico
