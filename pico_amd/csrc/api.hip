// C ABI: context, device memory, timing, scene upload, ray queries and the
// G-buffer pass.  The wavefront renderer lives in render.hip, the post
// filters in post.hip.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>

#include "../../include/pico_hip.h"
#include "ctx.h"

using namespace pico;

void* pico::scratch(pico_hip_ctx* c, const std::string& name, size_t bytes) {
    DevBuf& b = c->bufs[name];
    if (b.bytes >= bytes && b.p) return b.p;
    if (b.p) {
        hipStreamSynchronize(c->stream);
        hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    if (hipMalloc(&b.p, bytes ? bytes : 16) != hipSuccess) {
        b.p = nullptr;
        return nullptr;
    }
    b.bytes = bytes;
    return b.p;
}

namespace {

constexpr int kRayBlock = 128;

// One-ray-per-item kernels over persistent grids (as many workgroups as stay
// resident), each lane walking the items in grid stride.  Traversal stacks:
// kLds entries per lane in LDS, the rest (up to kSpill) in device memory.
constexpr int kLds = 32;
constexpr uint32_t kSpill = 32;
// Uploads refuse scenes whose traversal needs more than 64 entries, so these
// stacks never overflow; the kernels' overflow flags are a second guard
// (read back by the entry points that synchronise).
static_assert(kLds + kSpill >= 64, "ray-batch stacks hold the 64 entries an upload admits");

struct StackArgs { uint32_t* spill; uint32_t cap; };

__device__ __forceinline__ TraceStack<kLds> lane_stack(uint32_t* stk, StackArgs sa) {
    const uint32_t g = blockIdx.x * kRayBlock + threadIdx.x;
    return TraceStack<kLds>{stk + threadIdx.x, kRayBlock, sa.spill + g, gridDim.x * kRayBlock, sa.cap};
}

__global__ __launch_bounds__(kRayBlock) void k_closest_hit(DevScene s, uint32_t n, const float* __restrict__ o,
                                                            const float* __restrict__ d, float* t, uint32_t* inst,
                                                            uint32_t* prim, float* uv, uint32_t* overflow, StackArgs sa) {
    __shared__ uint32_t stk[kLds * kRayBlock];
    TraceStack<kLds> st = lane_stack(stk, sa);
    for (uint32_t i = blockIdx.x * kRayBlock + threadIdx.x; i < n; i += gridDim.x * kRayBlock) {
        Hit h = trace_closest<kLds, false>(s, ld3(o + 3 * i), ld3(d + 3 * i), st, nullptr, overflow);
        t[i] = h.t;
        inst[i] = h.inst;
        prim[i] = h.prim;
        uv[2 * i] = h.u;
        uv[2 * i + 1] = h.v;
    }
}

__global__ __launch_bounds__(kRayBlock) void k_primary_ids(DevScene s, DCam cam, float2 xi, uint32_t* inst, uint32_t* prim,
                                                            uint32_t* overflow, StackArgs sa) {
    __shared__ uint32_t stk[kLds * kRayBlock];
    TraceStack<kLds> st = lane_stack(stk, sa);
    for (uint32_t i = blockIdx.x * kRayBlock + threadIdx.x; i < cam.w * cam.h; i += gridDim.x * kRayBlock) {
        const f3 d = camera_dir(cam, xi, i % cam.w, i / cam.w);
        Hit h = trace_closest<kLds, false>(s, cam.pos, d, st, nullptr, overflow);
        inst[i] = h.inst;
        prim[i] = h.prim;
    }
}

// Scene::generate_denoiser_inputs, Source/Core/Scene.cpp:265-286.
__global__ __launch_bounds__(kRayBlock) void k_gbuffer(DevScene s, DCam cam, float4 v0, float4 v1, float4 v2, float4 v3,
                                                        float* normal, float* position, float* albedo, uint32_t* overflow,
                                                        StackArgs sa) {
    __shared__ uint32_t stk[kLds * kRayBlock];
    TraceStack<kLds> st = lane_stack(stk, sa);
    for (uint32_t i = blockIdx.x * kRayBlock + threadIdx.x; i < cam.w * cam.h; i += gridDim.x * kRayBlock) {
        const f3 d = camera_dir(cam, make_float2(0.0f, 0.0f), i % cam.w, i / cam.w);
        Hit h = trace_closest<kLds, false>(s, cam.pos, d, st, nullptr, overflow);
        f3 n, p, a;
        if (h.inst != 0xffffffffu) {
            const Frag f = hit_frag(s, h, cam.pos, d);
            n = f.n;
            // view matrix (glm::lookAt) times (p, 1): (c0*x + c1*y) + (c2*z + c3)
            p = mk((v0.x * f.p.x + v1.x * f.p.y) + (v2.x * f.p.z + v3.x),
                   (v0.y * f.p.x + v1.y * f.p.y) + (v2.y * f.p.z + v3.y),
                   (v0.z * f.p.x + v1.z * f.p.y) + (v2.z * f.p.z + v3.z));
            a = ld3(s.mats[s.inst[h.inst].material].diffuse);
        } else {
            n = -d;
            p = d * cam.far_plane;
            a = sky_radiance(s, d);
        }
        normal[3 * i] = n.x; normal[3 * i + 1] = n.y; normal[3 * i + 2] = n.z;
        position[3 * i] = p.x; position[3 * i + 1] = p.y; position[3 * i + 2] = p.z;
        albedo[3 * i] = a.x; albedo[3 * i + 1] = a.y; albedo[3 * i + 2] = a.z;
    }
}

// Copies v into a new device allocation of at least `capacity` elements.
template <class T>
int upload_vec(pico_hip_ctx* c, const std::vector<T>& v, const T** out, size_t capacity = 0) {
    void* p = nullptr;
    const size_t bytes = v.size() * sizeof(T);
    const size_t alloc = std::max(bytes, capacity * sizeof(T));
    PICO_HIP(c, hipMalloc(&p, alloc ? alloc : 16));
    c->scene_allocs.push_back(p);
    if (bytes) PICO_HIP(c, hipMemcpyAsync(p, v.data(), bytes, hipMemcpyHostToDevice, c->stream));
    *out = reinterpret_cast<const T*>(p);
    return PICO_OK;
}

}  // namespace

namespace pico {
// The sky image's texels (when there is one) into a scene allocation, and the
// sky and sun parameters into the device scene; the host copy keeps no
// pointer to the description's texels.
int upload_sky_sun(pico_hip_ctx* c, HostScene& hs, DevScene& s) {
    std::memcpy(s.sky, hs.sky, sizeof(s.sky));
    s.use_sun = hs.use_sun;
    std::memcpy(s.sun_to, hs.sun_to, sizeof(s.sun_to));
    std::memcpy(s.sun_col, hs.sun_col, sizeof(s.sun_col));
    s.sky_tex = nullptr;
    const pico_sky_image& k = hs.sky_image;
    if (k.width && k.texels) {
        const size_t bytes = sky_bytes(k);
        void* p = nullptr;
        PICO_HIP(c, hipMalloc(&p, bytes));
        c->scene_allocs.push_back(p);
        PICO_HIP(c, hipMemcpyAsync(p, k.texels, bytes, hipMemcpyHostToDevice, c->stream));
        PICO_HIP(c, hipStreamSynchronize(c->stream));
        s.sky_tex = static_cast<const uint8_t*>(p);
        s.sky_w = k.width;
        s.sky_h = k.height;
        s.sky_faces = k.faces;
        s.sky_fmt = k.format;
    }
    hs.sky_image.texels = nullptr;
    return PICO_OK;
}
}  // namespace pico

namespace {
void free_scene(pico_hip_ctx* c) {
    if (!c->scene_allocs.empty()) hipStreamSynchronize(c->stream);
    for (void* p : c->scene_allocs) hipFree(p);
    c->scene_allocs.clear();
    c->has_scene = false;
}

}  // namespace

namespace pico {
DCam make_dcam(const pico_camera& c, uint32_t w, uint32_t h) {
    DCam d;
    d.pos = {c.position[0], c.position[1], c.position[2]};
    d.dir = {c.direction[0], c.direction[1], c.direction[2]};
    d.up = {c.up[0], c.up[1], c.up[2]};
    // getRight: cross(normalize(D), up), evaluated on the host in float
    const float l = 1.0f / std::sqrt(d.dir.x * d.dir.x + d.dir.y * d.dir.y + d.dir.z * d.dir.z);
    const f3 nd = {d.dir.x * l, d.dir.y * l, d.dir.z * l};
    d.right = {nd.y * d.up.z - d.up.y * nd.z, nd.z * d.up.x - d.up.z * nd.x, nd.x * d.up.y - d.up.x * nd.y};
    d.aspect = c.aspect;
    d.far_plane = c.far_plane;
    d.w = w;
    d.h = h;
    return d;
}

int check_camera(pico_hip_ctx* c, const pico_camera* cam, uint32_t w, uint32_t h) {
    if (!cam) return fail(c, PICO_ERR_ARG, "camera is NULL");
    if (w == 0 || h == 0) return fail(c, PICO_ERR_ARG, "image size must be positive");
    if ((uint64_t)w * h > (1ull << 26)) return fail(c, PICO_ERR_LIMIT, "image larger than 2^26 pixels");
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(cam->position[k]) || !std::isfinite(cam->direction[k]) || !std::isfinite(cam->up[k]))
            return fail(c, PICO_ERR_ARG, "camera vectors must be finite");
    return PICO_OK;
}

// Launch a one-ray-per-item kernel on a persistent grid with its stack spill area.
#define PICO_RAY_LAUNCH(ctx, NAME, KERNEL, N, ...)                                                          \
    do {                                                                                                    \
        int occ_ = 0;                                                                                       \
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_, KERNEL, kRayBlock, 0);                           \
        const uint32_t want_ = ((N) + kRayBlock - 1) / kRayBlock;                                            \
        const uint32_t grid_ = std::max<uint32_t>(1, std::min<uint32_t>(want_, (uint32_t)(ctx)->num_cus *   \
                                                                       (uint32_t)std::max(1, std::min(occ_, 16)))); \
        StackArgs sa_{(uint32_t*)scratch(ctx, "ray_spill", (size_t)grid_ * kRayBlock * kSpill * 4 + 64),     \
                      (ctx)->stack_need > (uint32_t)kLds ? std::min((ctx)->stack_need - kLds, kSpill) : 0u};   \
        if (!sa_.spill) return fail(ctx, PICO_ERR_NOMEM, "device allocation failed");                        \
        PICO_LAUNCH(ctx, NAME, KERNEL, grid_, kRayBlock, 0, __VA_ARGS__, sa_);                               \
    } while (0)

}  // namespace pico

extern "C" {

int pico_hip_create(int device, pico_hip_ctx** out) {
    if (!out) return PICO_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return PICO_ERR_HIP;
    if (device < 0 || device >= n) return PICO_ERR_ARG;
    auto* c = new pico_hip_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return PICO_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    if (resolve_static_scratch(c) != PICO_OK) {
        hipStreamDestroy(c->stream);
        delete c;
        return PICO_ERR_HIP;
    }
    *out = c;
    return PICO_OK;
}

void pico_hip_destroy(pico_hip_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    const std::set<pico_cnn*> cnns = c->cnns;
    for (pico_cnn* n : cnns) pico_hip_cnn_destroy(n);
    free_scene(c);
    for (auto& kv : c->bufs)
        if (kv.second.p) hipFree(kv.second.p);
    if (c->pinned) hipHostFree(c->pinned);
    for (auto& kv : c->timers)
        for (auto& e : kv.second.pending) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    if (c->side) {
        hipStreamSynchronize(c->side);
        hipStreamDestroy(c->side);
    }
    hipStreamDestroy(c->stream);
    delete c;
}

const char* pico_hip_last_error(const pico_hip_ctx* c) { return c ? c->err.c_str() : "context is NULL"; }

int pico_hip_synchronize(pico_hip_ctx* c) {
    if (!c) return PICO_ERR_ARG;
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    return PICO_OK;
}

int pico_hip_device_name(pico_hip_ctx* c, char* buf, size_t n) {
    if (!c || !buf || n == 0) return PICO_ERR_ARG;
    hipDeviceProp_t prop;
    PICO_HIP(c, hipGetDeviceProperties(&prop, c->device));
    std::snprintf(buf, n, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return PICO_OK;
}

int pico_hip_malloc(pico_hip_ctx* c, size_t bytes, void** dptr) {
    if (!c || !dptr) return PICO_ERR_ARG;
    PICO_HIP(c, hipMalloc(dptr, bytes ? bytes : 16));
    return PICO_OK;
}
int pico_hip_free(pico_hip_ctx* c, void* dptr) {
    if (!c) return PICO_ERR_ARG;
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    PICO_HIP(c, hipFree(dptr));
    return PICO_OK;
}
int pico_hip_memcpy_h2d(pico_hip_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || (!dst && bytes) || (!src && bytes)) return PICO_ERR_ARG;
    PICO_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    return PICO_OK;
}
int pico_hip_memcpy_d2h(pico_hip_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || (!dst && bytes) || (!src && bytes)) return PICO_ERR_ARG;
    PICO_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    return PICO_OK;
}
int pico_hip_memset_d(pico_hip_ctx* c, void* dst, int value, size_t bytes) {
    if (!c || (!dst && bytes)) return PICO_ERR_ARG;
    PICO_HIP(c, hipMemsetAsync(dst, value, bytes, c->stream));
    return PICO_OK;
}

int pico_hip_enable_kernel_timing(pico_hip_ctx* c, int enable) {
    if (!c) return PICO_ERR_ARG;
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& kv : c->timers)
        for (auto& e : kv.second.pending) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    c->timers.clear();
    c->timing = enable != 0;
    return PICO_OK;
}

int pico_hip_kernel_time(pico_hip_ctx* c, const char* kernel, uint64_t* launches, double* mean_ms) {
    if (!c || !kernel) return PICO_ERR_ARG;
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    auto it = c->timers.find(kernel);
    if (it == c->timers.end()) {
        if (launches) *launches = 0;
        if (mean_ms) *mean_ms = 0.0;
        return PICO_OK;
    }
    KernelTimes& kt = it->second;
    for (auto& e : kt.pending) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e.first, e.second);
        kt.total_ms += ms;
        kt.launches++;
        hipEventDestroy(e.first);
        hipEventDestroy(e.second);
    }
    kt.pending.clear();
    if (launches) *launches = kt.launches;
    if (mean_ms) *mean_ms = kt.launches ? kt.total_ms / (double)kt.launches : 0.0;
    return PICO_OK;
}

int pico_hip_upload_scene(pico_hip_ctx* c, const pico_scene_desc* d) { return pico_hip_upload_scene_ex(c, d, 0, nullptr); }

int pico_hip_upload_scene_ex(pico_hip_ctx* c, const pico_scene_desc* d, uint32_t flags, pico_build_stats* stats) {
    if (!c || !d) return fail(c, PICO_ERR_ARG, "scene description is NULL");
    if (flags & ~PICO_BUILD_DEVICE) return fail(c, PICO_ERR_ARG, "unknown build flags");
    PICO_HIP(c, hipSetDevice(c->device));
    free_scene(c);
    if (flags & PICO_BUILD_DEVICE) return upload_scene_device(c, *d, stats);
    const auto t0 = std::chrono::steady_clock::now();
    HostScene hs;
    std::string e = build_scene(*d, 16, hs);
    const double build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (!e.empty()) return fail(c, PICO_ERR_ARG, e);
    const uint32_t need = hs.stack_bound;
    if (need > 64) return fail(c, PICO_ERR_LIMIT, "BVH too deep for the 64-entry traversal stack");
    DevScene s{};
    int rc;
    // a device TLAS rebuild emits at most max(n - 1, 1) nodes after tlas_base
    const size_t node_cap = hs.tlas_base + std::max<size_t>(hs.n_instances > 1 ? hs.n_instances - 1 : 1, 1);
    if ((rc = upload_vec(c, hs.nodes4, &s.nodes, node_cap))) return rc;
    const float* mesh_box = nullptr;
    const uint32_t* inst_mesh = nullptr;
    const uint32_t* mesh_vrange = nullptr;
    if ((rc = upload_vec(c, hs.mesh_box, &mesh_box))) return rc;
    if ((rc = upload_vec(c, hs.mesh_vrange, &mesh_vrange))) return rc;
    if ((rc = upload_vec(c, hs.instance_mesh, &inst_mesh))) return rc;
    if ((rc = upload_vec(c, hs.tris, &s.tris))) return rc;
    if ((rc = upload_vec(c, hs.instances, &s.inst))) return rc;
    if ((rc = upload_vec(c, hs.xforms, &s.xf))) return rc;
    if ((rc = upload_vec(c, hs.materials, &s.mats))) return rc;
    if ((rc = upload_vec(c, hs.lights, &s.lights))) return rc;
    if ((rc = upload_vec(c, hs.alias_weight, &s.alias_w))) return rc;
    if ((rc = upload_vec(c, hs.alias_index, &s.alias_i))) return rc;
    if ((rc = upload_vec(c, hs.positions, &s.pos))) return rc;
    if ((rc = upload_vec(c, hs.normals, &s.nrm))) return rc;
    if ((rc = upload_vec(c, hs.indices, &s.idx))) return rc;
    {
        const float* ts = nullptr;
        if ((rc = upload_vec(c, hs.tshade, &ts))) return rc;
        s.tshade = reinterpret_cast<const float4*>(ts);
    }
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    s.tlas_root = hs.tlas_root;
    s.n_lights = (uint32_t)hs.lights.size();
    s.microfacet = hs.microfacet;
    if ((rc = upload_sky_sun(c, hs, s))) return rc;
    c->dev = s;
    c->stack_need = need;
    c->d_mesh_box = mesh_box;
    c->d_mesh_vrange = mesh_vrange;
    c->max_mesh_vertices = max_range(hs.mesh_vrange);
    c->d_instance_mesh = inst_mesh;
    c->tlas_base = hs.tlas_base;
    c->blas_bound = hs.blas_bound;
    c->n_nodes4 = hs.nodes4.size();
    c->n_tris = hs.tris.size();
    c->tlas_depth = hs.tlas_depth;
    c->blas_depth = hs.blas_depth;
    if (stats) {
        *stats = pico_build_stats{};
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->bvh_ms = (float)build_ms;
        stats->binary_nodes = hs.blas_nodes2;
        stats->blas_nodes4 = hs.tlas_base;
        stats->blas_bound = hs.blas_bound;
        stats->stack_bound = hs.stack_bound;
        stats->levels = hs.blas_depth;
    }
    c->host = std::move(hs);
    c->has_scene = true;
    ++c->scene_serial;
    return PICO_OK;
}

int pico_hip_scene_info(pico_hip_ctx* c, uint64_t* bn, uint64_t* bt, uint64_t* bi, uint32_t* td, uint32_t* bd) {
    if (!c) return PICO_ERR_ARG;
    if (!c->has_scene) return fail(c, PICO_ERR_STATE, "no scene uploaded");
    if (bn) *bn = c->n_nodes4 * sizeof(DNode4);
    if (bt) *bt = c->n_tris * sizeof(DTri);
    if (bi) *bi = (uint64_t)c->host.n_instances * sizeof(DInstance);
    if (td) *td = c->tlas_depth;
    if (bd) *bd = c->blas_depth;
    return PICO_OK;
}

int pico_hip_closest_hit(pico_hip_ctx* c, uint32_t n, const float* o, const float* d, float* t, uint32_t* inst,
                         uint32_t* prim, float* uv) {
    if (!c) return PICO_ERR_ARG;
    if (!c->has_scene) return fail(c, PICO_ERR_STATE, "no scene uploaded");
    if (n == 0) return PICO_OK;
    if (!o || !d || !t || !inst || !prim || !uv) return fail(c, PICO_ERR_ARG, "NULL buffer");
    char* b = (char*)scratch(c, "rays", (size_t)n * 44 + 64);
    if (!b) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    float* dO = (float*)b;
    float* dD = dO + 3 * (size_t)n;
    float* dT = dD + 3 * (size_t)n;
    uint32_t* dI = (uint32_t*)(dT + n);
    uint32_t* dP = dI + n;
    float* dUV = (float*)(dP + n);
    uint32_t* ovf = (uint32_t*)(dUV + 2 * (size_t)n);
    PICO_HIP(c, hipMemcpyAsync(dO, o, 12 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    PICO_HIP(c, hipMemcpyAsync(dD, d, 12 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    PICO_HIP(c, hipMemsetAsync(ovf, 0, 4, c->stream));
    PICO_RAY_LAUNCH(c, "closest_hit", k_closest_hit, n, c->dev, n, dO, dD, dT, dI, dP, dUV, ovf);
    PICO_HIP(c, hipMemcpyAsync(t, dT, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(inst, dI, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(prim, dP, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(uv, dUV, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    uint32_t o_flag = 0;
    PICO_HIP(c, hipMemcpyAsync(&o_flag, ovf, 4, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    if (o_flag) return fail(c, PICO_ERR_LIMIT, "traversal stack overflow");
    return PICO_OK;
}

int pico_hip_primary_ids(pico_hip_ctx* c, const pico_camera* cam, uint32_t w, uint32_t h, float xx, float xy,
                         uint32_t* inst, uint32_t* prim) {
    if (!c) return PICO_ERR_ARG;
    if (!c->has_scene) return fail(c, PICO_ERR_STATE, "no scene uploaded");
    int rc = check_camera(c, cam, w, h);
    if (rc) return rc;
    if (!inst || !prim) return fail(c, PICO_ERR_ARG, "NULL buffer");
    const size_t n = (size_t)w * h;
    uint32_t* b = (uint32_t*)scratch(c, "ids", n * 8 + 64);
    if (!b) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    uint32_t* ovf = b + 2 * n;
    PICO_HIP(c, hipMemsetAsync(ovf, 0, 4, c->stream));
    const DCam dc = make_dcam(*cam, w, h);
    PICO_RAY_LAUNCH(c, "primary_ids", k_primary_ids, (uint32_t)n, c->dev, dc, make_float2(xx, xy), b, b + n, ovf);
    PICO_HIP(c, hipMemcpyAsync(inst, b, 4 * n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(prim, b + n, 4 * n, hipMemcpyDeviceToHost, c->stream));
    uint32_t o_flag = 0;
    PICO_HIP(c, hipMemcpyAsync(&o_flag, ovf, 4, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    if (o_flag) return fail(c, PICO_ERR_LIMIT, "traversal stack overflow");
    return PICO_OK;
}

int pico_hip_gbuffer_d(pico_hip_ctx* c, const pico_camera* cam, uint32_t w, uint32_t h, float* dn, float* dp, float* da) {
    if (!c) return PICO_ERR_ARG;
    if (!c->has_scene) return fail(c, PICO_ERR_STATE, "no scene uploaded");
    int rc = check_camera(c, cam, w, h);
    if (rc) return rc;
    if (!dn || !dp || !da) return fail(c, PICO_ERR_ARG, "NULL buffer");
    uint32_t* ovf = (uint32_t*)scratch(c, "gb_ovf", 64);
    if (!ovf) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    // glm::lookAt(eye, eye + dir, up), right-handed, in float on the host
    const float* E = cam->position;
    float ctr[3], f[3], up[3] = {cam->up[0], cam->up[1], cam->up[2]};
    for (int k = 0; k < 3; ++k) ctr[k] = E[k] + cam->direction[k];
    for (int k = 0; k < 3; ++k) f[k] = ctr[k] - E[k];
    float l = 1.0f / std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (int k = 0; k < 3; ++k) f[k] *= l;
    float s[3] = {f[1] * up[2] - up[1] * f[2], f[2] * up[0] - up[2] * f[0], f[0] * up[1] - up[0] * f[1]};
    l = 1.0f / std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int k = 0; k < 3; ++k) s[k] *= l;
    const float u[3] = {s[1] * f[2] - f[1] * s[2], s[2] * f[0] - f[2] * s[0], s[0] * f[1] - f[0] * s[1]};
    const float4 v0 = make_float4(s[0], u[0], -f[0], 0.0f), v1 = make_float4(s[1], u[1], -f[1], 0.0f),
                 v2 = make_float4(s[2], u[2], -f[2], 0.0f);
    const float4 v3 = make_float4(-(s[0] * E[0] + s[1] * E[1] + s[2] * E[2]), -(u[0] * E[0] + u[1] * E[1] + u[2] * E[2]),
                                  f[0] * E[0] + f[1] * E[1] + f[2] * E[2], 1.0f);
    PICO_HIP(c, hipMemsetAsync(ovf, 0, 4, c->stream));
    const DCam dc = make_dcam(*cam, w, h);
    PICO_RAY_LAUNCH(c, "gbuffer", k_gbuffer, w * h, c->dev, dc, v0, v1, v2, v3, dn, dp, da, ovf);
    return PICO_OK;
}

int pico_hip_gbuffer(pico_hip_ctx* c, const pico_camera* cam, uint32_t w, uint32_t h, float* n, float* p, float* a) {
    if (!c) return PICO_ERR_ARG;
    if (!n || !p || !a) return fail(c, PICO_ERR_ARG, "NULL buffer");
    const size_t px = (size_t)w * h;
    float* b = (float*)scratch(c, "gbuf", px * 36 + 64);
    if (!b) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    int rc = pico_hip_gbuffer_d(c, cam, w, h, b, b + 3 * px, b + 6 * px);
    if (rc) return rc;
    PICO_HIP(c, hipMemcpyAsync(n, b, 12 * px, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(p, b + 3 * px, 12 * px, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(a, b + 6 * px, 12 * px, hipMemcpyDeviceToHost, c->stream));
    uint32_t o_flag = 0;
    PICO_HIP(c, hipMemcpyAsync(&o_flag, scratch(c, "gb_ovf", 64), 4, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    if (o_flag) return fail(c, PICO_ERR_LIMIT, "traversal stack overflow");
    return PICO_OK;
}

}  // extern "C"
