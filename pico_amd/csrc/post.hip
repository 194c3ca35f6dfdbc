// Post-render filters: the a-trous denoiser, Reinhard tone mapping and the
// FP64 relMSE reduction.
//
// A-trous (Source/Util/Denoisers.cpp:18-126): one workgroup per tile and level.
// The workgroup first reduces the tile-local constants over exactly the set of
// pixels the tile's 25-tap kernels read (white point = max luminance of the
// level input, range = max - min of |position|), then filters its pixels.
// The tapped set is reduced tap by tap, not as a dense dilated rectangle: for
// tiles narrower than the dilation the two differ.  Per-pixel |position| is
// computed once per call.  HBM-bound: 4 x 12 B read + 12 B written per pixel
// and level, the tile reduce re-reads through L2.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/pico_hip.h"
#include "ctx.h"

using namespace pico;

// Scratch of pico_hip_relmse_noalloc_d: the per-block parts of the relMSE
// reduction (at most kRelmseBlocks doubles, 8 KiB).  One copy per device,
// held until the process ends; every call on a device shares it, so calls
// that overlap on one device race.  Host code must use the per-device
// address from hipGetSymbolAddress (pico_hip_ctx::static_scratch), never the
// name: in host code the name is only a stand-in for the symbol.
static __device__ __align__(256) double scratch[1024];

namespace {

constexpr int kTileThreads = 256;

__constant__ float kAtrousW[25] = {0.00292f, 0.01306f, 0.02154f, 0.01306f, 0.00292f, 0.01306f, 0.05855f,
                                   0.09653f, 0.05855f, 0.01306f, 0.02154f, 0.09653f, 0.15915f, 0.09653f,
                                   0.02154f, 0.01306f, 0.05855f, 0.09563f, 0.05855f, 0.01306f, 0.00292f,
                                   0.01306f, 0.02154f, 0.01306f, 0.00292f};

__global__ void k_plen(const float* __restrict__ pos, float* __restrict__ plen, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        plen[i] = length(ld3(pos + 3 * (size_t)i));
}

__device__ __forceinline__ uint32_t tap_index(int x, int y, int u, int scale, int w, int h) {
    const int ox = u % 5 - 2, oy = u / 5 - 2;
    const int ix = min(max(x + ox * scale, 0), w - 1);
    const int iy = min(max(y + oy * scale, 0), h - 1);
    return (uint32_t)iy * (uint32_t)w + (uint32_t)ix;
}

__global__ __launch_bounds__(kTileThreads) void k_atrous_level(const float* __restrict__ src, const float* __restrict__ nrm,
                                                               const float* __restrict__ pos, const float* __restrict__ alb,
                                                               const float* __restrict__ plen, float* __restrict__ dst,
                                                               int w, int h, int tw, int th, int tiles_y, int scale) {
    __shared__ float red[3][kTileThreads / 64];
    const int tx0 = (blockIdx.x / tiles_y) * tw, ty0 = (blockIdx.x % tiles_y) * th;
    const int cw = min(tw, w - tx0), ch = min(th, h - ty0);
    const int npx = cw * ch;
    float white = 0.0f, dmin = INFINITY, dmax = -INFINITY;
    for (int q = threadIdx.x; q < npx * 25; q += kTileThreads) {
        const int pidx = q / 25, u = q - pidx * 25;
        const int x = tx0 + pidx % cw, y = ty0 + pidx / cw;
        const uint32_t fi = tap_index(x, y, u, scale, w, h);
        const f3 c = ld3(src + 3 * (size_t)fi);
        const float lum = c.x * 0.2126f + c.y * 0.587f + c.z * 0.114f;
        white = fmaxf(white, lum);
        const float l = plen[fi];
        dmin = fminf(dmin, l);
        dmax = fmaxf(dmax, l);
    }
    // wave reduce, then across the workgroup's waves
    for (int o = 32; o > 0; o >>= 1) {
        white = fmaxf(white, __shfl_xor(white, o));
        dmin = fminf(dmin, __shfl_xor(dmin, o));
        dmax = fmaxf(dmax, __shfl_xor(dmax, o));
    }
    const int wave = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) { red[0][wave] = white; red[1][wave] = dmin; red[2][wave] = dmax; }
    __syncthreads();
    white = red[0][0]; dmin = red[1][0]; dmax = red[2][0];
    for (int k = 1; k < kTileThreads / 64; ++k) {
        white = fmaxf(white, red[0][k]);
        dmin = fminf(dmin, red[1][k]);
        dmax = fmaxf(dmax, red[2][k]);
    }
    const float range = dmax - dmin;
    const float r2 = range * range, w2 = white * white;
    for (int pidx = threadIdx.x; pidx < npx; pidx += kTileThreads) {
        const int x = tx0 + pidx % cw, y = ty0 + pidx / cw;
        const size_t ci = (size_t)y * w + x;
        const f3 cn = ld3(nrm + 3 * ci), cp = ld3(pos + 3 * ci), ca = ld3(alb + 3 * ci);
        f3 res = mk(0.0f, 0.0f, 0.0f);
        float wsum = 0.0f;
#pragma unroll
        for (int u = 0; u < 25; ++u) {
            const size_t fi = tap_index(x, y, u, scale, w, h);
            f3 t = cn - ld3(nrm + 3 * fi);
            float dist = dot(t, t);
            const float nw = smin(expf(-dist / 0.25f), 1.0f);
            t = cp - ld3(pos + 3 * fi);
            dist = dot(t, t);
            const float pw = smin(expf(-dist / r2), 1.0f);
            t = ca - ld3(alb + 3 * fi);
            dist = dot(t, t);
            const float aw = smin(expf(-dist / w2), 1.0f);
            const float tw_ = kAtrousW[u] * nw * pw * aw;
            res = res + ld3(src + 3 * fi) * tw_;
            wsum += tw_;
        }
        const f3 o = res / wsum;
        dst[3 * ci] = o.x; dst[3 * ci + 1] = o.y; dst[3 * ci + 2] = o.z;
    }
}

// Labelled mode with no counterpart in the reference: the edge-avoiding
// a-trous transform of Dammertz et al. (HPG 2010) with FIXED edge-stopping
// widths in place of the reference's tile-local constants, and the B3-spline
// taps (1, 4, 6, 4, 1) / 16 per axis.  Tap weight:
//   h(u) h(v) exp(-( |dn|^2 / sn^2 + (dd / d_c)^2 / sd^2 + |da|^2 / sa^2
//                   + |dc|^2 / (sc 2^-l)^2 ))
// with n the normal, d = |view-space position| (relative difference), a the
// albedo and c the level's input colour; a width of 0 drops its term.
// Neighbours are clamped to the image like the reference's taps.
__constant__ float kB3[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};

__global__ void k_guide(const float* __restrict__ nrm, const float* __restrict__ pos, float4* __restrict__ g, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const f3 nn = ld3(nrm + 3 * (size_t)i), p = ld3(pos + 3 * (size_t)i);
        g[i] = make_float4(nn.x, nn.y, nn.z, sqrtf(p.x * p.x + p.y * p.y + p.z * p.z));
    }
}

__global__ __launch_bounds__(256) void k_atrous_fixed(const float* __restrict__ src, const float4* __restrict__ guide,
                                                      const float* __restrict__ alb, float* __restrict__ dst, int w, int h,
                                                      int scale, float isn2, float isd2, float isa2, float isc2) {
    const uint32_t n = (uint32_t)w * (uint32_t)h;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int x = (int)(i % (uint32_t)w), y = (int)(i / (uint32_t)w);
        const f3 cc = ld3(src + 3 * (size_t)i), ca = ld3(alb + 3 * (size_t)i);
        const float4 cg = guide[i];
        const float idd2 = isd2 / (cg.w * cg.w + 1e-30f);
        f3 res = mk(0.0f, 0.0f, 0.0f);
        float wsum = 0.0f;
        for (int v = 0; v < 5; ++v) {
            const int yy = min(max(y + (v - 2) * scale, 0), h - 1);
            for (int u = 0; u < 5; ++u) {
                const int xx = min(max(x + (u - 2) * scale, 0), w - 1);
                const size_t j = (size_t)yy * w + xx;
                const float4 g = guide[j];
                const f3 a = ld3(alb + 3 * j), c = ld3(src + 3 * j);
                const f3 dn = mk(cg.x - g.x, cg.y - g.y, cg.z - g.z), da = ca - a, dc = cc - c;
                const float dd = cg.w - g.w;
                const float e = dot(dn, dn) * isn2 + dd * dd * idd2 + dot(da, da) * isa2 + dot(dc, dc) * isc2;
                const float wt = kB3[u] * kB3[v] * expf(-e);
                res = res + c * wt;
                wsum += wt;
            }
        }
        const f3 o = res / wsum;
        dst[3 * (size_t)i] = o.x; dst[3 * (size_t)i + 1] = o.y; dst[3 * (size_t)i + 2] = o.z;
    }
}

__global__ void k_lum_max(const float* __restrict__ rgb, uint32_t n, uint32_t* white_bits) {
    float m = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float l = rgb[3 * i] * 0.2126f + rgb[3 * i + 1] * 0.587f + rgb[3 * i + 2] * 0.114f;
        m = smax(m, l);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax(white_bits, __float_as_uint(m));
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Util::reinhard_tone_mapping, Source/Util/ToneMappers.cpp:12-44.
__global__ void k_reinhard(float* rgb, uint32_t n3, const uint32_t* white_bits) {
    const float white = __uint_as_float(*white_bits);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += gridDim.x * blockDim.x) {
        const float p = rgb[i];
        const float num = p * (1.0f + (p / (white * white)));
        rgb[i] = num / (1.0f + p);
    }
}

// relMSE in two deterministic stages: block b sums its grid-stride share of
// the terms into part[b] (every block writes its slot, so nothing needs
// clearing first); one block then adds the parts in block order and writes
// the mean.  The same grid for the same size gives the same bits.
constexpr uint32_t kRelmseBlocks = 1024;

__global__ void k_zero_f64(double* p) { *p = 0.0; }

__global__ __launch_bounds__(256) void k_relmse_part(const float* __restrict__ img, const double* __restrict__ ref,
                                                     double inv_count, size_t n3, double* __restrict__ part) {
    __shared__ double red[4];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n3; i += (size_t)gridDim.x * 256) {
        const double r = ref[i] * inv_count;
        const double d = (double)img[i] - r;
        acc += d * d / (r * r + 1e-2);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_relmse_final(const double* __restrict__ part, uint32_t n_part, double inv_n3,
                                                      double* __restrict__ out) {
    __shared__ double red[4];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n_part; i += 256) acc += part[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) *out = ((red[0] + red[1]) + (red[2] + red[3])) * inv_n3;
}

static_assert(sizeof(::scratch) / sizeof(double) >= kRelmseBlocks, "static scratch holds every relMSE part");

inline uint32_t relmse_grid(size_t n3) { return (uint32_t)std::min<size_t>((n3 + 255) / 256, kRelmseBlocks); }

// Both relMSE entry points: the parts in `part` (relmse_grid(n3) doubles), the
// mean in *d_out (device).
int relmse_launch(pico_hip_ctx* c, const float* img, const double* ref, double count, size_t n3, double* part,
                  double* d_out) {
    const uint32_t g = relmse_grid(n3);
    PICO_LAUNCH(c, "relmse", k_relmse_part, g, 256, 0, img, ref, 1.0 / count, n3, part);
    PICO_LAUNCH(c, "relmse_final", k_relmse_final, 1, 256, 0, part, g, 1.0 / (double)n3, d_out);
    return PICO_OK;
}

// 4 independent 16-byte loads in flight per lane, non-temporal stores.
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy(v4f* __restrict__ dst, const v4f* __restrict__ src, size_t n) {
    const size_t stride = (size_t)gridDim.x * 256;
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const v4f a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        __builtin_nontemporal_store(a, dst + i);
        __builtin_nontemporal_store(b, dst + i + stride);
        __builtin_nontemporal_store(c, dst + i + 2 * stride);
        __builtin_nontemporal_store(d, dst + i + 3 * stride);
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

namespace pico {

// Called by pico_hip_create with the context's device current: resolves
// this device's copy of the static scratch once.
int resolve_static_scratch(pico_hip_ctx* c) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || !ptr)
        return fail(c, PICO_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    c->static_scratch = ptr;
    return PICO_OK;
}

}  // namespace pico

extern "C" {

int pico_hip_copy_d(pico_hip_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return PICO_ERR_ARG;
    if (!dst || !src || (bytes & 15)) return fail(c, PICO_ERR_ARG, "copy needs non-NULL buffers and a multiple of 16 bytes");
    const size_t n = bytes / 16;
    const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, (size_t)c->num_cus * 32);
    if (n) PICO_LAUNCH(c, "copy", k_copy, grid, 256, 0, (v4f*)dst, (const v4f*)src, n);
    return PICO_OK;
}

int pico_hip_atrous_d(pico_hip_ctx* c, const float* rgb, const float* nrm, const float* pos, const float* alb, uint32_t w,
                      uint32_t h, uint32_t tw, uint32_t th, uint32_t levels, float* scratch_buf, float* out) {
    if (!c) return PICO_ERR_ARG;
    if (!rgb || !nrm || !pos || !alb || !scratch_buf || !out) return fail(c, PICO_ERR_ARG, "NULL buffer");
    if (w == 0 || h == 0 || tw == 0 || th == 0) return fail(c, PICO_ERR_ARG, "sizes must be positive");
    if ((uint64_t)w * h > (1ull << 28)) return fail(c, PICO_ERR_LIMIT, "image larger than 2^28 pixels");
    if (levels > 16) return fail(c, PICO_ERR_LIMIT, "at most 16 levels");
    const size_t n = (size_t)w * h;
    if (levels == 0) {
        PICO_HIP(c, hipMemcpyAsync(out, rgb, 12 * n, hipMemcpyDeviceToDevice, c->stream));
        return PICO_OK;
    }
    float* plen = (float*)pico::scratch(c, "atrous_plen", 4 * n + 64);
    if (!plen) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    PICO_LAUNCH(c, "atrous_plen", k_plen, std::min<size_t>((n + 255) / 256, 8192), 256, 0, pos, plen, (uint32_t)n);
    const int tiles_x = (int)((w + tw - 1) / tw), tiles_y = (int)((h + th - 1) / th);
    const float* src = rgb;
    for (uint32_t l = 0; l < levels; ++l) {
        float* dst = ((levels - 1 - l) % 2 == 0) ? out : scratch_buf;
        PICO_LAUNCH(c, "atrous", k_atrous_level, tiles_x * tiles_y, kTileThreads, 0, src, nrm, pos, alb, plen, dst, (int)w,
                    (int)h, (int)tw, (int)th, tiles_y, 1 << l);
        src = dst;
    }
    return PICO_OK;
}

int pico_hip_atrous_fixed_d(pico_hip_ctx* c, const float* rgb, const float* nrm, const float* pos, const float* alb,
                            uint32_t w, uint32_t h, uint32_t levels, float sigma_normal, float sigma_depth,
                            float sigma_albedo, float sigma_color, float* scratch_buf, float* out) {
    if (!c) return PICO_ERR_ARG;
    if (!rgb || !nrm || !pos || !alb || !scratch_buf || !out) return fail(c, PICO_ERR_ARG, "NULL buffer");
    if (w == 0 || h == 0) return fail(c, PICO_ERR_ARG, "sizes must be positive");
    if ((uint64_t)w * h > (1ull << 28)) return fail(c, PICO_ERR_LIMIT, "image larger than 2^28 pixels");
    if (levels > 16) return fail(c, PICO_ERR_LIMIT, "at most 16 levels");
    if (!(sigma_normal >= 0.0f && sigma_depth >= 0.0f && sigma_albedo >= 0.0f && sigma_color >= 0.0f))
        return fail(c, PICO_ERR_ARG, "edge-stopping widths must be >= 0");
    const size_t n = (size_t)w * h;
    if (levels == 0) {
        PICO_HIP(c, hipMemcpyAsync(out, rgb, 12 * n, hipMemcpyDeviceToDevice, c->stream));
        return PICO_OK;
    }
    float4* guide = (float4*)pico::scratch(c, "atrous_guide", 16 * n + 64);
    if (!guide) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, (size_t)c->num_cus * 16);
    PICO_LAUNCH(c, "atrous_guide", k_guide, grid, 256, 0, nrm, pos, guide, (uint32_t)n);
    auto inv2 = [](float s) { return s > 0.0f ? 1.0f / (s * s) : 0.0f; };
    const float isn2 = inv2(sigma_normal), isd2 = inv2(sigma_depth), isa2 = inv2(sigma_albedo);
    const float* src = rgb;
    for (uint32_t l = 0; l < levels; ++l) {
        float* dst = ((levels - 1 - l) % 2 == 0) ? out : scratch_buf;
        const float sc = sigma_color / (float)(1u << l);
        PICO_LAUNCH(c, "atrous_fixed", k_atrous_fixed, grid, 256, 0, src, guide, alb, dst, (int)w, (int)h, 1 << l, isn2,
                    isd2, isa2, inv2(sc));
        src = dst;
    }
    return PICO_OK;
}

int pico_hip_atrous(pico_hip_ctx* c, const float* rgb, const float* nrm, const float* pos, const float* alb, uint32_t w,
                    uint32_t h, uint32_t tw, uint32_t th, uint32_t levels, float* out) {
    if (!c) return PICO_ERR_ARG;
    if (!rgb || !nrm || !pos || !alb || !out) return fail(c, PICO_ERR_ARG, "NULL buffer");
    if ((uint64_t)w * h > (1ull << 28)) return fail(c, PICO_ERR_LIMIT, "image larger than 2^28 pixels");
    const size_t n = (size_t)w * h;
    float* b = (float*)pico::scratch(c, "atrous_io", 72 * n + 64);
    if (!b) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    float* d_rgb = b;
    float* d_n = b + 3 * n;
    float* d_p = b + 6 * n;
    float* d_a = b + 9 * n;
    float* d_s = b + 12 * n;
    float* d_o = b + 15 * n;
    PICO_HIP(c, hipMemcpyAsync(d_rgb, rgb, 12 * n, hipMemcpyHostToDevice, c->stream));
    PICO_HIP(c, hipMemcpyAsync(d_n, nrm, 12 * n, hipMemcpyHostToDevice, c->stream));
    PICO_HIP(c, hipMemcpyAsync(d_p, pos, 12 * n, hipMemcpyHostToDevice, c->stream));
    PICO_HIP(c, hipMemcpyAsync(d_a, alb, 12 * n, hipMemcpyHostToDevice, c->stream));
    int rc = pico_hip_atrous_d(c, d_rgb, d_n, d_p, d_a, w, h, tw, th, levels, d_s, d_o);
    if (rc) return rc;
    PICO_HIP(c, hipMemcpyAsync(out, d_o, 12 * n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    return PICO_OK;
}

int pico_hip_tonemap(pico_hip_ctx* c, float* rgb, uint32_t w, uint32_t h) {
    if (!c) return PICO_ERR_ARG;
    if (!rgb) return fail(c, PICO_ERR_ARG, "NULL buffer");
    if ((uint64_t)w * h > (1ull << 28)) return fail(c, PICO_ERR_LIMIT, "image larger than 2^28 pixels");
    const size_t n = (size_t)w * h;
    if (n == 0) return PICO_OK;
    float* d = (float*)pico::scratch(c, "tonemap", 12 * n + 64);
    if (!d) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    uint32_t* wb = (uint32_t*)(d + 3 * n);
    PICO_HIP(c, hipMemcpyAsync(d, rgb, 12 * n, hipMemcpyHostToDevice, c->stream));
    PICO_LAUNCH(c, "fill", k_fill_u32, 1, 64, 0, wb, 0u, 1u);
    PICO_LAUNCH(c, "lum_max", k_lum_max, std::min<size_t>((n + 255) / 256, 4096), 256, 0, d, (uint32_t)n, wb);
    PICO_LAUNCH(c, "reinhard", k_reinhard, std::min<size_t>((3 * n + 255) / 256, 8192), 256, 0, d, (uint32_t)(3 * n), wb);
    PICO_HIP(c, hipMemcpyAsync(rgb, d, 12 * n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    return PICO_OK;
}

int pico_hip_relmse_d(pico_hip_ctx* c, const float* img, const double* ref, double count, uint32_t n_pixels, double* out) {
    if (!c) return PICO_ERR_ARG;
    if (!img || !ref || !out) return fail(c, PICO_ERR_ARG, "NULL buffer");
    if (!(count > 0)) return fail(c, PICO_ERR_ARG, "reference count must be positive");
    if (n_pixels > (1u << 28)) return fail(c, PICO_ERR_LIMIT, "image larger than 2^28 pixels");
    if (n_pixels == 0) {
        *out = 0.0;
        return PICO_OK;
    }
    const size_t n3 = 3 * (size_t)n_pixels;
    double* part = (double*)pico::scratch(c, "relmse", 8 * (kRelmseBlocks + 1));
    if (!part) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    int rc = relmse_launch(c, img, ref, count, n3, part + 1, part);
    if (rc) return rc;
    double h = 0.0;
    PICO_HIP(c, hipMemcpyAsync(&h, part, 8, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    *out = h;
    return PICO_OK;
}

int pico_hip_relmse_noalloc_d(pico_hip_ctx* c, const float* img, const double* ref, double count, uint32_t n_pixels,
                              double* d_out) {
    if (!c) return PICO_ERR_ARG;
    if (!img || !ref || !d_out) return fail(c, PICO_ERR_ARG, "NULL buffer");
    if (!(count > 0)) return fail(c, PICO_ERR_ARG, "reference count must be positive");
    if (n_pixels > (1u << 28)) return fail(c, PICO_ERR_LIMIT, "image larger than 2^28 pixels");
    if (!c->static_scratch) return fail(c, PICO_ERR_STATE, "static scratch of this device not resolved");
    if (n_pixels == 0) {
        PICO_LAUNCH(c, "relmse_zero", k_zero_f64, 1, 1, 0, d_out);
        return PICO_OK;
    }
    // the parts live in this device's copy of `scratch` (relmse_grid <= kRelmseBlocks)
    return relmse_launch(c, img, ref, count, 3 * (size_t)n_pixels, (double*)c->static_scratch, d_out);
}

}  // extern "C"
