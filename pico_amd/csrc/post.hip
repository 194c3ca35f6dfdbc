// Post-render filters: the a-trous denoiser, Reinhard tone mapping and the
// FP64 relMSE reduction.
//
// A-trous (Source/Util/Denoisers.cpp:18-126): one workgroup per tile and level.
// The workgroup first reduces the tile-local constants (white point = max
// luminance of the level input, range = max - min of |position|) over the
// pixels the tile's 25-tap kernels read, then filters its pixels.  The tapped
// set is a product of a column set and a row set: along x it is
// {clamp(x0 + k + o s)} for k < tile width, o in -2..2, i.e. one run of
// width + 4s columns when the dilation s is at most the tile width, else five
// runs of the tile width; the same along y.  The reduce walks those runs once
// (max and min do not mind the repeats clamping makes) instead of 25 taps per
// pixel.  Luminance and |position| come from per-pixel arrays: |position| is
// computed once per call, the luminance of each level's output is written by
// the level that produces it.  The filter takes one exponential per tap:
// exp(-a) exp(-b) exp(-c) = exp2(-(a + b + c) log2 e), with the same NaN and
// zero cases as the reference's three (a width of 0 gives 0 / 0 = NaN on the
// tap at distance 0 and 0 elsewhere).
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

__device__ __forceinline__ float lum_of(f3 c) { return c.x * 0.2126f + c.y * 0.587f + c.z * 0.114f; }

// |position| and the luminance of the first level's input, per pixel.
__global__ void k_atrous_prep(const float* __restrict__ rgb, const float* __restrict__ pos, float* __restrict__ plen,
                              float* __restrict__ lum, uint32_t n, const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0u) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        plen[i] = length(ld3(pos + 3 * (size_t)i));
        lum[i] = lum_of(ld3(rgb + 3 * (size_t)i));
    }
}

// Runs of the tapped set along one axis for a tile [t0, t0 + c) at dilation s.
struct TapRuns {
    int n, len, start, step;   // n runs of len, run r starts at start + r * step
};
__device__ __forceinline__ TapRuns tap_runs(int t0, int c, int s) {
    if (s <= c) return {1, c + 4 * s, t0 - 2 * s, 0};
    return {5, c, t0 - 2 * s, s};
}

// Block index spread so that each of the 8 XCDs (each with its own L2) gets a
// contiguous run of blocks: hardware hands consecutive workgroups to the XCDs
// in turn, and neighbouring blocks share the halo of their taps.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t g) {
    const uint32_t xcd = b & 7u, k = b >> 3, per = g >> 3, rem = g & 7u;
    return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

// Tile-local constants of one level, one workgroup per tile: the exp2 factors
// log2(e) / range^2 and log2(e) / white^2 of the position and albedo terms.
__global__ __launch_bounds__(kTileThreads) void k_atrous_consts(const float* __restrict__ lum,
                                                                const float* __restrict__ plen, float2* __restrict__ consts,
                                                                int w, int h, int tw, int th, int tiles_y, int scale,
                                                                const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0u) return;
    __shared__ float red[3][kTileThreads / 64];
    const int t = (int)xcd_block(blockIdx.x, gridDim.x);
    const int tx0 = (t / tiles_y) * tw, ty0 = (t % tiles_y) * th;
    const int cw = min(tw, w - tx0), ch = min(th, h - ty0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float white = 0.0f, dmin = INFINITY, dmax = -INFINITY;
    const TapRuns rx = tap_runs(tx0, cw, scale), ry = tap_runs(ty0, ch, scale);
    for (int r = 0; r < ry.n; ++r)
        for (int j = wave; j < ry.len; j += kTileThreads / 64) {
            const uint32_t row = (uint32_t)min(max(ry.start + r * ry.step + j, 0), h - 1) * (uint32_t)w;
            for (int q = 0; q < rx.n; ++q)
                for (int k = lane; k < rx.len; k += 64) {
                    const uint32_t fi = row + (uint32_t)min(max(rx.start + q * rx.step + k, 0), w - 1);
                    white = fmaxf(white, lum[fi]);
                    const float l = plen[fi];
                    dmin = fminf(dmin, l);
                    dmax = fmaxf(dmax, l);
                }
        }
    for (int o = 32; o > 0; o >>= 1) {
        white = fmaxf(white, __shfl_xor(white, o));
        dmin = fminf(dmin, __shfl_xor(dmin, o));
        dmax = fmaxf(dmax, __shfl_xor(dmax, o));
    }
    if (lane == 0) { red[0][wave] = white; red[1][wave] = dmin; red[2][wave] = dmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kTileThreads / 64; ++k) {
            white = fmaxf(white, red[0][k]);
            dmin = fminf(dmin, red[1][k]);
            dmax = fmaxf(dmax, red[2][k]);
        }
        // exp(-d / r^2) exp(-d' / w^2) as exp2 factors; the reciprocals keep
        // 0 / 0 = NaN and d / 0 = inf for d > 0
        const float range = dmax - dmin;
        const float kLog2e = 1.44269504088896341f;
        consts[t] = make_float2(kLog2e / (range * range), kLog2e / (white * white));
    }
}

// One level of the filter.  A workgroup filters a block of BW x BH pixels
// (256 threads, one pixel each); wider blocks for wider dilations keep the
// union of the 25 taps small (it is (BW + 4s) x (BH + 4s) pixels, or 5 BW
// columns once s > BW).  Each pixel looks up its tile's constants.
template <int BW, int BH>
__global__ __launch_bounds__(256) void k_atrous_filter(
    const float* __restrict__ src, const float* __restrict__ nrm, const float* __restrict__ pos,
    const float* __restrict__ alb, const float2* __restrict__ consts, float* __restrict__ dst,
    float* __restrict__ lum_out, int w, int h, int tw, int th, int tiles_y, int blocks_x, int scale,
    const uint32_t* __restrict__ gate) {
    static_assert(BW * BH == 256, "one pixel per thread");
    if (gate && *gate == 0u) return;
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int x = (int)(b % (uint32_t)blocks_x) * BW + (int)threadIdx.x % BW;
    const int y = (int)(b / (uint32_t)blocks_x) * BH + (int)threadIdx.x / BW;
    if (x >= w || y >= h) return;
    const float2 kc = consts[(x / tw) * tiles_y + y / th];
    const float kn = 4.0f * 1.44269504088896341f;
    const uint32_t ci = (uint32_t)y * (uint32_t)w + (uint32_t)x;
    const f3 cn = ld3(nrm + 3 * ci), cp = ld3(pos + 3 * ci), ca = ld3(alb + 3 * ci);
    f3 res = mk(0.0f, 0.0f, 0.0f);
    float wsum = 0.0f;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const uint32_t row = (uint32_t)min(max(y + (v - 2) * scale, 0), h - 1) * (uint32_t)w;
#pragma unroll
        for (int u = 0; u < 5; ++u) {
            const uint32_t f3i = 3u * (row + (uint32_t)min(max(x + (u - 2) * scale, 0), w - 1));
            f3 t = cn - ld3(nrm + f3i);
            const float dn = fmaf(t.z, t.z, fmaf(t.y, t.y, t.x * t.x));
            t = cp - ld3(pos + f3i);
            const float dp = fmaf(t.z, t.z, fmaf(t.y, t.y, t.x * t.x));
            t = ca - ld3(alb + f3i);
            const float da = fmaf(t.z, t.z, fmaf(t.y, t.y, t.x * t.x));
            const float e = fmaf(dn, kn, fmaf(dp, kc.x, da * kc.y));
            const float tw_ = kAtrousW[5 * v + u] * __builtin_amdgcn_exp2f(-e);
            const f3 c = ld3(src + f3i);
            res = mk(fmaf(c.x, tw_, res.x), fmaf(c.y, tw_, res.y), fmaf(c.z, tw_, res.z));
            wsum += tw_;
        }
    }
    const f3 o = res / wsum;
    dst[3 * ci] = o.x; dst[3 * ci + 1] = o.y; dst[3 * ci + 2] = o.z;
    if (lum_out) lum_out[ci] = lum_of(o);
}

// One level at dilation S <= 16 with the taps staged in LDS.  A workgroup
// filters 64 columns x 4 rows of one row phase: rows ph + (4g + j) S, j < 4,
// so the 5 tap rows of its 4 rows are 8 rows of the same phase, and its tap
// columns are the 64 + 4S columns around its own.  Those 8 x (64 + 4S)
// pixels (normal, position, albedo, colour: 48 B each) are staged once; each
// tap is then four LDS reads at offsets fixed at compile time.
constexpr int kLdsCols = 128;  // 64 + 4 * 16

template <int S>
__global__ __launch_bounds__(256) void k_atrous_lds(
    const float* __restrict__ src, const float* __restrict__ nrm, const float* __restrict__ pos,
    const float* __restrict__ alb, const float2* __restrict__ consts, float* __restrict__ dst,
    float* __restrict__ lum_out, int w, int h, int tw, int th, int tiles_y, int blocks_x, int groups,
    const uint32_t* __restrict__ gate) {
    static_assert(64 + 4 * S <= kLdsCols, "staged columns");
    constexpr int FW = 64 + 4 * S;
    constexpr int kArr = 8 * kLdsCols * 3;  // floats per staged field
    if (gate && *gate == 0u) return;
    __shared__ float st[4 * kArr];
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int bx = (int)(b % (uint32_t)blocks_x);
    const int rest = (int)(b / (uint32_t)blocks_x);
    const int ph = rest % S, g = rest / S;
    const int x0 = bx * 64;
    // stage rows ph + (4g + r - 2) S, r < 8, columns x0 - 2S + k, k < FW (clamped)
    for (int i = threadIdx.x; i < 8 * FW; i += 256) {
        const int r = i / FW, k = i - r * FW;
        const int yy = min(max(ph + (4 * g + r - 2) * S, 0), h - 1);
        const int xx = min(max(x0 - 2 * S + k, 0), w - 1);
        const uint32_t gi = 3u * ((uint32_t)yy * (uint32_t)w + (uint32_t)xx);
        const int li = 3 * (r * kLdsCols + k);
        const f3 a = ld3(nrm + gi), p = ld3(pos + gi), d = ld3(alb + gi), c = ld3(src + gi);
        st[li] = a.x; st[li + 1] = a.y; st[li + 2] = a.z;
        st[kArr + li] = p.x; st[kArr + li + 1] = p.y; st[kArr + li + 2] = p.z;
        st[2 * kArr + li] = d.x; st[2 * kArr + li + 1] = d.y; st[2 * kArr + li + 2] = d.z;
        st[3 * kArr + li] = c.x; st[3 * kArr + li + 1] = c.y; st[3 * kArr + li + 2] = c.z;
    }
    __syncthreads();
    const int lx = threadIdx.x & 63, j = threadIdx.x >> 6;
    const int x = x0 + lx, y = ph + (4 * g + j) * S;
    if (x >= w || y >= h || g >= groups) return;
    const float2 kc = consts[(x / tw) * tiles_y + y / th];
    const float kn = 4.0f * 1.44269504088896341f;
    // this pixel's entry: row j + 2, column lx + 2S
    const float* base = st + 3 * (j * kLdsCols + lx);
    const int cidx = 3 * (2 * kLdsCols + 2 * S);
    const f3 cn = ld3(base + cidx), cp = ld3(base + kArr + cidx), ca = ld3(base + 2 * kArr + cidx);
    f3 res = mk(0.0f, 0.0f, 0.0f);
    float wsum = 0.0f;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
#pragma unroll
        for (int u = 0; u < 5; ++u) {
            const int o = 3 * (v * kLdsCols + u * S);
            f3 t = cn - ld3(base + o);
            const float dn = fmaf(t.z, t.z, fmaf(t.y, t.y, t.x * t.x));
            t = cp - ld3(base + kArr + o);
            const float dp = fmaf(t.z, t.z, fmaf(t.y, t.y, t.x * t.x));
            t = ca - ld3(base + 2 * kArr + o);
            const float da = fmaf(t.z, t.z, fmaf(t.y, t.y, t.x * t.x));
            const float e = fmaf(dn, kn, fmaf(dp, kc.x, da * kc.y));
            const float tw_ = kAtrousW[5 * v + u] * __builtin_amdgcn_exp2f(-e);
            const f3 c = ld3(base + 3 * kArr + o);
            res = mk(fmaf(c.x, tw_, res.x), fmaf(c.y, tw_, res.y), fmaf(c.z, tw_, res.z));
            wsum += tw_;
        }
    }
    const f3 out = res / wsum;
    const uint32_t ci = (uint32_t)y * (uint32_t)w + (uint32_t)x;
    dst[3 * ci] = out.x; dst[3 * ci + 1] = out.y; dst[3 * ci + 2] = out.z;
    if (lum_out) lum_out[ci] = lum_of(out);
}

__global__ void k_copy_gated(float* __restrict__ dst, const float* __restrict__ src, uint32_t n,
                             const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0u) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
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

// Largest luminance, one atomic per workgroup: atomics on one address
// serialise, so a grid-sized number of them (one per wave) cost more than
// reading the frame.
__global__ __launch_bounds__(256) void k_lum_max(const float* __restrict__ rgb, uint32_t n, uint32_t* white_bits) {
    __shared__ float wm[4];
    float m = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float l = rgb[3 * i] * 0.2126f + rgb[3 * i + 1] * 0.587f + rgb[3 * i + 2] * 0.114f;
        m = smax(m, l);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float b = fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
        if (b > 0.0f) atomicMax(white_bits, __float_as_uint(b));
    }
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

// Reference-mode a-trous; with a gate (device u32) every kernel does nothing
// unless *gate != 0, so the filter can follow a check on the device without
// a host synchronisation.
int atrous_gated(pico_hip_ctx* c, const float* rgb, const float* nrm, const float* pos, const float* alb, uint32_t w,
                 uint32_t h, uint32_t tw, uint32_t th, uint32_t levels, float* scratch_buf, float* out,
                 const uint32_t* gate) {
    if (!rgb || !nrm || !pos || !alb || !scratch_buf || !out) return fail(c, PICO_ERR_ARG, "NULL buffer");
    if (w == 0 || h == 0 || tw == 0 || th == 0) return fail(c, PICO_ERR_ARG, "sizes must be positive");
    if ((uint64_t)w * h > (1ull << 28)) return fail(c, PICO_ERR_LIMIT, "image larger than 2^28 pixels");
    if (levels > 16) return fail(c, PICO_ERR_LIMIT, "at most 16 levels");
    const size_t n = (size_t)w * h;
    const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, (size_t)c->num_cus * 16);
    if (levels == 0) {
        PICO_LAUNCH(c, "atrous_copy", k_copy_gated, std::min<size_t>((3 * n + 255) / 256, 8192), 256, 0, out, rgb,
                    (uint32_t)(3 * n), gate);
        return PICO_OK;
    }
    const int tiles_x = (int)((w + tw - 1) / tw), tiles_y = (int)((h + th - 1) / th);
    const size_t tiles = (size_t)tiles_x * tiles_y;
    // |position|, two luminance buffers, then the tile constants (16-B aligned);
    // at most 20 B per pixel + 80 B, whatever the tile size
    const size_t consts_off = (12 * n + 15) & ~(size_t)15;
    char* sb = (char*)pico::scratch(c, "atrous_plen", consts_off + 8 * tiles + 64);
    if (!sb) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    float* plen = (float*)sb;
    float* lum[2] = {plen + n, plen + 2 * n};
    float2* consts = (float2*)(sb + consts_off);
    PICO_LAUNCH(c, "atrous_prep", k_atrous_prep, grid, 256, 0, rgb, pos, plen, lum[0], (uint32_t)n, gate);
    static const char* const kLevelName[5] = {"atrous_l0", "atrous_l1", "atrous_l2", "atrous_l3", "atrous_l4"};
    const float* src = rgb;
    for (uint32_t l = 0; l < levels; ++l) {
        float* dst = ((levels - 1 - l) % 2 == 0) ? out : scratch_buf;
        float* lum_next = l + 1 < levels ? lum[(l + 1) & 1] : nullptr;
        const int s = 1 << l;
        PICO_LAUNCH(c, "atrous_consts", k_atrous_consts, (uint32_t)tiles, kTileThreads, 0, lum[l & 1], plen, consts,
                    (int)w, (int)h, (int)tw, (int)th, tiles_y, s, gate);
        const char* name = l < 5 ? kLevelName[l] : "atrous_l5+";
        if (s <= 16) {
            // taps staged in LDS: blocks of 64 columns x 4 rows of one row phase
            const int bx = (int)((w + 63) / 64), groups = (int)(((h + s - 1) / s + 3) / 4);
            const uint32_t nb = (uint32_t)(bx * s * groups);
#define PICO_ATROUS_LDS(S_)                                                                                        \
    PICO_LAUNCH(c, name, k_atrous_lds<S_>, nb, 256, 0, src, nrm, pos, alb, consts, dst, lum_next, (int)w, (int)h, \
                (int)tw, (int)th, tiles_y, bx, groups, gate)
            switch (s) {
                case 1: PICO_ATROUS_LDS(1); break;
                case 2: PICO_ATROUS_LDS(2); break;
                case 4: PICO_ATROUS_LDS(4); break;
                case 8: PICO_ATROUS_LDS(8); break;
                default: PICO_ATROUS_LDS(16); break;
            }
#undef PICO_ATROUS_LDS
        } else {
            // dilations of 32 and more (levels beyond the reference's 5): the
            // taps straight from memory, one row of 256 pixels per workgroup
            const int bx = (int)((w + 255) / 256), by = (int)h;
            PICO_LAUNCH(c, name, (k_atrous_filter<256, 1>), bx * by, 256, 0, src, nrm, pos, alb, consts, dst, lum_next,
                        (int)w, (int)h, (int)tw, (int)th, tiles_y, bx, s, gate);
        }
        src = dst;
    }
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
    return pico::atrous_gated(c, rgb, nrm, pos, alb, w, h, tw, th, levels, scratch_buf, out, nullptr);
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
    PICO_LAUNCH(c, "lum_max", k_lum_max, std::min<size_t>((n + 255) / 256, 1024), 256, 0, d, (uint32_t)n, wb);
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
