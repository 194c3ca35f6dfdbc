// Learned denoiser: kernel-predicting 3-level U-Net (32/64/128 channels, 3x3
// convolutions, 1x1 head predicting a 5x5 per-pixel kernel), trained on relMSE.
// No counterpart in the reference; it stands beside the a-trous filter.
//
// Device layout: activations are bf16 NHWC with the channel count padded to a
// multiple of 16 (the input's 10 channels live in 16).  A 3x3 convolution is
// an implicit GEMM (M = pixels, N = output channels, K = 9 x input channels)
// on v_mfma_f32_16x16x16_bf16 with FP32 accumulation: a wave owns 16
// consecutive pixels of one row and 32 output channels; each K step loads one
// 8-byte A fragment (4 channels of one pixel at the tap's offset, zero outside
// the image) and two 8-byte B fragments (weights stored [tap][out][in], K
// contiguous).  Backward-data is the same kernel with the taps mirrored and
// the weights stored [tap][in][out]; backward-weight stages 64-pixel tiles of
// the input and of the output gradient in LDS and reduces over pixels with
// MFMA (split over pixel chunks, FP32 atomics into the FP32 gradient).
// The softmax/kernel application, the relMSE loss (FP64 reduction) and Adam
// run in FP32/FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pico_hip.h"
#include "ctx.h"

using namespace pico;

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

struct LayerDef { int cin, cout, taps, kpad, npad; };
// e0a e0b e1a e1b ba bb d1a d1b d0a d0b head
constexpr LayerDef kL[11] = {{10, 32, 9, 16, 32},   {32, 32, 9, 32, 32},    {32, 64, 9, 32, 64},  {64, 64, 9, 64, 64},
                             {64, 128, 9, 64, 128}, {128, 128, 9, 128, 128}, {192, 64, 9, 192, 64}, {64, 64, 9, 64, 64},
                             {96, 32, 9, 96, 32},   {32, 32, 9, 32, 32},    {32, 25, 1, 32, 32}};

__host__ __device__ inline bf16 f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__device__ inline float bf2f(bf16 b) { return __uint_as_float((uint32_t)b << 16); }

// input channel k of the network from its planar value: log1p of the
// (non-negative) radiance and of the distance, the guides as given
__device__ inline float in_feature(float f, int k) {
    if (k < 3) f = log1pf(f > 0.0f ? f : 0.0f);
    if (k == 9) f = log1pf(f);
    return f;
}

struct Tensor { bf16* p; int stride, off; };

struct ConvArgs {
    const bf16* in; int in_stride, in_off;
    const bf16* w;          // [taps][npad][kpad]
    const float* bias;      // [npad] or null
    bf16* out; int out_stride, out_off;
    float* out_f32;         // [P][npad] (head logits) or null
    const bf16* mask; int mask_stride, mask_off;  // multiply by (mask > 0)
    int B, H, W, kpad, npad, taps, flip, relu;
    // optional: input channels [0, up_ch) read from `up`, a tensor of half the
    // resolution (stride up_stride), with the nearest-neighbour rule of the
    // upsampling (y / 2, x / 2): the upsample + concatenation fused into the
    // convolution's input staging
    const bf16* up; int up_stride, up_ch;
    // optional (the last 3x3 layer at inference): the 1x1 head, softmax and
    // 5x5 kernel apply on this layer's outputs, writing the filtered radiance
    // (rgb, [B][img_h][img_w][3]) from the planar network input (rad) instead
    // of the layer's activations
    const bf16* head_w; const float* head_b; const float* rad; float* rgb; int img_h, img_w;
    // optional: the 2x2 average pool of this layer's bf16 outputs, written to
    // `pool` (half resolution, stride pool_stride) from the output tile
    bf16* pool; int pool_stride;
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Implicit-GEMM convolution on MFMA (M = pixels, N = output channels,
// K = taps x input channels), FP32 accumulation.  KS is the K depth of one
// instruction: 32 (v_mfma_f32_16x16x32_bf16, 8 consecutive channels = 16
// bytes per lane and operand) where the input has a multiple of 32 channels,
// else 16 (v_mfma_f32_16x16x16_bf16, 8 bytes).  A workgroup of 4 waves owns a
// run of 4 * MT * 16 pixels of one image row and NT * 16 output channels;
// each wave computes MT x NT 16x16 tiles (8 accumulators, so every A
// fragment feeds NT instructions and every B fragment MT).  Per tap the
// workgroup stages that tap's weights for its channels in LDS (rows padded
// by 8 bf16 against bank conflicts); A fragments (channels of one pixel at
// the tap's offset, zero outside the image) come from memory.
// grid.x: row runs, grid.y: channel blocks of NT * 16.
template <int MT, int NT, int KS>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
    extern __shared__ bf16 wl[];  // [NT * 16][kpad + 8]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int kRun = 4 * MT * 16;
    constexpr int KL = KS / 4;  // channels per lane and fragment
    const int runs_x = (a.W + kRun - 1) / kRun;
    const int row = blockIdx.x / runs_x;  // b*H + y
    const int xw = (blockIdx.x % runs_x) * kRun + wave * MT * 16;
    const int b = row / a.H, y = row % a.H;
    const int n0 = blockIdx.y * NT * 16;
    const int mi = lane & 15, kq = lane >> 4;
    const int ldk = a.kpad + 8;
    f4v acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[m][t] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    const int kq4 = a.kpad / 4;
    for (int tap = 0; tap < a.taps; ++tap) {
        int dy = a.taps == 9 ? tap / 3 - 1 : 0, dx = a.taps == 9 ? tap % 3 - 1 : 0;
        if (a.flip) { dy = -dy; dx = -dx; }
        __syncthreads();
        {
            const bf16* wt = a.w + ((size_t)tap * a.npad + n0) * a.kpad;
            for (int c = threadIdx.x; c < NT * 16 * kq4; c += 256) {
                const int nn = c / kq4, k4 = c - nn * kq4;
                *reinterpret_cast<s4*>(wl + nn * ldk + 4 * k4) = *reinterpret_cast<const s4*>(wt + (size_t)nn * a.kpad + 4 * k4);
            }
        }
        __syncthreads();
        if (xw >= a.W) continue;
        const int yy = y + dy;
        const bool rok = yy >= 0 && yy < a.H;
        const bf16* ap[MT];
        bool ok[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int xx = xw + m * 16 + mi + dx;
            ok[m] = rok && xx >= 0 && xx < a.W;
            ap[m] = a.in + ((size_t)(b * a.H + (ok[m] ? yy : 0)) * a.W + (ok[m] ? xx : 0)) * a.in_stride + a.in_off + KL * kq;
        }
        const bf16* bp = wl + mi * ldk + KL * kq;
        for (int k0 = 0; k0 < a.kpad; k0 += KS) {
            if constexpr (KS == 32) {
                uint4 av[MT], bv[NT];
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    av[m] = make_uint4(0, 0, 0, 0);
                    if (ok[m]) av[m] = *reinterpret_cast<const uint4*>(ap[m] + k0);
                }
#pragma unroll
                for (int t = 0; t < NT; ++t) bv[t] = *reinterpret_cast<const uint4*>(bp + t * 16 * ldk + k0);
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av[m]),
                                                                            __builtin_bit_cast(bf16x8, bv[t]), acc[m][t], 0, 0, 0);
            } else {
                s4 av[MT], bv[NT];
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    av[m] = s4{0, 0, 0, 0};
                    if (ok[m]) av[m] = *reinterpret_cast<const s4*>(ap[m] + k0);
                }
#pragma unroll
                for (int t = 0; t < NT; ++t) bv[t] = *reinterpret_cast<const s4*>(bp + t * 16 * ldk + k0);
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int t = 0; t < NT; ++t) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[m], bv[t], acc[m][t], 0, 0, 0);
            }
        }
    }
    if (xw >= a.W) return;
    // D[i][j]: pixel i = 4*kq + r, channel j = mi
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int n = n0 + t * 16 + mi;
        const float bias = a.bias ? a.bias[n] : 0.0f;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int x = xw + m * 16 + 4 * kq + r;
                if (x >= a.W) continue;
                float v = acc[m][t][r] + bias;
                if (a.relu) v = v > 0.0f ? v : 0.0f;
                const size_t p = (size_t)(b * a.H + y) * a.W + x;
                if (a.mask && !(bf2f(a.mask[p * a.mask_stride + a.mask_off + n]) > 0.0f)) v = 0.0f;
                if (a.out_f32) a.out_f32[p * a.npad + n] = v;
                else a.out[p * a.out_stride + a.out_off + n] = f2bf(v);
            }
    }
}

// 3x3 convolution with four output rows per tile, one per wave, every
// operand from LDS.  Per K chunk of KS input channels the workgroup stages
// the six input rows its four output rows read (its MT * 16 pixels plus one
// on each side, zero outside the image) and the chunk's weights.  Input rows
// are read 1.5 times per layer (6 rows for 4) instead of 3 times with one
// output row per workgroup.  A workgroup takes the tiles of its XCD in turn
// (one tile unless PERSIST), and every global load of the next (tile, chunk)
// is issued into registers while the current one runs on MFMA and through
// its epilogue, so only a workgroup's first load is exposed.  A layer of one
// chunk with 32 outputs stages its weights once for the whole kernel.  Plain
// bf16 outputs go through LDS and leave as 16-byte stores of 8 channels.
// LDS: the weights, then one area for the band, the output tiles and (HEAD)
// the logits in turn.
template <int MT, int NT, int KS, bool HEAD, bool PERSIST>
__device__ __forceinline__ void conv3r_body(const ConvArgs& a, bf16* lds) {
    constexpr int kRun = MT * 16;       // pixels of one output row
    constexpr int kBand = kRun + 2;     // pixels of one staged row
    constexpr int kRows = 6;            // staged input rows: y0 - 1 .. y0 + 4
    constexpr int ldw = KS + 8;         // weight rows padded against bank conflicts
    constexpr int KL = KS / 4;          // channels per lane and fragment
    constexpr int kVec = KS / 8;        // 16-byte vectors per pixel (or weight row) and chunk
    constexpr int kBandVec = kRows * kBand * kVec;
    constexpr int kPer = (kBandVec + 255) / 256;
    constexpr int kCh = NT * 16, ldo = kCh + 8;
    // persistent (32 outputs, one chunk): the weights of all nine taps, staged
    // once; otherwise one row of three taps at a time, streamed through
    // registers (all nine per chunk measured slower there: their registers
    // cost a wave per SIMD)
    constexpr bool kAllW = PERSIST;
    static_assert(!PERSIST || NT == 2, "persistent layers have 32 outputs");
    constexpr int kWTaps = kAllW ? 9 : 3;
    bf16* wl = lds;                                // [kWTaps][NT * 16][ldw]
    bf16* area = lds + kWTaps * NT * 16 * ldw;     // band [6][kBand][KS] / output tiles / logits
    bf16* band = area;
    // 32-channel band: the four 16-byte chunks of pixel p are stored in the
    // order c ^ ((p >> 1) & 3).  An A fragment reads one chunk of 16
    // consecutive pixels (64 B apart); unswizzled, every 128-byte bank window
    // holds only two of the eight pixels a read cycle serves (4-way bank
    // conflicts), swizzled it holds eight distinct chunks.
    constexpr bool kSwz = KS == 32;
    auto swz = [](int px) { return (px >> 1) & 3; };
    auto band_slot = [&](int e) {  // bf16 offset of 16-byte vector e = (pixel, chunk) of the band
        if constexpr (kSwz) {
            const int px = e >> 2;
            return (px * 4 + ((e & 3) ^ swz(px))) * 8;
        } else {
            return e * 8;
        }
    };
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int runs_x = (a.W + kRun - 1) / kRun, groups = (a.H + 3) / 4;
    const int n_tiles = a.B * groups * runs_x;
    // consecutive workgroups go to the 8 XCDs in turn: each XCD gets a
    // contiguous range of tiles, which its workgroups take in turn, so the
    // tiles in flight on one XCD are neighbours sharing halo rows in its L2
    const uint32_t G = gridDim.x, xb = blockIdx.x & 7u, kb = blockIdx.x >> 3, per = G >> 3, remb = G & 7u;
    const uint32_t nwg = per + (xb < remb ? 1u : 0u);
    const uint32_t first = xb < remb ? xb * (per + 1) : remb * (per + 1) + (xb - remb) * per;
    const int hi = (int)(((uint64_t)n_tiles * (first + nwg)) / G);
    int tile = (int)(((uint64_t)n_tiles * first) / G) + (int)kb;
    if (tile >= hi) return;  // the whole workgroup, before any barrier
    const int n0 = blockIdx.y * NT * 16;
    const int mi = lane & 15, kq = lane >> 4;
    const int Hh = a.H >> 1, Wh = a.W >> 1;
    const int nchunk = a.kpad / KS;
    constexpr bool wresident = kAllW;  // the launch gives PERSIST only layers of one chunk
    auto coords = [&](int t, int& tb, int& ty0, int& tx0) {
        const int rest = t / runs_x;
        tx0 = (t - rest * runs_x) * kRun;
        tb = rest / groups;
        ty0 = (rest - tb * groups) * 4;
    };
    uint4 pre[kPer];
    auto load_band = [&](int tb, int ty0, int tx0, int k0) {
        const bool from_up = a.up && k0 < a.up_ch;
        // the thread's element indices are recomputed per call rather than
        // kept in registers across the tile loop (an empty asm hides the
        // thread index from loop-invariant hoisting)
        int tid = (int)threadIdx.x;
        if constexpr (PERSIST) asm volatile("" : "+v"(tid));
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int e = tid + 256 * j;
            pre[j] = make_uint4(0, 0, 0, 0);
            if (e < kBandVec) {
                const int r = e / (kBand * kVec), rem = e - r * kBand * kVec;
                const int px = rem / kVec, c8 = rem - px * kVec;
                const int yy = ty0 + r - 1, xx = tx0 + px - 1;
                if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
                    const bf16* src = from_up
                        ? a.up + ((size_t)(tb * Hh + (yy >> 1)) * Wh + (xx >> 1)) * a.up_stride + k0 + 8 * c8
                        : a.in + ((size_t)(tb * a.H + yy) * a.W + xx) * a.in_stride + a.in_off + k0 + 8 * c8;
                    pre[j] = *reinterpret_cast<const uint4*>(src);
                }
            }
        }
    };
    auto stage_weights = [&](int t0, int k0) {
        for (int e = threadIdx.x; e < kWTaps * NT * 16 * kVec; e += 256) {
            const int tg = e / (NT * 16 * kVec), rem = e - tg * NT * 16 * kVec;
            const int nn = rem / kVec, c8 = rem - nn * kVec;
            *reinterpret_cast<uint4*>(wl + (tg * NT * 16 + nn) * ldw + 8 * c8) = *reinterpret_cast<const uint4*>(
                a.w + ((size_t)(t0 + tg) * a.npad + n0 + nn) * a.kpad + k0 + 8 * c8);
        }
    };
    // streamed weights (not resident): item 3c + g is the row g of taps of
    // chunk c; the next item is loaded into registers while the current one
    // runs on MFMA, and stored to LDS after the barrier that ends it
    constexpr int kWVec = 3 * NT * 16 * kVec;
    constexpr int kWPer = (kWVec + 255) / 256;
    static_assert(kWPer <= 3, "three weight vectors per thread at most");
    uint4 wpre[3] = {};
    // (each vector by a literal index, so the array stays in registers)
    auto load_w1 = [&](int j, int cc, int t0) {
        const int e = (int)threadIdx.x + 256 * j;
        uint4 v = make_uint4(0, 0, 0, 0);  // assigned on every path, or the array is not promoted
        if (e < kWVec) {
            const int tg = e / (NT * 16 * kVec), rem = e - tg * NT * 16 * kVec;
            const int nn = rem / kVec, c8 = rem - nn * kVec;
            v = *reinterpret_cast<const uint4*>(a.w + ((size_t)(t0 + tg) * a.npad + n0 + nn) * a.kpad + cc * KS + 8 * c8);
        }
        wpre[j] = v;
    };
    auto load_w = [&](int item) {
        const int cc = item / 3, t0 = 3 * (item - 3 * cc);
        load_w1(0, cc, t0);
        if constexpr (kWPer > 1) load_w1(1, cc, t0);
        if constexpr (kWPer > 2) load_w1(2, cc, t0);
    };
    auto store_w1 = [&](int j) {
        const int e = (int)threadIdx.x + 256 * j;
        if (e < kWVec) {
            const int tg = e / (NT * 16 * kVec), rem = e - tg * NT * 16 * kVec;
            const int nn = rem / kVec, c8 = rem - nn * kVec;
            *reinterpret_cast<uint4*>(wl + (tg * NT * 16 + nn) * ldw + 8 * c8) = wpre[j];
        }
    };
    auto store_w = [&]() {
        store_w1(0);
        if constexpr (kWPer > 1) store_w1(1);
        if constexpr (kWPer > 2) store_w1(2);
    };
    int b, y0, x0;
    coords(tile, b, y0, x0);
    load_band(b, y0, x0, 0);
    if constexpr (wresident) stage_weights(0, 0);  // visible after the first chunk's barriers
    else load_w(0);
    for (;;) {
        const int y = y0 + wave;
        const bool live = y < a.H;  // every wave takes part in the barriers
        const int next = PERSIST ? tile + (int)nwg : hi;
        int nb = 0, ny0 = 0, nx0 = 0;
        if (next < hi) coords(next, nb, ny0, nx0);
        f4v acc[MT][NT];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[m][t] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
        for (int c = 0; c < nchunk; ++c) {
            const int k0 = c * KS;
            __syncthreads();  // the area's previous readers are done
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int e = threadIdx.x + 256 * j;
                if (e < kBandVec) *reinterpret_cast<uint4*>(band + band_slot(e)) = pre[j];
            }
            {
                // one call site for the prefetch of the next (tile, chunk)
                const bool same = c + 1 < nchunk;
                if (same || next < hi) load_band(same ? b : nb, same ? y0 : ny0, same ? x0 : nx0, same ? k0 + KS : 0);
            }
            // the taps in three groups of three (one row of the 3x3 kernel each)
#pragma unroll
            for (int g = 0; g < 3; ++g) {
                if constexpr (wresident) {
                    if (g == 0) __syncthreads();
                } else {
                    if (g) __syncthreads();  // the previous row of taps is done
                    store_w();
                    const int item = 3 * c + g + 1;
                    if (item < 3 * nchunk) load_w(item);
                    else if (next < hi) load_w(0);
                    __syncthreads();
                }
                if (!live) continue;
#pragma unroll
                for (int tg = 0; tg < 3; ++tg) {
                    const int tap = 3 * g + tg;
                    int dy = tap / 3 - 1, dx = tap % 3 - 1;
                    if (a.flip) { dy = -dy; dx = -dx; }
                    const int px0 = (wave + dy + 1) * kBand + mi + dx + 1;  // + 16 m: the same swizzle
                    const bf16* ab = band + px0 * KS + KL * (kSwz ? kq ^ swz(px0) : kq);
                    const bf16* bb = wl + ((kAllW ? tap : tg) * NT * 16 + mi) * ldw + KL * kq;
                    if constexpr (KS == 32) {
                        uint4 av[MT], bv[NT];
#pragma unroll
                        for (int m = 0; m < MT; ++m) av[m] = *reinterpret_cast<const uint4*>(ab + m * 16 * KS);
#pragma unroll
                        for (int t = 0; t < NT; ++t) bv[t] = *reinterpret_cast<const uint4*>(bb + t * 16 * ldw);
#pragma unroll
                        for (int m = 0; m < MT; ++m)
#pragma unroll
                            for (int t = 0; t < NT; ++t)
                                acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                    __builtin_bit_cast(bf16x8, av[m]), __builtin_bit_cast(bf16x8, bv[t]), acc[m][t], 0, 0, 0);
                    } else {
                        s4 av[MT], bv[NT];
#pragma unroll
                        for (int m = 0; m < MT; ++m) av[m] = *reinterpret_cast<const s4*>(ab + m * 16 * KS);
#pragma unroll
                        for (int t = 0; t < NT; ++t) bv[t] = *reinterpret_cast<const s4*>(bb + t * 16 * ldw);
#pragma unroll
                        for (int m = 0; m < MT; ++m)
#pragma unroll
                            for (int t = 0; t < NT; ++t)
                                acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[m], bv[t], acc[m][t], 0, 0, 0);
                    }
                }
            }
        }
        if (!a.mask && !a.out_f32) {
            // bf16 outputs: the wave's kRun x NT*16 tile through LDS (the
            // area, free once every wave is past its last chunk), then
            // 16-byte stores
            __syncthreads();
            bf16* ot = area + wave * kRun * ldo;
            if (live) {
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int ch = t * 16 + mi;
                    const float bias = a.bias ? a.bias[n0 + ch] : 0.0f;
#pragma unroll
                    for (int m = 0; m < MT; ++m)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            float v = acc[m][t][r] + bias;
                            if (a.relu) v = v > 0.0f ? v : 0.0f;
                            ot[(m * 16 + 4 * kq + r) * ldo + ch] = f2bf(v);
                        }
                }
            }
            __syncthreads();
            if (a.pool) {
                // 2x2 average pool of the four rows' tiles (rows 2pr, 2pr + 1
                // of the tile make pooled row pr): the four bf16 outputs
                // summed in FP32 in (y, x) order, times 1/4, rounded once
                constexpr int kC8 = kCh / 8, kHalf = kRun / 2;
                const int Hq = a.H >> 1, Wq = a.W >> 1;
                for (int e = threadIdx.x; e < 2 * kHalf * kC8; e += 256) {
                    const int c8 = e % kC8, rest = e / kC8;
                    const int px = rest % kHalf, pr = rest / kHalf;
                    const int yp = (y0 >> 1) + pr, xp = (x0 >> 1) + px;
                    if (yp >= Hq || xp >= Wq) continue;
                    const bf16* t0 = area + (2 * pr * kRun + 2 * px) * ldo + 8 * c8;
                    const bf16* t1 = t0 + kRun * ldo;
                    uint4 v[4] = {*reinterpret_cast<const uint4*>(t0), *reinterpret_cast<const uint4*>(t0 + ldo),
                                  *reinterpret_cast<const uint4*>(t1), *reinterpret_cast<const uint4*>(t1 + ldo)};
                    bf16 o[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        float sum = 0.0f;
#pragma unroll
                        for (int j = 0; j < 4; ++j) sum += bf2f(reinterpret_cast<const bf16*>(&v[j])[k]);
                        o[k] = f2bf(0.25f * sum);
                    }
                    *reinterpret_cast<uint4*>(a.pool + ((size_t)(b * Hq + yp) * Wq + xp) * a.pool_stride + n0 + 8 * c8) =
                        *reinterpret_cast<const uint4*>(o);
                }
            }
            if constexpr (HEAD) {
                // the head on this wave's tile: A fragments of 16 pixels x 32
                // channels straight from the tile, logits to LDS (the area is
                // free once every wave holds its fragments), then one pixel
                // per lane: softmax of its 25 logits and the clamped 5x5
                // window of the input radiance (k_kpn_apply's arithmetic)
                static_assert(NT == 2 && KS == 32 && MT == 4, "head after a 32-channel layer, 64 pixels per wave");
                uint4 hv[MT];
#pragma unroll
                for (int m = 0; m < MT; ++m)
                    hv[m] = live ? *reinterpret_cast<const uint4*>(ot + (m * 16 + mi) * ldo + 8 * kq) : make_uint4(0, 0, 0, 0);
                __syncthreads();
                if (live) {
                    float* lgt = reinterpret_cast<float*>(area) + wave * kRun * 25;
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const uint4 bw = *reinterpret_cast<const uint4*>(a.head_w + (t * 16 + mi) * 32 + 8 * kq);
                        const int ch = t * 16 + mi;
                        const float hb = a.head_b[ch];
#pragma unroll
                        for (int m = 0; m < MT; ++m) {
                            f4v l = f4v{0.0f, 0.0f, 0.0f, 0.0f};
                            l = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, hv[m]),
                                                                        __builtin_bit_cast(bf16x8, bw), l, 0, 0, 0);
                            if (ch < 25)
#pragma unroll
                                for (int r = 0; r < 4; ++r) lgt[(m * 16 + 4 * kq + r) * 25 + ch] = l[r] + hb;
                        }
                    }
                    // the wave's lanes exchange the logits through LDS
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const int x = x0 + lane;
                    if (x < a.img_w && y < a.img_h) {
                        float gl[25];
#pragma unroll
                        for (int k = 0; k < 25; ++k) gl[k] = lgt[lane * 25 + k];
                        float mx = gl[0];
#pragma unroll
                        for (int k = 1; k < 25; ++k) mx = fmaxf(mx, gl[k]);
                        // softmax with the hardware's base-2 exponential of
                        // the logits scaled by log2(e) (one v_exp_f32 and an
                        // FMA per logit instead of a full expf)
                        float sum = 0.0f;
                        const float kL2e = 1.44269504088896340736f, mxs = mx * kL2e;
#pragma unroll
                        for (int k = 0; k < 25; ++k) { gl[k] = __builtin_amdgcn_exp2f(fmaf(gl[k], kL2e, -mxs)); sum += gl[k]; }
                        const float is = 1.0f / sum;
                        const float* rad = a.rad + (size_t)b * 10 * a.img_h * a.img_w;
                        const size_t plane = (size_t)a.img_h * a.img_w;
                        float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f;
#pragma unroll
                        for (int k = 0; k < 25; ++k) {
                            const float wgt = gl[k] * is;
                            const int yy = min(max(y + k / 5 - 2, 0), a.img_h - 1), xx = min(max(x + k % 5 - 2, 0), a.img_w - 1);
                            const size_t q = (size_t)yy * a.img_w + xx;
                            r0 += wgt * rad[q];
                            r1 += wgt * rad[plane + q];
                            r2 += wgt * rad[2 * plane + q];
                        }
                        const size_t p = (size_t)(b * a.img_h + y) * a.img_w + x;
                        a.rgb[3 * p] = r0;
                        a.rgb[3 * p + 1] = r1;
                        a.rgb[3 * p + 2] = r2;
                    }
                }
            } else if (live) {
                constexpr int kOutVec = kRun * kCh / 8;
#pragma unroll
                for (int j = 0; j < (kOutVec + 63) / 64; ++j) {
                    const int e = lane + 64 * j;
                    if (e < kOutVec) {
                        const int px = e / (kCh / 8), c8 = e - px * (kCh / 8);
                        const int x = x0 + px;
                        if (x < a.W) {
                            const size_t p = (size_t)(b * a.H + y) * a.W + x;
                            *reinterpret_cast<uint4*>(a.out + p * a.out_stride + a.out_off + n0 + 8 * c8) =
                                *reinterpret_cast<const uint4*>(ot + px * ldo + 8 * c8);
                        }
                    }
                }
            }
        } else if (live) {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int n = n0 + t * 16 + mi;
                const float bias = a.bias ? a.bias[n] : 0.0f;
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int x = x0 + m * 16 + 4 * kq + r;
                        if (x >= a.W) continue;
                        float v = acc[m][t][r] + bias;
                        if (a.relu) v = v > 0.0f ? v : 0.0f;
                        const size_t p = (size_t)(b * a.H + y) * a.W + x;
                        if (a.mask && !(bf2f(a.mask[p * a.mask_stride + a.mask_off + n]) > 0.0f)) v = 0.0f;
                        if (a.out_f32) a.out_f32[p * a.npad + n] = v;
                        else a.out[p * a.out_stride + a.out_off + n] = f2bf(v);
                    }
            }
        }
        if (next >= hi) break;
        tile = next;
        b = nb;
        y0 = ny0;
        x0 = nx0;
    }
}

// The kernels: one tile per workgroup, or (PERSIST, layers of one K chunk
// whose weights stay in LDS) as many workgroups as are resident, each
// taking tiles in turn with the next tile's band loaded during the current
// one; three waves per SIMD keep three such workgroups on a CU.
template <int MT, int NT, int KS, bool HEAD = false>
__global__ __launch_bounds__(256) void k_conv3r(ConvArgs a) {
    extern __shared__ bf16 lds[];
    conv3r_body<MT, NT, KS, HEAD, false>(a, lds);
}
template <int MT, int NT, int KS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_conv3p(ConvArgs a) {
    extern __shared__ bf16 lds[];
    conv3r_body<MT, NT, KS, false, true>(a, lds);
}

struct WgradArgs {
    const bf16* x; int x_stride, x_off;   // layer input (kpad channels)
    const bf16* g; int g_stride, g_off;   // gradient of the pre-activation (npad channels)
    float* dw;   // [taps][cin][cout] FP32, logical
    float* db;   // [cout]
    int B, H, W, kpad, npad, cin, cout, taps, chunk;
};

constexpr int kWgTile = 64;  // pixels staged per step (a multiple of 32)
constexpr int kWgSlots = 8;  // 16x16 output tiles per wave (8 waves cover up to 64 tiles)

// dW[t][ci][co] = sum_p x[p + off_t][ci] g[p][co]; grid.x pixel chunks, grid.y taps.
__global__ __launch_bounds__(512) void k_conv_wgrad(WgradArgs a) {
    extern __shared__ bf16 lds[];
    bf16* xs = lds;                       // [64][kpad]
    bf16* gs = lds + kWgTile * a.kpad;    // [64][npad]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = blockIdx.y;
    const int dy = a.taps == 9 ? t / 3 - 1 : 0, dx = a.taps == 9 ? t % 3 - 1 : 0;
    const int P = a.B * a.H * a.W;
    const int p_begin = blockIdx.x * a.chunk, p_end = min(P, p_begin + a.chunk);
    const int tiles_n = a.npad / 16, tiles = (a.kpad / 16) * tiles_n;
    f4v acc[kWgSlots];
#pragma unroll
    for (int i = 0; i < kWgSlots; ++i) acc[i] = {0, 0, 0, 0};
    float bsum = 0.0f;
    const int mi = lane & 15, kq = lane >> 4;
    for (int p0 = p_begin; p0 < p_end; p0 += kWgTile) {
        __syncthreads();
        const int xv = a.kpad / 8, gv = a.npad / 8;  // 16-byte vectors per pixel
        for (int idx = threadIdx.x; idx < kWgTile * xv; idx += 512) {
            const int q = idx / xv, part = idx - q * xv;
            const int p = p0 + q;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (p < p_end) {
                const int b = p / (a.H * a.W), rem = p - b * a.H * a.W, yy = rem / a.W + dy, xx = rem % a.W + dx;
                if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                    v = *reinterpret_cast<const uint4*>(a.x + ((size_t)(b * a.H + yy) * a.W + xx) * a.x_stride + a.x_off + 8 * part);
            }
            *reinterpret_cast<uint4*>(xs + q * a.kpad + 8 * part) = v;
        }
        for (int idx = threadIdx.x; idx < kWgTile * gv; idx += 512) {
            const int q = idx / gv, part = idx - q * gv;
            const int p = p0 + q;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (p < p_end) v = *reinterpret_cast<const uint4*>(a.g + (size_t)p * a.g_stride + a.g_off + 8 * part);
            *reinterpret_cast<uint4*>(gs + q * a.npad + 8 * part) = v;
        }
        __syncthreads();
        if (t == 0 && threadIdx.x < a.npad)
            for (int q = 0; q < kWgTile; ++q) bsum += bf2f(gs[q * a.npad + threadIdx.x]);
        // K = pixels, 32 per instruction (v_mfma_f32_16x16x32_bf16): lane
        // (mi, kq) holds pixels kk + 8 kq .. + 7 of input channel ci (A) and
        // output channel co (B)
#pragma unroll 1
        for (int kk = 0; kk < kWgTile; kk += 32) {
#pragma unroll
            for (int slot = 0; slot < kWgSlots; ++slot) {
                const int ti = wave + 8 * slot;
                if (ti >= tiles) continue;
                const int ci = (ti / tiles_n) * 16 + mi, co = (ti % tiles_n) * 16 + mi;
                bf16x8 av, bv;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int q = kk + 8 * kq + e;
                    av[e] = __builtin_bit_cast(__bf16, xs[q * a.kpad + ci]);
                    bv[e] = __builtin_bit_cast(__bf16, gs[q * a.npad + co]);
                }
                acc[slot] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[slot], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int slot = 0; slot < kWgSlots; ++slot) {
        const int ti = wave + 8 * slot;
        if (ti >= tiles) continue;
        const int co = (ti % tiles_n) * 16 + mi;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int ci = (ti / tiles_n) * 16 + 4 * kq + r;
            if (ci < a.cin && co < a.cout) atomicAdd(&a.dw[((size_t)t * a.cin + ci) * a.cout + co], acc[slot][r]);
        }
    }
    if (t == 0 && threadIdx.x < a.cout) atomicAdd(&a.db[threadIdx.x], bsum);
}

// planar f32 [B][10][H][W] -> bf16 NHWC16 at the padded size (replicated edge)
__global__ void k_pack(const float* __restrict__ in, bf16* __restrict__ x, int B, int H, int W, int Hp, int Wp) {
    const int P = B * Hp * Wp;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        const int b = p / (Hp * Wp), rem = p - b * Hp * Wp;
        const int y = min(rem / Wp, H - 1), xx = min(rem % Wp, W - 1);
        const float* src = in + (size_t)b * 10 * H * W + (size_t)y * W + xx;
        bf16 v[16];
        for (int k = 0; k < 10; ++k) v[k] = f2bf(in_feature(src[(size_t)k * H * W], k));
        for (int k = 10; k < 16; ++k) v[k] = 0;
        uint4* dst = reinterpret_cast<uint4*>(x + (size_t)p * 16);
        dst[0] = *reinterpret_cast<uint4*>(&v[0]);
        dst[1] = *reinterpret_cast<uint4*>(&v[8]);
    }
}

// G-buffer -> the first layer's packed input at the padded size (k_pack of
// what k_inputs writes: the same values, transform and rounding) plus the
// three radiance planes the kernel apply reads, in one pass
__global__ void k_inputs_pack(const float* __restrict__ rgb, const float* __restrict__ nrm, const float* __restrict__ pos,
                              const float* __restrict__ alb, float* __restrict__ rad, bf16* __restrict__ x, int H, int W,
                              int Hp, int Wp) {
    const int P = Hp * Wp;
    const size_t plane = (size_t)H * W;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        const int yp = p / Wp, xp = p - yp * Wp;
        const size_t i = (size_t)min(yp, H - 1) * W + min(xp, W - 1);
        float f[10];
        for (int c = 0; c < 3; ++c) {
            f[c] = rgb[3 * i + c];
            f[3 + c] = alb[3 * i + c];
            f[6 + c] = nrm[3 * i + c];
        }
        const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        f[9] = sqrtf(px * px + py * py + pz * pz);
        bf16 v[16];
        for (int k = 0; k < 10; ++k) v[k] = f2bf(in_feature(f[k], k));
        for (int k = 10; k < 16; ++k) v[k] = 0;
        uint4* dst = reinterpret_cast<uint4*>(x + (size_t)p * 16);
        dst[0] = *reinterpret_cast<uint4*>(&v[0]);
        dst[1] = *reinterpret_cast<uint4*>(&v[8]);
        if (yp < H && xp < W)
            for (int c = 0; c < 3; ++c) rad[c * plane + i] = f[c];
    }
}

// G-buffer (packed RGB per pixel: radiance, world normal, view-space
// position, albedo) -> planar [10][H][W] network input: radiance, albedo,
// normal, distance to the camera (|position|; the far plane for misses)
__global__ void k_inputs(const float* __restrict__ rgb, const float* __restrict__ nrm, const float* __restrict__ pos,
                         const float* __restrict__ alb, float* __restrict__ out, size_t px) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < px; i += (size_t)gridDim.x * blockDim.x) {
        const float* src[3] = {rgb, alb, nrm};
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) out[(3 * k + c) * px + i] = src[k][3 * i + c];
        const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        out[9 * px + i] = sqrtf(x * x + y * y + z * z);
    }
}

// nearest 2x upsample of src (H/2, W/2, C) into the dst slice at (H, W)
__global__ void k_up(const bf16* __restrict__ s, bf16* __restrict__ d, int d_stride, int d_off, int C, int B, int H, int W) {
    // one thread per (output pixel, 8 channels): 16-byte loads and stores
    const uint32_t c8 = C / 8;
    const uint32_t n = (uint32_t)B * H * W * c8;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t g = i % c8, p = i / c8;
        const uint32_t x = p % W, q = p / W, y = q % H, b = q / H;
        const uint4 v = *reinterpret_cast<const uint4*>(s + ((size_t)(b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + 8 * g);
        *reinterpret_cast<uint4*>(d + (size_t)p * d_stride + d_off + 8 * g) = v;
    }
}

// backward of the upsample: dst (H/2, W/2, C) = sum of the 2x2 children of the slice
__global__ void k_up_bwd(const bf16* __restrict__ g, int g_stride, int g_off, bf16* __restrict__ d, int C, int B, int H, int W) {
    const int h = H / 2, w = W / 2;
    const size_t n = (size_t)B * h * w * C;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = i % C;
        const size_t p = i / C;
        const int b = p / (h * w), rem = p - b * h * w, y = rem / w, x = rem % w;
        auto at = [&](int yy, int xx) { return bf2f(g[((size_t)(b * H + yy) * W + xx) * g_stride + g_off + c]); };
        d[i] = f2bf(at(2 * y, 2 * x) + at(2 * y, 2 * x + 1) + at(2 * y + 1, 2 * x) + at(2 * y + 1, 2 * x + 1));
    }
}

// d[p][c] = (pool_bwd(gp)[p][c] + skip[p][c]) * (act[p][c] > 0)
__global__ void k_pool_bwd_add_mask(const bf16* __restrict__ gp, const bf16* __restrict__ sk, int sk_stride, int sk_off,
                                    const bf16* __restrict__ act, int act_stride, int act_off, bf16* __restrict__ d, int C,
                                    int B, int H, int W) {
    const size_t n = (size_t)B * H * W * C;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = i % C;
        const size_t p = i / C;
        const int b = p / (H * W), rem = p - b * H * W, y = rem / W, x = rem % W;
        float v = 0.25f * bf2f(gp[((size_t)(b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c]) +
                  bf2f(sk[p * sk_stride + sk_off + c]);
        if (!(bf2f(act[p * act_stride + act_off + c]) > 0.0f)) v = 0.0f;
        d[i] = f2bf(v);
    }
}

// softmax over the 25 logits and the 5x5 kernel applied to the radiance
// (planar input channels 0..2, clamped neighbours); out is packed RGB.
__global__ void k_kpn_apply(const float* __restrict__ logits, const float* __restrict__ in, float* __restrict__ wk,
                            float* __restrict__ out, int B, int H, int W, int Hp, int Wp) {
    const int P = B * H * W;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        const int b = p / (H * W), rem = p - b * H * W, y = rem / W, x = rem % W;
        const float* lg = logits + ((size_t)(b * Hp + y) * Wp + x) * 32;
        float m = lg[0];
        for (int k = 1; k < 25; ++k) m = fmaxf(m, lg[k]);
        float e[25], s = 0.0f;
        for (int k = 0; k < 25; ++k) { e[k] = expf(lg[k] - m); s += e[k]; }
        const float is = 1.0f / s;
        const float* rad = in + (size_t)b * 10 * H * W;
        float r0 = 0, r1 = 0, r2 = 0;
        for (int k = 0; k < 25; ++k) {
            const float wgt = e[k] * is;
            if (wk) wk[(size_t)p * 25 + k] = wgt;
            const int yy = min(max(y + k / 5 - 2, 0), H - 1), xx = min(max(x + k % 5 - 2, 0), W - 1);
            const size_t q = (size_t)yy * W + xx;
            r0 += wgt * rad[q];
            r1 += wgt * rad[(size_t)H * W + q];
            r2 += wgt * rad[(size_t)2 * H * W + q];
        }
        out[(size_t)p * 3] = r0; out[(size_t)p * 3 + 1] = r1; out[(size_t)p * 3 + 2] = r2;
    }
}

// relMSE loss (FP64 reduction, one atomic per workgroup: atomics on one
// address serialise) and its gradient with respect to the output.
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ out, const float* __restrict__ tgt,
                                              float* __restrict__ dout, size_t n3, double inv_count, double* loss) {
    __shared__ double wsum[4];
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        const double r = tgt[i], d = (double)out[i] - r, den = r * r + 1e-2;
        acc += d * d / den;
        dout[i] = (float)(2.0 * d / den * inv_count);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]));
}

// backward of the kernel application and softmax -> bf16 logit gradient [Pp][32]
__global__ void k_kpn_bwd(const float* __restrict__ dout, const float* __restrict__ wk, const float* __restrict__ in,
                          bf16* __restrict__ g, int B, int H, int W, int Hp, int Wp) {
    const int Pp = B * Hp * Wp;
    for (int pp = blockIdx.x * blockDim.x + threadIdx.x; pp < Pp; pp += gridDim.x * blockDim.x) {
        const int b = pp / (Hp * Wp), rem = pp - b * Hp * Wp, y = rem / Wp, x = rem % Wp;
        bf16* gd = g + (size_t)pp * 32;
        if (y >= H || x >= W) {
            for (int k = 0; k < 32; ++k) gd[k] = 0;
            continue;
        }
        const size_t p = (size_t)(b * H + y) * W + x;
        const float d0 = dout[p * 3], d1 = dout[p * 3 + 1], d2 = dout[p * 3 + 2];
        const float* rad = in + (size_t)b * 10 * H * W;
        float dw[25], dot = 0.0f;
        for (int k = 0; k < 25; ++k) {
            const int yy = min(max(y + k / 5 - 2, 0), H - 1), xx = min(max(x + k % 5 - 2, 0), W - 1);
            const size_t q = (size_t)yy * W + xx;
            dw[k] = d0 * rad[q] + d1 * rad[(size_t)H * W + q] + d2 * rad[(size_t)2 * H * W + q];
            dot += wk[p * 25 + k] * dw[k];
        }
        for (int k = 0; k < 25; ++k) gd[k] = f2bf(wk[p * 25 + k] * (dw[k] - dot));
        for (int k = 25; k < 32; ++k) gd[k] = 0;
    }
}

// fp32 logical parameters -> bf16 [taps][npad][kpad] (forward) and
// [taps][kpad][npad] (backward data) plus padded bias
__global__ void k_relayout(const float* __restrict__ w, const float* __restrict__ bias, bf16* wf, bf16* wb, float* bpad,
                           int cin, int cout, int taps, int kpad, int npad) {
    const int n = taps * kpad * npad;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int t = i / (kpad * npad), rem = i - t * kpad * npad, co = rem / kpad, ci = rem % kpad;
        const float v = (ci < cin && co < cout) ? w[((size_t)t * cin + ci) * cout + co] : 0.0f;
        const bf16 b = f2bf(v);
        wf[i] = b;
        wb[((size_t)t * kpad + ci) * npad + co] = b;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npad; i += gridDim.x * blockDim.x) bpad[i] = i < cout ? bias[i] : 0.0f;
}

// squared L2 norm of the gradient (FP64 reduction, one atomic per workgroup:
// atomics on one address serialise)
__global__ __launch_bounds__(256) void k_sumsq(const float* __restrict__ g, int n, double* out) {
    __shared__ double wsum[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += (double)g[i] * g[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]));
}

// Adam on the gradient scaled to a global L2 norm of at most `clip` (0: no clipping)
__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float c1, float c2,
                       float eps, const double* sumsq, float clip) {
    const double norm = sqrt(*sumsq);
    const float scale = (clip > 0.0f && norm > (double)clip) ? (float)((double)clip / norm) : 1.0f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i] * scale;
        const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ void k_zero_f32(float* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}
__global__ void k_zero1(double* p) { *p = 0.0; }

// Fallback check of the denoiser's output: flag = 1 when any value is not
// finite (NaN or inf), else 0.  The flag is written by the first block
// (cleared) and raised by any lane that sees a bad value; the one-block clear
// runs as its own launch first.
__global__ void k_flag_set(uint32_t* flag, uint32_t v) { *flag = v; }
__global__ void k_flag_nonfinite(const float* __restrict__ x, size_t n, uint32_t* flag) {
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        bad |= !isfinite(x[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

uint32_t grid_for(size_t n, int block = 256) { return (uint32_t)std::min<size_t>((n + block - 1) / block, 16384); }

}  // namespace

struct pico_cnn {
    pico_hip_ctx* c = nullptr;
    int maxW = 0, maxH = 0, maxB = 0, Hp = 0, Wp = 0;
    size_t nparams = 0;
    size_t off[11][2];
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    bf16 *wf[11], *wb[11];
    float* bpad[11];
    int step = 0;
    float clip = 1.0f;          // global gradient-norm clip of the Adam step (0: off)
    double* gnorm2 = nullptr;   // squared gradient norm of the last step
    // activations (bf16 NHWC)
    bf16 *x0, *e0a, *c0, *p0, *e1a, *c1, *p1, *ba, *bb, *d1a, *d1b, *d0a, *d0b;
    float *head, *wk, *out, *dout;
    double* loss;
    // gradients
    bf16 *g_head, *g_d0b, *g_d0a, *g_c0, *g_d1b, *g_d1a, *g_c1, *g_bb, *g_ba, *g_p1, *g_e1b, *g_e1a, *g_p0, *g_e0b, *g_e0a;
    std::vector<void*> allocs;
};

namespace {

template <class T>
T* dalloc(pico_cnn* n, size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count * sizeof(T), 64)) != hipSuccess) return nullptr;
    n->allocs.push_back(p);
    return (T*)p;
}

// The persistent 3x3 kernel: as many workgroups as are resident at once
// (per CU, at this LDS size; one size per instantiation), at most one per
// tile, per block of output channels.
template <int MT, int NT, int KS, bool HEAD>
int launch_conv3r(pico_cnn* n, const char* name, const ConvArgs& a, size_t lds) {
    constexpr int run4 = MT * 16;
    const long tiles = (long)a.B * ((a.H + 3) / 4) * ((a.W + run4 - 1) / run4);
    const unsigned ny = (unsigned)(a.npad / (NT * 16));
    if constexpr (NT == 2 && !HEAD) {
        if (a.kpad == KS) {
            static std::atomic<int> per_cu{0};
            int r = per_cu.load(std::memory_order_relaxed);
            if (r == 0) {
                int nb = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_conv3p<MT, NT, KS>, 256, lds) != hipSuccess || nb < 1)
                    nb = 1;
                per_cu.store(r = nb, std::memory_order_relaxed);
            }
            const long cap = std::max(1L, (long)r * std::max(1, n->c->num_cus) / (long)ny);
            PICO_LAUNCH(n->c, name, (k_conv3p<MT, NT, KS>), dim3((unsigned)std::min(tiles, cap), ny), 256, lds, a);
            return PICO_OK;
        }
    }
    PICO_LAUNCH(n->c, name, (k_conv3r<MT, NT, KS, HEAD>), dim3((unsigned)tiles, ny), 256, lds, a);
    return PICO_OK;
}

// Launches k_conv with the widest channel block (16, 32 or 64) that divides
// the output channels, 8 accumulators per wave, and the 32-deep instruction
// where the input channels allow it.
template <int KS>
int launch_conv_k(pico_cnn* n, const char* name, const ConvArgs& a) {
    const int nt = (a.npad % 64 == 0) ? 4 : (a.npad % 32 == 0) ? 2 : 1;
    const int run = 4 * (8 / nt) * 16;
    const dim3 grid((unsigned)(a.B * a.H * ((a.W + run - 1) / run)), (unsigned)(a.npad / (nt * 16)));
    if (a.pool && !(a.taps == 9 && nt >= 2)) return fail(n->c, PICO_ERR_STATE, "fused pool needs the 3x3 kernel");
    if (a.pool && (a.mask || a.out_f32 || a.head_w || ((a.H | a.W) & 1))) return fail(n->c, PICO_ERR_STATE, "fused pool needs bf16 outputs of even size");
    if (a.taps == 9 && nt >= 2) {
        // 3x3 with all operands in LDS, four output rows per tile: the
        // chunk's weights (k_conv3r's kWTaps taps), then one area for the six
        // staged rows of MT * 16 + 2 pixels, the output tiles and the logits
        const int mt = 8 / nt, run4 = mt * 16;
        // (nine taps for the persistent kernel: 32 outputs, one chunk, no head)
        const size_t wtaps = nt == 2 && a.kpad == KS && !a.head_w ? 9 : 3;
        const size_t wbytes = wtaps * nt * 16 * (KS + 8) * sizeof(bf16);
        const size_t area = std::max(6 * (size_t)(run4 + 2) * KS, 4 * (size_t)run4 * (nt * 16 + 8)) * sizeof(bf16);
        if (a.head_w) {
            if constexpr (KS == 32) {
                if (nt != 2 || a.mask || a.out_f32) return fail(n->c, PICO_ERR_STATE, "head fusion needs 32 channels");
                // logits of the four waves (64 x 25 floats each) share the area
                return launch_conv3r<4, 2, 32, true>(n, name, a, wbytes + std::max(area, 4 * (size_t)run4 * 25 * sizeof(float)));
            }
            return fail(n->c, PICO_ERR_STATE, "head fusion needs 32 input channels");
        }
        if (nt == 4) return launch_conv3r<2, 4, KS, false>(n, name, a, wbytes + area);
        return launch_conv3r<4, 2, KS, false>(n, name, a, wbytes + area);
    }
    if (a.up) return fail(n->c, PICO_ERR_STATE, "fused upsampling needs the 3x3 kernel");
    const size_t lds = (size_t)nt * 16 * (a.kpad + 8) * sizeof(bf16);
    if (nt == 4) PICO_LAUNCH(n->c, name, (k_conv<2, 4, KS>), grid, 256, lds, a);
    else if (nt == 2) PICO_LAUNCH(n->c, name, (k_conv<4, 2, KS>), grid, 256, lds, a);
    else PICO_LAUNCH(n->c, name, (k_conv<8, 1, KS>), grid, 256, lds, a);
    return PICO_OK;
}

int launch_conv(pico_cnn* n, const char* name, const ConvArgs& a) {
    return a.kpad % 32 == 0 ? launch_conv_k<32>(n, name, a) : launch_conv_k<16>(n, name, a);
}

int conv(pico_cnn* n, int l, const bf16* in, int in_stride, int in_off, bf16* out, int out_stride, int out_off, int B,
         int H, int W, bool relu, float* out_f32 = nullptr, const bf16* up = nullptr, int up_stride = 0, int up_ch = 0,
         bf16* pool = nullptr, int pool_stride = 0) {
    const LayerDef& L = kL[l];
    ConvArgs a{in, in_stride, in_off, n->wf[l], n->bpad[l], out, out_stride, out_off, out_f32, nullptr, 0, 0,
               B, H, W, L.kpad, L.npad, L.taps, 0, relu ? 1 : 0, up, up_stride, up_ch};
    a.pool = pool;
    a.pool_stride = pool_stride;
    return launch_conv(n, "cnn_conv_fwd", a);
}

// dX of layer l: gradient wrt its input (kpad channels), masked by `mask`
int conv_dx(pico_cnn* n, int l, const bf16* g, int g_stride, bf16* dx, int dx_stride, int dx_off, const bf16* mask,
            int mask_stride, int mask_off, int B, int H, int W) {
    const LayerDef& L = kL[l];
    ConvArgs a{g, g_stride, 0, n->wb[l], nullptr, dx, dx_stride, dx_off, nullptr, mask, mask_stride, mask_off,
               B, H, W, L.npad, L.kpad, L.taps, 1, 0};
    return launch_conv(n, "cnn_conv_bwd_data", a);
}

int conv_dw(pico_cnn* n, int l, const bf16* x, int x_stride, int x_off, const bf16* g, int g_stride, int B, int H, int W) {
    const LayerDef& L = kL[l];
    const int P = B * H * W;
    const int chunk = 2048;
    WgradArgs a{x, x_stride, x_off, g, g_stride, 0, n->grads + n->off[l][0], n->grads + n->off[l][1],
                B, H, W, L.kpad, L.npad, L.cin, L.cout, L.taps, chunk};
    dim3 grid((unsigned)((P + chunk - 1) / chunk), (unsigned)L.taps);
    const size_t lds = (size_t)kWgTile * (L.kpad + L.npad) * sizeof(bf16);
    PICO_LAUNCH(n->c, "cnn_conv_bwd_weight", k_conv_wgrad, grid, 512, lds, a);
    return PICO_OK;
}

int relayout(pico_cnn* n) {
    for (int l = 0; l < 11; ++l) {
        const LayerDef& L = kL[l];
        PICO_LAUNCH(n->c, "cnn_relayout", k_relayout, grid_for((size_t)L.taps * L.kpad * L.npad), 256, 0,
                    n->params + n->off[l][0], n->params + n->off[l][1], n->wf[l], n->wb[l], n->bpad[l], L.cin, L.cout, L.taps,
                    L.kpad, L.npad);
    }
    return PICO_OK;
}

int check_dims(pico_cnn* n, uint32_t B, uint32_t W, uint32_t H) {
    if (B == 0 || W == 0 || H == 0) return fail(n->c, PICO_ERR_ARG, "batch and image size must be positive");
    const int Hp = (int)((H + 3) / 4 * 4), Wp = (int)((W + 3) / 4 * 4);
    if ((int)B > n->maxB || Hp > n->Hp || Wp > n->Wp || (size_t)B * Hp * Wp > (size_t)n->maxB * n->Hp * n->Wp)
        return fail(n->c, PICO_ERR_LIMIT, "batch or image larger than the sizes the network was created for");
    return PICO_OK;
}

// Forward of a device-resident batch of B planar inputs; the filtered
// radiance lands in n->out, or at inference in d_out when given.  prepacked:
// the first layer's input (x0) is already written and d_in need only hold
// the radiance planes.
int forward(pico_cnn* n, const float* d_in, int B, int H, int W, bool keep_wk, bool prepacked = false,
            float* d_out = nullptr) {
    int rc;
    const int Hp = (H + 3) / 4 * 4, Wp = (W + 3) / 4 * 4;
    const int H1 = Hp / 2, W1 = Wp / 2, H2 = Hp / 4, W2 = Wp / 4;
    const size_t P0 = (size_t)B * Hp * Wp;
    if (!prepacked) PICO_LAUNCH(n->c, "cnn_pack", k_pack, grid_for(P0), 256, 0, d_in, n->x0, B, H, W, Hp, Wp);
    if ((rc = conv(n, 0, n->x0, 16, 0, n->e0a, 32, 0, B, Hp, Wp, true))) return rc;
    // e0b -> skip slice of c0, and its 2x2 pool -> p0
    if ((rc = conv(n, 1, n->e0a, 32, 0, n->c0, 96, 64, B, Hp, Wp, true, nullptr, nullptr, 0, 0, n->p0, 32))) return rc;
    if ((rc = conv(n, 2, n->p0, 32, 0, n->e1a, 64, 0, B, H1, W1, true))) return rc;
    // e1b -> skip slice of c1, and its 2x2 pool -> p1
    if ((rc = conv(n, 3, n->e1a, 64, 0, n->c1, 192, 128, B, H1, W1, true, nullptr, nullptr, 0, 0, n->p1, 64))) return rc;
    if ((rc = conv(n, 4, n->p1, 64, 0, n->ba, 128, 0, B, H2, W2, true))) return rc;
    if ((rc = conv(n, 5, n->ba, 128, 0, n->bb, 128, 0, B, H2, W2, true))) return rc;
    // The decoder's inputs are [nearest upsample of the level below, skip]:
    // a training pass materialises them (the backward pass reads them), an
    // inference pass reads the upsampled channels straight from the
    // half-resolution tensor while staging the convolution's input.
    if (keep_wk) {
        PICO_LAUNCH(n->c, "cnn_up", k_up, grid_for(P0 / 4 * 16), 256, 0, n->bb, n->c1, 192, 0, 128, B, H1, W1);
        if ((rc = conv(n, 6, n->c1, 192, 0, n->d1a, 64, 0, B, H1, W1, true))) return rc;
    } else if ((rc = conv(n, 6, n->c1, 192, 0, n->d1a, 64, 0, B, H1, W1, true, nullptr, n->bb, 128, 128))) {
        return rc;
    }
    if ((rc = conv(n, 7, n->d1a, 64, 0, n->d1b, 64, 0, B, H1, W1, true))) return rc;
    if (keep_wk) {
        PICO_LAUNCH(n->c, "cnn_up", k_up, grid_for(P0 * 8), 256, 0, n->d1b, n->c0, 96, 0, 64, B, Hp, Wp);
        if ((rc = conv(n, 8, n->c0, 96, 0, n->d0a, 32, 0, B, Hp, Wp, true))) return rc;
    } else if ((rc = conv(n, 8, n->c0, 96, 0, n->d0a, 32, 0, B, Hp, Wp, true, nullptr, n->d1b, 64, 64))) {
        return rc;
    }
    if (!keep_wk) {
        // inference: the last 3x3 layer with the head, softmax and kernel
        // apply on its output tile, so neither its activations nor the
        // logits reach memory
        const LayerDef& L = kL[9];
        ConvArgs a{n->d0a, 32, 0, n->wf[9], n->bpad[9], n->d0b, 32, 0, nullptr, nullptr, 0, 0,
                   B, Hp, Wp, L.kpad, L.npad, L.taps, 0, 1, nullptr, 0, 0,
                   n->wf[10], n->bpad[10], d_in, d_out ? d_out : n->out, H, W};
        return launch_conv(n, "cnn_conv_fwd", a);
    }
    if ((rc = conv(n, 9, n->d0a, 32, 0, n->d0b, 32, 0, B, Hp, Wp, true))) return rc;
    if ((rc = conv(n, 10, n->d0b, 32, 0, nullptr, 32, 0, B, Hp, Wp, false, n->head))) return rc;
    PICO_LAUNCH(n->c, "cnn_kpn_apply", k_kpn_apply, grid_for((size_t)B * H * W), 256, 0, n->head, d_in,
                n->wk, n->out, B, H, W, Hp, Wp);
    return PICO_OK;
}

// loss + backward (gradients accumulated into n->grads, zeroed first)
int backward(pico_cnn* n, const float* d_in, const float* d_tgt, int B, int H, int W) {
    int rc;
    const int Hp = (H + 3) / 4 * 4, Wp = (W + 3) / 4 * 4;
    const int H1 = Hp / 2, W1 = Wp / 2, H2 = Hp / 4, W2 = Wp / 4;
    const size_t P0 = (size_t)B * Hp * Wp;
    const size_t n3 = (size_t)B * H * W * 3;
    PICO_LAUNCH(n->c, "cnn_zero", k_zero_f32, grid_for(n->nparams), 256, 0, n->grads, n->nparams);
    PICO_LAUNCH(n->c, "cnn_zero", k_zero1, 1, 1, 0, n->loss);
    PICO_LAUNCH(n->c, "cnn_loss", k_loss, std::min<uint32_t>(grid_for(n3), 1024), 256, 0, n->out, d_tgt, n->dout, n3,
                1.0 / (double)n3, n->loss);
    PICO_LAUNCH(n->c, "cnn_kpn_bwd", k_kpn_bwd, grid_for(P0), 256, 0, n->dout, n->wk, d_in, n->g_head, B, H, W, Hp, Wp);
    // head
    if ((rc = conv_dw(n, 10, n->d0b, 32, 0, n->g_head, 32, B, Hp, Wp))) return rc;
    if ((rc = conv_dx(n, 10, n->g_head, 32, n->g_d0b, 32, 0, n->d0b, 32, 0, B, Hp, Wp))) return rc;
    // d0b, d0a
    if ((rc = conv_dw(n, 9, n->d0a, 32, 0, n->g_d0b, 32, B, Hp, Wp))) return rc;
    if ((rc = conv_dx(n, 9, n->g_d0b, 32, n->g_d0a, 32, 0, n->d0a, 32, 0, B, Hp, Wp))) return rc;
    if ((rc = conv_dw(n, 8, n->c0, 96, 0, n->g_d0a, 32, B, Hp, Wp))) return rc;
    if ((rc = conv_dx(n, 8, n->g_d0a, 32, n->g_c0, 96, 0, n->c0, 96, 0, B, Hp, Wp))) return rc;
    PICO_LAUNCH(n->c, "cnn_up_bwd", k_up_bwd, grid_for(P0 / 4 * 64), 256, 0, n->g_c0, 96, 0, n->g_d1b, 64, B, Hp, Wp);
    // d1b, d1a
    if ((rc = conv_dw(n, 7, n->d1a, 64, 0, n->g_d1b, 64, B, H1, W1))) return rc;
    if ((rc = conv_dx(n, 7, n->g_d1b, 64, n->g_d1a, 64, 0, n->d1a, 64, 0, B, H1, W1))) return rc;
    if ((rc = conv_dw(n, 6, n->c1, 192, 0, n->g_d1a, 64, B, H1, W1))) return rc;
    if ((rc = conv_dx(n, 6, n->g_d1a, 64, n->g_c1, 192, 0, n->c1, 192, 0, B, H1, W1))) return rc;
    PICO_LAUNCH(n->c, "cnn_up_bwd", k_up_bwd, grid_for(P0 / 16 * 128), 256, 0, n->g_c1, 192, 0, n->g_bb, 128, B, H1, W1);
    // bb, ba
    if ((rc = conv_dw(n, 5, n->ba, 128, 0, n->g_bb, 128, B, H2, W2))) return rc;
    if ((rc = conv_dx(n, 5, n->g_bb, 128, n->g_ba, 128, 0, n->ba, 128, 0, B, H2, W2))) return rc;
    if ((rc = conv_dw(n, 4, n->p1, 64, 0, n->g_ba, 128, B, H2, W2))) return rc;
    if ((rc = conv_dx(n, 4, n->g_ba, 128, n->g_p1, 64, 0, nullptr, 0, 0, B, H2, W2))) return rc;
    PICO_LAUNCH(n->c, "cnn_pool_bwd", k_pool_bwd_add_mask, grid_for(P0 / 4 * 64), 256, 0, n->g_p1, n->g_c1, 192, 128, n->c1,
                192, 128, n->g_e1b, 64, B, H1, W1);
    // e1b, e1a
    if ((rc = conv_dw(n, 3, n->e1a, 64, 0, n->g_e1b, 64, B, H1, W1))) return rc;
    if ((rc = conv_dx(n, 3, n->g_e1b, 64, n->g_e1a, 64, 0, n->e1a, 64, 0, B, H1, W1))) return rc;
    if ((rc = conv_dw(n, 2, n->p0, 32, 0, n->g_e1a, 64, B, H1, W1))) return rc;
    if ((rc = conv_dx(n, 2, n->g_e1a, 64, n->g_p0, 32, 0, nullptr, 0, 0, B, H1, W1))) return rc;
    PICO_LAUNCH(n->c, "cnn_pool_bwd", k_pool_bwd_add_mask, grid_for(P0 * 32), 256, 0, n->g_p0, n->g_c0, 96, 64, n->c0, 96, 64,
                n->g_e0b, 32, B, Hp, Wp);
    // e0b, e0a
    if ((rc = conv_dw(n, 1, n->e0a, 32, 0, n->g_e0b, 32, B, Hp, Wp))) return rc;
    if ((rc = conv_dx(n, 1, n->g_e0b, 32, n->g_e0a, 32, 0, n->e0a, 32, 0, B, Hp, Wp))) return rc;
    if ((rc = conv_dw(n, 0, n->x0, 16, 0, n->g_e0a, 32, B, Hp, Wp))) return rc;
    return PICO_OK;
}

}  // namespace

namespace pico {
bool cnn_belongs_to(const pico_cnn* n, const pico_hip_ctx* c) { return n && n->c == c; }
int cnn_check_size(pico_cnn* n, uint32_t w, uint32_t h) { return check_dims(n, 1, w, h); }
}  // namespace pico

extern "C" {

uint32_t pico_hip_cnn_num_params(void) {
    size_t s = 0;
    for (auto& L : kL) s += (size_t)L.taps * L.cin * L.cout + L.cout;
    return (uint32_t)s;
}

int pico_hip_cnn_create(pico_hip_ctx* c, uint32_t mw, uint32_t mh, uint32_t mb, pico_cnn** out) {
    if (!c || !out) return PICO_ERR_ARG;
    if (mw == 0 || mh == 0 || mb == 0) return fail(c, PICO_ERR_ARG, "sizes must be positive");
    if ((uint64_t)mw * mh * mb > (1ull << 27)) return fail(c, PICO_ERR_LIMIT, "at most 2^27 pixels per batch");
    auto* n = new pico_cnn();
    n->c = c;
    n->maxW = (int)mw; n->maxH = (int)mh; n->maxB = (int)mb;
    n->Hp = ((int)mh + 3) / 4 * 4; n->Wp = ((int)mw + 3) / 4 * 4;
    size_t o = 0;
    for (int l = 0; l < 11; ++l) {
        n->off[l][0] = o;
        o += (size_t)kL[l].taps * kL[l].cin * kL[l].cout;
        n->off[l][1] = o;
        o += kL[l].cout;
    }
    n->nparams = o;
    const size_t P0 = (size_t)mb * n->Hp * n->Wp, P1 = P0 / 4, P2 = P0 / 16;
    const size_t n3 = (size_t)mb * mw * mh * 3;
    bool ok = (n->params = dalloc<float>(n, o)) && (n->grads = dalloc<float>(n, o)) && (n->m = dalloc<float>(n, o)) &&
              (n->v = dalloc<float>(n, o));
    for (int l = 0; l < 11 && ok; ++l) {
        const size_t wsz = (size_t)kL[l].taps * kL[l].kpad * kL[l].npad;
        ok = (n->wf[l] = dalloc<bf16>(n, wsz)) && (n->wb[l] = dalloc<bf16>(n, wsz)) && (n->bpad[l] = dalloc<float>(n, kL[l].npad));
    }
    ok = ok && (n->x0 = dalloc<bf16>(n, P0 * 16)) && (n->e0a = dalloc<bf16>(n, P0 * 32)) && (n->c0 = dalloc<bf16>(n, P0 * 96)) &&
         (n->p0 = dalloc<bf16>(n, P1 * 32)) && (n->e1a = dalloc<bf16>(n, P1 * 64)) && (n->c1 = dalloc<bf16>(n, P1 * 192)) &&
         (n->p1 = dalloc<bf16>(n, P2 * 64)) && (n->ba = dalloc<bf16>(n, P2 * 128)) && (n->bb = dalloc<bf16>(n, P2 * 128)) &&
         (n->d1a = dalloc<bf16>(n, P1 * 64)) && (n->d1b = dalloc<bf16>(n, P1 * 64)) && (n->d0a = dalloc<bf16>(n, P0 * 32)) &&
         (n->d0b = dalloc<bf16>(n, P0 * 32)) && (n->head = dalloc<float>(n, P0 * 32)) && (n->wk = dalloc<float>(n, P0 * 25)) &&
         (n->out = dalloc<float>(n, n3)) && (n->dout = dalloc<float>(n, n3)) && (n->loss = dalloc<double>(n, 1)) &&
         (n->gnorm2 = dalloc<double>(n, 1)) &&
         (n->g_head = dalloc<bf16>(n, P0 * 32)) && (n->g_d0b = dalloc<bf16>(n, P0 * 32)) && (n->g_d0a = dalloc<bf16>(n, P0 * 32)) &&
         (n->g_c0 = dalloc<bf16>(n, P0 * 96)) && (n->g_d1b = dalloc<bf16>(n, P1 * 64)) && (n->g_d1a = dalloc<bf16>(n, P1 * 64)) &&
         (n->g_c1 = dalloc<bf16>(n, P1 * 192)) && (n->g_bb = dalloc<bf16>(n, P2 * 128)) && (n->g_ba = dalloc<bf16>(n, P2 * 128)) &&
         (n->g_p1 = dalloc<bf16>(n, P2 * 64)) && (n->g_e1b = dalloc<bf16>(n, P1 * 64)) && (n->g_e1a = dalloc<bf16>(n, P1 * 64)) &&
         (n->g_p0 = dalloc<bf16>(n, P1 * 32)) && (n->g_e0b = dalloc<bf16>(n, P0 * 32)) && (n->g_e0a = dalloc<bf16>(n, P0 * 32));
    if (!ok) {
        pico_hip_cnn_destroy(n);
        return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    }
    c->cnns.insert(n);
    *out = n;
    return pico_hip_cnn_init(n, 1);
}

void pico_hip_cnn_destroy(pico_cnn* n) {
    if (!n) return;
    hipStreamSynchronize(n->c->stream);
    for (void* p : n->allocs) hipFree(p);
    n->c->cnns.erase(n);
    delete n;
}

// He-normal initialisation (std sqrt(2 / fan_in)), zero biases; the head
// starts small so the first kernels are near-uniform 5x5 averages.
int pico_hip_cnn_init(pico_cnn* n, uint32_t seed) {
    if (!n) return PICO_ERR_ARG;
    std::vector<float> p(n->nparams, 0.0f);
    uint64_t s = mix64(seed ^ 0xC0FFEEull);
    auto u01 = [&]() { s = mix64(s); return ((s >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
    for (int l = 0; l < 11; ++l) {
        const double sd = l == 10 ? 0.01 : std::sqrt(2.0 / (kL[l].taps * kL[l].cin));
        for (size_t i = n->off[l][0]; i < n->off[l][1]; ++i)
            p[i] = (float)(sd * std::sqrt(-2.0 * std::log(u01())) * std::cos(6.283185307179586 * u01()));
    }
    n->step = 0;
    int rc = pico_hip_cnn_set_params(n, p.data());
    if (rc) return rc;
    PICO_HIP(n->c, hipMemsetAsync(n->m, 0, n->nparams * 4, n->c->stream));
    PICO_HIP(n->c, hipMemsetAsync(n->v, 0, n->nparams * 4, n->c->stream));
    PICO_HIP(n->c, hipStreamSynchronize(n->c->stream));
    return PICO_OK;
}

int pico_hip_cnn_set_params(pico_cnn* n, const float* host) {
    if (!n || !host) return PICO_ERR_ARG;
    PICO_HIP(n->c, hipMemcpyAsync(n->params, host, n->nparams * 4, hipMemcpyHostToDevice, n->c->stream));
    int rc = relayout(n);
    if (rc) return rc;
    PICO_HIP(n->c, hipStreamSynchronize(n->c->stream));
    return PICO_OK;
}

int pico_hip_cnn_get_params(pico_cnn* n, float* host) {
    if (!n || !host) return PICO_ERR_ARG;
    PICO_HIP(n->c, hipMemcpyAsync(host, n->params, n->nparams * 4, hipMemcpyDeviceToHost, n->c->stream));
    PICO_HIP(n->c, hipStreamSynchronize(n->c->stream));
    return PICO_OK;
}

int pico_hip_cnn_denoise_d(pico_cnn* n, const float* d_in, uint32_t w, uint32_t h, float* d_out) {
    if (!n) return PICO_ERR_ARG;
    if (!d_in || !d_out) return fail(n->c, PICO_ERR_ARG, "NULL buffer");
    int rc = check_dims(n, 1, w, h);
    if (rc) return rc;
    // the output is written in place unless it overlaps the radiance planes
    // the kernel apply reads
    const size_t px = (size_t)w * h;
    const bool overlap = (const char*)d_out < (const char*)(d_in + 3 * px) && (const char*)d_in < (const char*)(d_out + 3 * px);
    if ((rc = forward(n, d_in, 1, (int)h, (int)w, false, false, overlap ? nullptr : d_out))) return rc;
    if (overlap) PICO_HIP(n->c, hipMemcpyAsync(d_out, n->out, px * 12, hipMemcpyDeviceToDevice, n->c->stream));
    return PICO_OK;
}

int pico_hip_cnn_denoise(pico_cnn* n, const float* in, uint32_t w, uint32_t h, float* out) {
    if (!n) return PICO_ERR_ARG;
    if (!in || !out) return fail(n->c, PICO_ERR_ARG, "NULL buffer");
    int rc = check_dims(n, 1, w, h);
    if (rc) return rc;
    const size_t px = (size_t)w * h;
    float* d = (float*)scratch(n->c, "cnn_io", px * 52 + 64);
    if (!d) return fail(n->c, PICO_ERR_NOMEM, "device allocation failed");
    PICO_HIP(n->c, hipMemcpyAsync(d, in, px * 40, hipMemcpyHostToDevice, n->c->stream));
    if ((rc = pico_hip_cnn_denoise_d(n, d, w, h, d + 10 * px))) return rc;
    PICO_HIP(n->c, hipMemcpyAsync(out, d + 10 * px, px * 12, hipMemcpyDeviceToHost, n->c->stream));
    PICO_HIP(n->c, hipStreamSynchronize(n->c->stream));
    return PICO_OK;
}

int pico_hip_cnn_inputs_d(pico_hip_ctx* c, const float* d_rgb, const float* d_normal, const float* d_position,
                          const float* d_albedo, uint32_t w, uint32_t h, float* d_inputs) {
    if (!c) return PICO_ERR_ARG;
    if (!d_rgb || !d_normal || !d_position || !d_albedo || !d_inputs) return fail(c, PICO_ERR_ARG, "NULL buffer");
    const size_t px = (size_t)w * h;
    if (px) PICO_LAUNCH(c, "cnn_inputs", k_inputs, grid_for(px), 256, 0, d_rgb, d_normal, d_position, d_albedo, d_inputs, px);
    return PICO_OK;
}

int pico_hip_cnn_denoise_gbuffer_d(pico_cnn* n, const float* d_rgb, const float* d_normal, const float* d_position,
                                   const float* d_albedo, uint32_t w, uint32_t h, float* d_out) {
    if (!n) return PICO_ERR_ARG;
    int rc = check_dims(n, 1, w, h);
    if (rc) return rc;
    if (!d_rgb || !d_normal || !d_position || !d_albedo || !d_out) return fail(n->c, PICO_ERR_ARG, "NULL buffer");
    const size_t px = (size_t)w * h;
    // the packed input straight from the G-buffer (the planar input's
    // radiance planes for the kernel apply), then the network
    float* d = (float*)scratch(n->c, "cnn_frame_in", px * 12 + 64);
    if (!d) return fail(n->c, PICO_ERR_NOMEM, "device allocation failed");
    const int Hp = ((int)h + 3) / 4 * 4, Wp = ((int)w + 3) / 4 * 4;
    PICO_LAUNCH(n->c, "cnn_inputs_pack", k_inputs_pack, grid_for((size_t)Hp * Wp), 256, 0, d_rgb, d_normal, d_position,
                d_albedo, d, n->x0, (int)h, (int)w, Hp, Wp);
    const bool overlap = (const char*)d_out < (const char*)(d + 3 * px) && (const char*)d < (const char*)(d_out + 3 * px);
    if ((rc = forward(n, d, 1, (int)h, (int)w, false, true, overlap ? nullptr : d_out))) return rc;
    if (overlap) PICO_HIP(n->c, hipMemcpyAsync(d_out, n->out, px * 12, hipMemcpyDeviceToDevice, n->c->stream));
    return PICO_OK;
}

int pico_hip_cnn_denoise_gbuffer(pico_cnn* n, const float* rgb, const float* normal, const float* position,
                                 const float* albedo, uint32_t w, uint32_t h, float* out) {
    if (!n) return PICO_ERR_ARG;
    if (!rgb || !normal || !position || !albedo || !out) return fail(n->c, PICO_ERR_ARG, "NULL buffer");
    int rc = check_dims(n, 1, w, h);
    if (rc) return rc;
    const size_t px = (size_t)w * h;
    float* d = (float*)scratch(n->c, "cnn_frame_io", px * 60 + 64);
    if (!d) return fail(n->c, PICO_ERR_NOMEM, "device allocation failed");
    const float* src[4] = {rgb, normal, position, albedo};
    for (int k = 0; k < 4; ++k)
        PICO_HIP(n->c, hipMemcpyAsync(d + 3 * k * px, src[k], px * 12, hipMemcpyHostToDevice, n->c->stream));
    if ((rc = pico_hip_cnn_denoise_gbuffer_d(n, d, d + 3 * px, d + 6 * px, d + 9 * px, w, h, d + 12 * px))) return rc;
    PICO_HIP(n->c, hipMemcpyAsync(out, d + 12 * px, px * 12, hipMemcpyDeviceToHost, n->c->stream));
    PICO_HIP(n->c, hipStreamSynchronize(n->c->stream));
    return PICO_OK;
}

int pico_hip_denoise_gbuffer_d(pico_hip_ctx* c, pico_cnn* n, const float* d_rgb, const float* d_normal,
                               const float* d_position, const float* d_albedo, uint32_t w, uint32_t h, uint32_t tile_w,
                               uint32_t tile_h, float* d_scratch, float* d_out, uint32_t* d_fallback) {
    if (!c) return PICO_ERR_ARG;
    if (n && !cnn_belongs_to(n, c)) return fail(c, PICO_ERR_ARG, "network created on another context");
    if (!d_rgb || !d_normal || !d_position || !d_albedo || !d_scratch || !d_out || !d_fallback)
        return fail(c, PICO_ERR_ARG, "NULL buffer");
    if (w == 0 || h == 0 || tile_w == 0 || tile_h == 0) return fail(c, PICO_ERR_ARG, "sizes must be positive");
    const size_t n3 = (size_t)w * h * 3;
    // the network, unless there is none or it refuses the call (its size
    // limit); then the a-trous filter unconditionally
    const int rc = n ? pico_hip_cnn_denoise_gbuffer_d(n, d_rgb, d_normal, d_position, d_albedo, w, h, d_out) : PICO_ERR_STATE;
    if (rc) {
        const std::string why = n ? c->err : std::string("no network");
        PICO_LAUNCH(c, "denoise_flag", k_flag_set, 1, 1, 0, d_fallback, 1u);
        const int rc2 = atrous_gated(c, d_rgb, d_normal, d_position, d_albedo, w, h, tile_w, tile_h, 5, d_scratch, d_out,
                                     nullptr);
        if (rc2) return rc2;
        c->err = "learned denoiser not used (" + why + "): a-trous output";
        return PICO_OK;
    }
    // a non-finite value anywhere in the network's output: the a-trous filter
    // replaces the whole image (kernels gated on the device flag)
    PICO_LAUNCH(c, "denoise_flag", k_flag_set, 1, 1, 0, d_fallback, 0u);
    PICO_LAUNCH(c, "denoise_check", k_flag_nonfinite, grid_for(n3), 256, 0, d_out, n3, d_fallback);
    return atrous_gated(c, d_rgb, d_normal, d_position, d_albedo, w, h, tile_w, tile_h, 5, d_scratch, d_out, d_fallback);
}

int pico_hip_denoise_gbuffer(pico_hip_ctx* c, pico_cnn* n, const float* rgb, const float* normal, const float* position,
                             const float* albedo, uint32_t w, uint32_t h, uint32_t tile_w, uint32_t tile_h, float* out,
                             int32_t* used_fallback) {
    if (!c) return PICO_ERR_ARG;
    if (!rgb || !normal || !position || !albedo || !out) return fail(c, PICO_ERR_ARG, "NULL buffer");
    if ((uint64_t)w * h > (1ull << 28)) return fail(c, PICO_ERR_LIMIT, "image larger than 2^28 pixels");
    const size_t px = (size_t)w * h;
    float* d = (float*)scratch(c, "denoise_io", px * 72 + 64);
    if (!d) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    uint32_t* flag = (uint32_t*)(d + 18 * px);
    const float* src[4] = {rgb, normal, position, albedo};
    for (int k = 0; k < 4; ++k)
        PICO_HIP(c, hipMemcpyAsync(d + 3 * k * px, src[k], px * 12, hipMemcpyHostToDevice, c->stream));
    int rc = pico_hip_denoise_gbuffer_d(c, n, d, d + 3 * px, d + 6 * px, d + 9 * px, w, h, tile_w, tile_h, d + 12 * px,
                                        d + 15 * px, flag);
    if (rc) return rc;
    uint32_t hf = 0;
    PICO_HIP(c, hipMemcpyAsync(out, d + 15 * px, px * 12, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(&hf, flag, 4, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    if (used_fallback) *used_fallback = hf ? 1 : 0;
    return PICO_OK;
}

int pico_hip_cnn_loss_grad(pico_cnn* n, const float* in, const float* tgt, uint32_t B, uint32_t w, uint32_t h, double* loss,
                           float* grads) {
    if (!n) return PICO_ERR_ARG;
    if (!in || !tgt) return fail(n->c, PICO_ERR_ARG, "NULL buffer");
    int rc = check_dims(n, B, w, h);
    if (rc) return rc;
    const size_t px = (size_t)B * w * h;
    float* d = (float*)scratch(n->c, "cnn_train_io", px * 52 + 64);
    if (!d) return fail(n->c, PICO_ERR_NOMEM, "device allocation failed");
    PICO_HIP(n->c, hipMemcpyAsync(d, in, px * 40, hipMemcpyHostToDevice, n->c->stream));
    PICO_HIP(n->c, hipMemcpyAsync(d + 10 * px, tgt, px * 12, hipMemcpyHostToDevice, n->c->stream));
    if ((rc = forward(n, d, (int)B, (int)h, (int)w, true))) return rc;
    if ((rc = backward(n, d, d + 10 * px, (int)B, (int)h, (int)w))) return rc;
    double l = 0.0;
    PICO_HIP(n->c, hipMemcpyAsync(&l, n->loss, 8, hipMemcpyDeviceToHost, n->c->stream));
    if (grads) PICO_HIP(n->c, hipMemcpyAsync(grads, n->grads, n->nparams * 4, hipMemcpyDeviceToHost, n->c->stream));
    PICO_HIP(n->c, hipStreamSynchronize(n->c->stream));
    if (loss) *loss = l / (double)(px * 3);
    return PICO_OK;
}

int pico_hip_cnn_set_clip(pico_cnn* n, float max_grad_norm) {
    if (!n) return PICO_ERR_ARG;
    if (!(max_grad_norm >= 0.0f)) return fail(n->c, PICO_ERR_ARG, "clip norm must be >= 0");
    n->clip = max_grad_norm;
    return PICO_OK;
}

int pico_hip_cnn_last_grad_norm(pico_cnn* n, double* out) {
    if (!n || !out) return PICO_ERR_ARG;
    double h = 0.0;
    PICO_HIP(n->c, hipMemcpyAsync(&h, n->gnorm2, 8, hipMemcpyDeviceToHost, n->c->stream));
    PICO_HIP(n->c, hipStreamSynchronize(n->c->stream));
    *out = std::sqrt(h);
    return PICO_OK;
}

int pico_hip_cnn_train_step_d(pico_cnn* n, const float* d_in, const float* d_tgt, uint32_t B, uint32_t w, uint32_t h,
                              float lr, double* loss) {
    if (!n) return PICO_ERR_ARG;
    if (!d_in || !d_tgt) return fail(n->c, PICO_ERR_ARG, "NULL buffer");
    int rc = check_dims(n, B, w, h);
    if (rc) return rc;
    if ((rc = forward(n, d_in, (int)B, (int)h, (int)w, true))) return rc;
    if ((rc = backward(n, d_in, d_tgt, (int)B, (int)h, (int)w))) return rc;
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)n->step), c2 = 1.0f - std::pow(b2, (float)n->step);
    PICO_LAUNCH(n->c, "cnn_zero", k_zero1, 1, 1, 0, n->gnorm2);
    PICO_LAUNCH(n->c, "cnn_sumsq", k_sumsq, std::min<uint32_t>(grid_for(n->nparams), 512), 256, 0, n->grads, (int)n->nparams, n->gnorm2);
    PICO_LAUNCH(n->c, "cnn_adam", k_adam, grid_for(n->nparams), 256, 0, n->params, n->grads, n->m, n->v, (int)n->nparams, lr,
                b1, b2, c1, c2, 1e-8f, n->gnorm2, n->clip);
    if ((rc = relayout(n))) return rc;
    if (loss) {
        double l = 0.0;
        PICO_HIP(n->c, hipMemcpyAsync(&l, n->loss, 8, hipMemcpyDeviceToHost, n->c->stream));
        PICO_HIP(n->c, hipStreamSynchronize(n->c->stream));
        *loss = l / ((double)B * w * h * 3);
    }
    return PICO_OK;
}

}  // extern "C"
