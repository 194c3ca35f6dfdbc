// Wavefront path tracer: the MI355X replacement of the reference's tiled
// thread pool, per-pixel sample loop and recursive integrator
// (Source/Core/Scene.cpp:143-201, Source/Render/Integrators.cpp:46-209).
//
// A render runs in chunks.  A chunk takes every still-active (view, pixel)
// item and S samples of it, i.e. P = n_active * S paths, and runs them
// through persistent kernels that pull work from device-resident queues:
//   generate  camera rays, per-(pixel, sample) PCG32 streams   -> extend queue 0
//   trace     closest hits (TLAS -> BLAS) of the extension rays of bounce b
//             -> hit records, and of the shadow rays of bounce b - 1 -> a light
//             hit adds radiance
//   shade     emission, BSDF sample, NEE light sample, RR       -> shadow queue b, extend queue b+1
//   finalize  (views that return their samples) per-path radiance (NaN -> pink)
//   accumulate  the reference's running mean / variance and early-out, in
//             sample order, from each path's radiance (NaN -> pink)
//   compact   surviving items -> next chunk's active list
// Every count lives in device memory; the host never synchronises inside a
// render.  Queue appends are wave-aggregated (ballot + popcount, one atomic per
// wave and queue); persistent waves fetch 64 or more queue entries per atomic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/pico_hip.h"
#include "ctx.h"

using namespace pico;

namespace pico {
DCam make_dcam(const pico_camera& c, uint32_t w, uint32_t h);
int check_camera(pico_hip_ctx* c, const pico_camera* cam, uint32_t w, uint32_t h);
}

namespace {

constexpr uint32_t kMaxViews = 64;
constexpr uint32_t kViewShift = 26;
constexpr uint32_t kPixMask = (1u << kViewShift) - 1u;
constexpr uint32_t kMaxBounce = 64;
constexpr int kBlock = 128;

enum : uint32_t { MODE_ACCUM = 0, MODE_SAMPLES = 1, MODE_SUM64 = 2 };

struct DView {
    DCam cam;
    uint32_t spp, max_depth, frame_seed, view_id;
    float max_variance;
    uint32_t mode, resume, sample_base;
    float* pixels;
    uint32_t* counts;
    float* variance;
    double* sum64;
    float* samples_out;
    uint32_t* count0;   // per-pixel sample count at render start
    uint8_t* done;      // per-pixel
    unsigned long long* stats;  // kCtrSlots counter slots of (rays, samples)
};

// Counter slots: atomics on one address serialise, so counters that every
// workgroup or wave of a grid adds to at the end are spread over kCtrSlots
// slots kCtrStride 64-bit words (256 B) apart and summed by their reader.
constexpr uint32_t kCtrSlots = 16;
// chunks of a render whose hand-over point is recorded for later renders
constexpr uint32_t kLogChunks = 64;
constexpr uint32_t kCtrStride = 32;
constexpr uint32_t kNoFinish = 0xffffffffu;

struct Ctl {
    uint32_t n_active;
    uint32_t n_next;
    uint32_t gen_head;
    uint32_t overflow;
    uint32_t ext_count[kMaxBounce + 2];
    uint32_t ext_head[kMaxBounce + 2];
    uint32_t shade_head[kMaxBounce + 2];
    uint32_t sh_count[kMaxBounce + 2];
    unsigned long long rays_cam, rays_ext, rays_sh, rays_acc, samples_acc;
    unsigned long long nodes, tris, insts;           // extend (camera + extension rays)
    unsigned long long nodes_sh, tris_sh, insts_sh;  // shadow rays
    unsigned long long spill_push, spill_pop;        // traversal-stack entries past LDS (both kernels)
    unsigned long long nodes_fin, tris_fin, insts_fin;  // k_finish's share of the node / triangle / instance counts
    unsigned long long rays_fin;                        // rays traced by k_finish, every chunk
    uint32_t finish_from;                            // first bounce left to k_finish, or kNoFinish
    // whether this chunk's hand-over, forced by the launch count, took
    // kFinishPaths or more paths; and per chunk (the first kLogChunks) its
    // finish_from and that flag, which the host reads after the render
    uint32_t forced_big;
    uint32_t chunk_finish[kLogChunks], chunk_forced[kLogChunks];
    // k_finish's extension rays, shadow rays (both: those no queue count
    // holds) and every ray it traced: words 0, 1, 2 of a counter slot
    unsigned long long fin_rays[kCtrSlots * kCtrStride];
};

struct Pool {
    float4* ray_o;   // origin, w: 1 when the camera ray missed
    float4* ray_d;
    float4* thr;
    float4* rad;     // radiance, w: rays traced by the path
    uint4* rng;      // hs state, ms state
    uint32_t* key;   // (view << 26) | pixel
    float4* hit;     // t, u, v, instance bits, by extension-queue position
    uint32_t* hprim;
    float4* sh_o;
    float4* sh_d;
    float4* sh_c;    // (throughput * direct pdf) * BSDF sample energy; w: 1 towards the sun
    uint32_t* q0;
    uint32_t* q1;
    uint32_t* shq;
    float4* ior;     // index-of-refraction stack ring; its count is in thr.w (bits)
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

constexpr uint32_t kDeadPath = 0xffffffffu;

// Work distribution of the persistent kernels over the n entries of a queue,
// `chunk` entries per claim.  Device-scope atomics on one address serialise
// (about 11.5 ns each on MI355X), so a claim per wave at the start of a launch
// kept a grid of 6144 waves some 70 us on its counter before the last wave
// began, and another 70 us at the end, however short the queue.  Here a wave's
// first batch is fixed by its index (entries [w chunk, (w + 1) chunk)) and
// later batches are claimed with an atomic past the first batches of all
// waves; a wave whose fixed batch lies past the end of the queue stops without
// touching the counter, so a short queue costs atomics only from the waves
// that had work.  (Checking the counter with an atomic load before claiming,
// to spare the final claims, made the launches three times slower: such a
// load bypasses L2 and waits behind the atomics on the same address.)
// claim() returns the batch's first entry, n or more when none is left;
// wave-uniform.  Every lane of the wave calls it.
struct WaveClaim {
    uint32_t* head;
    uint32_t n, chunk, first;  // first: the wave's fixed batch, kDeadPath once taken
    __device__ __forceinline__ WaveClaim(uint32_t* head_, uint32_t n_, uint32_t chunk_)
        : head(head_), n(n_), chunk(chunk_), first((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * chunk_) {}
    __device__ __forceinline__ uint32_t claim() {
        if (first != kDeadPath) {
            const uint32_t b = first;
            first = kDeadPath;
            return b;
        }
        const uint32_t past = gridDim.x * (blockDim.x >> 6) * chunk;
        uint32_t base = n;
        if (lane_id() == 0 && past < n) base = past + atomicAdd(head, chunk);
        return __shfl(base, 0);
    }
};

// Workgroup-aggregated, order-preserving append of R rounds of items with one
// atomic per workgroup and call: bit r of `bits` says whether this thread's
// item of round r is kept; pos[r] receives its slot (valid where kept).  Slots
// follow round, then wave, then lane order.  Every thread of the workgroup
// must call it.  (One atomic per round of a 4096-workgroup grid left a
// compaction waiting on its counter: same-address atomics serialise.)
template <int NW, int R>
__device__ __forceinline__ void block_append(uint32_t* counter, uint32_t bits, uint32_t (&pos)[R]) {
    __shared__ uint32_t cnt[R][NW];
    __shared__ uint32_t wbase;
    const uint32_t wave = threadIdx.x >> 6;
    uint64_t m[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        m[r] = __ballot((bits >> r) & 1u);
        if (lane_id() == 0) cnt[r][wave] = (uint32_t)__popcll(m[r]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int r = 0; r < R; ++r)
            for (int w = 0; w < NW; ++w) { const uint32_t c = cnt[r][w]; cnt[r][w] = tot; tot += c; }
        wbase = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane_id()) - 1ull;
#pragma unroll
    for (int r = 0; r < R; ++r) pos[r] = wbase + cnt[r][wave] + (uint32_t)__popcll(m[r] & below);
    __syncthreads();  // the next call reuses cnt / wbase
}

__global__ void k_reset_chunk(Ctl* ctl, uint32_t first, uint32_t S) {
    const uint32_t t = threadIdx.x;
    __syncthreads();
    const uint32_t na = first ? ctl->n_active : ctl->n_next;
    __syncthreads();
    for (uint32_t b = t; b < kMaxBounce + 2; b += blockDim.x) {
        ctl->ext_count[b] = 0; ctl->ext_head[b] = 0; ctl->shade_head[b] = 0;
        ctl->sh_count[b] = 0;
    }
    __syncthreads();
    if (t == 0) {
        ctl->n_active = na;
        ctl->n_next = 0;
        ctl->gen_head = 0;
        ctl->finish_from = kNoFinish;
        ctl->forced_big = 0;
        // bounce 0 uses identity slots: path i sits at queue entry i (dead
        // paths of views with fewer samples are marked kDeadPath)
        ctl->ext_count[0] = na * S;
    }
}

// Per-view set-up: sample-count snapshot, cleared outputs (unless resuming),
// done flags, and the initial active list.  The active list enumerates the
// pixels in tiles of tw x th (row-major tile order, row-major inside a tile),
// which decides which pixels neighbouring lanes trace; tw = width, th = 1 is
// plain image order.
constexpr int kInitRounds = 8;

__global__ __launch_bounds__(256) void k_init_view(DView* views, uint32_t vi, uint32_t* active, Ctl* ctl, uint32_t tw, uint32_t th) {
    const DView& V = views[vi];
    const uint32_t w = V.cam.w, h = V.cam.h;
    const uint32_t tiles_x = (w + tw - 1) / tw, tiles_y = (h + th - 1) / th, tsz = tw * th;
    const uint64_t total = (uint64_t)tiles_x * tiles_y * tsz;
    const uint64_t span = (uint64_t)kInitRounds * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * span; base < total; base += (uint64_t)gridDim.x * span) {
        uint32_t bits = 0, item[kInitRounds];
#pragma unroll
        for (int r = 0; r < kInitRounds; ++r) {
            const uint64_t t = base + (uint64_t)r * blockDim.x + threadIdx.x;
            item[r] = 0;
            if (t < total) {
                const uint32_t tile = (uint32_t)(t / tsz), rr = (uint32_t)(t - (uint64_t)tile * tsz);
                const uint32_t x = (tile % tiles_x) * tw + rr % tw, y = (tile / tiles_x) * th + rr / tw;
                if (x < w && y < h) {
                    const uint32_t p = y * w + x;
                    uint32_t c0 = 0;
                    if (V.mode == MODE_ACCUM) {
                        if (V.resume) {
                            c0 = V.counts[p];
                        } else {
                            V.counts[p] = 0;
                            for (int k = 0; k < 3; ++k) { V.pixels[3 * p + k] = 0.0f; V.variance[3 * p + k] = 0.0f; }
                        }
                    }
                    V.count0[p] = c0;
                    const bool keep = V.mode != MODE_ACCUM || c0 < V.spp;
                    V.done[p] = keep ? 0 : 1;
                    if (keep) bits |= 1u << r;
                    item[r] = (vi << kViewShift) | p;
                }
            }
        }
        uint32_t pos[kInitRounds];
        block_append<4, kInitRounds>(&ctl->n_active, bits, pos);
#pragma unroll
        for (int r = 0; r < kInitRounds; ++r)
            if ((bits >> r) & 1u) active[pos[r]] = item[r];
    }
}

__global__ __launch_bounds__(kBlock) void k_generate(const DView* __restrict__ views, const uint32_t* __restrict__ active,
                                                     Ctl* ctl, uint32_t S, uint32_t off, Pool P) {
    const uint32_t na = ctl->n_active;
    const uint32_t total = na * S;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        bool live = false;
        {
            const uint32_t j = i / na, k = i - j * na;
            const uint32_t key = active[k];
            const DView& V = views[key >> kViewShift];
            const uint32_t p = key & kPixMask;
            const uint32_t rel = off + j;
            if (rel < V.spp) {
                live = true;
                const uint32_t sample = (V.mode == MODE_ACCUM ? V.count0[p] : V.sample_base) + rel;
                Rng rng;
                rng.seed(V.frame_seed, V.view_id, sample, p);
                const float2 xi = rng.next2d();
                const f3 d = camera_dir(V.cam, xi, p % V.cam.w, p / V.cam.w);
                P.ray_o[i] = f4(V.cam.pos, 0.0f);
                P.ray_d[i] = f4(d, 0.0f);
                P.rng[i] = make_uint4((uint32_t)rng.hs, (uint32_t)(rng.hs >> 32), (uint32_t)rng.ms, (uint32_t)(rng.ms >> 32));
                P.key[i] = key;
            }
        }
        P.q0[i] = live ? i : kDeadPath;
    }
}

// Traversal stacks of the persistent trace kernels: kTraceLds entries per
// lane in LDS, up to kTraceSpill more in device memory (rarely touched).
// With 16 LDS entries the register file, not LDS, bounds the resident waves;
// the kernels ask for 6 waves per SIMD.
constexpr int kTraceLds = 16;
constexpr uint32_t kTraceSpill = 48;
static_assert(kTraceLds + kTraceSpill >= 64, "trace stacks hold the 64 entries an upload admits");

// A point on a mesh light: AliasTable::sample (Source/Util/AliasTable.cpp:89-104)
// then LowerLevelMeshBVH::sample_geometry (Source/Core/LowerLevelMeshBVH.cpp:119-132).
__device__ __forceinline__ f3 mesh_light_point(const DevScene& s, const DLight& L, Rng& rng) {
    const uint32_t bucket = rng.next_u32() % L.n_tris;
    const float coin = (float)rng.next_u32() / 4294967296.0f;
    const uint32_t tri = coin <= s.alias_w[L.alias_offset + bucket] ? bucket : s.alias_i[L.alias_offset + bucket];
    const float2 xi = rng.next2d();
    const float su0 = sqrtf(xi.x);
    const float bx = 1 - su0, by = xi.y * su0;
    const float4* r = s.tshade + (size_t)kShadeRec * (L.tri_offset + tri);
    const float4 a = r[0], b = r[1], c = r[2];
    const f3 p0 = mk(a.x, a.y, a.z), p1 = mk(a.w, b.x, b.y), p2 = mk(b.z, b.w, c.x);
    return (((1.0f - bx - by) * p0) + (bx * p1)) + (by * p2);
}

// A point on a built-in shape light, sample_geometry of
// Source/Core/LowerLevelImplicitShapesBVH.cpp:57-66 (sphere: 0.5 *
// uniform_sample_sphere, Source/Core/RandUtils.cpp:43-49) and 90-110 (cube:
// a corner, a face axis and two offsets from three draws).
__device__ __forceinline__ f3 implicit_light_point(uint32_t shape, Rng& rng) {
    if (shape == PICO_SHAPE_SPHERE) {
        const float2 xi = rng.next2d();
        const float z = 1 - 2 * xi.x;
        const float r = sqrtf(smax(0.0f, 1.0f - z * z));
        const float phi = (float)(2.0 * PICO_PI_D * (double)xi.y);
        return mk(0.5f * (r * cosf(phi)), 0.5f * (r * sinf(phi)), 0.5f * z);
    }
    const float2 xi = rng.next2d(), xj = rng.next2d(), xk = rng.next2d();
    auto sgn = [](float x) { return (float)((0.0f < x) - (x < 0.0f)); };
    const f3 corner = mk(sgn(xi.x - 0.5f), sgn(xi.y - 0.5f), sgn(xj.x - 0.5f));
    uint32_t zero = (uint32_t)(xj.y * 3.0f);
    if (zero == 3) zero = 2;
    f3 off;
    off.x = zero == 0 ? 0.0f : (zero == 1 ? xk.y : xk.x);   // axis (zero + 1) % 3 gets xk.x, (zero + 2) % 3 gets xk.y
    off.y = zero == 1 ? 0.0f : (zero == 2 ? xk.y : xk.x);
    off.z = zero == 2 ? 0.0f : (zero == 0 ? xk.y : xk.x);
    return (corner * 0.5f) + (-corner * off);
}

constexpr uint32_t kTraceClaim = 64;

// Lane refill for the persistent trace kernels: a wave keeps a reservoir of
// queue entries (claimed kTraceClaim at a time) and hands them to idle lanes
// between traversal steps, so that lanes whose ray finished early trace the
// next one instead of waiting for the wave's slowest ray.  Returns the queue
// position this lane takes, or kDeadPath; wave-uniform state in res_base /
// res_left / drained.
__device__ __forceinline__ uint32_t refill_slot(bool idle, WaveClaim& wc, uint32_t& res_base, uint32_t& res_left,
                                                bool& drained) {
    if (drained) return kDeadPath;
    const uint64_t im = __ballot(idle);
    if (!im) return kDeadPath;
    if (res_left == 0) {
        const uint32_t base = wc.claim();
        if (base >= wc.n) {
            drained = true;
            return kDeadPath;
        }
        res_base = base;
        res_left = min(kTraceClaim, wc.n - base);
    }
    const uint32_t take = min((uint32_t)__popcll(im), res_left);
    const uint32_t rank = (uint32_t)__popcll(im & ((1ull << lane_id()) - 1ull));
    const uint32_t e = (idle && rank < take) ? res_base + rank : kDeadPath;
    res_base += take;
    res_left -= take;
    return e;
}

// Refill only once kRefillMin lanes of the wave are idle (or all are): a
// refill costs the wave the ray set-up (three divisions and the state)
// whichever lanes take part.  Instanced 1080p frame, Mrays/s: 1 / 4 / 8 / 16
// idle lanes: 1114 / 1135 / 1146 / 1139.
constexpr uint32_t kRefillMin = 8;
__device__ __forceinline__ bool refill_due(bool idle) {
    const uint64_t im = __ballot(idle);
    return (uint32_t)__popcll(im) >= kRefillMin || im == __ballot(true);
}

// The closest hit h of path i's shadow ray: towards the sun the point is lit
// when the ray met nothing, towards a light when the closest hit is any light
// (Integrators.cpp:61-77, 101-117); a lit point adds the ray's contribution
// times the emission to the path's radiance.
__device__ __forceinline__ void shadow_hit(const DevScene& s, const Pool& P, const Hit& h, uint32_t i) {
    const float4 tc = P.sh_c[i];
    bool lit = false;
    f3 em;
    if (tc.w != 0.0f) {
        lit = h.inst == 0xffffffffu;
        em = mk(s.sun_col[0], s.sun_col[1], s.sun_col[2]);
    } else if (h.inst != 0xffffffffu) {
        const DInstance& I = s.inst[h.inst];
        lit = I.bsdf == PICO_BSDF_LIGHT;
        if (lit) em = ld3(s.mats[I.material].emissive);
    }
    if (lit) {
        const float4 r = P.rad[i];
        const f3 add = xyz(tc) * em;
        P.rad[i] = make_float4(r.x + add.x, r.y + add.y, r.z + add.z, r.w);
    }
}

// Starts the traversal of path i's extension ray, or of its shadow ray (which
// stops short of the sampled light point).
template <int STACK>
__device__ __forceinline__ void begin_path_ray(const DevScene& s, const Pool& P, TraceState& t, TraceStack<STACK>& st,
                                               uint32_t i, bool shadow) {
    const float4 o4 = (shadow ? P.sh_o : P.ray_o)[i];
    const float4 d4 = (shadow ? P.sh_d : P.ray_d)[i];
    trace_begin(s, t, st, xyz(o4), xyz(d4), shadow ? d4.w : INFINITY);
}

// The closest-hit queries of bounce b in one launch: the extension rays of
// bounce b (entries [0, n_ext) of the combined queue) and the shadow rays that
// the shading of bounce b - 1 made (entries [n_ext, n_ext + n_sh)).  Neither
// kind reads what the other writes (an extension ray reads its path's ray and
// writes a hit record, a shadow ray reads its shadow ray and adds to its
// path's radiance), so a bounce pays one tail of slow rays instead of two.
// Extension rays come first: unbounded, they are the longer ones.
template <bool COUNT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_trace(DevScene s, Ctl* ctl, uint32_t b, Pool P, uint32_t* spill, uint32_t spill_cap) {
    if (ctl->finish_from <= b) return;
    __shared__ uint32_t stk[kTraceLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    TraceStack<kTraceLds> st{stk + threadIdx.x, kBlock, spill + gtid, gridDim.x * kBlock, spill_cap};
    const uint32_t* q = (b & 1) ? P.q1 : P.q0;
    const uint32_t n_ext = ctl->ext_count[b];
    const uint32_t n = n_ext + (b ? ctl->sh_count[b - 1] : 0u);
    TraceCounters tc{0, 0, 0, 0, 0}, tce{0, 0, 0, 0, 0}, tcs{0, 0, 0, 0, 0};
    TraceState t;
    // slot: the extension ray's queue entry (its hit record goes there), or
    // kDeadPath for a shadow ray
    uint32_t item = kDeadPath, slot = 0, res_base = 0, res_left = 0;
    bool drained = false;
    WaveClaim wc(&ctl->ext_head[b], n, kTraceClaim);
    for (;;) {
        // idle lanes take queue entries until none is idle or the queue is empty
        // (dead entries, samples a view does not need, leave the lane idle),
        // once at least kRefillMin lanes are idle
        for (;;) {
            if (!refill_due(item == kDeadPath)) break;
            const uint32_t e = refill_slot(item == kDeadPath, wc, res_base, res_left, drained);
            if (e != kDeadPath) {
                const bool sh = e >= n_ext;
                const uint32_t i = sh ? P.shq[e - n_ext] : q[e];
                if (i != kDeadPath) {
                    item = i;
                    slot = sh ? kDeadPath : e;
                    begin_path_ray(s, P, t, st, i, sh);
                }
            }
            if (drained || !__ballot(item == kDeadPath)) break;
        }
        if (!__ballot(item != kDeadPath)) break;
        if (item != kDeadPath && trace_step<kTraceLds, COUNT>(s, t, st, &tc, &ctl->overflow)) {
            if (slot != kDeadPath) {
                // hit records in queue order: the lanes of a wave hold neighbouring
                // queue entries, so their records share cache lines (the shading
                // pass reads them in the same order)
                P.hit[slot] = make_float4(t.h.t, t.h.u, t.h.v, __uint_as_float(t.h.inst));
                P.hprim[slot] = t.h.prim;
            } else {
                shadow_hit(s, P, t.h, item);
            }
            if (COUNT) {
                TraceCounters& d = slot != kDeadPath ? tce : tcs;
                d.nodes += tc.nodes; d.tris += tc.tris; d.insts += tc.insts;
                d.spill_push += tc.spill_push; d.spill_pop += tc.spill_pop;
                tc = TraceCounters{0, 0, 0, 0, 0};
            }
            item = kDeadPath;
        }
    }
    if (COUNT) {
        atomicAdd(&ctl->nodes, (unsigned long long)tce.nodes);
        atomicAdd(&ctl->tris, (unsigned long long)tce.tris);
        atomicAdd(&ctl->insts, (unsigned long long)tce.insts);
        atomicAdd(&ctl->nodes_sh, (unsigned long long)tcs.nodes);
        atomicAdd(&ctl->tris_sh, (unsigned long long)tcs.tris);
        atomicAdd(&ctl->insts_sh, (unsigned long long)tcs.insts);
        atomicAdd(&ctl->spill_push, (unsigned long long)(tce.spill_push + tcs.spill_push));
        atomicAdd(&ctl->spill_pop, (unsigned long long)(tce.spill_pop + tcs.spill_pop));
    }
}

// Shading of path i, whose extension ray of bounce b found h: emission,
// BSDF sample, next-event estimation's shadow ray, Russian roulette and the
// next extension ray (Integrators.cpp:130-209).  Returns bit 0 when the path
// continues (next extension queue) and bit 1 when it has a shadow ray.
__device__ __forceinline__ uint32_t shade_path(const DevScene& s, const DView* __restrict__ views, uint32_t b,
                                               const Pool& P, const Hit& h, uint32_t i) {
    bool cont = false, shadow = false;
    {
        const uint32_t inst = h.inst;
        const uint32_t key = P.key[i];
        const DView& V = views[key >> kViewShift];
        const f3 dir = xyz(P.ray_d[i]);
        // a camera ray's path starts with no radiance (one ray traced) and
        // unit throughput outside any transmissive object: not stored
        float4 r4 = b ? P.rad[i] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        f3 rad = xyz(r4);
        const float4 t4 = b ? P.thr[i] : make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        f3 thr = xyz(t4);
        IorStack ist;
        ist.count = __float_as_uint(t4.w);
        if (inst == 0xffffffffu) {
            const f3 sky = sky_radiance(s, dir);
            if (b == 0) {
                rad = sky;  // Integrators.cpp:152 (no NaN rule on this path)
                P.ray_o[i].w = 1.0f;
            } else {
                rad = rad + thr * sky;  // Integrators.cpp:208
            }
        } else if (b < V.max_depth) {
            const Frag f = s.inst[inst].blas_root >= kImplicitRoot ? implicit_frag(s, h, xyz(P.ray_o[i]), dir)
                                                                   : make_frag(s, h);
            const DInstance& I = s.inst[inst];
            const uint32_t kind = I.bsdf;
            const DMaterial& M = s.mats[I.material];
            const uint4 rs = P.rng[i];
            Rng rng;
            rng.set_streams(key & kPixMask);
            rng.hs = (uint64_t)rs.x | ((uint64_t)rs.y << 32);
            rng.ms = (uint64_t)rs.z | ((uint64_t)rs.w << 32);
            const bool btdf = kind == PICO_BSDF_TRANSPARENT || kind == PICO_BSDF_FRESNEL;
            BSample smp;
            if (btdf) {
                const float4 iv = P.ior[i];
                ist.v[0] = iv.x; ist.v[1] = iv.y; ist.v[2] = iv.z; ist.v[3] = iv.w;
                smp = kind == PICO_BSDF_TRANSPARENT ? btdf_transparent(s.microfacet, M, f.n, dir, rng, ist)
                                                    : btdf_fresnel(s.microfacet, M, f.n, dir, rng, ist);
            } else {
                smp = bsdf_sample(kind, s.microfacet, M, f.n, dir, rng);
            }
            if (kind == PICO_BSDF_LIGHT) rad = rad + thr * smp.energy;  // Integrators.cpp:164-168
            // next-event estimation, Integrators.cpp:46-128 (none at lights and BTDFs)
            if (kind != PICO_BSDF_LIGHT && !btdf) {
                const float u = rng.uniform();
                // the sun, when there is one, is light n_lights (Integrators.cpp:55-62)
                const uint32_t count = s.n_lights + s.use_sun;
                const uint32_t pick = (uint32_t)(u * (float)count);
                if (s.use_sun && pick >= s.n_lights) {
                    // Integrators.cpp:61-77: lit when the ray towards the sun meets
                    // nothing; no facing test and no length bound
                    const f3 to_sun = mk(s.sun_to[0], s.sun_to[1], s.sun_to[2]);
                    const float pdf = direct_pdf(kind, s.microfacet, M, f.n, -dir, to_sun);
                    P.sh_o[i] = f4(f.p + 0.01f * to_sun, 0.0f);
                    P.sh_d[i] = f4(to_sun, __int_as_float(0x7f800000));
                    P.sh_c[i] = f4((thr * pdf) * smp.energy, 1.0f);
                    shadow = true;
                    r4.w += 1.0f;
                } else if (s.n_lights) {
                    // the reference only decrements its count when u * count
                    // rounds up to count; the index is clamped here
                    const uint32_t li = pick >= s.n_lights ? s.n_lights - 1 : pick;
                    const DLight L = s.lights[li];
                    const f3 sp = L.n_tris ? mesh_light_point(s, L, rng) : implicit_light_point(L.tri_offset, rng);
                    const float* LM = s.xf[L.instance].m;
                    const f3 wp = mk((LM[0] * sp.x + LM[4] * sp.y) + (LM[8] * sp.z + LM[12]),
                                     (LM[1] * sp.x + LM[5] * sp.y) + (LM[9] * sp.z + LM[13]),
                                     (LM[2] * sp.x + LM[6] * sp.y) + (LM[10] * sp.z + LM[14]));
                    const f3 to_light = normalize(wp - f.p);
                    if (!(dot(to_light, f.n) < 0.0f)) {
                        const float pdf = direct_pdf(kind, s.microfacet, M, f.n, -dir, to_light);
                        // the shadow ray's closest hit lies before the sampled
                        // light point (the light itself, or a blocker), so the
                        // query stops a margin past it
                        const f3 dl = wp - f.p;
                        P.sh_o[i] = f4(f.p + 0.01f * to_light, 0.0f);
                        P.sh_d[i] = f4(to_light, 1.01f * sqrtf(dot(dl, dl)) + 1e-3f);
                        // the contribution without the light's emission, which the
                        // shadow ray's hit supplies: Integrators.cpp:173-176 forms
                        // ((throughput pdf) L_e) energy, here ((throughput pdf)
                        // energy) L_e (the same products, rounded in another order)
                        P.sh_c[i] = f4((thr * pdf) * smp.energy, 0.0f);  // w = 0: a mesh or shape light
                        shadow = true;
                        r4.w += 1.0f;
                    }
                }
            }
            if (smp.P != 0.0f) {
                // Russian roulette then throughput update, Integrators.cpp:185-197
                const float qk = smax(smax(thr.x, thr.y), thr.z);
                if (!(rng.uniform() > qk)) {
                    thr = thr * (1.0f / qk);
                    thr = thr * (smp.P * smp.energy);
                    // Integrators.cpp:199: below the surface while inside geometry
                    P.ray_o[i] = f4(f.p + 0.01f * (ist.inside() ? -f.n : f.n), 0.0f);
                    P.ray_d[i] = f4(smp.L, 0.0f);
                    P.thr[i] = make_float4(thr.x, thr.y, thr.z, __uint_as_float(ist.count));
                    if (btdf) P.ior[i] = make_float4(ist.v[0], ist.v[1], ist.v[2], ist.v[3]);
                    cont = true;
                    r4.w += 1.0f;
                }
            }
            P.rng[i] = make_uint4((uint32_t)rng.hs, (uint32_t)(rng.hs >> 32), (uint32_t)rng.ms, (uint32_t)(rng.ms >> 32));
        }
        P.rad[i] = make_float4(rad.x, rad.y, rad.z, r4.w);
    }
    return (cont ? 1u : 0u) | (shadow ? 2u : 0u);
}

// Appends, with one atomic on `counter`, the queue entries q[base + 64 k +
// lane] whose bit k is set in the lane's `bits` (k < nb), in batch then lane
// order, with bit 31 set where bit k of `flags` is.  Every lane of the wave
// calls it.
__device__ __forceinline__ void append_batches(uint32_t* counter, uint32_t bits, uint32_t flags, uint32_t nb,
                                               uint32_t* __restrict__ out, const uint32_t* __restrict__ q,
                                               uint32_t base) {
    uint32_t tot = 0;
    for (uint32_t k = 0; k < nb; ++k) tot += (uint32_t)__popcll(__ballot((bits >> k) & 1u));
    if (tot == 0) return;
    uint32_t o = 0;
    if (lane_id() == 0) o = atomicAdd(counter, tot);
    o = __shfl(o, 0);
    const uint64_t below = (1ull << lane_id()) - 1ull;
    for (uint32_t k = 0; k < nb; ++k) {
        const uint64_t m = __ballot((bits >> k) & 1u);
        if ((bits >> k) & 1u)
            out[o + (uint32_t)__popcll(m & below)] = q[base + 64u * k + lane_id()] | (((flags >> k) & 1u) << 31);
        o += (uint32_t)__popcll(m);
    }
}

// Device-scope atomics on one address serialise (about 11.5 ns each on
// MI355X, whatever the number of waves), so with one claim and two appends per
// 64 paths the shading pass could not exceed one batch of 64 paths per 11.5 ns
// (5.4 G paths/s).  A wave claims up to kShadeBatches batches with one atomic
// and appends the survivors of all of them with one atomic per queue.  Short
// queues (the last bounces) take one batch per claim, so that their few
// entries spread over as many waves as possible.
constexpr uint32_t kShadeBatches = 4;

// Paths entering a shading pass below which the remaining bounces of the
// chunk go to k_finish.
constexpr uint32_t kFinishPaths = 1u << 20;

// `last`: the chunk's last bounce.  With fewer than kFinishPaths paths
// entering bounce b < last, or at bounce force - 1 (the last bounce the host
// launched), the pass hands the rest of the chunk over to k_finish: a
// continuing path's queue entry carries bit 31 when the path also has a
// shadow ray (k_finish traces it first), and only the shadow rays of paths
// that end here go to the shadow queue.
// Five waves per SIMD (96 VGPRs and 116 B of spills instead of 139 VGPRs at
// three): the pass waits on chains of dependent loads (queue entry, path
// state, instance, shading record, light), and more waves cover more of them;
// instanced 1080p frame, Mrays/s, 3 / 4 / 5 / 6 / 8 waves: 3106 / 3132 /
// 3169 / 3096 / 3010.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_shade(DevScene s, const DView* __restrict__ views, Ctl* ctl, uint32_t b,
                                                  uint32_t last, uint32_t force, Pool P) {
    if (ctl->finish_from <= b) return;
    const uint32_t* q = (b & 1) ? P.q1 : P.q0;
    uint32_t* qn = (b & 1) ? P.q0 : P.q1;
    const uint32_t n = ctl->ext_count[b];
    const bool forced = b + 1 >= force;
    const bool handoff = b < last && (n < kFinishPaths || forced);
    if (handoff && blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->finish_from = b + 1;
        if (n >= kFinishPaths) ctl->forced_big = 1;
    }
    const uint32_t waves = gridDim.x * (kBlock / 64);
    const uint32_t nb = n / (64u * kShadeBatches) >= 2u * waves ? kShadeBatches : 1u;
    WaveClaim wc(&ctl->shade_head[b], n, 64u * nb);
    for (;;) {
        const uint32_t base = wc.claim();
        if (base >= n) break;
        uint32_t cbits = 0, sbits = 0;
        for (uint32_t k = 0; k < nb; ++k) {
            const uint32_t e = base + 64u * k + lane_id();
            const uint32_t i = e < n ? q[e] : kDeadPath;
            if (i != kDeadPath) {
                const float4 h4 = P.hit[e];
                const Hit h{h4.x, h4.y, h4.z, __float_as_uint(h4.w), P.hprim[e]};
                const uint32_t r = shade_path(s, views, b, P, h, i);
                cbits |= (r & 1u) << k;
                sbits |= (r >> 1) << k;
            }
        }
        append_batches(&ctl->ext_count[b + 1], cbits, handoff ? sbits : 0u, nb, qn, q, base);
        append_batches(&ctl->sh_count[b], handoff ? sbits & ~cbits : sbits, 0u, nb, P.shq, q, base);
    }
}

// The bounces of a chunk from ctl->finish_from on, path by path.  Once few
// paths are left, a launch per bounce lasts as long as its slowest ray, and
// those times add up over the remaining bounces; here a lane takes a path and
// runs its remaining work in order (a pending shadow ray, then the extension
// ray, its shading point, that point's shadow ray, the next extension ray
// ...), so the chunk ends with its slowest path instead.  Work items: the
// entries of the extension queue of bounce b0 (bit 31: the path has a shadow
// ray of bounce b0 - 1 to trace first), then the shadow queue of bounce
// b0 - 1 (paths that end there).  Per path the arithmetic is that of k_trace
// and k_shade, so the output does not depend on where the hand-over happens.
enum : uint32_t { kPhExt = 1, kPhShadowThenExt = 2, kPhShadowOnly = 3 };

template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_finish(DevScene s, const DView* __restrict__ views, Ctl* ctl, Pool P,
                                                   uint32_t* spill, uint32_t spill_cap) {
    const uint32_t b0 = ctl->finish_from;
    if (b0 == kNoFinish) return;
    __shared__ uint32_t stk[kTraceLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    TraceStack<kTraceLds> st{stk + threadIdx.x, kBlock, spill + gtid, gridDim.x * kBlock, spill_cap};
    const uint32_t* q = (b0 & 1) ? P.q1 : P.q0;
    const uint32_t n_ext = ctl->ext_count[b0];
    const uint32_t n = n_ext + ctl->sh_count[b0 - 1];
    TraceCounters tc{0, 0, 0, 0, 0}, tce{0, 0, 0, 0, 0}, tcs{0, 0, 0, 0, 0};
    TraceState t;
    // bb: the bounce of the path's next extension ray
    uint32_t item = kDeadPath, bb = 0, ph = 0, res_base = 0, res_left = 0;
    uint32_t rays_ext = 0, rays_sh = 0;  // rays no queue count holds
    uint32_t traced = 0;
    bool drained = false;
    WaveClaim wc(&ctl->ext_head[b0], n, kTraceClaim);
    for (;;) {
        for (;;) {
            if (!refill_due(item == kDeadPath)) break;
            const uint32_t e = refill_slot(item == kDeadPath, wc, res_base, res_left, drained);
            if (e != kDeadPath) {
                if (e < n_ext) {
                    const uint32_t v = q[e];
                    item = v & 0x7fffffffu;
                    bb = b0;
                    ph = (v >> 31) ? kPhShadowThenExt : kPhExt;
                    rays_sh += v >> 31;
                } else {
                    item = P.shq[e - n_ext];
                    ph = kPhShadowOnly;
                }
                begin_path_ray(s, P, t, st, item, ph != kPhExt);
                ++traced;
            }
            if (drained || !__ballot(item == kDeadPath)) break;
        }
        if (!__ballot(item != kDeadPath)) break;
        if (item != kDeadPath && trace_step<kTraceLds, COUNT>(s, t, st, &tc, &ctl->overflow)) {
            if (COUNT) {
                TraceCounters& d = ph == kPhExt ? tce : tcs;
                d.nodes += tc.nodes; d.tris += tc.tris; d.insts += tc.insts;
                d.spill_push += tc.spill_push; d.spill_pop += tc.spill_pop;
                tc = TraceCounters{0, 0, 0, 0, 0};
            }
            if (ph != kPhExt) {
                shadow_hit(s, P, t.h, item);
                if (ph == kPhShadowThenExt) {
                    ph = kPhExt;
                    begin_path_ray(s, P, t, st, item, false);
                    ++traced;
                } else {
                    item = kDeadPath;
                }
            } else {
                const uint32_t r = shade_path(s, views, bb, P, t.h, item);
                ++bb;
                rays_ext += r & 1u;
                rays_sh += r >> 1;
                if (r & 2u) {
                    ph = (r & 1u) ? kPhShadowThenExt : kPhShadowOnly;
                    begin_path_ray(s, P, t, st, item, true);
                    ++traced;
                } else if (r & 1u) {
                    begin_path_ray(s, P, t, st, item, false);
                    ++traced;
                } else {
                    item = kDeadPath;
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        rays_ext += __shfl_xor(rays_ext, o);
        rays_sh += __shfl_xor(rays_sh, o);
        traced += __shfl_xor(traced, o);
    }
    if (lane_id() == 0 && traced) {
        unsigned long long* slot = ctl->fin_rays + kCtrStride * ((gtid >> 6) % kCtrSlots);
        atomicAdd(slot, (unsigned long long)rays_ext);
        atomicAdd(slot + 1, (unsigned long long)rays_sh);
        atomicAdd(slot + 2, (unsigned long long)traced);
    }
    if (COUNT) {
        atomicAdd(&ctl->nodes_fin, (unsigned long long)(tce.nodes + tcs.nodes));
        atomicAdd(&ctl->tris_fin, (unsigned long long)(tce.tris + tcs.tris));
        atomicAdd(&ctl->insts_fin, (unsigned long long)(tce.insts + tcs.insts));
        atomicAdd(&ctl->nodes, (unsigned long long)tce.nodes);
        atomicAdd(&ctl->tris, (unsigned long long)tce.tris);
        atomicAdd(&ctl->insts, (unsigned long long)tce.insts);
        atomicAdd(&ctl->nodes_sh, (unsigned long long)tcs.nodes);
        atomicAdd(&ctl->tris_sh, (unsigned long long)tcs.tris);
        atomicAdd(&ctl->insts_sh, (unsigned long long)tcs.insts);
        atomicAdd(&ctl->spill_push, (unsigned long long)(tce.spill_push + tcs.spill_push));
        atomicAdd(&ctl->spill_pop, (unsigned long long)(tce.spill_pop + tcs.spill_pop));
    }
}

// Per-workgroup (rays, samples) counters of every view in LDS: items add to
// them (one LDS atomic pair per wave when the wave is on one view), and the
// workgroup adds its non-zero totals to the views' counters once at the end,
// into counter slot (rays, samples) workgroup mod kCtrSlots of each view.

struct BlockViewStats {
    unsigned long long v[kMaxViews][2];
    __device__ __forceinline__ void init() {
        for (uint32_t i = threadIdx.x; i < 2 * kMaxViews; i += blockDim.x) (&v[0][0])[i] = 0ull;
        __syncthreads();
    }
    __device__ __forceinline__ void add(uint32_t view, unsigned long long rays, unsigned long long ns) {
        const uint64_t act = __ballot(view != 0xffffffffu);
        if (!act) return;
        const uint32_t vv = __shfl(view, (int)(__ffsll((unsigned long long)act) - 1));
        if (__ballot(view != vv && view != 0xffffffffu) == 0ull) {
            for (int o = 32; o > 0; o >>= 1) { rays += __shfl_xor(rays, o); ns += __shfl_xor(ns, o); }
            if (lane_id() == 0) { atomicAdd(&v[vv][0], rays); atomicAdd(&v[vv][1], ns); }
        } else if (view != 0xffffffffu) {
            atomicAdd(&v[view][0], rays);
            atomicAdd(&v[view][1], ns);
        }
    }
    // every workgroup of a grid flushes at about the same time: one slot per
    // view would hold the grid's last workgroups for 4096 atomics
    __device__ __forceinline__ void flush(const DView* views) {
        __syncthreads();
        const uint32_t slot = kCtrStride * (blockIdx.x % kCtrSlots);
        for (uint32_t i = threadIdx.x; i < kMaxViews; i += blockDim.x)
            if (v[i][0] | v[i][1]) {
                atomicAdd(&views[i].stats[slot], v[i][0]);
                atomicAdd(&views[i].stats[slot + 1], v[i][1]);
            }
    }
};

// A path's finished radiance and its rays traced: a camera ray that hit and
// came back NaN or infinite is shown pink (Integrators.cpp:145-146).
__device__ __forceinline__ float4 path_sample(const Pool& P, size_t i) {
    const float4 r = P.rad[i];
    f3 res = xyz(r);
    if (P.ray_o[i].w == 0.0f && bad(res)) res = mk(1.0f, 0.4f, 0.7f);
    return make_float4(res.x, res.y, res.z, r.w);
}

// Views that return their samples (MODE_SAMPLES): every sample to the
// caller's array.  Accumulating views read their samples in k_accumulate.
__global__ __launch_bounds__(kBlock) void k_finalize(const DView* __restrict__ views, const uint32_t* __restrict__ active,
                                                     const Ctl* ctl, uint32_t S, uint32_t off, Pool P) {
    __shared__ BlockViewStats bvs;
    bvs.init();
    const uint32_t na = ctl->n_active;
    const uint32_t total = na * S;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        uint32_t vstat = 0xffffffffu;
        unsigned long long rays = 0;
        if (i < total) {
            const uint32_t j = i / na, k = i - j * na;
            const uint32_t key = active[k];
            const DView& V = views[key >> kViewShift];
            const uint32_t rel = off + j;
            if (rel < V.spp && V.mode == MODE_SAMPLES) {
                const float4 x = path_sample(P, i);
                const uint32_t p = key & kPixMask;
                float* o = V.samples_out + ((size_t)rel * V.cam.w * V.cam.h + p) * 3;
                o[0] = x.x; o[1] = x.y; o[2] = x.z;
                vstat = key >> kViewShift;
                rays = (unsigned long long)x.w;
            }
        }
        bvs.add(vstat, rays, vstat != 0xffffffffu ? 1ull : 0ull);
    }
    bvs.flush(views);
}

// Running mean / variance in sample order with the early-out,
// Source/Core/Scene.cpp:173-197.
__global__ __launch_bounds__(kBlock) void k_accumulate(const DView* __restrict__ views, const uint32_t* __restrict__ active,
                                                       Ctl* ctl, uint32_t S, uint32_t off, Pool P) {
    __shared__ BlockViewStats bvs;
    bvs.init();
    const uint32_t na = ctl->n_active;
    for (uint32_t base = blockIdx.x * blockDim.x; base < na; base += gridDim.x * blockDim.x) {
        const uint32_t k = base + threadIdx.x;
        uint32_t vstat = 0xffffffffu;
        unsigned long long rays = 0, ns = 0;
        if (k < na) {
            const uint32_t key = active[k];
            const DView& V = views[key >> kViewShift];
            const uint32_t p = key & kPixMask;
            if (V.mode == MODE_SAMPLES) {
                if (off + S >= V.spp) V.done[p] = 1;
            } else {
                vstat = key >> kViewShift;
                bool stop = false;
                if (V.mode == MODE_SUM64) {
                    double sx = V.sum64[3 * p], sy = V.sum64[3 * p + 1], sz = V.sum64[3 * p + 2];
                    for (uint32_t j = 0; j < S && off + j < V.spp; ++j) {
                        const float4 x = path_sample(P, (size_t)j * na + k);
                        sx += (double)x.x; sy += (double)x.y; sz += (double)x.z;
                        rays += (unsigned long long)x.w;
                        ++ns;
                    }
                    V.sum64[3 * p] = sx; V.sum64[3 * p + 1] = sy; V.sum64[3 * p + 2] = sz;
                } else {
                    uint32_t cnt = V.counts[p];
                    f3 pix = ld3(V.pixels + 3 * p), var = ld3(V.variance + 3 * p);
                    for (uint32_t j = 0; j < S && off + j < V.spp; ++j) {
                        const float4 x4 = path_sample(P, (size_t)j * na + k);
                        f3 x = xyz(x4);
                        const uint32_t prev = cnt;
                        if (prev < 1) {
                            cnt = 1;
                            var = mk(0.0f, 0.0f, 0.0f);
                        } else {
                            const f3 np = pix + ((x - pix) * (1.0f / (float)prev));
                            var = ((var * (float)prev) + ((x - pix) * (x - np))) / (float)(prev + 1);
                            cnt = prev + 1;
                            x = np;
                        }
                        pix = x;
                        rays += (unsigned long long)x4.w;
                        ++ns;
                        if (prev > 4 && var.x <= V.max_variance && var.y <= V.max_variance && var.z <= V.max_variance) {
                            stop = true;
                            break;
                        }
                    }
                    V.counts[p] = cnt;
                    V.pixels[3 * p] = pix.x; V.pixels[3 * p + 1] = pix.y; V.pixels[3 * p + 2] = pix.z;
                    V.variance[3 * p] = var.x; V.variance[3 * p + 1] = var.y; V.variance[3 * p + 2] = var.z;
                }
                if (stop || off + S >= V.spp) V.done[p] = 1;
            }
        }
        bvs.add(vstat, rays, ns);
    }
    bvs.flush(views);
}

constexpr int kCompactRounds = 16;

__global__ __launch_bounds__(kBlock) void k_compact(const DView* __restrict__ views, const uint32_t* __restrict__ active,
                                                    uint32_t* __restrict__ next, Ctl* ctl) {
    const uint32_t na = ctl->n_active;
    const uint32_t span = (uint32_t)kCompactRounds * kBlock;
    for (uint32_t base = blockIdx.x * span; base < na; base += gridDim.x * span) {
        uint32_t bits = 0, key[kCompactRounds];
#pragma unroll
        for (int r = 0; r < kCompactRounds; ++r) {
            const uint32_t k = base + (uint32_t)r * kBlock + threadIdx.x;
            key[r] = 0;
            if (k < na) {
                key[r] = active[k];
                if (views[key[r] >> kViewShift].done[key[r] & kPixMask] == 0) bits |= 1u << r;
            }
        }
        uint32_t pos[kCompactRounds];
        block_append<kBlock / 64, kCompactRounds>(&ctl->n_next, bits, pos);
#pragma unroll
        for (int r = 0; r < kCompactRounds; ++r)
            if ((bits >> r) & 1u) next[pos[r]] = key[r];
    }
}

__global__ void k_sum_rays(Ctl* ctl, uint32_t bounces, uint32_t chunk) {
    if (threadIdx.x == 0) {
        unsigned long long e = 0, sh = 0;
        // queues of the bounces that ran as launches, and k_finish's rays
        const uint32_t last = ctl->finish_from == kNoFinish ? bounces : ctl->finish_from;
        if (chunk < kLogChunks) {
            ctl->chunk_finish[chunk] = ctl->finish_from;
            ctl->chunk_forced[chunk] = ctl->forced_big;
        }
        for (uint32_t b = 1; b <= last; ++b) e += ctl->ext_count[b];
        for (uint32_t b = 0; b < last; ++b) sh += ctl->sh_count[b];
        for (uint32_t k = 0; k < kCtrSlots; ++k) {
            e += ctl->fin_rays[kCtrStride * k];
            sh += ctl->fin_rays[kCtrStride * k + 1];
            ctl->rays_fin += ctl->fin_rays[kCtrStride * k + 2];
            ctl->fin_rays[kCtrStride * k] = 0;
            ctl->fin_rays[kCtrStride * k + 1] = 0;
            ctl->fin_rays[kCtrStride * k + 2] = 0;
        }
        ctl->rays_cam += ctl->ext_count[0];
        ctl->rays_ext += e;
        ctl->rays_sh += sh;
    }
}

struct Plan {
    uint32_t max_spp = 0, max_depth = 0, total_items = 0;
};

// Bounce b: the extension rays of bounce b with the shadow rays of bounce
// b - 1, then the shading of bounce b (which makes the extension rays of
// bounce b + 1 and the shadow rays of bounce b).  The shading of the last
// bounce makes no shadow rays.
int launch_bounce_kernels(pico_hip_ctx* c, Ctl* ctl, uint32_t b, uint32_t last, uint32_t force, const Pool& P,
                          const DView* dviews, bool count, uint32_t grid_trace, uint32_t grid_shade, uint32_t* spill,
                          uint32_t spill_cap) {
    if (count) PICO_LAUNCH(c, "trace", (k_trace<true>), grid_trace, kBlock, 0, c->dev, ctl, b, P, spill, spill_cap);
    else PICO_LAUNCH(c, "trace", (k_trace<false>), grid_trace, kBlock, 0, c->dev, ctl, b, P, spill, spill_cap);
    PICO_LAUNCH(c, "shade", k_shade, grid_shade, kBlock, 0, c->dev, dviews, ctl, b, last, force, P);
    return PICO_OK;
}

// Optional per-view follow-up of a batched launch: the learned denoiser once
// the view's last chunk is done.
struct ViewHooks {
    pico_cnn* cnn = nullptr;
    float* const* d_out = nullptr;
    const pico_view* views = nullptr;
    // progressive display of a single view: after every chunk of samples the
    // view's pixels and counts are copied to the host staging buffers and the
    // callback is run (a non-zero return stops the render as a cancel does)
    pico_pass_callback pass_cb = nullptr;
    void* pass_user = nullptr;
    float* h_pixels = nullptr;
    uint32_t* h_counts = nullptr;
};

// Runs the chunked wavefront over `nv` views already described on the host.
int run_views(pico_hip_ctx* c, std::vector<DView>& hv, const pico_tuning& tune, const volatile int32_t* cancel,
              bool count, pico_stats* stats, uint32_t n_stats, const ViewHooks* hooks = nullptr,
              hipEvent_t epoch = nullptr) {
    if (!c->has_scene) return fail(c, PICO_ERR_STATE, "no scene uploaded");
    const uint32_t nv = (uint32_t)hv.size();
    if (nv == 0 || nv > kMaxViews) return fail(c, PICO_ERR_LIMIT, "between 1 and 64 views per launch");
    Plan pl;
    size_t pix_total = 0;
    for (auto& v : hv) {
        pl.max_spp = std::max(pl.max_spp, v.spp);
        pl.max_depth = std::max(pl.max_depth, v.max_depth);
        pix_total += (size_t)v.cam.w * v.cam.h;
    }
    if (pix_total > (1ull << 31)) return fail(c, PICO_ERR_LIMIT, "too many pixels in one launch");
    if (pl.max_depth > kMaxBounce - 1) return fail(c, PICO_ERR_LIMIT, "max_depth above 63");
    if (pl.max_spp == 0) return PICO_OK;
    uint32_t cap = tune.queue_paths ? tune.queue_paths : kDefaultQueuePaths;
    cap = std::max<uint32_t>(cap, 64);
    const bool accum_mode = hv[0].mode == MODE_ACCUM;
    bool samples_mode = false;
    for (auto& v : hv) samples_mode = samples_mode || v.mode == MODE_SAMPLES;
    // Chunk schedule, fixed on the host: a chunk gives every pixel of the
    // views still short of their sample count S more samples, S = path budget
    // / those pixels (views with fewer samples drop out, so later chunks take
    // more samples at once).  The first chunk of an accumulating render ends
    // right after the sample at which the reference's early-out is first
    // tested (prev_count > 4).
    std::vector<std::pair<uint32_t, uint32_t>> sched;  // (first sample, samples)
    size_t paths = 64;
    for (uint32_t off = 0; off < pl.max_spp;) {
        size_t live = 0;
        for (auto& v : hv)
            if (v.spp > off) live += (size_t)v.cam.w * v.cam.h;
        uint32_t S = (uint32_t)std::max<size_t>(1, std::min<size_t>(cap / live, pl.max_spp - off));
        if (off == 0 && accum_mode) S = std::min<uint32_t>(S, 6);
        sched.emplace_back(off, S);
        paths = std::max(paths, (size_t)S * live);
        off += S;
    }
    // buffers
    char* pb = (char*)scratch(c, "pool", paths * (16 * 10 + 4 * 6) + 8192);
    uint32_t* act = (uint32_t*)scratch(c, "active", pix_total * 8 + 64);
    Ctl* ctl = (Ctl*)scratch(c, "ctl", sizeof(Ctl));
    uint32_t* count0 = (uint32_t*)scratch(c, "count0", pix_total * 4 + 64);
    uint8_t* done = (uint8_t*)scratch(c, "done", pix_total + 64);
    const size_t vstat_words = (size_t)kCtrSlots * kCtrStride;  // per view
    unsigned long long* vstats = (unsigned long long*)scratch(c, "vstats", nv * vstat_words * 8 + 64);
    DView* dv = (DView*)scratch(c, "views", sizeof(DView) * nv);
    if (!pb || !act || !ctl || !count0 || !done || !vstats || !dv) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    // the follow-up's buffers, sized before the first launch so that no
    // allocation happens while the two streams run
    float *gn = nullptr, *gp = nullptr, *ga = nullptr, *gs = nullptr;
    uint32_t* fallback = nullptr;
    if (hooks && hooks->cnn) {
        size_t maxpx = 0;
        for (auto& v : hv) maxpx = std::max(maxpx, (size_t)v.cam.w * v.cam.h);
        float* g = (float*)scratch(c, "views_gbuffer", maxpx * 48 + 64);
        fallback = (uint32_t*)scratch(c, "views_fallback", 4 * nv + 64);
        // the network's input, and the a-trous fallback's per-pixel arrays
        if (!g || !fallback || !scratch(c, "cnn_frame_in", maxpx * 40 + 64) || !scratch(c, "gb_ovf", 64) ||
            !scratch(c, "atrous_plen", maxpx * 20 + 128))
            return fail(c, PICO_ERR_NOMEM, "device allocation failed");
        gn = g; gp = g + 3 * maxpx; ga = g + 6 * maxpx; gs = g + 9 * maxpx;
        if (!c->side) PICO_HIP(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    }
    Pool P;
    {
        char* q = pb;
        auto take = [&](size_t bytes) { char* r = q; q += (bytes + 255) & ~(size_t)255; return r; };
        P.ray_o = (float4*)take(paths * 16); P.ray_d = (float4*)take(paths * 16);
        P.thr = (float4*)take(paths * 16); P.rad = (float4*)take(paths * 16);
        P.rng = (uint4*)take(paths * 16); P.hit = (float4*)take(paths * 16);
        P.sh_o = (float4*)take(paths * 16); P.sh_d = (float4*)take(paths * 16);
        P.sh_c = (float4*)take(paths * 16);
        P.ior = (float4*)take(paths * 16);
        P.key = (uint32_t*)take(paths * 4); P.hprim = (uint32_t*)take(paths * 4);
        P.q0 = (uint32_t*)take(paths * 4); P.q1 = (uint32_t*)take(paths * 4); P.shq = (uint32_t*)take(paths * 4);
    }
    uint32_t* act_a = act;
    uint32_t* act_b = act + pix_total;
    size_t off_px = 0;
    for (uint32_t v = 0; v < nv; ++v) {
        hv[v].count0 = count0 + off_px;
        hv[v].done = done + off_px;
        hv[v].stats = vstats + vstat_words * v;
        off_px += (size_t)hv[v].cam.w * hv[v].cam.h;
    }
    PICO_HIP(c, hipMemcpyAsync(dv, hv.data(), sizeof(DView) * nv, hipMemcpyHostToDevice, c->stream));
    PICO_HIP(c, hipMemsetAsync(ctl, 0, sizeof(Ctl), c->stream));
    PICO_HIP(c, hipMemsetAsync(vstats, 0, nv * vstat_words * 8, c->stream));
    // events: launch start/end, the end of every chunk (batched launches) and
    // each view's final output (follow-up on the second stream)
    std::vector<hipEvent_t> evs;
    auto new_event = [&](hipEvent_t* e) -> bool {
        if (hipEventCreate(e) != hipSuccess) return false;
        evs.push_back(*e);
        return true;
    };
    struct EventsGuard {
        std::vector<hipEvent_t>& v;
        ~EventsGuard() { for (hipEvent_t e : v) hipEventDestroy(e); }
    } guard{evs};
    hipEvent_t e0, e1;
    if (!new_event(&e0) || !new_event(&e1)) return fail(c, PICO_ERR_HIP, "hipEventCreate failed");
    PICO_HIP(c, hipEventRecord(e0, c->stream));
    for (uint32_t v = 0; v < nv; ++v) {
        const uint32_t n = hv[v].cam.w * hv[v].cam.h;
        uint32_t tw = tune.tile_w ? std::min(tune.tile_w, hv[v].cam.w) : hv[v].cam.w;
        uint32_t th = tune.tile_w ? std::max<uint32_t>(1, std::min(tune.tile_h, hv[v].cam.h)) : 1;
        const uint32_t g = std::min<uint32_t>((n + 256 * kInitRounds - 1) / (256 * kInitRounds), 4096);
        PICO_LAUNCH(c, "init_view", k_init_view, g, 256, 0, dv, v, act_a, ctl, tw, th);
    }
    // persistent grids: as many workgroups as stay resident at once
    int occ_t = 0, occ_s = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_t, k_trace<false>, kBlock, 0);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_s, k_shade, kBlock, 0);
    const uint32_t grid_trace = (uint32_t)c->num_cus * (uint32_t)std::max(1, std::min(occ_t, 16));
    // stack entries beyond the LDS part, per trace lane, in device memory
    const uint32_t spill_cap = c->stack_need > (uint32_t)kTraceLds ? std::min(c->stack_need - kTraceLds, kTraceSpill) : 0;
    uint32_t* spill = (uint32_t*)scratch(c, "trace_spill", (size_t)grid_trace * kBlock * kTraceSpill * 4 + 64);
    if (!spill) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    const uint32_t grid_shade = (uint32_t)c->num_cus * (uint32_t)std::max(1, std::min(occ_s, 16));
    int occ_f = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_f, k_finish<false>, kBlock, 0);
    const uint32_t grid_finish = std::min(grid_trace, (uint32_t)c->num_cus * (uint32_t)std::max(1, std::min(occ_f, 16)));
    const uint32_t grid_lin = (uint32_t)c->num_cus * 16;
    // per-view completion events: several views, a follow-up per view, or a
    // group of a shortest-first batch (its times count from the batch's epoch)
    const bool per_view_times = n_stats > 1 || hooks || (epoch && n_stats > 0);
    std::vector<hipEvent_t> view_render(nv, nullptr), view_done(nv, nullptr);
    // Bounce launches of a chunk: up to the hand-over point that an earlier
    // render's chunk of the same shape (live pixels, first sample, samples,
    // depth) reached, with the hand-over forced at the last of them (the
    // launches after a hand-over would only return); every bounce for a shape
    // not seen yet, or whose forced hand-over took kFinishPaths or more paths.
    // The output does not depend on where the hand-over happens.
    std::vector<FinishKey> chunk_key;
    std::vector<uint32_t> chunk_force;
    bool first = true;
    bool cancelled = false;
    for (const auto& chunk : sched) {
        const uint32_t off = chunk.first, S = chunk.second;
        size_t live = 0;
        for (auto& v : hv)
            if (v.spp > off) live += (size_t)v.cam.w * v.cam.h;
        const FinishKey key{(uint64_t)live, off, S, pl.max_depth};
        const auto hit = c->finish_hints.find(key);
        const uint32_t force = hit != c->finish_hints.end() ? hit->second : kNoFinish;
        const uint32_t n_launch = force == kNoFinish ? pl.max_depth + 1 : force;
        const uint32_t ci = (uint32_t)chunk_key.size();
        chunk_key.push_back(key);
        chunk_force.push_back(force);
        if (cancel) {
            // the flag is host memory written by another thread: look at it once
            // the previous chunk has finished on the device
            if (off) PICO_HIP(c, hipStreamSynchronize(c->stream));
            if (*cancel) { cancelled = true; break; }
        }
        PICO_LAUNCH(c, "reset_chunk", k_reset_chunk, 1, 256, 0, ctl, first ? 1u : 0u, S);
        first = false;
        PICO_LAUNCH(c, "generate", k_generate, grid_lin, kBlock, 0, dv, act_a, ctl, S, off, P);
        for (uint32_t b = 0; b < n_launch; ++b) {
            int rc;
            rc = launch_bounce_kernels(c, ctl, b, pl.max_depth, force, P, dv, count, grid_trace, grid_shade, spill,
                                       spill_cap);
            if (rc) return rc;
        }
        if (count) PICO_LAUNCH(c, "finish", (k_finish<true>), grid_finish, kBlock, 0, c->dev, dv, ctl, P, spill, spill_cap);
        else PICO_LAUNCH(c, "finish", (k_finish<false>), grid_finish, kBlock, 0, c->dev, dv, ctl, P, spill, spill_cap);
        PICO_LAUNCH(c, "sum_rays", k_sum_rays, 1, 64, 0, ctl, pl.max_depth, ci);
        if (samples_mode) PICO_LAUNCH(c, "finalize", k_finalize, grid_lin, kBlock, 0, dv, act_a, ctl, S, off, P);
        PICO_LAUNCH(c, "accumulate", k_accumulate, grid_lin, kBlock, 0, dv, act_a, ctl, S, off, P);
        PICO_LAUNCH(c, "compact", k_compact, grid_lin, kBlock, 0, dv, act_a, act_b, ctl);
        std::swap(act_a, act_b);
        if (hooks && hooks->pass_cb) {
            // a consistent snapshot for the caller between two chunks (the
            // reference's viewer reads the live buffers while tiles write them)
            const size_t npx = (size_t)hv[0].cam.w * hv[0].cam.h;
            PICO_HIP(c, hipMemcpyAsync(hooks->h_pixels, hv[0].pixels, 12 * npx, hipMemcpyDeviceToHost, c->stream));
            PICO_HIP(c, hipMemcpyAsync(hooks->h_counts, hv[0].counts, 4 * npx, hipMemcpyDeviceToHost, c->stream));
            PICO_HIP(c, hipStreamSynchronize(c->stream));
            if (hooks->pass_cb(hooks->pass_user, hooks->h_pixels, hooks->h_counts, hv[0].cam.w, hv[0].cam.h,
                               std::min(off + S, hv[0].spp))) {
                cancelled = true;
                break;
            }
        }
        if (per_view_times) {
            hipEvent_t ce = nullptr;
            for (uint32_t v = 0; v < nv; ++v) {
                if (view_render[v] || off + S < hv[v].spp) continue;
                if (!ce) {
                    if (!new_event(&ce)) return fail(c, PICO_ERR_HIP, "hipEventCreate failed");
                    PICO_HIP(c, hipEventRecord(ce, c->stream));
                }
                view_render[v] = ce;
                if (hooks && hooks->cnn) {
                    // G-buffer and network of this view, either between this
                    // chunk and the next on the main stream, or on the second
                    // stream overlapping the passes still to come
                    const bool side = (tune.flags & PICO_TUNE_SIDE_STREAM) != 0;
                    if (side) PICO_HIP(c, hipStreamWaitEvent(c->side, ce, 0));
                    StreamSwap sw(c, side ? c->side : c->stream);
                    const uint32_t w = hv[v].cam.w, h = hv[v].cam.h;
                    const pico_render_params& vp = hooks->views[v].params;
                    const uint32_t ftw = vp.tile_w ? vp.tile_w : 64, fth = vp.tile_h ? vp.tile_h : 64;
                    int rc = pico_hip_gbuffer_d(c, &hooks->views[v].camera, w, h, gn, gp, ga);
                    if (!rc)
                        rc = pico_hip_denoise_gbuffer_d(c, hooks->cnn, hv[v].pixels, gn, gp, ga, w, h, ftw, fth, gs,
                                                        hooks->d_out[v], fallback + v);
                    if (rc) return rc;
                    if (!new_event(&view_done[v])) return fail(c, PICO_ERR_HIP, "hipEventCreate failed");
                    PICO_HIP(c, hipEventRecord(view_done[v], c->stream));
                } else {
                    view_done[v] = ce;
                }
            }
        }
    }
    if (hooks && hooks->cnn && (tune.flags & PICO_TUNE_SIDE_STREAM)) {
        // the caller's stream sees the follow-up's outputs
        hipEvent_t se;
        if (!new_event(&se)) return fail(c, PICO_ERR_HIP, "hipEventCreate failed");
        PICO_HIP(c, hipEventRecord(se, c->side));
        PICO_HIP(c, hipStreamWaitEvent(c->stream, se, 0));
    }
    PICO_HIP(c, hipEventRecord(e1, c->stream));
    Ctl hc;
    std::vector<unsigned long long> hsl(nv * vstat_words), hs(2 * nv, 0ull);
    PICO_HIP(c, hipMemcpyAsync(&hc, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(hsl.data(), vstats, nv * vstat_words * 8, hipMemcpyDeviceToHost, c->stream));
    std::vector<uint32_t> hfb(nv, 0u);
    if (fallback) PICO_HIP(c, hipMemcpyAsync(hfb.data(), fallback, 4 * nv, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    for (uint32_t v = 0; v < nv; ++v)
        for (uint32_t k = 0; k < kCtrSlots; ++k) {
            hs[2 * v] += hsl[v * vstat_words + k * kCtrStride];
            hs[2 * v + 1] += hsl[v * vstat_words + k * kCtrStride + 1];
        }
    float ms = 0.0f;
    hipEventElapsedTime(&ms, e0, e1);
    {
        // the latest hand-over of each chunk shape in this render; none when a
        // chunk of it ran every bounce or was forced with many paths left
        std::map<FinishKey, uint32_t> seen;
        for (size_t k = 0; k < chunk_key.size() && k < kLogChunks; ++k) {
            const uint32_t f = hc.chunk_finish[k];
            const bool none = f == kNoFinish || hc.chunk_forced[k] != 0;
            auto it = seen.find(chunk_key[k]);
            if (it == seen.end()) seen[chunk_key[k]] = none ? kNoFinish : f;
            else if (it->second != kNoFinish) it->second = none ? kNoFinish : std::max(it->second, f);
        }
        if (c->finish_hints.size() > 256) c->finish_hints.clear();
        for (const auto& kv : seen) {
            if (kv.second == kNoFinish || kv.second < 1 || kv.second > kv.first.depth) c->finish_hints.erase(kv.first);
            else c->finish_hints[kv.first] = kv.second;
        }
    }
    if (hc.overflow) return fail(c, PICO_ERR_LIMIT, "traversal stack overflow");
    if (stats) {
        for (uint32_t v = 0; v < n_stats; ++v) {
            pico_stats& st = stats[v];
            std::memset(&st, 0, sizeof(st));
            st.render_ms = ms;
            st.done_ms = ms;
            st.cancelled = cancelled ? 1 : 0;
            st.denoise_fallback = hfb[v] ? 1 : 0;
            if (n_stats == 1) {
                // the whole launch: rays of every pass (including samples the
                // early-out discarded), and rays of the accumulated samples
                st.camera_rays = hc.rays_cam;
                st.extension_rays = hc.rays_ext;
                st.shadow_rays = hc.rays_sh;
                st.node_visits = hc.nodes;
                st.tri_tests = hc.tris;
                st.instance_visits = hc.insts;
                st.shadow_node_visits = hc.nodes_sh;
                st.shadow_tri_tests = hc.tris_sh;
                st.shadow_instance_visits = hc.insts_sh;
                st.spill_pushes = hc.spill_push;
                st.spill_pops = hc.spill_pop;
                st.finish_rays = hc.rays_fin;
                st.finish_node_visits = hc.nodes_fin;
                st.finish_tri_tests = hc.tris_fin;
                st.finish_instance_visits = hc.insts_fin;
                for (uint32_t k = 0; k < nv; ++k) {
                    st.accumulated_rays += hs[2 * k];
                    st.samples += hs[2 * k + 1];
                }
            } else {
                st.accumulated_rays = hs[2 * v];
                st.samples = hs[2 * v + 1];
            }
            if (per_view_times && view_render[v]) {
                float a = 0.0f, b = 0.0f;
                hipEventElapsedTime(&a, epoch ? epoch : e0, view_render[v]);
                st.render_ms = a;
                if (view_done[v] && hipEventElapsedTime(&b, epoch ? epoch : e0, view_done[v]) == hipSuccess) st.done_ms = b;
                else st.done_ms = a;
            }
        }
    }
    return cancelled ? fail(c, PICO_ERR_CANCELLED, "cancelled") : PICO_OK;
}

// Batched views with PICO_TUNE_SHORTEST_FIRST: the views, ordered by work
// (pixels x samples), run as successive launches whose cumulative work
// doubles (1/16, 1/8, 1/4, 1/2, all of it), so that the cheap views are done
// early; per-view times count from the first launch.  Without the flag: one
// launch.
int run_batched(pico_hip_ctx* c, const pico_view* views, uint32_t n, std::vector<DView>& hv, const pico_tuning& tune,
                pico_stats* stats, const ViewHooks* hooks) {
    if (!(tune.flags & PICO_TUNE_SHORTEST_FIRST) || n < 2)
        return run_views(c, hv, tune, nullptr, false, stats, stats ? n : 0, hooks);
    std::vector<uint32_t> order(n);
    std::vector<double> work(n);
    double total = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        order[i] = i;
        work[i] = (double)hv[i].cam.w * hv[i].cam.h * hv[i].spp;
        total += work[i];
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return work[a] < work[b]; });
    hipEvent_t epoch;
    PICO_HIP(c, hipEventCreate(&epoch));
    struct Guard { hipEvent_t e; ~Guard() { hipEventDestroy(e); } } g{epoch};
    PICO_HIP(c, hipEventRecord(epoch, c->stream));
    double done = 0.0, bound = total / 16.0;
    size_t k = 0;
    while (k < n) {
        std::vector<uint32_t> grp;
        while (k < n && (grp.empty() || done + work[order[k]] <= bound * (1.0 + 1e-9))) {
            done += work[order[k]];
            grp.push_back(order[k++]);
        }
        bound = std::min(total, bound * 2.0);
        std::vector<DView> gv;
        std::vector<pico_view> gpv;
        std::vector<float*> gout;
        for (uint32_t i : grp) {
            gv.push_back(hv[i]);
            gpv.push_back(views[i]);
            if (hooks && hooks->d_out) gout.push_back(hooks->d_out[i]);
        }
        ViewHooks gh;
        if (hooks) {
            gh.cnn = hooks->cnn;
            gh.d_out = gout.data();
            gh.views = gpv.data();
        }
        std::vector<pico_stats> gs(grp.size());
        int rc = run_views(c, gv, tune, nullptr, false, stats ? gs.data() : nullptr, stats ? (uint32_t)grp.size() : 0,
                           hooks ? &gh : nullptr, epoch);
        if (rc) return rc;
        if (stats)
            for (size_t j = 0; j < grp.size(); ++j) stats[grp[j]] = gs[j];
    }
    return PICO_OK;
}

DView base_view(const pico_camera* cam, const pico_render_params* p) {
    DView v{};
    v.cam = make_dcam(*cam, p->width, p->height);
    v.spp = p->spp;
    v.max_depth = p->max_depth;
    v.frame_seed = p->frame_seed;
    v.view_id = p->view_id;
    v.max_variance = p->max_variance;
    v.resume = (p->flags & PICO_RENDER_RESUME) ? 1u : 0u;
    return v;
}

int check_params(pico_hip_ctx* c, const pico_camera* cam, const pico_render_params* p);

// Host descriptions of a caller's views (accumulating into its buffers).
int make_views(pico_hip_ctx* c, const pico_view* views, uint32_t n, std::vector<DView>& hv) {
    if (!views || n == 0) return fail(c, PICO_ERR_ARG, "no views");
    if (n > kMaxViews) return fail(c, PICO_ERR_LIMIT, "at most 64 views per launch");
    for (uint32_t i = 0; i < n; ++i) {
        int rc = check_params(c, &views[i].camera, &views[i].params);
        if (rc) return rc;
        if (!views[i].d_pixels || !views[i].d_sample_count || !views[i].d_variance) return fail(c, PICO_ERR_ARG, "NULL buffer");
        DView v = base_view(&views[i].camera, &views[i].params);
        v.mode = MODE_ACCUM;
        v.pixels = views[i].d_pixels;
        v.counts = views[i].d_sample_count;
        v.variance = views[i].d_variance;
        hv.push_back(v);
    }
    return PICO_OK;
}

int check_params(pico_hip_ctx* c, const pico_camera* cam, const pico_render_params* p) {
    if (!p) return fail(c, PICO_ERR_ARG, "params is NULL");
    int rc = check_camera(c, cam, p->width, p->height);
    if (rc) return rc;
    if (p->max_depth > kMaxBounce - 1) return fail(c, PICO_ERR_LIMIT, "max_depth above 63");
    return PICO_OK;
}

}  // namespace

extern "C" {

int pico_hip_render_d(pico_hip_ctx* c, const pico_camera* cam, const pico_render_params* p, float* dpix, uint32_t* dcnt,
                      float* dvar, const volatile int32_t* cancel, pico_stats* stats) {
    if (!c) return PICO_ERR_ARG;
    int rc = check_params(c, cam, p);
    if (rc) return rc;
    if (!dpix || !dcnt || !dvar) return fail(c, PICO_ERR_ARG, "NULL buffer");
    std::vector<DView> v(1, base_view(cam, p));
    v[0].mode = MODE_ACCUM;
    v[0].pixels = dpix;
    v[0].counts = dcnt;
    v[0].variance = dvar;
    return run_views(c, v, c->tuning, cancel, (p->flags & PICO_RENDER_COUNT) != 0, stats, stats ? 1 : 0);
}

int pico_hip_render(pico_hip_ctx* c, const pico_camera* cam, const pico_render_params* p, float* pix, uint32_t* cnt,
                    float* var, const volatile int32_t* cancel, pico_stats* stats) {
    if (!c) return PICO_ERR_ARG;
    int rc = check_params(c, cam, p);
    if (rc) return rc;
    if (!pix || !cnt || !var) return fail(c, PICO_ERR_ARG, "NULL buffer");
    const size_t n = (size_t)p->width * p->height;
    char* b = (char*)scratch(c, "frame", n * 28 + 64);
    if (!b) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    float* dp = (float*)b;
    float* dv = dp + 3 * n;
    uint32_t* dc = (uint32_t*)(dv + 3 * n);
    if (p->flags & PICO_RENDER_RESUME) {
        PICO_HIP(c, hipMemcpyAsync(dp, pix, 12 * n, hipMemcpyHostToDevice, c->stream));
        PICO_HIP(c, hipMemcpyAsync(dv, var, 12 * n, hipMemcpyHostToDevice, c->stream));
        PICO_HIP(c, hipMemcpyAsync(dc, cnt, 4 * n, hipMemcpyHostToDevice, c->stream));
    }
    rc = pico_hip_render_d(c, cam, p, dp, dc, dv, cancel, stats);
    if (rc && rc != PICO_ERR_CANCELLED) return rc;
    PICO_HIP(c, hipMemcpyAsync(pix, dp, 12 * n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(var, dv, 12 * n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(cnt, dc, 4 * n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    return rc;
}

int pico_hip_render_samples(pico_hip_ctx* c, const pico_camera* cam, const pico_render_params* p, uint32_t sample_base,
                            float* out) {
    if (!c) return PICO_ERR_ARG;
    int rc = check_params(c, cam, p);
    if (rc) return rc;
    if (!out) return fail(c, PICO_ERR_ARG, "NULL buffer");
    const size_t n = (size_t)p->width * p->height * p->spp * 3;
    float* d = (float*)scratch(c, "samples_out", n * 4 + 64);
    if (!d) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    std::vector<DView> v(1, base_view(cam, p));
    v[0].mode = MODE_SAMPLES;
    v[0].sample_base = sample_base;
    v[0].samples_out = d;
    rc = run_views(c, v, c->tuning, nullptr, false, nullptr, 0);
    if (rc) return rc;
    PICO_HIP(c, hipMemcpyAsync(out, d, n * 4, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    return PICO_OK;
}

int pico_hip_render_accum64_d(pico_hip_ctx* c, const pico_camera* cam, const pico_render_params* p, uint32_t sample_base,
                              double* d_sum, pico_stats* stats) {
    if (!c) return PICO_ERR_ARG;
    int rc = check_params(c, cam, p);
    if (rc) return rc;
    if (!d_sum) return fail(c, PICO_ERR_ARG, "NULL buffer");
    std::vector<DView> v(1, base_view(cam, p));
    v[0].mode = MODE_SUM64;
    v[0].sample_base = sample_base;
    v[0].sum64 = d_sum;
    return run_views(c, v, c->tuning, nullptr, false, stats, stats ? 1 : 0);
}

int pico_hip_render_progressive(pico_hip_ctx* c, const pico_camera* cam, const pico_render_params* p, float* pix,
                                uint32_t* cnt, float* var, pico_pass_callback cb, void* user, pico_stats* stats) {
    if (!c) return PICO_ERR_ARG;
    int rc = check_params(c, cam, p);
    if (rc) return rc;
    if (!pix || !cnt || !var || !cb) return fail(c, PICO_ERR_ARG, "NULL buffer or callback");
    if (p->flags & PICO_RENDER_COUNT) return fail(c, PICO_ERR_ARG, "no counting run with a pass callback");
    const size_t n = (size_t)p->width * p->height;
    char* b = (char*)scratch(c, "frame", n * 28 + 64);
    if (!b) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    float* dp = (float*)b;
    float* dv = dp + 3 * n;
    uint32_t* dc = (uint32_t*)(dv + 3 * n);
    if (c->pinned_bytes < n * 16) {
        if (c->pinned) hipHostFree(c->pinned);
        c->pinned = nullptr;
        c->pinned_bytes = 0;
        PICO_HIP(c, hipHostMalloc(&c->pinned, n * 16));
        c->pinned_bytes = n * 16;
    }
    if (p->flags & PICO_RENDER_RESUME) {
        PICO_HIP(c, hipMemcpyAsync(dp, pix, 12 * n, hipMemcpyHostToDevice, c->stream));
        PICO_HIP(c, hipMemcpyAsync(dv, var, 12 * n, hipMemcpyHostToDevice, c->stream));
        PICO_HIP(c, hipMemcpyAsync(dc, cnt, 4 * n, hipMemcpyHostToDevice, c->stream));
    }
    std::vector<DView> v(1, base_view(cam, p));
    v[0].mode = MODE_ACCUM;
    v[0].pixels = dp;
    v[0].counts = dc;
    v[0].variance = dv;
    ViewHooks hk;
    hk.pass_cb = cb;
    hk.pass_user = user;
    hk.h_pixels = (float*)c->pinned;
    hk.h_counts = (uint32_t*)((char*)c->pinned + 12 * n);
    rc = run_views(c, v, c->tuning, nullptr, false, stats, stats ? 1 : 0, &hk);
    if (rc && rc != PICO_ERR_CANCELLED) return rc;
    PICO_HIP(c, hipMemcpyAsync(pix, dp, 12 * n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(var, dv, 12 * n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(cnt, dc, 4 * n, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    return rc;
}

int pico_hip_render_views(pico_hip_ctx* c, const pico_view_host* views, uint32_t n, uint32_t queue_paths,
                          pico_stats* stats) {
    if (!c) return PICO_ERR_ARG;
    if (!views || n == 0 || n > kMaxViews) return fail(c, PICO_ERR_LIMIT, "between 1 and 64 views per launch");
    std::vector<size_t> off(n + 1, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const pico_render_params& p = views[i].params;
        if (!views[i].pixels || !views[i].sample_count || !views[i].variance) return fail(c, PICO_ERR_ARG, "NULL buffer");
        if ((uint64_t)p.width * p.height > (1ull << 28)) return fail(c, PICO_ERR_LIMIT, "view larger than 2^28 pixels");
        off[i + 1] = off[i] + (size_t)p.width * p.height;
    }
    char* b = (char*)scratch(c, "views_host", off[n] * 28 + 256 * (size_t)n + 64);
    if (!b) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    std::vector<pico_view> dv(n);
    for (uint32_t i = 0; i < n; ++i) {
        const size_t px = off[i + 1] - off[i];
        char* vb = b + 28 * off[i] + 256 * (size_t)i;  // each view's block 4-byte aligned, apart
        dv[i].camera = views[i].camera;
        dv[i].params = views[i].params;
        dv[i].d_pixels = (float*)vb;
        dv[i].d_variance = (float*)(vb + 12 * px);
        dv[i].d_sample_count = (uint32_t*)(vb + 24 * px);
        if (views[i].params.flags & PICO_RENDER_RESUME) {
            PICO_HIP(c, hipMemcpyAsync(dv[i].d_pixels, views[i].pixels, 12 * px, hipMemcpyHostToDevice, c->stream));
            PICO_HIP(c, hipMemcpyAsync(dv[i].d_variance, views[i].variance, 12 * px, hipMemcpyHostToDevice, c->stream));
            PICO_HIP(c, hipMemcpyAsync(dv[i].d_sample_count, views[i].sample_count, 4 * px, hipMemcpyHostToDevice,
                                       c->stream));
        }
    }
    int rc = pico_hip_render_views_d(c, dv.data(), n, queue_paths, stats);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        const size_t px = off[i + 1] - off[i];
        PICO_HIP(c, hipMemcpyAsync(views[i].pixels, dv[i].d_pixels, 12 * px, hipMemcpyDeviceToHost, c->stream));
        PICO_HIP(c, hipMemcpyAsync(views[i].variance, dv[i].d_variance, 12 * px, hipMemcpyDeviceToHost, c->stream));
        PICO_HIP(c, hipMemcpyAsync(views[i].sample_count, dv[i].d_sample_count, 4 * px, hipMemcpyDeviceToHost, c->stream));
    }
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    return PICO_OK;
}

int pico_hip_render_views_d(pico_hip_ctx* c, const pico_view* views, uint32_t n, uint32_t queue_paths, pico_stats* stats) {
    if (!c) return PICO_ERR_ARG;
    std::vector<DView> hv;
    int rc = make_views(c, views, n, hv);
    if (rc) return rc;
    pico_tuning t = c->tuning;
    if (queue_paths) t.queue_paths = queue_paths;
    return run_batched(c, views, n, hv, t, stats, nullptr);
}

int pico_hip_render_views_denoise_d(pico_hip_ctx* c, const pico_view* views, uint32_t n, pico_cnn* cnn,
                                    float* const* d_denoised, pico_stats* stats) {
    if (!c) return PICO_ERR_ARG;
    if (!cnn || !d_denoised) return fail(c, PICO_ERR_ARG, "NULL network or output list");
    if (!cnn_belongs_to(cnn, c)) return fail(c, PICO_ERR_ARG, "the network was created on another context");
    std::vector<DView> hv;
    int rc = make_views(c, views, n, hv);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; ++i) {
        if (!d_denoised[i]) return fail(c, PICO_ERR_ARG, "NULL output buffer");
        if ((rc = cnn_check_size(cnn, views[i].params.width, views[i].params.height))) return rc;
    }
    ViewHooks hk;
    hk.cnn = cnn;
    hk.d_out = d_denoised;
    hk.views = views;
    return run_batched(c, views, n, hv, c->tuning, stats, &hk);
}

int pico_hip_set_tuning(pico_hip_ctx* c, const pico_tuning* t) {
    if (!c) return PICO_ERR_ARG;
    if (!t) return fail(c, PICO_ERR_ARG, "NULL tuning");
    if ((t->tile_w == 0) != (t->tile_h == 0)) return fail(c, PICO_ERR_ARG, "tile_w and tile_h are both 0 or both > 0");
    if ((uint64_t)t->tile_w * t->tile_h > (1u << 20)) return fail(c, PICO_ERR_LIMIT, "tile of more than 2^20 pixels");
    c->tuning = *t;
    if (c->tuning.queue_paths) c->tuning.queue_paths = std::max<uint32_t>(c->tuning.queue_paths, 1024);
    return PICO_OK;
}

int pico_hip_get_tuning(pico_hip_ctx* c, pico_tuning* t) {
    if (!c || !t) return PICO_ERR_ARG;
    *t = c->tuning;
    return PICO_OK;
}

int pico_hip_autotune_views(pico_hip_ctx* c, const pico_view* views, uint32_t n, uint32_t calib_spp, pico_tuning* chosen,
                            float* chosen_ms, float* default_ms) {
    if (!c) return PICO_ERR_ARG;
    if (calib_spp == 0) return fail(c, PICO_ERR_ARG, "calib_spp must be positive");
    std::vector<DView> hv;
    int rc = make_views(c, views, n, hv);
    if (rc) return rc;
    // key: scene upload, view sizes, samples, depths
    uint64_t key = 0xcbf29ce484222325ull;
    auto mixin = [&](uint64_t x) { key = (key ^ x) * 0x100000001b3ull; };
    mixin(c->scene_serial);
    mixin(n);
    mixin(calib_spp);
    uint32_t max_spp = 1;
    for (uint32_t i = 0; i < n; ++i) {
        mixin(views[i].params.width);
        mixin(views[i].params.height);
        mixin(views[i].params.spp);
        mixin(views[i].params.max_depth);
        max_spp = std::max(max_spp, views[i].params.spp);
    }
    auto it = c->tune_cache.find(key);
    if (it == c->tune_cache.end()) {
        for (uint32_t i = 0; i < n; ++i)
            hv[i].spp = std::max<uint32_t>(1, (uint32_t)(((uint64_t)views[i].params.spp * calib_spp + max_spp - 1) / max_spp));
        const pico_tuning defaults{kDefaultQueuePaths, 0, 0, 0};
        std::vector<pico_tuning> cand{defaults};
        const uint32_t queues[] = {4u << 20, 8u << 20, 16u << 20, 32u << 20, 64u << 20, 128u << 20};
        const uint32_t tiles[][2] = {{0, 0}, {8, 8}, {16, 16}, {32, 4}};
        for (uint32_t q : queues)
            for (auto& t : tiles) {
                pico_tuning x{q, t[0], t[1], 0};
                if (!(x.queue_paths == defaults.queue_paths && x.tile_w == 0)) cand.push_back(x);
            }
        // warm-up at the deepest queue: the work buffers reach their largest size
        std::vector<DView> w = hv;
        if ((rc = run_views(c, w, pico_tuning{64u << 20, 0, 0, 0}, nullptr, false, nullptr, 0))) return rc;
        std::vector<float> best(cand.size(), 1e30f);
        for (int rep = 0; rep < 2; ++rep)
            for (size_t k = 0; k < cand.size(); ++k) {
                std::vector<DView> v = hv;
                pico_stats st;
                if ((rc = run_views(c, v, cand[k], nullptr, false, &st, 1))) return rc;
                best[k] = std::min(best[k], (float)st.render_ms);
            }
        size_t kb = 0;
        for (size_t k = 1; k < cand.size(); ++k)
            if (best[k] < best[kb]) kb = k;
        it = c->tune_cache.emplace(key, TuneEntry{cand[kb], best[kb], best[0]}).first;
    }
    c->tuning = it->second.tuning;
    if (chosen) *chosen = it->second.tuning;
    if (chosen_ms) *chosen_ms = it->second.chosen_ms;
    if (default_ms) *default_ms = it->second.default_ms;
    return PICO_OK;
}

}  // extern "C"
