// Device build of a whole scene: every mesh BVH on the GPU, all meshes at
// once, then the TLAS (tlas.hip).  The host keeps only the checks of the
// description and the small tables (materials, instances, lights).
//
// Binary build, level by level (one launch per level, one workgroup per
// node of the level):
//   - bin the node's triangles by centroid, 16 bins on each axis, counts and
//     triangle / centroid boxes in LDS (exact float min/max through mapped
//     integer atomics, so the bins do not depend on the order of arrival);
//   - one lane evaluates the surface-area cost of the 45 candidate planes in
//     the host builder's order and arithmetic, so both builders take the same
//     decisions and make the same tree; a node becomes a leaf when it holds
//     one triangle, or at most 4 and the leaf is no dearer than the split;
//   - a stable partition of the node's range of the triangle permutation
//     (wave ballots and a workgroup prefix) writes the children's ranges into
//     the other permutation buffer; the children's boxes come from the bins.
// Leaves write their triangle records in place (range position = record
// index), so leaves never move again.  The binary tree is then collapsed to
// 4-wide nodes level by level with the host's greedy rule (open the child of
// largest area until four), quantised with the host's function, and the
// traversal-stack bound is the largest sum of (children - 1) down any path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "../../include/pico_hip.h"
#include "ctx.h"
#include "dev_atomics.h"
#include "scene_math.h"

using namespace pico;

namespace {

constexpr uint32_t kBins = 16;
constexpr uint32_t kMaxLeaf = 4;
constexpr uint32_t kThreads = 256;
constexpr uint32_t kRootSlot = 0x80000000u;  // slot | mesh: the reference goes to the mesh's root entry
constexpr uint32_t kMaxLevels = 1u << 16;
constexpr uint32_t kLarge = 8192;  // nodes of more triangles are binned and partitioned by many workgroups
constexpr uint32_t kChunk = 2048;  // triangles per workgroup of such a node
constexpr uint32_t kSmall = 64;    // nodes of at most this many triangles: the whole subtree by one wave

struct BuildTask {
    uint32_t begin, end;  // range of the triangle permutation
    uint32_t slot;        // binary node * 2 + child that receives this node's reference, or kRootSlot | mesh
    uint32_t mesh;
    float box[6];         // union of the triangle boxes: lo, hi
    float cbox[6];        // bounds of the triangle centroids: lo, hi
};
static_assert(sizeof(BuildTask) == 64, "task is one 64-byte record");

struct WideTask {
    uint32_t bnode;  // binary node the wide node is made from
    uint32_t slot;   // wide node * 4 + child that receives this node's index, or kRootSlot | mesh
    uint32_t acc;    // stack entries the wide nodes above can leave
    uint32_t pad;
};

struct Geo {
    const float* pos;       // vertex positions (3 per vertex, all meshes)
    const uint32_t* lidx;   // mesh-local triangle indices as given
    uint32_t* gidx;         // global vertex indices (written)
    const uint32_t* vofs;   // n_meshes + 1 vertex offsets
    const uint32_t* tofs;   // n_meshes + 1 triangle offsets
    uint32_t n_meshes, n_tris, n_verts;
    float* pbox;            // 6 per triangle
    float* pcen;            // 3 per triangle
    uint32_t* perm;         // identity permutation (written)
    uint32_t* mbox;         // 6 per mesh, mapped: vertex bounds (the TLAS uses them)
    uint32_t* rbox;         // 12 per mesh, mapped: triangle box union, centroid bounds
    uint32_t* err;
    const uint32_t* shape;  // per mesh PICO_SHAPE_*, or null
    const float* nrm;       // vertex normals
    float4* tshade;         // per triangle shading record (written)
};

struct Bins;
struct LargeTask;
struct Chunk;

// A subtree of at most kSmall triangles, built by one wave: its root task,
// the permutation buffer that holds its range, and the level its root would
// have in the level-by-level build.
struct SmallTask {
    BuildTask t;
    uint32_t parity, level, pad[2];
};

struct Level {
    const BuildTask* in;
    BuildTask* out;
    uint32_t* out_count;
    const uint32_t* perm_in;
    uint32_t* perm_out;
    const float* pbox;
    const float* pcen;
    const float* pos;
    const uint32_t* gidx;
    const uint32_t* tofs;
    DNode* bnodes;
    uint32_t* bcount;
    uint32_t* mesh_root2;
    DTri* tris;
    // nodes of more than kLarge triangles
    LargeTask* large;
    uint32_t* large_count;
    Chunk* chunks;
    uint32_t* chunk_count;
    Bins* gbins;
    uint32_t* chunk_left;
    uint32_t* next_large;   // children of more than kLarge triangles, and their chunks
    uint32_t* next_chunks;
    // children of at most kSmall triangles: built later, one subtree per wave
    SmallTask* small;
    uint32_t* small_count;
    uint32_t out_parity;    // which permutation buffer perm_out is (0 or 1)
    uint32_t level;         // this level's number (1 for the roots)
};

// host Box arithmetic (bvh_build.cpp), replayed on the device
struct DBox {
    float lo[3], hi[3];
    __device__ void clear() {
        for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    }
    __device__ void grow(const float* l, const float* h) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = l[k] < lo[k] ? l[k] : lo[k];  // std::min(lo, l)
            hi[k] = hi[k] < h[k] ? h[k] : hi[k];  // std::max(hi, h)
        }
    }
    __device__ float area() const {
        float d[3];
        for (int k = 0; k < 3; ++k) {
            const float x = hi[k] - lo[k];
            d[k] = 0.0f < x ? x : 0.0f;
        }
        return 2.0f * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2]);
    }
    __device__ bool empty() const { return !(lo[0] <= hi[0]); }
};

__device__ __forceinline__ uint32_t find_mesh(const uint32_t* ofs, uint32_t n_meshes, uint32_t x) {
    // the last m with ofs[m] <= x (skips empty meshes, whose offsets repeat)
    uint32_t lo = 0, hi = n_meshes;  // ofs[lo] <= x < ofs[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (ofs[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void k_geo_clear(Geo g) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.n_meshes; i += gridDim.x * blockDim.x) {
        for (int k = 0; k < 3; ++k) {
            g.mbox[6 * i + k] = kMappedPosInf;
            g.mbox[6 * i + 3 + k] = kMappedNegInf;
            g.rbox[12 * i + k] = kMappedPosInf;
            g.rbox[12 * i + 3 + k] = kMappedNegInf;
            g.rbox[12 * i + 6 + k] = kMappedPosInf;
            g.rbox[12 * i + 9 + k] = kMappedNegInf;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *g.err = 0;
}

// Per triangle: global indices, box, centroid, identity permutation.
__global__ void k_geo_tris(Geo g) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < g.n_tris; t += gridDim.x * blockDim.x) {
        const uint32_t m = find_mesh(g.tofs, g.n_meshes, t);
        const uint32_t v0 = g.vofs[m], nv = g.vofs[m + 1] - v0;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = 0; c < 3; ++c) {
            uint32_t li = g.lidx[3 * (size_t)t + c];
            if (li >= nv) {
                atomicOr(g.err, 1u);
                li = 0;
            }
            const uint32_t v = v0 + li;
            g.gidx[3 * (size_t)t + c] = v;
            const float* p = g.pos + 3 * (size_t)v;
            for (int k = 0; k < 3; ++k) {
                lo[k] = p[k] < lo[k] ? p[k] : lo[k];
                hi[k] = hi[k] < p[k] ? p[k] : hi[k];
            }
        }
        {
            // shading record: positions, then normals, of the three vertices
            float r[20];
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < 3; ++k) {
                    r[3 * c + k] = g.pos[3 * (size_t)g.gidx[3 * (size_t)t + c] + k];
                    r[9 + 3 * c + k] = g.nrm[3 * (size_t)g.gidx[3 * (size_t)t + c] + k];
                }
            r[18] = 0.0f;
            r[19] = 0.0f;
            for (uint32_t q = 0; q < kShadeRec; ++q)
                g.tshade[kShadeRec * (size_t)t + q] = make_float4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
        }
        for (int k = 0; k < 3; ++k) {
            g.pbox[6 * (size_t)t + k] = lo[k];
            g.pbox[6 * (size_t)t + 3 + k] = hi[k];
            g.pcen[3 * (size_t)t + k] = 0.5f * (lo[k] + hi[k]);
        }
        g.perm[t] = t;
    }
}

// Per-mesh bounds of D values over the items of each mesh (items grouped by
// mesh through the offsets `ofs`): every lane folds a run of kRun items in
// order, a wave whose lanes all end in one mesh combines them by shuffles,
// and the results reach device memory through mapped-float atomics, so the
// atomics per mesh are about one per 2048 items.
constexpr uint32_t kRun = 32;

template <int D, class Load, class Flush>
__device__ void run_bounds(uint32_t n, const uint32_t* ofs, uint32_t n_meshes, Load load, Flush flush) {
    const uint32_t r0 = (blockIdx.x * blockDim.x + threadIdx.x) * kRun;
    const bool active = r0 < n;
    const uint32_t r1 = active ? min(n, r0 + kRun) : r0;
    float lo[D], hi[D];
    for (int k = 0; k < D; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    uint32_t m = active ? find_mesh(ofs, n_meshes, r0) : 0u;
    uint32_t next = active ? ofs[m + 1] : 0u;
    bool any = false;
    for (uint32_t i = r0; i < r1; ++i) {
        while (i >= next) {  // the run crosses into a later mesh (empty meshes skipped)
            if (any) flush(m, lo, hi);
            for (int k = 0; k < D; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
            any = false;
            ++m;
            next = ofs[m + 1];
        }
        float vl[D], vh[D];
        load(i, vl, vh);
        for (int k = 0; k < D; ++k) {
            lo[k] = fminf(lo[k], vl[k]);
            hi[k] = fmaxf(hi[k], vh[k]);
        }
        any = true;
    }
    const uint32_t m0 = __shfl(m, 0);
    if (__all(!any || m == m0)) {
        for (int k = 0; k < D; ++k)
            for (int o = 32; o > 0; o >>= 1) {
                lo[k] = fminf(lo[k], __shfl_xor(lo[k], o));
                hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o));
            }
        if ((threadIdx.x & 63u) == 0 && __any(any)) flush(m0, lo, hi);
    } else if (any) {
        flush(m, lo, hi);
    }
}

// Per mesh: the union of its triangle boxes and the bounds of its centroids.
__global__ void k_tri_bounds(Geo g) {
    run_bounds<6>(
        g.n_tris, g.tofs, g.n_meshes,
        [&](uint32_t t, float* vl, float* vh) {
            for (int k = 0; k < 3; ++k) {
                vl[k] = g.pbox[6 * (size_t)t + k];
                vh[k] = g.pbox[6 * (size_t)t + 3 + k];
                vl[3 + k] = vh[3 + k] = g.pcen[3 * (size_t)t + k];
            }
        },
        [&](uint32_t m, const float* lo, const float* hi) {
            uint32_t* rb = g.rbox + 12 * (size_t)m;
            for (int k = 0; k < 3; ++k) {
                atomic_min_f(rb + k, lo[k]);
                atomic_max_f(rb + 3 + k, hi[k]);
                atomic_min_f(rb + 6 + k, lo[3 + k]);
                atomic_max_f(rb + 9 + k, hi[3 + k]);
            }
        });
}

// Per mesh: the bounds of its vertices (the host's mesh box for the TLAS).
__global__ void k_vert_bounds(Geo g) {
    run_bounds<3>(
        g.n_verts, g.vofs, g.n_meshes,
        [&](uint32_t v, float* vl, float* vh) {
            for (int k = 0; k < 3; ++k) vl[k] = vh[k] = g.pos[3 * (size_t)v + k];
        },
        [&](uint32_t m, const float* lo, const float* hi) {
            for (int k = 0; k < 3; ++k) {
                atomic_min_f(g.mbox + 6 * (size_t)m + k, lo[k]);
                atomic_max_f(g.mbox + 6 * (size_t)m + 3 + k, hi[k]);
            }
        });
}

__device__ void set_child(DNode& nd, int c, const DBox& b, uint32_t ref) {
    float* x = c == 0 ? nd.c0x : nd.c1x;
    float* y = c == 0 ? nd.c0y : nd.c1y;
    float* z = c == 0 ? nd.c0z : nd.c1z;
    x[0] = b.lo[0]; x[1] = b.hi[0];
    y[0] = b.lo[1]; y[1] = b.hi[1];
    z[0] = b.lo[2]; z[1] = b.hi[2];
    nd.ref[c] = ref;
}

// One root task per mesh with triangles; a mesh without triangles gets a
// node with two empty children, as on the host.
__global__ void k_roots(Geo g, BuildTask* out, uint32_t* out_count, DNode* bnodes, uint32_t* bcount, uint32_t* mesh_root2,
                        float* mesh_box, uint32_t* n_large, uint32_t* n_chunks) {
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < g.n_meshes; m += gridDim.x * blockDim.x) {
        if (g.shape && g.shape[m] != PICO_SHAPE_TRIANGLES) {
            // a built-in shape: no BLAS, bounds [-0.5, 0.5]^3
            for (int k = 0; k < 3; ++k) { mesh_box[6 * (size_t)m + k] = -0.5f; mesh_box[6 * (size_t)m + 3 + k] = 0.5f; }
            mesh_root2[m] = kImplicitRoot | g.shape[m];
            continue;
        }
        for (int k = 0; k < 6; ++k) mesh_box[6 * (size_t)m + k] = unmap_f(g.mbox[6 * (size_t)m + k]);
        const uint32_t b = g.tofs[m], e = g.tofs[m + 1];
        if (b == e) {
            const uint32_t id = atomicAdd(bcount, 1u);
            DNode nd{};
            DBox none;
            none.clear();
            set_child(nd, 0, none, kEmptyRef);
            set_child(nd, 1, none, kEmptyRef);
            bnodes[id] = nd;
            mesh_root2[m] = id;
            continue;
        }
        BuildTask t;
        t.begin = b;
        t.end = e;
        t.slot = kRootSlot | m;
        t.mesh = m;
        for (int k = 0; k < 6; ++k) {
            t.box[k] = unmap_f(g.rbox[12 * (size_t)m + k]);
            t.cbox[k] = unmap_f(g.rbox[12 * (size_t)m + 6 + k]);
        }
        out[atomicAdd(out_count, 1u)] = t;
        if (e - b > kLarge) {
            atomicAdd(n_large, 1u);
            atomicAdd(n_chunks, (e - b + kChunk - 1) / kChunk);
        }
    }
}

__device__ __forceinline__ uint32_t bin_of(float c, float lo, float sc) {
    const float x = (c - lo) * sc;
    return (uint32_t)(x < (float)(kBins - 1) ? x : (float)(kBins - 1));  // std::min<float>(bins - 1, x)
}

// Bins of one node: counts and mapped triangle / centroid boxes per bin and axis.
struct Bins {
    uint32_t cnt[3][kBins];
    uint32_t box[3][kBins][6];
    uint32_t cen[3][kBins][6];
};
constexpr uint32_t kBinCells = 3 * kBins;

// What a node does: 0 leaf, 1 split at a bin plane, 2 split in the middle of
// its range (all centroids coincide); children's boxes and centroid bounds.
struct Decision {
    int kind;
    int axis;
    uint32_t bin, nleft;
    float lbox[6], rbox[6], lcen[6], rcen[6];
};

// A node of more than kLarge triangles: binned, decided and partitioned by
// one workgroup per chunk of kChunk triangles.
struct LargeTask {
    uint32_t task;                   // index in the level's task list
    uint32_t first_chunk, n_chunks;
    uint32_t pad;
    Decision dec;
};

struct Chunk {
    uint32_t large;                  // LargeTask slot
    uint32_t begin, end;             // range of the permutation
    uint32_t pad;
};

__device__ void write_ref(const Level& L, uint32_t slot, uint32_t ref) {
    if (slot & kRootSlot) L.mesh_root2[slot & ~kRootSlot] = ref;
    else L.bnodes[slot >> 1].ref[slot & 1] = ref;
}

__device__ void task_scales(const BuildTask& t, float ext[3], float sc[3]) {
    for (int a = 0; a < 3; ++a) {
        ext[a] = t.cbox[3 + a] - t.cbox[a];
        sc[a] = ext[a] > 0 ? (float)kBins / ext[a] : 0.0f;
    }
}

__device__ void bins_clear(Bins* b, uint32_t tid, uint32_t nthreads) {
    for (uint32_t i = tid; i < kBinCells; i += nthreads) {
        (&b->cnt[0][0])[i] = 0;
        for (int k = 0; k < 3; ++k) {
            (&b->box[0][0][0])[6 * i + k] = kMappedPosInf;
            (&b->box[0][0][0])[6 * i + 3 + k] = kMappedNegInf;
            (&b->cen[0][0][0])[6 * i + k] = kMappedPosInf;
            (&b->cen[0][0][0])[6 * i + 3 + k] = kMappedNegInf;
        }
    }
}

// One triangle (centroid c, box bx) of task t into the bins b (LDS).
__device__ __forceinline__ void bin_tri(Bins* b, const BuildTask& t, const float* c, const float* bx, const float* ext,
                                        const float* sc) {
    for (int a = 0; a < 3; ++a) {
        if (!(ext[a] > 0)) continue;
        const uint32_t k = bin_of(c[a], t.cbox[a], sc[a]);
        atomicAdd(&b->cnt[a][k], 1u);
        for (int j = 0; j < 3; ++j) {
            atomicMin(&b->box[a][k][j], map_f(bx[j]));
            atomicMax(&b->box[a][k][3 + j], map_f(bx[3 + j]));
            atomicMin(&b->cen[a][k][j], map_f(c[j]));
            atomicMax(&b->cen[a][k][3 + j], map_f(c[j]));
        }
    }
}

// Bins the triangles perm_in[begin, end) of task t into b (LDS).
__device__ void bin_range(Bins* b, const Level& L, const BuildTask& t, uint32_t begin, uint32_t end, const float* ext,
                          const float* sc, uint32_t tid) {
    for (uint32_t i = begin + tid; i < end; i += kThreads) {
        const uint32_t p = L.perm_in[i];
        float c[3], bx[6];
        for (int k = 0; k < 3; ++k) c[k] = L.pcen[3 * (size_t)p + k];
        for (int k = 0; k < 6; ++k) bx[k] = L.pbox[6 * (size_t)p + k];
        bin_tri(b, t, c, bx, ext, sc);
    }
}

// The rest of the host builder's split() once the best plane is known
// (best_axis -1: none): leaf, plane or middle split, and the children's
// boxes (one lane).
__device__ void finish_decision(const Bins* b, const BuildTask& t, float best_cost, int best_axis, uint32_t best_bin,
                                Decision& d);

// The host builder's split() on the bins, by a whole wave (every lane of
// it calls this): lane 15 * axis + k prices the plane
// after bin k of `axis`, sweeping the bins in the host's order (right part
// from the last bin down, left part from the first bin up), and the wave
// takes the host's choice, the first of the cheapest planes in axis-major
// order (a cost that is not below infinity is never chosen).  Lane 0 then
// finishes the decision.
__device__ void decide_wave(const Bins* b, const BuildTask& t, const float* ext, Decision& d, uint32_t lane) {
    const uint32_t n = t.end - t.begin;
    float cost = INFINITY;
    uint32_t idx = 0xffffffffu;
    const int axis = (int)(lane / (kBins - 1));
    const uint32_t k = lane % (kBins - 1);
    if (n > 1 && axis < 3 && ext[axis] > 0) {
        DBox acc;
        acc.clear();
        uint32_t rc = 0;
        for (int j = (int)kBins - 1; j > (int)k; --j) {
            float l[3], h[3];
            for (int q = 0; q < 3; ++q) { l[q] = unmap_f(b->box[axis][j][q]); h[q] = unmap_f(b->box[axis][j][3 + q]); }
            acc.grow(l, h);
            rc += b->cnt[axis][j];
        }
        const float rarea = acc.empty() ? 0.0f : acc.area();
        acc.clear();
        uint32_t lc = 0;
        for (uint32_t j = 0; j <= k; ++j) {
            float l[3], h[3];
            for (int q = 0; q < 3; ++q) { l[q] = unmap_f(b->box[axis][j][q]); h[q] = unmap_f(b->box[axis][j][3 + q]); }
            acc.grow(l, h);
            lc += b->cnt[axis][j];
        }
        if (lc != 0 && rc != 0) {
            const float c = (float)lc * acc.area() + (float)rc * rarea;
            if (c < INFINITY) { cost = c; idx = lane; }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float oc = __shfl_xor(cost, o);
        const uint32_t oi = __shfl_xor(idx, o);
        if (oc < cost || (oc == cost && oi < idx)) { cost = oc; idx = oi; }
    }
    if (lane == 0) {
        const bool found = idx != 0xffffffffu;
        finish_decision(b, t, cost, found ? (int)(idx / (kBins - 1)) : -1, found ? idx % (kBins - 1) : 0u, d);
    }
}

__device__ void finish_decision(const Bins* b, const BuildTask& t, float best_cost, int best_axis, uint32_t best_bin,
                                Decision& d) {
    const uint32_t n = t.end - t.begin;
    int kind = 0;
    if (n > 1) {
        DBox bounds;
        for (int j = 0; j < 3; ++j) { bounds.lo[j] = t.box[j]; bounds.hi[j] = t.box[3 + j]; }
        const float parent = bounds.area();
        const float leaf_cost = (float)n * parent;
        const float split_cost = kSahTraversal * parent + best_cost;
        if (best_axis < 0) kind = n <= kMaxLeaf ? 0 : 2;
        else kind = (n <= kMaxLeaf && leaf_cost <= split_cost) ? 0 : 1;
    }
    d.kind = kind;
    d.axis = best_axis;
    d.bin = best_bin;
    d.nleft = 0;
    if (kind == 1) {
        DBox lb, rb, lc, rc;
        lb.clear(); rb.clear(); lc.clear(); rc.clear();
        uint32_t nl = 0;
        for (uint32_t k = 0; k < kBins; ++k) {
            float l[3], h[3], cl[3], ch[3];
            for (int j = 0; j < 3; ++j) {
                l[j] = unmap_f(b->box[best_axis][k][j]); h[j] = unmap_f(b->box[best_axis][k][3 + j]);
                cl[j] = unmap_f(b->cen[best_axis][k][j]); ch[j] = unmap_f(b->cen[best_axis][k][3 + j]);
            }
            if (k <= best_bin) { lb.grow(l, h); lc.grow(cl, ch); nl += b->cnt[best_axis][k]; }
            else { rb.grow(l, h); rc.grow(cl, ch); }
        }
        d.nleft = nl;
        for (int j = 0; j < 3; ++j) {
            d.lbox[j] = lb.lo[j]; d.lbox[3 + j] = lb.hi[j];
            d.rbox[j] = rb.lo[j]; d.rbox[3 + j] = rb.hi[j];
            d.lcen[j] = lc.lo[j]; d.lcen[3 + j] = lc.hi[j];
            d.rcen[j] = rc.lo[j]; d.rcen[3 + j] = rc.hi[j];
        }
    } else if (kind == 2) {
        d.nleft = n / 2;  // boxes of the halves are reduced by the caller
        for (int j = 0; j < 6; ++j) { d.lcen[j] = t.cbox[j]; d.rcen[j] = t.cbox[j]; }
    }
}

// Box reduction of the two halves of a middle split, by the whole workgroup;
// `half` is LDS of 12 words.  Ends with the boxes in d.
__device__ void median_boxes(const Level& L, const BuildTask& t, Decision& d, uint32_t* half, uint32_t tid) {
    if (tid == 0)
        for (int h = 0; h < 2; ++h)
            for (int j = 0; j < 3; ++j) { half[6 * h + j] = kMappedPosInf; half[6 * h + 3 + j] = kMappedNegInf; }
    __syncthreads();
    for (uint32_t i = t.begin + tid; i < t.end; i += kThreads) {
        const uint32_t p = L.perm_in[i];
        const int h = (i - t.begin) < d.nleft ? 0 : 1;
        for (int j = 0; j < 3; ++j) {
            atomicMin(&half[6 * h + j], map_f(L.pbox[6 * (size_t)p + j]));
            atomicMax(&half[6 * h + 3 + j], map_f(L.pbox[6 * (size_t)p + 3 + j]));
        }
    }
    __syncthreads();
    if (tid == 0)
        for (int j = 0; j < 6; ++j) {
            d.lbox[j] = unmap_f(half[j]);
            d.rbox[j] = unmap_f(half[6 + j]);
        }
}

__device__ void count_large_child(const Level& L, uint32_t n) {
    if (n > kLarge) {
        atomicAdd(L.next_large, 1u);
        atomicAdd(L.next_chunks, (n + kChunk - 1) / kChunk);
    }
}

// One lane: the binary node `id` of a split and its reference in the parent.
__device__ void write_node(const Level& L, const BuildTask& t, const Decision& d, uint32_t id) {
    DNode nd{};
    DBox lb, rb;
    for (int j = 0; j < 3; ++j) {
        lb.lo[j] = d.lbox[j]; lb.hi[j] = d.lbox[3 + j];
        rb.lo[j] = d.rbox[j]; rb.hi[j] = d.rbox[3 + j];
    }
    set_child(nd, 0, lb, kEmptyRef);  // references written by the children
    set_child(nd, 1, rb, kEmptyRef);
    L.bnodes[id] = nd;
    write_ref(L, t.slot, id);
}

// One lane: write_node() at a new index; returns it.
__device__ uint32_t emit_node(const Level& L, const BuildTask& t, const Decision& d) {
    const uint32_t id = atomicAdd(L.bcount, 1u);
    write_node(L, t, d, id);
    return id;
}

// The two child tasks of node `id`, split as d says.
__device__ void child_tasks(const BuildTask& t, const Decision& d, uint32_t id, BuildTask& c0, BuildTask& c1) {
    c0.begin = t.begin; c0.end = t.begin + d.nleft; c0.slot = 2 * id; c0.mesh = t.mesh;
    c1.begin = t.begin + d.nleft; c1.end = t.end; c1.slot = 2 * id + 1; c1.mesh = t.mesh;
    for (int j = 0; j < 6; ++j) {
        c0.box[j] = d.lbox[j]; c0.cbox[j] = d.lcen[j];
        c1.box[j] = d.rbox[j]; c1.cbox[j] = d.rcen[j];
    }
}

// One lane: the node of a split, its reference in the parent and the two
// child tasks, for the next level or, when small, for the subtree kernel.
__device__ void emit_split(const Level& L, const BuildTask& t, const Decision& d) {
    const uint32_t id = emit_node(L, t, d);
    BuildTask ch[2];
    child_tasks(t, d, id, ch[0], ch[1]);
    for (int k = 0; k < 2; ++k) {
        const uint32_t n = ch[k].end - ch[k].begin;
        if (n <= kSmall) {
            L.small[atomicAdd(L.small_count, 1u)] = SmallTask{ch[k], L.out_parity, L.level + 1, {0u, 0u}};
        } else {
            L.out[atomicAdd(L.out_count, 1u)] = ch[k];
            count_large_child(L, n);
        }
    }
}

__device__ __forceinline__ bool goes_left(const Level& L, const BuildTask& t, const Decision& d, const float* sc, uint32_t i,
                                          uint32_t p) {
    if (d.kind == 2) return i - t.begin < d.nleft;
    return bin_of(L.pcen[3 * (size_t)p + d.axis], t.cbox[d.axis], sc[d.axis]) <= d.bin;
}

// Stable partition of perm_in[begin, end): the left ones go to perm_out
// from `left_at` on, the right ones from `right_at` on, each in range order.
// Block-uniform (barriers inside); `wsum` is LDS of kThreads / 64 words.
__device__ void partition_range(const Level& L, const BuildTask& t, const Decision& d, const float* sc, uint32_t begin,
                                uint32_t end, uint32_t left_at, uint32_t right_at, uint32_t* wsum, uint32_t tid) {
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    for (uint32_t s = begin; s < end; s += kThreads) {
        const uint32_t i = s + tid;
        const bool valid = i < end;
        const uint32_t p = valid ? L.perm_in[i] : 0u;
        const bool left = valid && goes_left(L, t, d, sc, i, p);
        const uint64_t mask = __ballot(left);
        const uint32_t below = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t woff = 0, total = 0;
        for (uint32_t w = 0; w < kThreads / 64; ++w) {
            if (w < wave) woff += wsum[w];
            total += wsum[w];
        }
        if (valid) {
            const uint32_t lrank = woff + below;
            if (left) L.perm_out[left_at + lrank] = p;
            else L.perm_out[right_at + (tid - lrank)] = p;
        }
        const uint32_t chunk = min(kThreads, end - s);
        left_at += total;
        right_at += chunk - total;
        __syncthreads();
    }
}

// The record of triangle p (of mesh `mesh`) at leaf position i.
__device__ __forceinline__ void write_tri(const Level& L, uint32_t i, uint32_t p, uint32_t mesh) {
    const float* p0 = L.pos + 3 * (size_t)L.gidx[3 * (size_t)p];
    const float* p1 = L.pos + 3 * (size_t)L.gidx[3 * (size_t)p + 1];
    const float* p2 = L.pos + 3 * (size_t)L.gidx[3 * (size_t)p + 2];
    DTri tr{};
    for (int c = 0; c < 3; ++c) { tr.v0[c] = p0[c]; tr.e1[c] = p1[c] - p0[c]; tr.e2[c] = p2[c] - p0[c]; }
    tr.prim = p - L.tofs[mesh];
    L.tris[i] = tr;
}

// One lane: the reference of leaf t in the parent's slot.
__device__ void leaf_ref(const Level& L, const BuildTask& t);

// Leaf: triangle records in place and the reference in the parent's slot.
__device__ void emit_leaf(const Level& L, const BuildTask& t, uint32_t tid) {
    for (uint32_t i = t.begin + tid; i < t.end; i += kThreads) write_tri(L, i, L.perm_in[i], t.mesh);
    if (tid == 0) leaf_ref(L, t);
}

__device__ void leaf_ref(const Level& L, const BuildTask& t) {
    const uint32_t n = t.end - t.begin;
    {
        const uint32_t ref = make_leaf(t.begin, n);
        if (t.slot & kRootSlot) {
            // a mesh that is one leaf: a node with the leaf and an empty child, as on the host
            const uint32_t id = atomicAdd(L.bcount, 1u);
            DNode nd{};
            DBox b, none;
            for (int j = 0; j < 3; ++j) { b.lo[j] = t.box[j]; b.hi[j] = t.box[3 + j]; }
            none.clear();
            set_child(nd, 0, b, ref);
            set_child(nd, 1, none, kEmptyRef);
            L.bnodes[id] = nd;
            L.mesh_root2[t.slot & ~kRootSlot] = id;
        } else {
            write_ref(L, t.slot, ref);
        }
    }
}

// One workgroup per node of the level.  Nodes of more than kLarge
// triangles are only registered here (their chunks and cleared global bins)
// and handled by the k_*_large kernels.
__global__ void __launch_bounds__(kThreads) k_build_level(Level L) {
    __shared__ Bins bins;
    __shared__ Decision dec;
    __shared__ uint32_t half[12];
    __shared__ uint32_t wsum[kThreads / 64];
    __shared__ uint32_t reg[2];
    const BuildTask t = L.in[blockIdx.x];
    const uint32_t n = t.end - t.begin;
    const uint32_t tid = threadIdx.x;
    if (n > kLarge) {
        const uint32_t nch = (n + kChunk - 1) / kChunk;
        if (tid == 0) {
            reg[0] = atomicAdd(L.large_count, 1u);
            reg[1] = atomicAdd(L.chunk_count, nch);
            LargeTask lt{};
            lt.task = blockIdx.x;
            lt.first_chunk = reg[1];
            lt.n_chunks = nch;
            L.large[reg[0]] = lt;
        }
        __syncthreads();
        for (uint32_t k = tid; k < nch; k += kThreads)
            L.chunks[reg[1] + k] = Chunk{reg[0], t.begin + k * kChunk, min(t.end, t.begin + (k + 1) * kChunk), 0u};
        bins_clear(L.gbins + reg[0], tid, kThreads);
        return;
    }
    float ext[3], sc[3];
    task_scales(t, ext, sc);
    if (n > 1) {
        bins_clear(&bins, tid, kThreads);
        __syncthreads();
        bin_range(&bins, L, t, t.begin, t.end, ext, sc, tid);
        __syncthreads();
    }
    if (tid < 64) decide_wave(&bins, t, ext, dec, tid);
    __syncthreads();
    if (dec.kind == 0) {
        emit_leaf(L, t, tid);
        return;
    }
    if (dec.kind == 2) median_boxes(L, t, dec, half, tid);
    partition_range(L, t, dec, sc, t.begin, t.end, t.begin, t.begin + dec.nleft, wsum, tid);
    if (tid == 0) emit_split(L, t, dec);
}

// Subtrees of at most kSmall triangles, one wave each, depth first: lane j
// holds the triangle at position root.begin + j (its index, centroid and
// box in registers); every node is binned in the wave's LDS, decided by
// the wave with decide_wave(), and partitioned stably by moving the
// lanes' triangles (ballot ranks, then a gather through LDS), so the
// subtree is the one the level-by-level kernels would build.  The triangle
// records are written at their final positions once the subtree is done.  `max_level` receives
// the deepest level reached (in the level-by-level numbering).
// Node indices: every split divides at least two triangles into two
// non-empty parts, so a subtree of n triangles has at most n - 1 binary
// nodes; the wave reserves that many with one atomic (one per node on the
// shared counter serialises the whole launch) and adds the ones it leaves
// unused to `unused`, so the node count stays exact.  Unused indices are
// never referenced.
struct SubtreeLds {
    Bins bins;
    Decision dec;
    uint32_t half[12];
    uint32_t src[64];
    SmallTask stack[kSmall];
};
constexpr uint32_t kWaves = kThreads / 64;

// A task that every lane of the wave holds alike, moved to scalar registers.
__device__ __forceinline__ void to_scalar(SmallTask& s) {
    static_assert(sizeof(SmallTask) % 4 == 0, "SmallTask is whole words");
    uint32_t w[sizeof(SmallTask) / 4];
    __builtin_memcpy(w, &s, sizeof(SmallTask));
    for (uint32_t k = 0; k < sizeof(SmallTask) / 4; ++k) w[k] = __builtin_amdgcn_readfirstlane(w[k]);
    __builtin_memcpy(&s, w, sizeof(SmallTask));
}

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) k_build_small(Level L, const SmallTask* tasks, uint32_t n_tasks,
                                                         const uint32_t* perm0, const uint32_t* perm1,
                                                         uint32_t* max_level, uint32_t* unused) {
    __shared__ SubtreeLds lds[kWaves];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t ti = blockIdx.x * kWaves + wave;
    if (ti >= n_tasks) return;  // the whole wave; no workgroup barrier below
    SubtreeLds& W = lds[wave];
    SmallTask cur = tasks[ti];
    const uint32_t base = cur.t.begin, n0 = cur.t.end - cur.t.begin, mesh = cur.t.mesh;
    const uint32_t* pin = cur.parity ? perm1 : perm0;
    const bool has = lane < n0;
    uint32_t p = has ? pin[base + lane] : 0u;
    float c[3] = {0.0f, 0.0f, 0.0f}, bx[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (has) {
        for (int k = 0; k < 3; ++k) c[k] = L.pcen[3 * (size_t)p + k];
        for (int k = 0; k < 6; ++k) bx[k] = L.pbox[6 * (size_t)p + k];
    }
    uint32_t first_id = 0, used = 0;
    if (lane == 0 && n0 > 1) first_id = atomicAdd(L.bcount, n0 - 1);
    first_id = __shfl(first_id, 0);
    uint32_t sp = 0, deepest = 0;
    for (;;) {
        to_scalar(cur);
        const BuildTask& t = cur.t;
        deepest = max(deepest, cur.level);
        const uint32_t lo = t.begin - base, hi = t.end - base, n = hi - lo;
        const bool mine = lane >= lo && lane < hi;
        float ext[3], sc[3];
        task_scales(t, ext, sc);
        if (n > 1) {
            bins_clear(&W.bins, lane, 64u);
            __builtin_amdgcn_wave_barrier();
            if (mine) bin_tri(&W.bins, t, c, bx, ext, sc);
            __builtin_amdgcn_wave_barrier();
        }
        decide_wave(&W.bins, t, ext, W.dec, lane);
        __builtin_amdgcn_wave_barrier();
        const int kind = W.dec.kind;
        if (kind == 0) {
            if (lane == 0) leaf_ref(L, t);
            if (sp == 0) break;
            cur = W.stack[--sp];
            continue;
        }
        const uint32_t nleft = W.dec.nleft;
        if (kind == 2) {
            // boxes of the two halves of a middle split
            if (lane < 12) W.half[lane] = (lane % 6) < 3 ? kMappedPosInf : kMappedNegInf;
            __builtin_amdgcn_wave_barrier();
            if (mine) {
                const int h = (lane - lo) < nleft ? 0 : 1;
                for (int j = 0; j < 3; ++j) {
                    atomicMin(&W.half[6 * h + j], map_f(bx[j]));
                    atomicMax(&W.half[6 * h + 3 + j], map_f(bx[3 + j]));
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (lane == 0)
                for (int j = 0; j < 6; ++j) {
                    W.dec.lbox[j] = unmap_f(W.half[j]);
                    W.dec.rbox[j] = unmap_f(W.half[6 + j]);
                }
            __builtin_amdgcn_wave_barrier();
        }
        // stable partition of the lanes [lo, hi): left ones first
        // (selects rather than a run-time index keep c, cbox and sc in registers)
        const int axis = W.dec.axis;
        const uint32_t bin = W.dec.bin;
        const float ca = axis == 0 ? c[0] : (axis == 1 ? c[1] : c[2]);
        const float la = axis == 0 ? t.cbox[0] : (axis == 1 ? t.cbox[1] : t.cbox[2]);
        const float sa = axis == 0 ? sc[0] : (axis == 1 ? sc[1] : sc[2]);
        const bool left = mine && (kind == 2 ? (lane - lo) < nleft : bin_of(ca, la, sa) <= bin);
        const uint64_t lmask = __ballot(left), rmask = __ballot(mine && !left);
        const uint64_t below = (1ull << lane) - 1ull;
        uint32_t dest = lane;
        if (mine) dest = left ? lo + (uint32_t)__popcll(lmask & below) : lo + nleft + (uint32_t)__popcll(rmask & below);
        W.src[dest] = lane;
        __builtin_amdgcn_wave_barrier();
        const int from = (int)W.src[lane];
        p = __shfl(p, from);
        for (int k = 0; k < 3; ++k) c[k] = __shfl(c[k], from);
        for (int k = 0; k < 6; ++k) bx[k] = __shfl(bx[k], from);
        // the node, then its left child now and its right child later
        const uint32_t id = first_id + used++;
        if (lane == 0) write_node(L, t, W.dec, id);
        BuildTask c0, c1;
        child_tasks(t, W.dec, id, c0, c1);
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) W.stack[sp] = SmallTask{c1, cur.parity, cur.level + 1, {0u, 0u}};
        ++sp;
        cur = SmallTask{c0, cur.parity, cur.level + 1, {0u, 0u}};
        __builtin_amdgcn_wave_barrier();
    }
    // no partition moves a lane's triangle once its leaf is emitted, so lane j
    // now holds the triangle of position root.begin + j: all records at once
    if (has) write_tri(L, base + lane, p, mesh);
    if (lane == 0) {
        atomicMax(max_level, deepest);
        if (n0 > 1 + used) atomicAdd(unused, n0 - 1 - used);
    }
}

// Large nodes, step 1: one workgroup bins a chunk in LDS and merges the
// bins into the node's global bins.
__global__ void __launch_bounds__(kThreads) k_bin_large(Level L) {
    __shared__ Bins bins;
    const Chunk ch = L.chunks[blockIdx.x];
    const BuildTask t = L.in[L.large[ch.large].task];
    const uint32_t tid = threadIdx.x;
    float ext[3], sc[3];
    task_scales(t, ext, sc);
    bins_clear(&bins, tid, kThreads);
    __syncthreads();
    bin_range(&bins, L, t, ch.begin, ch.end, ext, sc, tid);
    __syncthreads();
    Bins* g = L.gbins + ch.large;
    for (uint32_t i = tid; i < kBinCells; i += kThreads) {
        const uint32_t c = (&bins.cnt[0][0])[i];
        if (c == 0) continue;
        atomicAdd(&g->cnt[0][0] + i, c);
        for (int k = 0; k < 3; ++k) {
            atomicMin(&g->box[0][0][0] + 6 * i + k, (&bins.box[0][0][0])[6 * i + k]);
            atomicMax(&g->box[0][0][0] + 6 * i + 3 + k, (&bins.box[0][0][0])[6 * i + 3 + k]);
            atomicMin(&g->cen[0][0][0] + 6 * i + k, (&bins.cen[0][0][0])[6 * i + k]);
            atomicMax(&g->cen[0][0][0] + 6 * i + 3 + k, (&bins.cen[0][0][0])[6 * i + 3 + k]);
        }
    }
}

// Step 2: one workgroup per large node takes the decision (a large node is
// never a leaf) and emits the node and its children.
__global__ void __launch_bounds__(kThreads) k_split_large(Level L) {
    __shared__ Decision dec;
    __shared__ uint32_t half[12];
    LargeTask* lt = L.large + blockIdx.x;
    const BuildTask t = L.in[lt->task];
    const uint32_t tid = threadIdx.x;
    float ext[3], sc[3];
    task_scales(t, ext, sc);
    if (tid < 64) decide_wave(L.gbins + blockIdx.x, t, ext, dec, tid);
    __syncthreads();
    if (dec.kind == 2) median_boxes(L, t, dec, half, tid);
    __syncthreads();
    if (tid == 0) {
        lt->dec = dec;
        emit_split(L, t, dec);
    }
}

// Step 3: left count of every chunk.
__global__ void __launch_bounds__(kThreads) k_count_large(Level L) {
    __shared__ uint32_t wsum[kThreads / 64];
    const Chunk ch = L.chunks[blockIdx.x];
    const LargeTask& lt = L.large[ch.large];
    const BuildTask t = L.in[lt.task];
    const Decision d = lt.dec;
    const uint32_t tid = threadIdx.x;
    float ext[3], sc[3];
    task_scales(t, ext, sc);
    uint32_t cnt = 0;
    for (uint32_t i = ch.begin + tid; i < ch.end; i += kThreads) cnt += goes_left(L, t, d, sc, i, L.perm_in[i]) ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((tid & 63u) == 0) wsum[tid >> 6] = cnt;
    __syncthreads();
    if (tid == 0) {
        uint32_t s = 0;
        for (uint32_t w = 0; w < kThreads / 64; ++w) s += wsum[w];
        L.chunk_left[blockIdx.x] = s;
    }
}

// Step 4: stable scatter of every chunk after the earlier chunks of its node.
__global__ void __launch_bounds__(kThreads) k_scatter_large(Level L) {
    __shared__ uint32_t wsum[kThreads / 64];
    __shared__ uint32_t before[2];
    const Chunk ch = L.chunks[blockIdx.x];
    const LargeTask& lt = L.large[ch.large];
    const BuildTask t = L.in[lt.task];
    const Decision d = lt.dec;
    const uint32_t tid = threadIdx.x;
    float ext[3], sc[3];
    task_scales(t, ext, sc);
    uint32_t s = 0;
    for (uint32_t k = lt.first_chunk + tid; k < blockIdx.x; k += kThreads) s += L.chunk_left[k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((tid & 63u) == 0) wsum[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        uint32_t a = 0;
        for (uint32_t w = 0; w < kThreads / 64; ++w) a += wsum[w];
        before[0] = a;                          // left ones of the earlier chunks
        before[1] = (ch.begin - t.begin) - a;   // right ones of the earlier chunks
    }
    __syncthreads();
    const uint32_t lb = before[0], rb = before[1];
    partition_range(L, t, d, sc, ch.begin, ch.end, t.begin + lb, t.begin + d.nleft + rb, wsum, tid);
}

// ---- collapse to 4-wide nodes ------------------------------------------
struct Wide {
    const WideTask* in;
    uint32_t n_in;
    WideTask* out;
    uint32_t* out_count;
    const DNode* bnodes;
    DNode4* nodes4;
    uint32_t* wcount;
    uint32_t* mesh_root4;
    uint32_t* bound;
};

__device__ void child_box(const DNode& n, int c, float lo[3], float hi[3]) {
    const float* x = c == 0 ? n.c0x : n.c1x;
    const float* y = c == 0 ? n.c0y : n.c1y;
    const float* z = c == 0 ? n.c0z : n.c1z;
    lo[0] = x[0]; hi[0] = x[1]; lo[1] = y[0]; hi[1] = y[1]; lo[2] = z[0]; hi[2] = z[1];
}

__device__ float wide_area(const float lo[3], const float hi[3]) {
    float d[3];
    for (int k = 0; k < 3; ++k) {
        const float x = hi[k] - lo[k];
        d[k] = 0.0f < x ? x : 0.0f;  // std::max(0.0f, x)
    }
    return d[0] * d[1] + d[0] * d[2] + d[1] * d[2];
}

__global__ void k_roots4(const uint32_t* mesh_root2, uint32_t n_meshes, WideTask* out, uint32_t* out_count) {
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < n_meshes; m += gridDim.x * blockDim.x)
        out[m] = WideTask{mesh_root2[m], kRootSlot | m, 0u, 0u};
    if (blockIdx.x == 0 && threadIdx.x == 0) *out_count = n_meshes;
}

// the host WideBuilder::collapse, one wide node per lane
__global__ void k_wide_level(Wide W) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < W.n_in; i += gridDim.x * blockDim.x) {
        const WideTask w = W.in[i];
        if (w.bnode >= kImplicitRoot) {  // a built-in shape's root passes through
            W.mesh_root4[w.slot & ~kRootSlot] = w.bnode;
            continue;
        }
        uint32_t ref[4];
        float lo[4][3], hi[4][3];
        int nc = 0;
        auto add_children = [&](uint32_t id) {
            const DNode nd = W.bnodes[id];
            for (int c = 0; c < 2; ++c) {
                if (nd.ref[c] == kEmptyRef) continue;
                ref[nc] = nd.ref[c];
                child_box(nd, c, lo[nc], hi[nc]);
                ++nc;
            }
        };
        add_children(w.bnode);
        while (nc < 4) {
            int best = -1;
            float ba = -1.0f;
            for (int k = 0; k < nc; ++k)
                if (!(ref[k] & kLeafBit) && wide_area(lo[k], hi[k]) > ba) { ba = wide_area(lo[k], hi[k]); best = k; }
            if (best < 0) break;
            const uint32_t open = ref[best];
            --nc;
            ref[best] = ref[nc];
            for (int k = 0; k < 3; ++k) { lo[best][k] = lo[nc][k]; hi[best][k] = hi[nc][k]; }
            add_children(open);
        }
        const uint32_t me = atomicAdd(W.wcount, 1u);
        DNode4 nd;
        quantise_children(lo, hi, nc, nd);
        const uint32_t acc = w.acc + (nc > 0 ? (uint32_t)nc - 1u : 0u);
        for (int k = 0; k < nc; ++k) {
            if (ref[k] & kLeafBit) {
                nd.ref[k] = ref[k];
            } else {
                nd.ref[k] = kEmptyRef;  // written by the child's task
                W.out[atomicAdd(W.out_count, 1u)] = WideTask{ref[k], me * 4u + (uint32_t)k, acc, 0u};
            }
        }
        W.nodes4[me] = nd;
        if (w.slot & kRootSlot) W.mesh_root4[w.slot & ~kRootSlot] = me;
        else W.nodes4[w.slot >> 2].ref[w.slot & 3] = me;
        atomicMax(W.bound, acc);
    }
}

__global__ void k_instance_roots(DInstance* inst, const uint32_t* inst_mesh, const uint32_t* mesh_root4, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        inst[i].blas_root = mesh_root4[inst_mesh[i]];
}

uint32_t grid_for(size_t n) { return (uint32_t)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }
// one lane per run of kRun items, every item covered
uint32_t runs_grid(size_t n) { return (uint32_t)std::max<size_t>(1, (n + 256 * (size_t)kRun - 1) / (256 * (size_t)kRun)); }

// Device allocations of the build itself, freed when it ends.
struct Temp {
    std::vector<void*> p;
    ~Temp() {
        for (void* x : p) hipFree(x);
    }
    template <class T>
    T* get(size_t count) {
        void* x = nullptr;
        if (hipMalloc(&x, std::max<size_t>(count * sizeof(T), 64)) != hipSuccess) return nullptr;
        p.push_back(x);
        return static_cast<T*>(x);
    }
};

}  // namespace

namespace pico {

// The scene arrays live in the context's scene allocations (freed with the scene).
template <class T>
static T* scene_alloc(pico_hip_ctx* c, size_t count) {
    void* x = nullptr;
    if (hipMalloc(&x, std::max<size_t>(count * sizeof(T), 64)) != hipSuccess) return nullptr;
    c->scene_allocs.push_back(x);
    return static_cast<T*>(x);
}

template <class T>
static T* scene_upload(pico_hip_ctx* c, const T* src, size_t count, hipError_t& e) {
    T* d = scene_alloc<T>(c, count);
    if (!d) { e = hipErrorOutOfMemory; return nullptr; }
    if (count) e = hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, c->stream);
    return d;
}

static void drop_scene(pico_hip_ctx* c) {
    hipStreamSynchronize(c->stream);
    for (void* p : c->scene_allocs) hipFree(p);
    c->scene_allocs.clear();
    c->has_scene = false;
}

static int build_device(pico_hip_ctx* c, const pico_scene_desc& d, pico_build_stats* stats) {
    const auto t0 = std::chrono::steady_clock::now();
    std::string e = check_scene(d);
    if (!e.empty()) return fail(c, PICO_ERR_ARG, e);
    HostScene hs;
    e = scene_tables(d, hs);
    if (!e.empty()) return fail(c, PICO_ERR_ARG, e);
    const uint32_t nm = d.n_meshes, nv = d.mesh_vertex_offset[nm], nt = d.mesh_triangle_offset[nm];
    const uint32_t ni = d.n_instances;
    // scene arrays
    DevScene s{};
    hipError_t he = hipSuccess;
    const size_t node_cap = (size_t)nt + nm + std::max<uint32_t>(ni > 1 ? ni - 1 : 1, 1);
    DNode4* nodes4 = scene_alloc<DNode4>(c, node_cap);
    DTri* tris = scene_alloc<DTri>(c, nt);
    uint32_t* gidx = scene_alloc<uint32_t>(c, 3 * (size_t)nt);
    float* mesh_box = scene_alloc<float>(c, 6 * (size_t)nm);
    if (!nodes4 || !tris || !gidx || !mesh_box) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    s.nodes = nodes4;
    s.tris = tris;
    s.idx = gidx;
    s.pos = scene_upload(c, d.positions, 3 * (size_t)nv, he);
    if (he == hipSuccess) s.nrm = scene_upload(c, d.normals, 3 * (size_t)nv, he);
    if (he == hipSuccess) s.inst = scene_upload(c, hs.instances.data(), hs.instances.size(), he);
    if (he == hipSuccess) s.xf = scene_upload(c, hs.xforms.data(), hs.xforms.size(), he);
    if (he == hipSuccess) s.mats = scene_upload(c, hs.materials.data(), hs.materials.size(), he);
    if (he == hipSuccess) s.lights = scene_upload(c, hs.lights.data(), hs.lights.size(), he);
    if (he == hipSuccess) s.alias_w = scene_upload(c, hs.alias_weight.data(), hs.alias_weight.size(), he);
    if (he == hipSuccess) s.alias_i = scene_upload(c, hs.alias_index.data(), hs.alias_index.size(), he);
    const uint32_t* inst_mesh = he == hipSuccess ? scene_upload(c, hs.instance_mesh.data(), hs.instance_mesh.size(), he) : nullptr;
    const std::vector<uint32_t> vrange = mesh_vertex_ranges(d);
    const uint32_t* mesh_vrange = he == hipSuccess ? scene_upload(c, vrange.data(), vrange.size(), he) : nullptr;
    if (he != hipSuccess) return fail(c, PICO_ERR_HIP, std::string("scene upload: ") + hipGetErrorString(he));
    // build inputs and work arrays
    Temp tmp;
    uint32_t* lidx = tmp.get<uint32_t>(3 * (size_t)nt);
    uint32_t* vofs = tmp.get<uint32_t>(nm + 1);
    uint32_t* tofs = tmp.get<uint32_t>(nm + 1);
    float* xform = tmp.get<float>(16 * (size_t)ni);
    uint32_t* shapes = d.mesh_shape ? tmp.get<uint32_t>(nm) : nullptr;
    float* pbox = tmp.get<float>(6 * (size_t)nt);
    float* pcen = tmp.get<float>(3 * (size_t)nt);
    uint32_t* perm[2] = {tmp.get<uint32_t>(nt), tmp.get<uint32_t>(nt)};
    BuildTask* tasks[2] = {tmp.get<BuildTask>(std::max<uint32_t>(nt, nm)), tmp.get<BuildTask>(std::max<uint32_t>(nt, nm))};
    DNode* bnodes = tmp.get<DNode>((size_t)nt + nm);
    uint32_t* mbox = tmp.get<uint32_t>(6 * (size_t)nm);
    uint32_t* rbox = tmp.get<uint32_t>(12 * (size_t)nm);
    uint32_t* roots = tmp.get<uint32_t>(2 * (size_t)nm);
    WideTask* wtasks[2] = {tmp.get<WideTask>((size_t)nt + nm), tmp.get<WideTask>((size_t)nt + nm)};
    // [0] err [1] binary nodes [2,3] task counts [4] wide nodes [5] stack bound [6,7] wide task counts
    // [8] large nodes registered [9] their chunks [10] large children [11] their chunks
    // [12] small subtrees [13] deepest level they reach [14] binary node indices they reserved and left unused
    uint32_t* counters = tmp.get<uint32_t>(16);
    SmallTask* small = tmp.get<SmallTask>(std::max<uint32_t>(nt, 1));
    const size_t large_cap = nt / kLarge + 2, chunk_cap = nt / kChunk + large_cap;
    LargeTask* large = tmp.get<LargeTask>(large_cap);
    Bins* gbins = tmp.get<Bins>(large_cap);
    Chunk* chunks = tmp.get<Chunk>(chunk_cap);
    uint32_t* chunk_left = tmp.get<uint32_t>(chunk_cap);
    if (!lidx || !vofs || !tofs || !xform || !pbox || !pcen || !perm[0] || !perm[1] || !tasks[0] || !tasks[1] || !bnodes ||
        !mbox || !rbox || !roots || !wtasks[0] || !wtasks[1] || !counters || !large || !gbins || !chunks || !chunk_left ||
        !small)
        return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    PICO_HIP(c, hipMemcpyAsync(lidx, d.indices, 12 * (size_t)nt, hipMemcpyHostToDevice, c->stream));
    PICO_HIP(c, hipMemcpyAsync(vofs, d.mesh_vertex_offset, 4 * ((size_t)nm + 1), hipMemcpyHostToDevice, c->stream));
    PICO_HIP(c, hipMemcpyAsync(tofs, d.mesh_triangle_offset, 4 * ((size_t)nm + 1), hipMemcpyHostToDevice, c->stream));
    PICO_HIP(c, hipMemcpyAsync(xform, d.transforms, 64 * (size_t)ni, hipMemcpyHostToDevice, c->stream));
    if (d.mesh_shape) {
        if (!shapes) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
        PICO_HIP(c, hipMemcpyAsync(shapes, d.mesh_shape, 4 * (size_t)nm, hipMemcpyHostToDevice, c->stream));
    }
    PICO_HIP(c, hipMemsetAsync(counters, 0, 64, c->stream));
    hipEvent_t e0, e1;
    PICO_HIP(c, hipEventCreate(&e0));
    PICO_HIP(c, hipEventCreate(&e1));
    struct Ev { hipEvent_t a, b; ~Ev() { hipEventDestroy(a); hipEventDestroy(b); } } ev{e0, e1};
    PICO_HIP(c, hipEventRecord(e0, c->stream));
    float4* tshade = scene_alloc<float4>(c, kShadeRec * (size_t)nt);
    if (!tshade) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    s.tshade = tshade;
    Geo g{s.pos, lidx, gidx, vofs, tofs, nm, nt, nv, pbox, pcen, perm[0], mbox, rbox, counters, shapes, s.nrm, tshade};
    PICO_LAUNCH(c, "bvh_geo_clear", k_geo_clear, grid_for(nm), 256, 0, g);
    PICO_LAUNCH(c, "bvh_geo_tris", k_geo_tris, grid_for(nt), 256, 0, g);
    PICO_LAUNCH(c, "bvh_tri_bounds", k_tri_bounds, runs_grid(nt), 256, 0, g);
    PICO_LAUNCH(c, "bvh_vert_bounds", k_vert_bounds, runs_grid(nv), 256, 0, g);
    uint32_t* mesh_root2 = roots;
    uint32_t* mesh_root4 = roots + nm;
    PICO_LAUNCH(c, "bvh_roots", k_roots, grid_for(nm), 256, 0, g, tasks[0], counters + 2, bnodes, counters + 1, mesh_root2,
                mesh_box, counters + 10, counters + 11);
    uint32_t h[12] = {};
    PICO_HIP(c, hipMemcpyAsync(h, counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    if (h[0]) return fail(c, PICO_ERR_ARG, "triangle index out of range");
    uint32_t n_tasks = h[2], n_large = h[10], n_chunks = h[11];
    // binary levels
    uint32_t levels = 0, cur = 0;
    while (n_tasks > 0) {
        if (++levels > kMaxLevels) return fail(c, PICO_ERR_LIMIT, "BVH build did not converge");
        if (n_large > large_cap || n_chunks > chunk_cap) return fail(c, PICO_ERR_LIMIT, "BVH build: work lists overflow");
        uint32_t* next_count = counters + 2 + (1 - cur);
        PICO_HIP(c, hipMemsetAsync(next_count, 0, 4, c->stream));
        PICO_HIP(c, hipMemsetAsync(counters + 8, 0, 16, c->stream));
        Level L{tasks[cur], tasks[1 - cur], next_count, perm[cur], perm[1 - cur], pbox, pcen, s.pos, gidx, tofs,
                bnodes, counters + 1, mesh_root2, tris, large, counters + 8, chunks, counters + 9, gbins, chunk_left,
                counters + 10, counters + 11, small, counters + 12, 1u - cur, levels};
        PICO_LAUNCH(c, "bvh_build_level", k_build_level, n_tasks, kThreads, 0, L);
        if (n_large) {
            PICO_LAUNCH(c, "bvh_bin_large", k_bin_large, n_chunks, kThreads, 0, L);
            PICO_LAUNCH(c, "bvh_split_large", k_split_large, n_large, kThreads, 0, L);
            PICO_LAUNCH(c, "bvh_count_large", k_count_large, n_chunks, kThreads, 0, L);
            PICO_LAUNCH(c, "bvh_scatter_large", k_scatter_large, n_chunks, kThreads, 0, L);
        }
        PICO_HIP(c, hipMemcpyAsync(&n_tasks, next_count, 4, hipMemcpyDeviceToHost, c->stream));
        PICO_HIP(c, hipMemcpyAsync(h + 10, counters + 10, 8, hipMemcpyDeviceToHost, c->stream));
        PICO_HIP(c, hipStreamSynchronize(c->stream));
        n_large = h[10];
        n_chunks = h[11];
        cur = 1 - cur;
    }
    // the small subtrees, one wave each
    uint32_t hsm[2] = {0, 0};
    PICO_HIP(c, hipMemcpyAsync(hsm, counters + 12, 8, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    if (hsm[0] > std::max<uint32_t>(nt, 1)) return fail(c, PICO_ERR_LIMIT, "BVH build: work lists overflow");
    if (hsm[0]) {
        Level L{nullptr, nullptr, nullptr, nullptr, nullptr, pbox, pcen, s.pos, gidx, tofs, bnodes, counters + 1,
                mesh_root2, tris, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                nullptr, 0u, 0u};
        PICO_LAUNCH(c, "bvh_build_small", k_build_small, (hsm[0] + kWaves - 1) / kWaves, kThreads, 0, L, small, hsm[0],
                    perm[0], perm[1], counters + 13, counters + 14);
        PICO_HIP(c, hipMemcpyAsync(hsm + 1, counters + 13, 4, hipMemcpyDeviceToHost, c->stream));
        PICO_HIP(c, hipStreamSynchronize(c->stream));
        levels = std::max(levels, hsm[1]);
    }
    uint32_t nb[14] = {};
    PICO_HIP(c, hipMemcpyAsync(nb, counters + 1, sizeof(nb), hipMemcpyDeviceToHost, c->stream));
    // 4-wide collapse
    PICO_LAUNCH(c, "bvh_roots4", k_roots4, grid_for(nm), 256, 0, mesh_root2, nm, wtasks[0], counters + 6);
    uint32_t n_wide_tasks = nm;
    cur = 0;
    while (n_wide_tasks > 0) {
        uint32_t* next_count = counters + 6 + (1 - cur);
        PICO_HIP(c, hipMemsetAsync(next_count, 0, 4, c->stream));
        Wide W{wtasks[cur], n_wide_tasks, wtasks[1 - cur], next_count, bnodes, nodes4, counters + 4, mesh_root4,
               counters + 5};
        PICO_LAUNCH(c, "bvh_wide_level", k_wide_level, grid_for(n_wide_tasks), 256, 0, W);
        PICO_HIP(c, hipMemcpyAsync(&n_wide_tasks, next_count, 4, hipMemcpyDeviceToHost, c->stream));
        PICO_HIP(c, hipStreamSynchronize(c->stream));
        cur = 1 - cur;
    }
    PICO_LAUNCH(c, "bvh_instance_roots", k_instance_roots, grid_for(ni), 256, 0, const_cast<DInstance*>(s.inst), inst_mesh,
                mesh_root4, ni);
    PICO_HIP(c, hipEventRecord(e1, c->stream));
    uint32_t wb[2] = {0, 0};
    PICO_HIP(c, hipMemcpyAsync(wb, counters + 4, 8, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    float blas_ms = 0.0f;
    hipEventElapsedTime(&blas_ms, e0, e1);
    const uint32_t n_blas4 = wb[0], blas_bound = wb[1];
    if (blas_bound + 1 > 64) return fail(c, PICO_ERR_LIMIT, "BVH too deep for the 64-entry traversal stack");
    // the TLAS, by the rebuild path, into the node array after the BLAS
    s.n_lights = (uint32_t)hs.lights.size();
    s.microfacet = hs.microfacet;
    if (int rc = upload_sky_sun(c, hs, s)) return rc;
    c->dev = s;
    c->d_mesh_box = mesh_box;
    c->d_mesh_vrange = mesh_vrange;
    c->max_mesh_vertices = max_range(vrange);
    c->d_instance_mesh = inst_mesh;
    c->tlas_base = n_blas4;
    c->blas_bound = blas_bound;
    pico_rebuild_stats ts{};
    int rc = tlas_rebuild(c, ni, xform, &ts);
    if (rc) return rc;
    c->n_tris = nt;
    c->blas_depth = std::max(levels, 2u) - 1;  // levels of internal nodes (the last level holds only leaves)
    hs.n_instances = ni;
    c->host = std::move(hs);
    if (stats) {
        *stats = pico_build_stats{};
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->bvh_ms = blas_ms + ts.device_ms;
        stats->binary_nodes = nb[0] - nb[13];  // allocated less reserved-but-unused
        stats->blas_nodes4 = n_blas4;
        stats->blas_bound = blas_bound;
        stats->stack_bound = ts.stack_bound;
        stats->levels = c->blas_depth;
    }
    return PICO_OK;
}

int upload_scene_device(pico_hip_ctx* c, const pico_scene_desc& d, pico_build_stats* stats) {
    const int rc = build_device(c, d, stats);
    if (rc) {
        drop_scene(c);
        return rc;
    }
    c->has_scene = true;
    ++c->scene_serial;
    return PICO_OK;
}

}  // namespace pico
