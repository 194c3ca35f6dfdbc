// Per-frame TLAS rebuild on the device for animated instances: new
// transforms for every instance replace the instance records and the top
// level is rebuilt as a linear BVH (Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG 2012) over the instance
// world boxes, then collapsed into the 4-wide quantised nodes the traversal
// reads.  The reference builds its TLAS once, serially, on the host
// (Source/Core/UpperLevelBVH.cpp:27-57); the mesh BVHs are unchanged.
//
// Steps (one stream, one host synchronisation at the end):
//   1  per instance: affine check, float cofactor inverse (bit-identical to the
//      host build) (staged, not yet visible); world box of its mesh's vertices
//      (chunks of vertices, one workgroup each; a built-in shape: its box
//      corners), as the host build bounds an instance
//   2  centroid bounds, 30-bit Morton code | instance id as a unique 64-bit key
//   3  rocPRIM radix sort of the keys
//   4  Karras hierarchy: n - 1 binary nodes, children and parents
//   5  depth parity of every binary node: even-depth nodes become the 4-wide
//      nodes, whose children are their grandchildren (or leaf children)
//   6  bottom-up boxes (the second child to arrive at a node forms the union)
//      and traversal-stack bounds of the 4-wide nodes
//   7  wide nodes numbered by a rocPRIM exclusive scan of the parity flags,
//      then quantised into staged DNode4 records
//   8  commit: when every transform was valid and the traversal stack bound
//      fits, copy the staged instances and nodes into the scene
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "../../include/pico_hip.h"
#include "ctx.h"
#include "dev_atomics.h"
#include "scene_math.h"

using namespace pico;

namespace {

struct TlasBuf {
    const float* xform;       // n x 16, column-major
    const uint32_t* mesh;     // instance -> mesh
    const float* mesh_box;    // mesh -> lo[3], hi[3]
    const uint32_t* vrange;   // mesh -> first, end vertex (empty for a built-in shape)
    const float* pos;         // vertex positions
    float* inv;               // staged inverse rows, n x 12
    float* box;               // instance world boxes, n x 6
    uint32_t* wbox;           // the same as order-preserving integers while chunks combine
    unsigned long long* key;  // n (unsorted), then key_s sorted
    unsigned long long* key_s;
    uint32_t* child;          // (n - 1) x 2: internal index or kLeafBit | sorted leaf position
    uint32_t* parent;         // internal nodes [0, n-1), leaves [n-1, 2n-1)
    float* nbox;              // (n - 1) x 6
    uint32_t* visits;         // n - 1
    uint32_t* bnd;            // n - 1: even node: stack entries below it; odd node: max over its wide children
    uint32_t* flag;           // n - 1: 1 for even depth
    uint32_t* wide;           // n - 1: exclusive scan of flag
    DNode4* staged;           // wide nodes
    uint32_t* info;           // [0] bad transform, [1] max depth, [2..7] centroid bounds (float bits)
    uint32_t n;
    uint32_t base;            // DNode4 index of the TLAS region
};

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void k_tlas_init(TlasBuf t) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        t.info[0] = 0;
        t.info[1] = 0;
        for (int k = 0; k < 3; ++k) { t.info[2 + k] = kMappedPosInf; t.info[5 + k] = kMappedNegInf; }
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i + 1 < t.n; i += gridDim.x * blockDim.x) t.visits[i] = 0;
}

// step 1: affine check and inverse
__global__ void k_tlas_instances(TlasBuf t) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += gridDim.x * blockDim.x) {
        const float* M = t.xform + 16 * (size_t)i;
        float inv[3][4];
        const bool affine = M[3] == 0.0f && M[7] == 0.0f && M[11] == 0.0f && M[15] == 1.0f;
        if (!affine || !affine_inverse_rows(M, inv)) {
            atomicOr(&t.info[0], 1u);
            continue;
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) t.inv[12 * (size_t)i + 4 * r + c] = inv[r][c];
    }
}

// step 1: world boxes.  An instance's vertices are split into chunks of
// kBoxChunk, one workgroup of 256 each (a single instance of a 2M-triangle
// mesh is 245 chunks); each chunk's box goes into the instance's box through
// order-preserving integer atomics, so the result is the exact minimum and
// maximum, the host build's box, whatever the order.
constexpr uint32_t kBoxChunk = 4096;

__global__ void k_tlas_wbox_init(TlasBuf t) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; ++k) { t.wbox[6 * (size_t)i + k] = kMappedPosInf; t.wbox[6 * (size_t)i + 3 + k] = kMappedNegInf; }
}

// workgroup w: instance w / chunks, chunk w % chunks of its mesh's vertices
__global__ __launch_bounds__(256) void k_tlas_world_boxes(TlasBuf t, uint32_t chunks) {
    __shared__ float red[6][4];
    const uint32_t wave = threadIdx.x >> 6;
    for (size_t w = blockIdx.x; w < (size_t)t.n * chunks; w += gridDim.x) {
        const uint32_t i = (uint32_t)(w / chunks), c = (uint32_t)(w % chunks);
        const float* M = t.xform + 16 * (size_t)i;
        const uint32_t m = t.mesh[i];
        const uint32_t v0 = t.vrange[2 * m], v1 = t.vrange[2 * m + 1];
        const uint32_t c0 = v0 + c * kBoxChunk;
        if (c0 >= v1) continue;  // uniform over the workgroup
        const uint32_t c1 = min(v1, c0 + kBoxChunk);
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t v = c0 + threadIdx.x; v < c1; v += blockDim.x) grow_world_box(M, t.pos + 3 * (size_t)v, lo, hi);
        for (int o = 32; o > 0; o >>= 1)
            for (int k = 0; k < 3; ++k) {
                lo[k] = fminf(lo[k], __shfl_xor(lo[k], o));
                hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o));
            }
        if ((threadIdx.x & 63) == 0)
            for (int k = 0; k < 3; ++k) { red[k][wave] = lo[k]; red[3 + k][wave] = hi[k]; }
        __syncthreads();
        if (threadIdx.x < 3) {
            const int k = threadIdx.x;
            atomic_min_f(&t.wbox[6 * (size_t)i + k], fminf(fminf(red[k][0], red[k][1]), fminf(red[k][2], red[k][3])));
            atomic_max_f(&t.wbox[6 * (size_t)i + 3 + k],
                         fmaxf(fmaxf(red[3 + k][0], red[3 + k][1]), fmaxf(red[3 + k][2], red[3 + k][3])));
        }
        __syncthreads();  // red is reused by the next chunk
    }
}

// per instance: the widened box of its vertices, a built-in shape's box
// corners, or an empty box at the origin for an invalid transform (the
// rebuild is then refused)
__global__ void k_tlas_world_boxes_done(TlasBuf t) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += gridDim.x * blockDim.x) {
        const float* M = t.xform + 16 * (size_t)i;
        const uint32_t m = t.mesh[i];
        float inv[3][4];
        const bool ok = M[3] == 0.0f && M[7] == 0.0f && M[11] == 0.0f && M[15] == 1.0f && affine_inverse_rows(M, inv);
        float lo[3], hi[3];
        if (!ok) {
            for (int k = 0; k < 3; ++k) { lo[k] = 0.0f; hi[k] = 0.0f; }
        } else if (t.vrange[2 * m] < t.vrange[2 * m + 1]) {
            for (int k = 0; k < 3; ++k) {
                lo[k] = unmap_f(t.wbox[6 * (size_t)i + k]);
                hi[k] = unmap_f(t.wbox[6 * (size_t)i + 3 + k]);
            }
            widen_world_box(lo, hi);
        } else {
            const float* mb = t.mesh_box + 6 * (size_t)m;
            instance_world_box(M, mb, mb + 3, lo, hi);
        }
        for (int k = 0; k < 3; ++k) {
            t.box[6 * (size_t)i + k] = lo[k];
            t.box[6 * (size_t)i + 3 + k] = hi[k];
        }
    }
}

// step 2: centroid bounds of the world boxes, one atomic per workgroup and
// bound (atomics on one address serialise)
__global__ __launch_bounds__(256) void k_tlas_centroid_bounds(TlasBuf t) {
    __shared__ float red[6][4];
    const uint32_t wave = threadIdx.x >> 6;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; ++k) {
            const float c = 0.5f * (t.box[6 * (size_t)i + k] + t.box[6 * (size_t)i + 3 + k]);
            lo[k] = fminf(lo[k], c);
            hi[k] = fmaxf(hi[k], c);
        }
    for (int o = 32; o > 0; o >>= 1)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], o));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o));
        }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) { red[k][wave] = lo[k]; red[3 + k][wave] = hi[k]; }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        const float l = fminf(fminf(red[k][0], red[k][1]), fminf(red[k][2], red[k][3]));
        const float h = fmaxf(fmaxf(red[3 + k][0], red[3 + k][1]), fmaxf(red[3 + k][2], red[3 + k][3]));
        if (l <= h) {
            atomic_min_f(&t.info[2 + k], l);
            atomic_max_f(&t.info[5 + k], h);
        }
    }
}

__global__ void k_tlas_keys(TlasBuf t) {
    float lo[3], ext[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = unmap_f(t.info[2 + k]);
        ext[k] = unmap_f(t.info[5 + k]) - lo[k];
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += gridDim.x * blockDim.x) {
        uint32_t code = 0;
        for (int k = 0; k < 3; ++k) {
            const float c = 0.5f * (t.box[6 * (size_t)i + k] + t.box[6 * (size_t)i + 3 + k]);
            const float u = ext[k] > 0.0f ? (c - lo[k]) / ext[k] : 0.0f;
            const uint32_t q = (uint32_t)fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
            code |= expand_bits(q) << (2 - k);
        }
        t.key[i] = ((unsigned long long)code << 32) | i;
    }
}

__device__ __forceinline__ int delta(const unsigned long long* k, uint32_t n, int i, int j) {
    if (j < 0 || j >= (int)n) return -1;
    return __clzll(k[i] ^ k[j]);  // keys are unique
}

// step 4: Karras 2012, Figure 4
__global__ void k_tlas_hierarchy(TlasBuf t) {
    const unsigned long long* k = t.key_s;
    const int n = (int)t.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
        const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = delta(k, n, i, i - d);
        int lmax = 2;
        while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int s = lmax / 2; s >= 1; s /= 2)
            if (delta(k, n, i, i + (l + s) * d) > dmin) l += s;
        const int j = i + l * d;
        const int dnode = delta(k, n, i, j);
        int s = 0;
        int w = l;
        do {
            w = (w + 1) / 2;
            if (delta(k, n, i, i + (s + w) * d) > dnode) s += w;
        } while (w > 1);
        const int gamma = i + s * d + min(d, 0);
        const int lo = min(i, j), hi = max(i, j);
        const uint32_t left = lo == gamma ? (kLeafBit | (uint32_t)gamma) : (uint32_t)gamma;
        const uint32_t right = hi == gamma + 1 ? (kLeafBit | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
        t.child[2 * i] = left;
        t.child[2 * i + 1] = right;
        t.parent[(left & kLeafBit) ? (n - 1) + (left & ~kLeafBit) : left] = (uint32_t)i;
        t.parent[(right & kLeafBit) ? (n - 1) + (right & ~kLeafBit) : right] = (uint32_t)i;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) t.parent[0] = 0xffffffffu;
}

__device__ __forceinline__ void ref_box(const TlasBuf& t, uint32_t r, float* b) {
    const float* src = (r & kLeafBit) ? t.box + 6 * (size_t)(t.key_s[r & ~kLeafBit] & 0xffffffffu) : t.nbox + 6 * (size_t)r;
    for (int k = 0; k < 6; ++k) b[k] = src[k];
}

// Contribution of a child of a wide node: its number of wide-node children
// (a leaf or an odd-depth binary node's two children) and the deepest stack
// below them.
__device__ __forceinline__ void wide_child(const TlasBuf& t, uint32_t r, uint32_t& count, uint32_t& below) {
    if (r & kLeafBit) {
        count += 1;
    } else {
        count += 2;
        below = max(below, t.bnd[r]);
    }
}

// step 6: bottom-up boxes and stack bounds.  A wide node leaves at most
// (children - 1) siblings on the stack when the traversal enters one child,
// the accounting of the host collapse.
__global__ void k_tlas_boxes(TlasBuf t) {
    const uint32_t n = t.n;
    for (uint32_t leaf = blockIdx.x * blockDim.x + threadIdx.x; leaf < n; leaf += gridDim.x * blockDim.x) {
        uint32_t node = t.parent[(n - 1) + leaf];
        while (node != 0xffffffffu) {
            __threadfence();  // release: this path's boxes are visible before the count
            if (atomicAdd(&t.visits[node], 1u) == 0u) break;  // the sibling is not done yet
            __threadfence();  // acquire: the sibling's box is visible to the loads below
            float a[6], b[6];
            ref_box(t, t.child[2 * node], a);
            ref_box(t, t.child[2 * node + 1], b);
            float* o = t.nbox + 6 * (size_t)node;
            for (int k = 0; k < 3; ++k) {
                o[k] = fminf(a[k], b[k]);
                o[3 + k] = fmaxf(a[3 + k], b[3 + k]);
            }
            const uint32_t c0 = t.child[2 * node], c1 = t.child[2 * node + 1];
            if (t.flag[node]) {
                uint32_t count = 0, below = 0;
                wide_child(t, c0, count, below);
                wide_child(t, c1, count, below);
                t.bnd[node] = count - 1 + below;
            } else {
                // the children of an odd-depth node are leaves or wide nodes
                const uint32_t b0 = (c0 & kLeafBit) ? 0u : t.bnd[c0];
                const uint32_t b1 = (c1 & kLeafBit) ? 0u : t.bnd[c1];
                t.bnd[node] = max(b0, b1);
            }
            node = t.parent[node];
        }
    }
}

// step 5: depth parity of every binary node (and the deepest node)
__global__ void k_tlas_depth(TlasBuf t) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i + 1 < t.n; i += gridDim.x * blockDim.x) {
        uint32_t dd = 0, p = t.parent[i];
        while (p != 0xffffffffu) { ++dd; p = t.parent[p]; }
        t.flag[i] = (dd & 1u) ? 0u : 1u;
        atomicMax(&t.info[1], dd);
    }
}

// step 7: 4-wide nodes over the even-depth binary nodes
__global__ void k_tlas_wide(TlasBuf t) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i + 1 < t.n; i += gridDim.x * blockDim.x) {
        if (!t.flag[i]) continue;
        uint32_t refs[4];
        float lo[4][3], hi[4][3];
        int nc = 0;
        for (int c = 0; c < 2; ++c) {
            const uint32_t r = t.child[2 * i + c];
            uint32_t sub[2] = {r, 0xffffffffu};
            if (!(r & kLeafBit)) { sub[0] = t.child[2 * r]; sub[1] = t.child[2 * r + 1]; }
            for (int g = 0; g < 2; ++g) {
                const uint32_t x = sub[g];
                if (x == 0xffffffffu) continue;
                float b[6];
                ref_box(t, x, b);
                for (int k = 0; k < 3; ++k) { lo[nc][k] = b[k]; hi[nc][k] = b[3 + k]; }
                refs[nc++] = (x & kLeafBit) ? make_leaf((uint32_t)(t.key_s[x & ~kLeafBit] & 0xffffffffu), 1)
                                            : t.base + t.wide[x];
            }
        }
        DNode4 nd;
        quantise_children(lo, hi, nc, nd);
        for (int c = 0; c < nc; ++c) nd.ref[c] = refs[c];
        t.staged[t.wide[i]] = nd;
    }
}

// a single instance: one 4-wide node with one leaf child
__global__ void k_tlas_single(TlasBuf t) {
    if (threadIdx.x || blockIdx.x) return;
    float lo[1][3], hi[1][3];
    for (int k = 0; k < 3; ++k) { lo[0][k] = t.box[k]; hi[0][k] = t.box[3 + k]; }
    DNode4 nd;
    quantise_children(lo, hi, 1, nd);
    nd.ref[0] = make_leaf(0, 1);
    t.staged[0] = nd;
}

// step 8: publish when valid
__global__ void k_tlas_commit(TlasBuf t, uint32_t n_wide_max, uint32_t blas_bound, DNode4* nodes, DInstance* inst,
                              DXform* xf, uint32_t* out_ok) {
    const uint32_t need = (t.n > 1 ? t.bnd[0] : 0u) + blas_bound;  // TLAS entries, then BLAS entries
    const bool ok = t.info[0] == 0u && need <= 64u;
    const uint32_t n_wide = t.n > 1 ? t.wide[t.n - 2] + t.flag[t.n - 2] : 1u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out_ok[0] = ok ? 1u : 0u;
        out_ok[1] = n_wide;
        out_ok[2] = need;
    }
    if (!ok) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < max(t.n, n_wide_max); i += gridDim.x * blockDim.x) {
        if (i < n_wide) nodes[t.base + i] = t.staged[i];
        if (i < t.n) {
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 4; ++c) inst[i].inv[r][c] = t.inv[12 * (size_t)i + 4 * r + c];
            for (int k = 0; k < 16; ++k) xf[i].m[k] = t.xform[16 * (size_t)i + k];
        }
    }
}

uint32_t grid_of(size_t n) { return (uint32_t)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 4096)); }

}  // namespace

namespace pico {

// Rebuilds the TLAS of the context's scene (n instances) from device
// transforms; also the last step of the device scene build.
int tlas_rebuild(pico_hip_ctx* c, uint32_t n, const float* d_transforms, pico_rebuild_stats* stats) {
    TlasBuf t{};
    t.n = n;
    t.base = c->tlas_base;
    t.xform = d_transforms;
    t.mesh = c->d_instance_mesh;
    t.mesh_box = c->d_mesh_box;
    t.vrange = c->d_mesh_vrange;
    t.pos = c->dev.pos;
    const size_t m = std::max<uint32_t>(n, 2);
    t.inv = (float*)scratch(c, "tlas_inv", m * 48);
    t.box = (float*)scratch(c, "tlas_box", m * 24);
    t.wbox = (uint32_t*)scratch(c, "tlas_wbox", m * 24);
    t.key = (unsigned long long*)scratch(c, "tlas_key", m * 8);
    t.key_s = (unsigned long long*)scratch(c, "tlas_key_s", m * 8);
    t.child = (uint32_t*)scratch(c, "tlas_child", m * 8);
    t.parent = (uint32_t*)scratch(c, "tlas_parent", m * 8);
    t.nbox = (float*)scratch(c, "tlas_nbox", m * 24);
    t.visits = (uint32_t*)scratch(c, "tlas_visits", m * 4);
    t.bnd = (uint32_t*)scratch(c, "tlas_bnd", m * 4);
    t.flag = (uint32_t*)scratch(c, "tlas_flag", m * 4);
    t.wide = (uint32_t*)scratch(c, "tlas_wide", m * 4);
    t.staged = (DNode4*)scratch(c, "tlas_staged", m * sizeof(DNode4));
    t.info = (uint32_t*)scratch(c, "tlas_info", 64);
    uint32_t* ok_flag = (uint32_t*)scratch(c, "tlas_ok", 64);
    if (!t.inv || !t.box || !t.wbox || !t.key || !t.key_s || !t.child || !t.parent || !t.nbox || !t.visits || !t.bnd || !t.flag ||
        !t.wide || !t.staged || !t.info || !ok_flag)
        return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    hipEvent_t e0, e1;
    PICO_HIP(c, hipEventCreate(&e0));
    PICO_HIP(c, hipEventCreate(&e1));
    struct Ev { hipEvent_t a, b; ~Ev() { hipEventDestroy(a); hipEventDestroy(b); } } ev{e0, e1};
    PICO_HIP(c, hipEventRecord(e0, c->stream));
    PICO_LAUNCH(c, "tlas_init", k_tlas_init, grid_of(n), 256, 0, t);
    PICO_LAUNCH(c, "tlas_instances", k_tlas_instances, grid_of(n), 256, 0, t);
    const uint32_t chunks = std::max<uint32_t>(1, (c->max_mesh_vertices + kBoxChunk - 1) / kBoxChunk);
    PICO_LAUNCH(c, "tlas_wbox_init", k_tlas_wbox_init, grid_of(n), 256, 0, t);
    PICO_LAUNCH(c, "tlas_world_boxes", k_tlas_world_boxes,
                (uint32_t)std::min<size_t>(std::max<size_t>((size_t)n * chunks, 1), 65535), 256, 0, t, chunks);
    PICO_LAUNCH(c, "tlas_world_boxes_done", k_tlas_world_boxes_done, grid_of(n), 256, 0, t);
    PICO_LAUNCH(c, "tlas_centroid_bounds", k_tlas_centroid_bounds, std::min<uint32_t>(grid_of(n), 64), 256, 0, t);
    PICO_LAUNCH(c, "tlas_keys", k_tlas_keys, grid_of(n), 256, 0, t);
    size_t tmp = 0;
    PICO_HIP(c, rocprim::radix_sort_keys(nullptr, tmp, t.key, t.key_s, n, 0, 64, c->stream));
    void* tsort = scratch(c, "tlas_sort_tmp", tmp + 64);
    if (!tsort) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    PICO_HIP(c, rocprim::radix_sort_keys(tsort, tmp, t.key, t.key_s, n, 0, 64, c->stream));
    if (n > 1) {
        PICO_LAUNCH(c, "tlas_hierarchy", k_tlas_hierarchy, grid_of(n), 256, 0, t);
        PICO_LAUNCH(c, "tlas_depth", k_tlas_depth, grid_of(n), 256, 0, t);
        PICO_LAUNCH(c, "tlas_boxes", k_tlas_boxes, grid_of(n), 256, 0, t);
        size_t stmp = 0;
        PICO_HIP(c, rocprim::exclusive_scan(nullptr, stmp, t.flag, t.wide, 0u, (size_t)(n - 1), rocprim::plus<uint32_t>(),
                                            c->stream));
        void* tscan = scratch(c, "tlas_scan_tmp", stmp + 64);
        if (!tscan) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
        PICO_HIP(c, rocprim::exclusive_scan(tscan, stmp, t.flag, t.wide, 0u, (size_t)(n - 1), rocprim::plus<uint32_t>(),
                                            c->stream));
        PICO_LAUNCH(c, "tlas_wide", k_tlas_wide, grid_of(n), 256, 0, t);
    } else {
        PICO_LAUNCH(c, "tlas_single", k_tlas_single, 1, 64, 0, t);
    }
    PICO_LAUNCH(c, "tlas_commit", k_tlas_commit, grid_of(n), 256, 0, t, n, c->blas_bound, const_cast<DNode4*>(c->dev.nodes),
                const_cast<DInstance*>(c->dev.inst), const_cast<DXform*>(c->dev.xf), ok_flag);
    PICO_HIP(c, hipEventRecord(e1, c->stream));
    uint32_t h[5] = {0, 0, 0, 0, 0};
    PICO_HIP(c, hipMemcpyAsync(h, t.info, 8, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipMemcpyAsync(h + 2, ok_flag, 12, hipMemcpyDeviceToHost, c->stream));
    PICO_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    hipEventElapsedTime(&ms, e0, e1);
    if (stats) {
        stats->device_ms = ms;
        stats->binary_depth = h[1];
        stats->stack_bound = h[4];
        stats->wide_nodes = h[3];
    }
    if (h[0]) return fail(c, PICO_ERR_ARG, "an instance transform is not affine or is singular (scene unchanged)");
    if (!h[2]) return fail(c, PICO_ERR_LIMIT, "rebuilt TLAS too deep for the 64-entry traversal stack (scene unchanged)");
    c->dev.tlas_root = c->tlas_base;
    c->stack_need = h[4];
    c->n_nodes4 = (uint64_t)c->tlas_base + h[3];
    c->tlas_depth = h[1] + 1;  // levels of internal nodes, as the host builder counts them
    return PICO_OK;
}

}  // namespace pico

extern "C" {

int pico_hip_set_instance_transforms_d(pico_hip_ctx* c, const float* d_transforms, pico_rebuild_stats* stats) {
    if (!c) return PICO_ERR_ARG;
    if (!c->has_scene) return fail(c, PICO_ERR_STATE, "no scene uploaded");
    if (!d_transforms) return fail(c, PICO_ERR_ARG, "NULL transforms");
    return tlas_rebuild(c, c->host.n_instances, d_transforms, stats);
}

int pico_hip_set_instance_transforms(pico_hip_ctx* c, const float* transforms, pico_rebuild_stats* stats) {
    if (!c) return PICO_ERR_ARG;
    if (!c->has_scene) return fail(c, PICO_ERR_STATE, "no scene uploaded");
    if (!transforms) return fail(c, PICO_ERR_ARG, "NULL transforms");
    const size_t bytes = (size_t)c->host.n_instances * 64;
    float* d = (float*)scratch(c, "tlas_xform_in", bytes + 64);
    if (!d) return fail(c, PICO_ERR_NOMEM, "device allocation failed");
    PICO_HIP(c, hipMemcpyAsync(d, transforms, bytes, hipMemcpyHostToDevice, c->stream));
    return pico_hip_set_instance_transforms_d(c, d, stats);
}

}  // extern "C"
