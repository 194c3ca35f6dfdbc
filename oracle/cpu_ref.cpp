/*
 * cpu_ref.cpp — TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A clean-room CPU restatement of Pico's rendering hot path: reference-style
 * binary BVH with the reference's binned SAH builder, recursive closest-hit
 * traversal, Moller-Trumbore intersection with the reference's epsilons, the
 * Monte-Carlo integrator with its next-event estimation (mesh, built-in
 * shape and sun lights), the BSDFs and distributions (the Transparent and
 * Fresnel BTDFs with the ray's index-of-refraction stack included), the
 * built-in sphere and cube, the sky image lookup, the per-pixel accumulation
 * loop, the G-buffer pass, the a-trous denoiser and the Reinhard tone mapper.  Each function cites the
 * reference file:line it follows.  Arithmetic is written in the reference's
 * evaluation order (glm evaluates dot products as (x*x + y*y) + z*z and
 * matrix-vector products column by column) so that, with the reference's own
 * generators, results agree with the compiled reference to rounding.
 *
 * Build with -ffp-contract=off.  Only tests/, the smoke check and the bench's
 * CPU-baseline leg may load this library.
 *
 * Parity status of each part (how it is pinned) is stated in the tests that
 * exercise it: against the compiled reference subset (oracle/_ref), against
 * closed forms, and against brute force.
 */
#include "cpu_ref.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <random>
#include <thread>
#include <vector>

namespace cr {

// ---------------------------------------------------------------- math ----
struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };

static inline V3 v3(float x, float y, float z) { return {x, y, z}; }
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
static inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline V4 operator+(V4 a, V4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
static inline V4 operator-(V4 a, V4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
static inline V4 operator*(V4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
static inline V3 xyz(V4 a) { return {a.x, a.y, a.z}; }
static inline V4 v4(V3 a, float w) { return {a.x, a.y, a.z, w}; }

// glm::dot: (x*x + y*y) + z*z
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// glm::normalize: v * (1 / sqrt(dot(v, v)))
static inline V3 normalize(V3 a) { return a * (1.0f / std::sqrt(dot(a, a))); }
static inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// glm::reflect(I, N) = I - N * dot(N, I) * 2
static inline V3 reflect(V3 i, V3 n) { return i - (n * dot(n, i)) * 2.0f; }
static inline bool bad(V3 a) {
    return std::isnan(a.x) || std::isnan(a.y) || std::isnan(a.z) || std::isinf(a.x) || std::isinf(a.y) ||
           std::isinf(a.z);
}
// std::max / std::min semantics (NaN in the second operand is dropped, in the first kept)
static inline float smax(float a, float b) { return (a < b) ? b : a; }
static inline float smin(float a, float b) { return (b < a) ? b : a; }

// Column-major 4x4 (glm::mat4 layout).  M*v = (c0*x + c1*y) + (c2*z + c3*w).
struct M4 { V4 c[4]; };
static inline V4 mul(const M4& m, V4 v) {
    V4 a = {m.c[0].x * v.x + m.c[1].x * v.y, m.c[0].y * v.x + m.c[1].y * v.y, m.c[0].z * v.x + m.c[1].z * v.y,
            m.c[0].w * v.x + m.c[1].w * v.y};
    V4 b = {m.c[2].x * v.z + m.c[3].x * v.w, m.c[2].y * v.z + m.c[3].y * v.w, m.c[2].z * v.z + m.c[3].z * v.w,
            m.c[2].w * v.z + m.c[3].w * v.w};
    return a + b;
}
// mat3(M) * v, glm column order ((c0*x + c1*y) + c2*z)
static inline V3 mul3(const M4& m, V3 v) {
    return {m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z, m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z,
            m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z};
}
// General 4x4 inverse by cofactor expansion (textbook adjugate / determinant).
static M4 inverse(const M4& m) {
    float a[16], inv[16];
    for (int c = 0; c < 4; ++c) {
        a[c * 4 + 0] = m.c[c].x; a[c * 4 + 1] = m.c[c].y; a[c * 4 + 2] = m.c[c].z; a[c * 4 + 3] = m.c[c].w;
    }
    // element (row r, col c) = a[c*4 + r]
    auto e = [&](int r, int c) { return a[c * 4 + r]; };
    auto minor3 = [&](int r0, int r1, int r2, int c0, int c1, int c2) {
        return e(r0, c0) * (e(r1, c1) * e(r2, c2) - e(r2, c1) * e(r1, c2)) -
               e(r0, c1) * (e(r1, c0) * e(r2, c2) - e(r2, c0) * e(r1, c2)) +
               e(r0, c2) * (e(r1, c0) * e(r2, c1) - e(r2, c0) * e(r1, c1));
    };
    int rs[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    float cof[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            float mn = minor3(rs[r][0], rs[r][1], rs[r][2], rs[c][0], rs[c][1], rs[c][2]);
            cof[r][c] = ((r + c) & 1) ? -mn : mn;
        }
    float det = e(0, 0) * cof[0][0] + e(0, 1) * cof[0][1] + e(0, 2) * cof[0][2] + e(0, 3) * cof[0][3];
    float id = 1.0f / det;
    // inverse(r, c) = cof(c, r) / det
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[c * 4 + r] = cof[c][r] * id;
    M4 o;
    for (int c = 0; c < 4; ++c) o.c[c] = {inv[c * 4 + 0], inv[c * 4 + 1], inv[c * 4 + 2], inv[c * 4 + 3]};
    return o;
}

// -------------------------------------------------------------- random ----
// xoshiro128++ output function with the reference's rotl seeding,
// Source/Core/RandUtils.cpp:52-81.
struct XorShift {
    uint32_t s[4];
    static uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
    explicit XorShift(uint32_t seed) { s[0] = rotl(seed, 3); s[1] = rotl(seed, 7); s[2] = rotl(seed, 11); s[3] = rotl(seed, 13); }
    uint32_t next() {
        const uint32_t result = rotl(s[0] + s[3], 7) + s[0];
        const uint32_t t = s[1] << 9;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 11);
        return result;
    }
};

// Van der Corput radical inverse, Source/Core/RandUtils.cpp:13-21 (bit reversal,
// scaled in double, narrowed to float).
static float radical_inverse(uint32_t b) {
    b = (b << 16u) | (b >> 16u);
    b = ((b & 0x55555555u) << 1u) | ((b & 0xAAAAAAAAu) >> 1u);
    b = ((b & 0x33333333u) << 2u) | ((b & 0xCCCCCCCCu) >> 2u);
    b = ((b & 0x0F0F0F0Fu) << 4u) | ((b & 0xF0F0F0F0u) >> 4u);
    b = ((b & 0x00FF00FFu) << 8u) | ((b & 0xFF00FF00u) >> 8u);
    return (float)((double)(float)b * 2.3283064365386963e-10);
}
// Hammersley_Generator::next, Source/Core/RandUtils.cpp:23-27,87-90:
// (i / float(UINT32_MAX), VdC(i)) from one 32-bit draw.
static inline V2 hammersley(uint32_t i) {
    return {(float)i / (float)std::numeric_limits<uint32_t>::max(), radical_inverse(i)};
}

// PCG32 (O'Neill 2014, pcg32_random_r / pcg32_srandom_r).
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u; inc = (initseq << 1u) | 1u; next(); state += initstate; next();
    }
    // Jump ahead by delta steps (Brown 1994, arbitrary-stride LCG skip).
    void advance(uint64_t delta) {
        uint64_t cur_mult = 6364136223846793005ULL, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
        while (delta > 0) {
            if (delta & 1) { acc_mult *= cur_mult; acc_plus = acc_plus * cur_mult + cur_plus; }
            cur_plus = (cur_mult + 1) * cur_plus;
            cur_mult *= cur_mult;
            delta /= 2;
        }
        state = acc_mult * state + acc_plus;
    }
};

static inline uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

// The integrator's random source: `hs` stands for the reference's
// xorshift/Hammersley generator, `mt` for its mt19937_64 + uniform float.
struct Rng {
    uint32_t mode;
    XorShift xs{0};
    std::mt19937_64 mt{0};
    std::uniform_real_distribution<float> ud{0.0f, 1.0f};
    Pcg32 hs, ms;
    // Reference mode: Monte_Carlo_Integrator ctor, Source/Render/Integrators.cpp:36-44.
    void seed_reference(uint32_t seed) { mode = CPUREF_RNG_REFERENCE; xs = XorShift(seed); mt.seed(seed); ud.reset(); }
    // PCG mode: fresh streams per (frame, view, sample, pixel).
    void seed_pcg(uint32_t frame, uint32_t view, uint32_t sample, uint32_t pixel) {
        mode = CPUREF_RNG_PCG;
        uint64_t base = mix64(((uint64_t)frame << 32) | sample) + (uint64_t)view * 0xD1B54A32D192ED03ULL;
        hs.seed(base, 2ull * pixel);
        ms.seed(base ^ 0xA0761D6478BD642FULL, 2ull * pixel + 1ull);
    }
    uint32_t next_u32() { return mode == CPUREF_RNG_PCG ? hs.next() : xs.next(); }
    V2 next2d() { return hammersley(next_u32()); }
    float uniform() {
        if (mode == CPUREF_RNG_PCG) return (float)(ms.next() >> 8) * (1.0f / 16777216.0f);
        return ud(mt);
    }
};

// ------------------------------------------------------------ geometry ----
struct Ray { V4 o; V3 d; V3 inv; float len; };

struct AABB {
    V4 mn, mx;
    // AABB::intersection_distance, Source/Core/AABB.cpp:108-147.
    float hit(const Ray& r) const {
        float t1 = (mn.x - r.o.x) * r.inv.x, t2 = (mx.x - r.o.x) * r.inv.x;
        float t3 = (mn.y - r.o.y) * r.inv.y, t4 = (mx.y - r.o.y) * r.inv.y;
        float t5 = (mn.z - r.o.z) * r.inv.z, t6 = (mx.z - r.o.z) * r.inv.z;
        float a1 = t1 > t2 ? t2 : t1, b1 = t1 > t2 ? t1 : t2;
        float a2 = t3 > t4 ? t4 : t3, b2 = t3 > t4 ? t3 : t4;
        float a3 = t5 > t6 ? t6 : t5, b3 = t5 > t6 ? t5 : t6;
        float tmin = smax(smax(a1, a2), a3);
        float tmax = smin(smin(b1, b2), b3);
        if (tmax < 0 || tmin > tmax) return INFINITY;
        return tmin;
    }
    V4 centre() const { return mn + (mx - mn) * 0.5f; }
    void add(V4 p) {
        // glm::min/max are (y < x ? y : x) / (x < y ? y : x)
        mn = {p.x < mn.x ? p.x : mn.x, p.y < mn.y ? p.y : mn.y, p.z < mn.z ? p.z : mn.z, p.w < mn.w ? p.w : mn.w};
        mx = {mx.x < p.x ? p.x : mx.x, mx.y < p.y ? p.y : mx.y, mx.z < p.z ? p.z : mx.z, mx.w < p.w ? p.w : mx.w};
    }
    void unite(const AABB& b) {
        mn = {b.mn.x < mn.x ? b.mn.x : mn.x, b.mn.y < mn.y ? b.mn.y : mn.y, b.mn.z < mn.z ? b.mn.z : mn.z,
              b.mn.w < mn.w ? b.mn.w : mn.w};
        mx = {mx.x < b.mx.x ? b.mx.x : mx.x, mx.y < b.mx.y ? b.mx.y : mx.y, mx.z < b.mx.z ? b.mx.z : mx.z,
              mx.w < b.mx.w ? b.mx.w : mx.w};
    }
    V3 sides() const { return {std::fabs(mx.x - mn.x), std::fabs(mx.y - mn.y), std::fabs(mx.z - mn.z)}; }
    float area() const { V3 s = sides(); return 2.0f * ((s.x * s.y) + (s.x * s.z) + (s.y * s.z)); }
    // AABB::get_offset, Source/Core/AABB.cpp:194-201.
    V3 offset(V4 p) const {
        V3 o = {p.x - mn.x, p.y - mn.y, p.z - mn.z};
        if (mx.x > mn.x) o.x /= mx.x - mn.x;
        if (mx.y > mn.y) o.y /= mx.y - mn.y;
        if (mx.z > mn.z) o.z /= mx.z - mn.z;
        return o;
    }
    static AABB empty() { return {{INFINITY, INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY, -INFINITY}}; }
};

static inline float comp(V3 v, uint32_t i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
static inline float comp(V4 v, uint32_t i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
// maximum_component_index, Source/Core/AABB.cpp:33-41.
static uint32_t max_component(V3 v) {
    V3 a = {std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)};
    if ((a.x >= a.y) && (a.x >= a.z)) return 0;
    if ((a.y >= a.x) && (a.y >= a.z)) return 1;
    return 2;
}
// AABB * mat4 via the 8 corners, divided by w, Source/Core/AABB.cpp:253-270
// (corners in get_cube_as_vertex_array order, :81-96).
static AABB transform_box(const AABB& b, const M4& m) {
    V4 cs[8] = {b.mn, {b.mn.x, b.mn.y, b.mx.z, 1.0f}, {b.mx.x, b.mn.y, b.mx.z, 1.0f}, {b.mx.x, b.mn.y, b.mn.z, 1.0f},
                {b.mn.x, b.mx.y, b.mn.z, 1.0f}, {b.mn.x, b.mx.y, b.mx.z, 1.0f}, b.mx, {b.mx.x, b.mx.y, b.mn.z, 1.0f}};
    AABB o = {{1e7f, 1e7f, 1e7f, 1e7f}, {-1e7f, -1e7f, -1e7f, -1e7f}};
    for (auto& c : cs) {
        V4 p = mul(m, c);
        float w = p.w;
        p = {p.x / w, p.y / w, p.z / w, p.w / w};
        o.add(p);
    }
    return o;
}

// ----------------------------------------------------------------- BVH ----
// BVH<T,2> with the reference's node layout semantics (Source/Core/BVH.hpp:52-66)
// and builder (Source/Core/BVH.cpp:65-197).
struct BV { AABB box; uint32_t value; };
struct Node { AABB box; std::vector<BV> values; uint32_t child[2]; };
static constexpr uint32_t kInvalid = ~0u;

struct Builder {
    std::vector<Node> nodes;
    uint32_t buckets;
    uint32_t max_depth_seen = 0;
    using It = std::vector<BV>::iterator;

    // SAH_parition_Scheme::partition, Source/Core/BVH.cpp:126-197.
    It partition(It start, It end) {
        if (start == end) return end;
        AABB bounds = AABB::empty();
        for (It i = start; i != end; ++i) bounds.add(i->box.centre());
        const uint32_t axis = max_component(bounds.sides());
        const uint32_t nb = buckets;
        std::vector<uint32_t> cnt(nb, 0);
        std::vector<AABB> bb(nb, AABB::empty());
        auto bucket_of = [&](const BV& v) {
            uint32_t b = (uint32_t)((float)nb * comp(bounds.offset(v.box.centre()), axis));
            if (b == nb) b = nb - 1;
            return b;
        };
        for (It i = start; i < end; ++i) {
            uint32_t b = bucket_of(*i);
            cnt[b]++;
            bb[b].unite(i->box);
        }
        const float leaf_cost = (float)std::distance(start, end);
        std::vector<float> cost(nb - 1);
        for (uint32_t i = 0; i < nb - 1; ++i) {
            AABB b0 = AABB::empty(), b1 = AABB::empty();
            uint32_t c0 = 0, c1 = 0;
            for (uint32_t j = 0; j <= i; ++j) { b0.unite(bb[j]); c0 += cnt[j]; }
            for (uint32_t j = i + 1; j < nb; ++j) { b1.unite(bb[j]); c1 += cnt[j]; }
            cost[i] = .125f + ((float)c0 * b0.area() + (float)c1 * b1.area()) / (leaf_cost * bounds.area());
        }
        auto me = std::min_element(cost.begin(), cost.end());
        const float min_cost = *me;
        const uint32_t split = (uint32_t)std::distance(cost.begin(), me);
        if (leaf_cost > 2 || min_cost < leaf_cost)
            return std::partition(start, end, [&](const BV pi) { return bucket_of(pi) <= split; });
        return end;
    }

    // BVHFactory::split_primitives, Source/Core/BVH.cpp:73-107.
    uint32_t split(It start, It end, uint32_t depth, uint32_t level) {
        if (start == end) return kInvalid;
        max_depth_seen = std::max(max_depth_seen, level);
        AABB bounds = AABB::empty();
        for (It i = start; i != end; ++i) bounds.unite(i->box);
        Node n;
        n.box = bounds;
        It pivot = partition(start, end);
        if (depth == 0 || pivot == end || std::distance(start, end) == 2) {
            for (It i = start; i != end; ++i) n.values.push_back(*i);
            n.child[0] = n.child[1] = kInvalid;
            nodes.push_back(n);
            return (uint32_t)nodes.size() - 1;
        }
        n.child[0] = split(start, pivot, depth - 1, level + 1);
        n.child[1] = split(pivot, end, depth - 1, level + 1);
        nodes.push_back(n);
        return (uint32_t)nodes.size() - 1;
    }
};

// -------------------------------------------------------------- scene -----
struct Hit {
    V4 pos; V3 nrm; float t; uint32_t inst, prim; float u, v;
};

struct Mesh {
    uint32_t shape = PICO_SHAPE_TRIANGLES;  // a built-in shape has no triangles
    std::vector<V3> pos, nrm;
    std::vector<uint32_t> idx;
    std::vector<Node> nodes;
    uint32_t root = kInvalid;
    AABB bounds;
    // light sampling data (LowerLevelMeshBVH::generate_sampling_data)
    std::vector<float> alias_w;
    std::vector<uint32_t> alias_i;
    uint32_t depth = 0;

    // Mesh_Intersector::intersects, Source/Core/LowerLevelMeshBVH.cpp:165-207
    // (Moller-Trumbore; the reciprocal is taken in double).
    bool intersect(const Ray& r, uint32_t prim, float& t, float& u, float& v) const {
        const float EPS = 0.0000001f;
        const V3 p0 = pos[idx[prim * 3]], p1 = pos[idx[prim * 3 + 1]], p2 = pos[idx[prim * 3 + 2]];
        const V3 e1 = p1 - p0, e2 = p2 - p0;
        const V3 h = cross(r.d, e2);
        const float a = dot(e1, h);
        if (a > -EPS && a < EPS) return false;
        const float f = (float)(1.0 / (double)a);
        const V3 s = xyz(r.o) - p0;
        const float uu = f * dot(s, h);
        if (uu < 0.0 || uu > 1.0) return false;
        const V3 q = cross(s, e1);
        const float vv = f * dot(r.d, q);
        if (vv < 0.0 || uu + vv > 1.0) return false;
        const float tt = f * dot(e2, q);
        if (tt > EPS) { t = tt; u = uu; v = vv; return true; }
        return false;
    }
    // interpolate_fragment, Source/Core/LowerLevelMeshBVH.cpp:134-163.
    void interpolate(uint32_t prim, float u, float v, V4& P, V3& N) const {
        const uint32_t i0 = idx[prim * 3], i1 = idx[prim * 3 + 1], i2 = idx[prim * 3 + 2];
        const float w = 1.0f - v - u;
        P = v4(((w * pos[i0]) + (u * pos[i1])) + (v * pos[i2]), 1.0f);
        N = normalize(((w * nrm[i0]) + (u * nrm[i1])) + (v * nrm[i2]));
    }
    // BVH<uint32_t,2>::get_closest_intersections, Source/Core/BVH.cpp:27-63.
    void closest(const Ray& r, const Node& n, float& best, uint32_t& best_prim, float& bu, float& bv) const {
        if (n.box.hit(r) < best) {
            if (!n.values.empty()) {
                for (const BV& bvv : n.values) {
                    float t = INFINITY, u, v;
                    if (intersect(r, bvv.value, t, u, v) && t < best) { best = t; best_prim = bvv.value; bu = u; bv = v; }
                }
            } else {
                for (uint32_t c : n.child)
                    if (c != kInvalid) closest(r, nodes[c], best, best_prim, bu, bv);
            }
        }
    }
};

struct Instance { M4 m, minv; uint32_t mesh, material, bsdf; };

// LowerLevelSphereBVH(0.5) / LowerLevelCube::calculate_intersection,
// Source/Core/LowerLevelImplicitShapesBVH.cpp:21-48, 75-88, on the object-space
// ray (direction not renormalised): hit point, normal and parameter.
static bool implicit_intersect(uint32_t shape, const Ray& o, V4& P, V3& N, float& t) {
    if (shape == PICO_SHAPE_SPHERE) {
        const V3 m = xyz(o.o);
        const float b = dot(m, o.d);
        const float c = dot(m, m) - 0.5f * 0.5f;
        if (c > 0.0f && b > 0.0f) return false;
        const float discr = b * b - c;
        if (discr < 0.0f) return false;
        t = -b - std::sqrt(discr);
        if (t < 0.0f) t = 0.0f;  // the ray starts inside
        P = {o.o.x + t * o.d.x, o.o.y + t * o.d.y, o.o.z + t * o.d.z, 1.0f};
        N = normalize(xyz(P));
        return true;
    }
    const AABB box{{-0.5f, -0.5f, -0.5f, 1.0f}, {0.5f, 0.5f, 0.5f, 1.0f}};
    t = box.hit(o);
    if (t == INFINITY) return false;
    P = {o.o.x + o.d.x * t, o.o.y + o.d.y * t, o.o.z + o.d.z * t, 1.0f};
    // maximum_component_index, Source/Core/AABB.cpp:31-41, and glm::sign
    const float ax = std::fabs(P.x), ay = std::fabs(P.y), az = std::fabs(P.z);
    const int k = (ax >= ay && ax >= az) ? 0 : ((ay >= ax && ay >= az) ? 1 : 2);
    const float c = k == 0 ? P.x : (k == 1 ? P.y : P.z);
    const float sg = (float)((0.0f < c) - (c < 0.0f));
    N = {k == 0 ? sg : 0.0f, k == 1 ? sg : 0.0f, k == 2 ? sg : 0.0f};
    return true;
}

// sample_geometry of the built-in shapes, Source/Core/LowerLevelImplicitShapesBVH.cpp:57-66
// (0.5 * uniform_sample_sphere, Source/Core/RandUtils.cpp:43-49) and 90-110.
static V3 implicit_sample(uint32_t shape, Rng& rng) {
    if (shape == PICO_SHAPE_SPHERE) {
        const V2 xi = rng.next2d();
        const float z = 1 - 2 * xi.x;
        const float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
        const float phi = (float)(2 * M_PI * (double)xi.y);
        return {0.5f * (r * std::cos(phi)), 0.5f * (r * std::sin(phi)), 0.5f * z};
    }
    const V2 xi = rng.next2d(), xj = rng.next2d(), xk = rng.next2d();
    auto sgn = [](float x) { return (float)((0.0f < x) - (x < 0.0f)); };
    const V3 corner = {sgn(xi.x - 0.5f), sgn(xi.y - 0.5f), sgn(xj.x - 0.5f)};
    uint32_t zero = (uint32_t)(xj.y * 3.0f);
    if (zero == 3) zero = 2;
    float off[3];
    off[zero] = 0.0f;
    off[(zero + 1) % 3] = xk.x;
    off[(zero + 2) % 3] = xk.y;
    return {corner.x * 0.5f + (-corner.x * off[0]), corner.y * 0.5f + (-corner.y * off[1]),
            corner.z * 0.5f + (-corner.z * off[2])};
}

struct Scene {
    std::vector<Mesh> meshes;
    std::vector<Instance> inst;
    std::vector<Node> tlas;
    uint32_t tlas_root = kInvalid, tlas_depth = 0;
    std::vector<pico_material> mats;
    std::vector<uint32_t> lights;  // instance ids
    V3 sky;
    uint32_t microfacet;
    bool use_sun = false;
    V3 sun_dir{0, -1, 0}, sun_col{1, 1, 1};
    // sky image (width 0: the constant sky)
    uint32_t sky_w = 0, sky_h = 0, sky_faces = 0, sky_fmt = 0;
    std::vector<uint8_t> sky_tex;
};

// ImageCube::sample4 with resolve_cubemap_UV and get_data_ptr,
// Source/Core/Image.cpp:361-466: the nearest texel in direction d.  The
// equirectangular form's asin is the double overload (the reference calls
// asin on a float, which resolves to ::asin(double) there); its atan2 is
// glm::atan, the float overload.
static V3 sky_lookup(const Scene& s, V3 d) {
    if (!s.sky_w) return s.sky;
    uint32_t face = 0;
    float u, v;
    if (s.sky_faces == 6) {
        const V3 a = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
        float ma;
        V2 uv;
        if (a.z >= a.x && a.z >= a.y) {
            face = d.z < 0.0f ? 5 : 4;
            ma = 0.5f / a.z;
            uv = {d.z < 0.0f ? -d.x : d.x, -d.y};
        } else if (a.y >= a.x) {
            face = d.y < 0.0f ? 3 : 2;
            ma = 0.5f / a.y;
            uv = {d.x, d.y < 0.0f ? -d.z : d.z};
        } else {
            face = d.x < 0.0f ? 1 : 0;
            ma = 0.5f / a.x;
            uv = {d.x < 0.0f ? d.z : -d.z, -d.y};
        }
        u = uv.x * ma + 0.5f;
        v = uv.y * ma + 0.5f;
    } else {
        u = std::atan2(d.z, d.x);
        v = (float)std::asin((double)d.y);
        u *= (float)(1.0 / (3.14159265358979323846 * 2.0f));
        v *= (float)(1.0 / 3.14159265358979323846);
        u += 0.5f;
        v += 0.5f;
        v = 1.0f - v;
    }
    // uint32_t(x) of a NaN is 0 on the reference's x86-64 build (the 64-bit
    // truncating conversion gives 2^63, whose low word is 0)
    auto texel = [](float c, uint32_t n) {
        const float x = c * (float)n;
        const uint32_t i = std::isnan(x) ? 0u : (uint32_t)(int64_t)x;
        return i % n;
    };
    const size_t t = ((size_t)face * s.sky_h + texel(v, s.sky_h)) * s.sky_w + texel(u, s.sky_w);
    if (s.sky_fmt == PICO_TEXEL_RGBA32F) {
        float f[3];
        std::memcpy(f, s.sky_tex.data() + 16 * t, 12);
        return {f[0], f[1], f[2]};
    }
    const uint8_t* c = s.sky_tex.data() + 4 * t;
    return {c[0] / 255.0f, c[1] / 255.0f, c[2] / 255.0f};
}

// UpperLevelBVH::lower_level_intersector::intersects, Source/Core/UpperLevelBVH.cpp:59-80,
// and transform_ray, Source/Core/AABB.cpp:43-52.
static bool instance_hit(const Scene& s, const Ray& r, uint32_t ii, float& dist, Hit& h) {
    const Instance& e = s.inst[ii];
    Ray o;
    o.o = mul(e.minv, r.o);
    o.d = mul3(e.minv, r.d);
    o.inv = v3(1.0f, 1.0f, 1.0f) / o.d;
    o.len = r.len;
    const Mesh& m = s.meshes[e.mesh];
    if (m.shape != PICO_SHAPE_TRIANGLES) {
        V4 P; V3 N; float t;
        if (!implicit_intersect(m.shape, o, P, N, t)) return false;
        h.pos = mul(e.m, P);
        h.nrm = normalize(mul3(e.m, N));
        V4 dd = h.pos - r.o;
        dist = std::sqrt((dd.x * dd.x + dd.y * dd.y) + (dd.z * dd.z + dd.w * dd.w));
        h.inst = ii; h.prim = 0; h.u = t; h.v = 0.0f;
        return true;
    }
    // LowerLevelMeshBVH::calculate_intersection -> BVH::get_first_intersection (t from +inf)
    float best = INFINITY, bu = 0, bv = 0;
    uint32_t bp = kInvalid;
    o.inv = v3(1.0f, 1.0f, 1.0f) / o.d;
    if (m.root != kInvalid) m.closest(o, m.nodes[m.root], best, bp, bu, bv);
    if (best == INFINITY) return false;
    V4 P; V3 N;
    m.interpolate(bp, bu, bv, P, N);
    h.pos = mul(e.m, P);
    h.nrm = normalize(mul3(e.m, N));
    V4 dd = h.pos - r.o;
    dist = std::sqrt((dd.x * dd.x + dd.y * dd.y) + (dd.z * dd.z + dd.w * dd.w));
    h.inst = ii; h.prim = bp; h.u = bu; h.v = bv;
    return true;
}

static void tlas_closest(const Scene& s, const Ray& r, const Node& n, float& best, Hit& out) {
    if (n.box.hit(r) < best) {
        if (!n.values.empty()) {
            for (const BV& bvv : n.values) {
                float d = INFINITY; Hit h;
                if (instance_hit(s, r, bvv.value, d, h) && d < best) { best = d; out = h; out.t = d; }
            }
        } else {
            for (uint32_t c : n.child)
                if (c != kInvalid) tlas_closest(s, r, s.tlas[c], best, out);
        }
    }
}

// UpperLevelBVH::get_closest_intersection, Source/Core/UpperLevelBVH.cpp:12-20.
static bool closest_hit(const Scene& s, Ray r, Hit& h) {
    r.inv = v3(1.0f, 1.0f, 1.0f) / r.d;
    float best = INFINITY;
    if (s.tlas_root == kInvalid) return false;
    tlas_closest(s, r, s.tlas[s.tlas_root], best, h);
    return best != INFINITY;
}

// ------------------------------------------------------- alias table -----
// AliasTable::build_table, Source/Util/AliasTable.cpp:12-87.
static void alias_build(const std::vector<float>& w, std::vector<float>& bw, std::vector<uint32_t>& ba) {
    const uint32_t n = (uint32_t)w.size();
    bw.assign(n, 0.0f);
    ba.assign(n, 0u);
    const float total = (float)std::accumulate(w.begin(), w.end(), 0.0);
    for (uint32_t i = 0; i < n; ++i) bw[i] = w[i] / total;
    struct O { float p; uint32_t i; };
    std::vector<O> over, under;
    for (uint32_t i = 0; i < n; ++i) {
        const float ph = bw[i] * (float)n;
        if (ph < 1.0f) under.push_back({ph, i}); else over.push_back({ph, i});
    }
    while (!over.empty() && !under.empty()) {
        O ov = over.back(); over.pop_back();
        O un = under.back(); under.pop_back();
        bw[un.i] = un.p;
        ba[un.i] = ov.i;
        const float excess = un.p + ov.p - 1.0f;
        if (excess < 1.0f) under.push_back({excess, ov.i}); else over.push_back({excess, ov.i});
    }
    while (!under.empty()) { O un = under.back(); under.pop_back(); ba[un.i] = 0; bw[un.i] = 1.0f; }
    while (!over.empty()) { O ov = over.back(); over.pop_back(); ba[ov.i] = 0; bw[ov.i] = 1.0f; }
}
// AliasTable::sample, Source/Util/AliasTable.cpp:89-104 (pdf output unused by callers).
static uint32_t alias_sample(const std::vector<float>& bw, const std::vector<uint32_t>& ba, Rng& rng) {
    const uint32_t b = rng.next_u32() % (uint32_t)bw.size();
    const float coin = (float)rng.next_u32() / (float)std::numeric_limits<uint32_t>::max();
    if (coin <= bw[b]) return b;
    return ba[b];
}

// Triangle "areas" of generate_sampling_data, Source/Core/LowerLevelMeshBVH.cpp:92-117
// (the sine takes the angle scaled by 180/pi).
static float tri_area(V3 a, V3 b, V3 c) {
    const V3 ac = a - c, bc = b - c;
    const float bc_len = length(bc);
    const float base = length(ac);
    const float theta = (float)std::sin((double)std::acos(dot(normalize(ac), normalize(bc))) * (180.0f / M_PI));
    const float height = theta * bc_len;
    return (base * height) / 2.0f;
}

// ---------------------------------------------------------- shading ------
// construct_world_to_tangent_transform, Source/Core/vectorUtils.cpp:9-35:
// rows (T, B, N).
struct Frame { V3 t, b, n; };
static Frame tangent_frame(V3 V, V3 N) {
    V3 t;
    if (std::fabs(dot(V, N)) > 0.95f) {
        V3 vv = std::fabs(N.z) > 0.99f ? v3(1, 0, 0) : v3(0, 0, 1);
        t = normalize(cross(N, vv));
    } else {
        t = normalize(cross(V, N));
    }
    V3 b = normalize(cross(t, N));
    return {t, b, N};
}
// world -> tangent: glm mat3 * v with rows T, B, N
static inline V3 to_tangent(const Frame& f, V3 v) {
    return {f.t.x * v.x + f.t.y * v.y + f.t.z * v.z, f.b.x * v.x + f.b.y * v.y + f.b.z * v.z,
            f.n.x * v.x + f.n.y * v.y + f.n.z * v.z};
}
// tangent -> world: inverse of the frame matrix (adjugate / determinant).
static inline V3 to_world(const Frame& f, V3 v) {
    const V3 c0 = cross(f.b, f.n), c1 = cross(f.n, f.t), c2 = cross(f.t, f.b);
    const float det = dot(f.t, c0);
    const float id = 1.0f / det;
    // inverse has columns c0/det, c1/det, c2/det
    return {(c0.x * id) * v.x + (c1.x * id) * v.y + (c2.x * id) * v.z,
            (c0.y * id) * v.x + (c1.y * id) * v.y + (c2.y * id) * v.z,
            (c0.z * id) * v.x + (c1.z * id) * v.y + (c2.z * id) * v.z};
}

// TangentSpace helpers, Source/Core/vectorUtils.hpp:24-75.
static inline float cos2t(V3 w) { return w.z * w.z; }
static inline float sin2t(V3 w) { return smax(0.0f, 1.0f - cos2t(w)); }
static inline float sint(V3 w) { return std::sqrt(sin2t(w)); }
static inline float tan2t(V3 w) { return sin2t(w) / cos2t(w); }
static inline float cosp(V3 w) { float s = sint(w); return (s == 0) ? 1 : std::clamp(w.x / s, -1.0f, 1.0f); }
static inline float sinp(V3 w) { float s = sint(w); return (s == 0) ? 0 : std::clamp(w.y / s, -1.0f, 1.0f); }

enum { DIST_COS = 0, DIST_BECKMANN = 1, DIST_GGX = 2 };

// Cos_Weighted_Hemisphere_Distribution, Source/Render/Distributions.cpp:14-29.
static V3 cos_sample(V2 xi) {
    const float r = std::sqrt(xi.x);
    const float phi = (float)((double)xi.y * M_PI * 2.0f);
    return normalize(v3(r * std::cos(phi), r * std::sin(phi), std::sqrt(smax(0.0f, 1.0f - xi.x))));
}
static float cos_pdf(V3 h) { return smax(0.0f, h.z / (float)M_PI); }

// Beckmann_All_Microfacet_Distribution, Source/Render/Distributions.cpp:31-78.
static V3 beckmann_sample(V2 xi, V3 V, float R) {
    const float alpha = R * R;
    float ls = std::log(1.0f - xi.x);
    if (std::isinf(ls)) ls = 0;
    const float tan2 = -alpha * alpha * ls;
    const float phi = (float)((double)(xi.y * 2.0f) * M_PI);
    const float ct = 1.0f / std::sqrt(1.0f + tan2);
    const float st = std::sqrt(smax(0.0f, 1.0f - ct * ct));
    V3 H = v3(st * std::cos(phi), st * std::sin(phi), ct);
    if (!(V.z * H.z > 0.0f)) H = -H;
    return normalize(H);
}
static float beckmann_alpha(float r) {
    r = smax(r, (float)1e-3);
    const float x = std::log(r);
    return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}
static float beckmann_D(V3 wh, float R) {
    const float a = beckmann_alpha(R);
    const float tan2 = tan2t(wh);
    if (std::isinf(tan2)) return 0.0f;
    const float cos4 = cos2t(wh) * cos2t(wh);
    const float c = cosp(wh), s = sinp(wh);
    const float num = std::exp(-tan2 * ((c * c) / (a * a) + (s * s) / (a * a)));
    return (float)((double)num / (M_PI * a * a * cos4));
}
static float beckmann_pdf(V3 H, float R) {
    float d = beckmann_D(H, R);
    d = std::clamp(d, 0.0f, 1.0f);
    return d * std::fabs(H.z);
}

// GGX (Trowbridge-Reitz) distribution of normals, Walter et al. 2007:
// D(h) = a^2 / (pi cos^4 (a^2 + tan^2)^2), h sampled by tan^2 = a^2 u/(1-u),
// pdf_h = D |cos|.  a = R^2 (the reference sampler's roughness convention).
// New Distribution subclass; sampled h is flipped into V's hemisphere as the
// reference's microfacet sampler does.
static V3 ggx_sample(V2 xi, V3 V, float R) {
    const float a = R * R;
    const float tan2 = a * a * xi.x / (1.0f - xi.x);
    const float phi = (float)((double)(xi.y * 2.0f) * M_PI);
    const float ct = 1.0f / std::sqrt(1.0f + tan2);
    const float st = std::sqrt(smax(0.0f, 1.0f - ct * ct));
    V3 H = v3(st * std::cos(phi), st * std::sin(phi), ct);
    if (!(V.z * H.z > 0.0f)) H = -H;
    return normalize(H);
}
static float ggx_D(V3 wh, float R) {
    const float a = R * R;
    const float a2 = a * a;
    const float c2 = cos2t(wh);
    const float tan2 = tan2t(wh);
    if (std::isinf(tan2)) return 0.0f;
    const float den = a2 + tan2;
    return a2 / ((float)M_PI * (c2 * c2) * (den * den));
}
static float ggx_pdf(V3 H, float R) { return ggx_D(H, R) * std::fabs(H.z); }

static V3 spec_sample(uint32_t mf, V2 xi, V3 V, float R) {
    return mf == PICO_MICROFACET_GGX ? ggx_sample(xi, V, R) : beckmann_sample(xi, V, R);
}
static float spec_pdf(uint32_t mf, V3 H, float R) { return mf == PICO_MICROFACET_GGX ? ggx_pdf(H, R) : beckmann_pdf(H, R); }

static inline V3 mdiff(const pico_material& m) { return {m.diffuse[0], m.diffuse[1], m.diffuse[2]}; }
static inline V3 mspec(const pico_material& m) { return {m.specular[0], m.specular[1], m.specular[2]}; }
static inline V3 memis(const pico_material& m) { return {m.emissive[0], m.emissive[1], m.emissive[2]}; }
// EvaluatedMaterial::get_reflectance, Source/Core/MaterialManager.hpp:40-45.
static inline float reflectance(const pico_material& m) {
    const float s = length(mspec(m));
    return s / (s + length(mdiff(m)));
}

struct Sample { V3 L; float P; V3 energy; };

// BSRDF::sample for Diffuse / Specular / Dielectric / Light / Delta,
// Source/Render/BSRDF.cpp:11-238.
static Sample bsdf_sample(uint32_t kind, uint32_t mf, const pico_material& m, V3 N, V3 raydir, Rng& rng) {
    const V3 V = -raydir;
    Sample s{};
    if (kind == PICO_BSDF_DELTA) {  // BSRDF.cpp:216-227
        s.energy = mspec(m);
        s.L = reflect(-V, N);
        s.P = 1.0f;
        return s;
    }
    const Frame f = tangent_frame(V, N);
    const V3 vt = normalize(to_tangent(f, V));
    if (kind == PICO_BSDF_DIFFUSE || kind == PICO_BSDF_LIGHT) {  // BSRDF.cpp:11-40, 174-202
        const V2 xi = rng.next2d();
        const V3 H = cos_sample(xi);
        const float pdf = cos_pdf(H);
        s.L = normalize(to_world(f, H));
        s.P = pdf;
        s.energy = kind == PICO_BSDF_LIGHT ? memis(m) : mdiff(m);
        return s;
    }
    if (kind == PICO_BSDF_SPECULAR) {  // BSRDF.cpp:54-85
        const V2 xi = rng.next2d();
        const V3 H = spec_sample(mf, xi, vt, m.roughness);
        const V3 L = normalize(reflect(-vt, H));
        s.L = normalize(to_world(f, L));
        s.P = spec_pdf(mf, H, m.roughness) / (2.0f * dot(vt, H));
        s.energy = mspec(m);
        return s;
    }
    // Dielectric, BSRDF.cpp:100-149
    const float sp = reflectance(m);
    V2 xi = rng.next2d();
    if (xi.y >= sp) {
        const V2 x2 = rng.next2d();
        const V3 H = cos_sample(x2);
        s.L = normalize(to_world(f, H));
        s.P = cos_pdf(H);
        s.energy = mdiff(m);
    } else {
        xi = rng.next2d();
        const V3 H = spec_sample(mf, xi, vt, m.roughness);
        const V3 L = normalize(reflect(-vt, H));
        s.L = normalize(to_world(f, L));
        s.P = spec_pdf(mf, H, m.roughness) / (2.0f * dot(vt, H));
        s.energy = mspec(m);
    }
    return s;
}

// The ray's index-of-refraction stack, Source/Core/AABB.cpp:54-79; the
// camera pushes 1.0 (Source/Core/Camera.cpp:23).  A pop that would remove
// the camera's entry keeps it (the reference asserts there).
struct IorStack {
    std::vector<float> v{1.0f};
    void push(float x) { v.push_back(x); }
    void pop() { if (v.size() > 1) v.pop_back(); }
    float current() const { return v.back(); }
    bool inside() const { return v.size() > 1; }  // Ray::inside_geometry
};

static inline bool is_btdf(uint32_t kind) { return kind == PICO_BSDF_TRANSPARENT || kind == PICO_BSDF_FRESNEL; }

// glm::refract (glm/detail/func_geometric.inl): NaN where k < 0, since the
// square root of a negative is multiplied by 0.
static V3 refract_glm(V3 I, V3 N, float eta) {
    const float d = dot(N, I);
    const float k = 1.0f - eta * eta * (1.0f - d * d);
    const float s = eta * d + std::sqrt(k);
    const float sel = (float)(k >= 0.0f);
    return {(eta * I.x - s * N.x) * sel, (eta * I.y - s * N.y) * sel, (eta * I.z - s * N.z) * sel};
}

// Transparent_BTDF::sample, Source/Render/BSRDF.cpp:240-309.
static Sample btdf_transparent(uint32_t mf, const pico_material& m, V3 N, V3 raydir, Rng& rng, IorStack& st) {
    const V3 V = -raydir;
    const Frame f = tangent_frame(V, N);
    const V3 vt = normalize(to_tangent(f, V));
    const bool enter = vt.z >= 0.0f;  // TangentSpace::cos_theta
    const V2 xi = rng.next2d();
    const V3 H = spec_sample(mf, xi, vt, m.roughness);
    const float eta = 1.0f / m.index_of_refraction;
    if (enter) st.push(m.index_of_refraction);
    else st.pop();
    const V3 L = refract_glm(vt, H, eta);  // Transparent_BTDF::refract always succeeds
    Sample s{};
    s.L = normalize(to_world(f, L));
    const float pdf = spec_pdf(mf, H, m.roughness);
    const V3 wh = normalize(vt + L * eta);
    const float den = dot(vt, wh) + eta * dot(L, wh);
    s.P = pdf * std::abs((eta * eta * dot(L, wh)) / (den * den));
    s.energy = mdiff(m);
    return s;
}

// Fresnel_BTDF::fresnel_factor, Source/Render/BSRDF.cpp:400-421.
static float fresnel_factor(float cos_i, float eta_i, float eta_t) {
    cos_i = std::clamp(cos_i, -1.0f, 1.0f);
    if (!(cos_i > 0.f)) {
        std::swap(eta_i, eta_t);
        cos_i = std::abs(cos_i);
    }
    const float sin_i = std::sqrt(std::max(0.f, 1.0f - cos_i * cos_i));
    const float sin_t = eta_i / eta_t * sin_i;
    const float cos_t = std::sqrt(std::max(0.0f, 1.0f - sin_t * sin_t));
    const float rpa = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
    const float rpe = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
    return (rpa * rpa + rpe * rpe) / 2;
}

// Fresnel_BTDF::sample, Source/Render/BSRDF.cpp:335-362.
static Sample btdf_fresnel(uint32_t mf, const pico_material& m, V3 N, V3 raydir, Rng& rng, IorStack& st) {
    const V3 V = -raydir;
    const Frame f = tangent_frame(V, N);
    const V3 vt = normalize(to_tangent(f, V));
    const V2 xi = rng.next2d();
    const float F = fresnel_factor(vt.z, m.index_of_refraction, st.current());
    Sample s;
    if (xi.x < F) {
        s = bsdf_sample(PICO_BSDF_SPECULAR, mf, m, N, raydir, rng);
        s.P *= F;
    } else {
        s = btdf_transparent(mf, m, N, raydir, rng, st);
        s.P *= 1.0f - F;
    }
    return s;
}

static Sample any_sample(uint32_t kind, uint32_t mf, const pico_material& m, V3 N, V3 raydir, Rng& rng, IorStack& st) {
    if (kind == PICO_BSDF_TRANSPARENT) return btdf_transparent(mf, m, N, raydir, rng, st);
    if (kind == PICO_BSDF_FRESNEL) return btdf_fresnel(mf, m, N, raydir, rng, st);
    return bsdf_sample(kind, mf, m, N, raydir, rng);
}

// BSRDF::pdf, Source/Render/BSRDF.cpp:42-45, 87-91, 151-155, 229-232.
static float bsdf_pdf(uint32_t kind, uint32_t mf, const pico_material& m, V3 wo, V3 H) {
    switch (kind) {
        case PICO_BSDF_DIFFUSE: return cos_pdf(H);
        case PICO_BSDF_LIGHT: return cos_pdf(H);
        case PICO_BSDF_SPECULAR: return spec_pdf(mf, H, m.roughness) / (2.0f * dot(wo, H));
        case PICO_BSDF_DIELECTRIC: {
            const float r = reflectance(m);
            return (cos_pdf(H) * (1 - r)) + ((spec_pdf(mf, H, m.roughness) / (2.0f * dot(wo, H))) * r);
        }
        default: return 0.0f;
    }
}

// direct_lighting_pdf, Source/Render/Integrators.cpp:13-22.
static float direct_pdf(uint32_t kind, uint32_t mf, const pico_material& m, V3 N, V3 wi, V3 wo) {
    const Frame f = tangent_frame(wi, N);
    const V3 two = to_tangent(f, wo);
    V3 H = normalize(wi + wo);
    H = to_tangent(f, H);
    return bsdf_pdf(kind, mf, m, two, H);
}

struct Counters { uint64_t camera = 0, ext = 0, shadow = 0; };

// Monte_Carlo_Integrator::sample_direct_lighting, Source/Render/Integrators.cpp:46-128.
static bool direct_light(const Scene& s, const Hit& frag, V3 wi, V3& rad, float& pdf, Rng& rng, Counters& c) {
    const Instance& fi = s.inst[frag.inst];
    if (fi.bsdf == PICO_BSDF_LIGHT || is_btdf(fi.bsdf)) return false;
    const pico_material& mat = s.mats[fi.material];
    const uint32_t n_lights = (uint32_t)s.lights.size();
    uint32_t count = n_lights + (s.use_sun ? 1u : 0u);
    const float u = rng.uniform();
    const uint32_t pick = (uint32_t)(u * (float)count);
    if (pick == count) count--;
    // the sun, Integrators.cpp:61-77: light index >= count - 1 after the
    // decrement, i.e. pick >= n_lights; lit when nothing is hit at all
    if (s.use_sun && pick >= count - 1) {
        const V3 to_sun = -s.sun_dir;
        Ray r;
        r.d = to_sun;
        r.o = frag.pos + v4(0.01f * to_sun, 0.0f);
        r.len = 10000.0f;
        Hit h;
        c.shadow++;
        if (closest_hit(s, r, h)) return false;
        pdf = direct_pdf(fi.bsdf, s.microfacet, mat, frag.nrm, wi, to_sun);
        rad = s.sun_col;
        return true;
    }
    if (n_lights == 0) return false;  // the reference indexes an empty light list here
    // reference decrements the count only when u * count rounds up; index clamped here
    const uint32_t li = pick >= n_lights ? n_lights - 1 : pick;
    const Instance& L = s.inst[s.lights[li]];
    const Mesh& lm = s.meshes[L.mesh];
    V3 sp;
    if (lm.shape != PICO_SHAPE_TRIANGLES) {
        sp = implicit_sample(lm.shape, rng);
    } else {
        // LowerLevelMeshBVH::sample_geometry, Source/Core/LowerLevelMeshBVH.cpp:119-132
        const uint32_t tri = alias_sample(lm.alias_w, lm.alias_i, rng);
        const V2 xi = rng.next2d();
        const float su0 = std::sqrt(xi.x);
        const V2 b = {1 - su0, xi.y * su0};
        const uint32_t i0 = lm.idx[3 * tri], i1 = lm.idx[3 * tri + 1], i2 = lm.idx[3 * tri + 2];
        sp = (((1.0f - b.x - b.y) * lm.pos[i0]) + (b.x * lm.pos[i1])) + (b.y * lm.pos[i2]);
    }
    sp = xyz(mul(L.m, v4(sp, 1.0f)));
    const V3 to_light = normalize(sp - xyz(frag.pos));
    if (dot(to_light, frag.nrm) < 0.0f) return false;
    Ray r;
    r.d = to_light;
    r.o = frag.pos + v4(0.01f * to_light, 0.0f);
    r.len = 10000.0f;
    Hit h;
    c.shadow++;
    if (closest_hit(s, r, h)) {
        const Instance& hi = s.inst[h.inst];
        if (hi.bsdf == PICO_BSDF_LIGHT) {
            pdf = direct_pdf(fi.bsdf, s.microfacet, mat, frag.nrm, wi, to_light);
            rad = memis(s.mats[hi.material]);
            return true;
        }
    }
    return false;
}

// Camera::generate_ray, Source/Core/Camera.cpp:11-28.
static Ray camera_ray(const pico_camera& cam, uint32_t W, uint32_t H, V2 rnd, uint32_t px, uint32_t py) {
    const V2 rr = {rnd.x / 2.0f, rnd.y / 2.0f};
    V3 d = {(((float)px + rr.x) / (float)W - 0.5f) * cam.aspect, ((float)py + rr.y) / (float)H - 0.5f, 1.0f};
    const V3 D = {cam.direction[0], cam.direction[1], cam.direction[2]};
    const V3 U = {cam.up[0], cam.up[1], cam.up[2]};
    const V3 R = cross(normalize(D), U);
    d = normalize(((d.z * D) + (d.y * U)) + (d.x * R));
    Ray r;
    r.d = d;
    r.o = {cam.position[0], cam.position[1], cam.position[2], 1.0f};
    r.len = cam.far_plane;
    return r;
}

// Monte_Carlo_Integrator::integrate_ray + trace_ray, Source/Render/Integrators.cpp:130-209
// (recursion unrolled; the order of random draws is the reference's).
static V3 integrate(const Scene& s, const pico_camera& cam, uint32_t W, uint32_t H, uint32_t px, uint32_t py,
                    uint32_t max_depth, Rng& rng, Counters& c) {
    Ray ray = camera_ray(cam, W, H, rng.next2d(), px, py);
    Hit frag;
    c.camera++;
    if (!closest_hit(s, ray, frag)) return sky_lookup(s, ray.d);
    V3 payload = {0, 0, 0}, thr = {1, 1, 1};
    IorStack ior;
    for (uint32_t depth = 0;; ++depth) {
        if (depth == max_depth) break;
        const Instance& fi = s.inst[frag.inst];
        const pico_material& m = s.mats[fi.material];
        Sample smp = any_sample(fi.bsdf, s.microfacet, m, frag.nrm, ray.d, rng, ior);
        if (fi.bsdf == PICO_BSDF_LIGHT) payload = payload + thr * smp.energy;
        V3 rad; float pdf;
        if (direct_light(s, frag, -ray.d, rad, pdf, rng, c)) payload = payload + ((thr * pdf) * rad) * smp.energy;
        if (smp.P == 0.0f) break;
        const float q = smax(smax(thr.x, thr.y), thr.z);
        if (rng.uniform() > q) break;
        thr = thr * (1.0f / q);
        thr = thr * (smp.P * smp.energy);
        ray.o = frag.pos + v4(0.01f * (ior.inside() ? -frag.nrm : frag.nrm), 0.0f);
        ray.d = smp.L;
        Hit nh;
        c.ext++;
        if (closest_hit(s, ray, nh)) { frag = nh; continue; }
        payload = payload + thr * sky_lookup(s, smp.L);
        break;
    }
    if (bad(payload)) payload = {1.0f, 0.4f, 0.7f};
    return payload;
}

struct SceneBox { Scene s; };

}  // namespace cr

using namespace cr;

struct cpuref_scene { Scene s; };

static uint32_t cpu_threads(uint32_t t) {
    if (t) return t;
    uint32_t h = std::thread::hardware_concurrency();
    return h ? h : 1;
}

template <class F>
static void parallel_for(uint32_t n, uint32_t threads, F f) {
    threads = std::min(cpu_threads(threads), std::max(1u, n));
    if (threads <= 1) { for (uint32_t i = 0; i < n; ++i) f(i); return; }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> ts;
    for (uint32_t t = 0; t < threads; ++t)
        ts.emplace_back([&]() { for (uint32_t i; (i = next.fetch_add(1)) < n;) f(i); });
    for (auto& t : ts) t.join();
}

extern "C" {

cpuref_scene* cpuref_scene_create(const pico_scene_desc* d, uint32_t buckets) {
    if (!d) return nullptr;
    auto* out = new cpuref_scene();
    Scene& s = out->s;
    if (buckets < 2) buckets = 12;
    s.meshes.resize(d->n_meshes);
    for (uint32_t mi = 0; mi < d->n_meshes; ++mi) {
        Mesh& m = s.meshes[mi];
        m.shape = d->mesh_shape ? d->mesh_shape[mi] : PICO_SHAPE_TRIANGLES;
        if (m.shape != PICO_SHAPE_TRIANGLES) {
            m.bounds = {{-0.5f, -0.5f, -0.5f, 1.0f}, {0.5f, 0.5f, 0.5f, 1.0f}};  // get_bounds
            continue;
        }
        const uint32_t v0 = d->mesh_vertex_offset[mi], v1 = d->mesh_vertex_offset[mi + 1];
        const uint32_t t0 = d->mesh_triangle_offset[mi], t1 = d->mesh_triangle_offset[mi + 1];
        for (uint32_t v = v0; v < v1; ++v) {
            m.pos.push_back({d->positions[3 * v], d->positions[3 * v + 1], d->positions[3 * v + 2]});
            m.nrm.push_back({d->normals[3 * v], d->normals[3 * v + 1], d->normals[3 * v + 2]});
        }
        for (uint32_t t = t0; t < t1; ++t)
            for (int k = 0; k < 3; ++k) m.idx.push_back(d->indices[3 * t + k]);
        // mesh bounds as assimp's aiProcess_GenBoundingBoxes gives them (vertex min/max)
        AABB mb = AABB::empty();
        for (auto& p : m.pos) mb.add(v4(p, 1.0f));
        m.bounds = mb;
        // per-triangle bounds, Source/Core/LowerLevelMeshBVH.cpp:48-64
        std::vector<BV> vals;
        for (uint32_t i = 0; i < m.idx.size(); i += 3) {
            V3 mn = {INFINITY, INFINITY, INFINITY}, mx = {-INFINITY, -INFINITY, -INFINITY};
            for (int k = 0; k < 3; ++k) {
                V3 p = m.pos[m.idx[i + k]];
                mn = {p.x < mn.x ? p.x : mn.x, p.y < mn.y ? p.y : mn.y, p.z < mn.z ? p.z : mn.z};
                mx = {mx.x < p.x ? p.x : mx.x, mx.y < p.y ? p.y : mx.y, mx.z < p.z ? p.z : mx.z};
            }
            vals.push_back({{v4(mn, 1.0f), v4(mx, 1.0f)}, i / 3});
        }
        Builder b; b.buckets = buckets;
        m.root = b.split(vals.begin(), vals.end(), 32, 0);
        m.nodes = std::move(b.nodes);
        m.depth = b.max_depth_seen;
    }
    s.mats.assign(d->materials, d->materials + d->n_materials);
    s.inst.resize(d->n_instances);
    for (uint32_t i = 0; i < d->n_instances; ++i) {
        Instance& e = s.inst[i];
        for (int c = 0; c < 4; ++c)
            e.m.c[c] = {d->transforms[16 * i + 4 * c], d->transforms[16 * i + 4 * c + 1], d->transforms[16 * i + 4 * c + 2],
                        d->transforms[16 * i + 4 * c + 3]};
        e.minv = inverse(e.m);
        e.mesh = d->instance_mesh[i];
        e.material = d->instance_material[i];
        e.bsdf = d->instance_bsdf[i];
    }
    // lights: generate_sampling_data on the light's mesh, Source/Core/Scene.cpp:405-411
    for (uint32_t l = 0; l < d->n_lights; ++l) {
        const uint32_t ii = d->light_instances[l];
        s.lights.push_back(ii);
        Mesh& m = s.meshes[s.inst[ii].mesh];
        if (m.shape == PICO_SHAPE_TRIANGLES && m.alias_w.empty()) {
            std::vector<float> areas;
            for (uint32_t t = 0; t < m.idx.size(); t += 3) areas.push_back(tri_area(m.pos[m.idx[t]], m.pos[m.idx[t + 1]], m.pos[m.idx[t + 2]]));
            alias_build(areas, m.alias_w, m.alias_i);
        }
    }
    // TLAS over world-space instance bounds, Source/Core/UpperLevelBVH.cpp:27-57
    std::vector<BV> vals;
    for (uint32_t i = 0; i < s.inst.size(); ++i) vals.push_back({transform_box(s.meshes[s.inst[i].mesh].bounds, s.inst[i].m), i});
    Builder b; b.buckets = buckets;
    s.tlas_root = b.split(vals.begin(), vals.end(), 32, 0);
    s.tlas = std::move(b.nodes);
    s.tlas_depth = b.max_depth_seen;
    s.sky = {d->sky_rgb[0], d->sky_rgb[1], d->sky_rgb[2]};
    s.microfacet = d->microfacet;
    s.use_sun = d->use_sun != 0;
    s.sun_dir = {d->sun_direction[0], d->sun_direction[1], d->sun_direction[2]};
    s.sun_col = {d->sun_colour[0], d->sun_colour[1], d->sun_colour[2]};
    if (d->sky_image) {
        const pico_sky_image& k = *d->sky_image;
        s.sky_w = k.width; s.sky_h = k.height; s.sky_faces = k.faces; s.sky_fmt = k.format;
        const size_t bytes = (size_t)k.width * k.height * k.faces * (k.format == PICO_TEXEL_RGBA32F ? 16 : 4);
        const uint8_t* p = static_cast<const uint8_t*>(k.texels);
        s.sky_tex.assign(p, p + bytes);
    }
    return out;
}

void cpuref_scene_destroy(cpuref_scene* s) { delete s; }

uint32_t cpuref_scene_node_count(const cpuref_scene* sc, uint32_t* tlas_nodes, uint32_t* max_depth) {
    uint32_t n = 0, dmax = 0;
    for (auto& m : sc->s.meshes) { n += (uint32_t)m.nodes.size(); dmax = std::max(dmax, m.depth); }
    if (tlas_nodes) *tlas_nodes = (uint32_t)sc->s.tlas.size();
    if (max_depth) *max_depth = dmax + sc->s.tlas_depth;
    return n;
}

void cpuref_closest_hit(const cpuref_scene* sc, uint32_t n, const float* o, const float* d, float* t, uint32_t* inst,
                        uint32_t* prim, float* uv, float* pos, float* nrm) {
    parallel_for(n, 0, [&](uint32_t i) {
        Ray r;
        r.o = {o[3 * i], o[3 * i + 1], o[3 * i + 2], 1.0f};
        r.d = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        r.len = 10000.0f;
        Hit h;
        bool ok = closest_hit(sc->s, r, h);
        if (t) t[i] = ok ? h.t : INFINITY;
        if (inst) inst[i] = ok ? h.inst : kInvalid;
        if (prim) prim[i] = ok ? h.prim : kInvalid;
        if (uv) { uv[2 * i] = ok ? h.u : 0; uv[2 * i + 1] = ok ? h.v : 0; }
        if (pos) { pos[3 * i] = ok ? h.pos.x : 0; pos[3 * i + 1] = ok ? h.pos.y : 0; pos[3 * i + 2] = ok ? h.pos.z : 0; }
        if (nrm) { nrm[3 * i] = ok ? h.nrm.x : 0; nrm[3 * i + 1] = ok ? h.nrm.y : 0; nrm[3 * i + 2] = ok ? h.nrm.z : 0; }
    });
}

void cpuref_closest_hit_brute(const cpuref_scene* sc, uint32_t n, const float* o, const float* d, float* t,
                              uint32_t* inst, uint32_t* prim) {
    const Scene& s = sc->s;
    parallel_for(n, 0, [&](uint32_t i) {
        Ray r;
        r.o = {o[3 * i], o[3 * i + 1], o[3 * i + 2], 1.0f};
        r.d = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        float best = INFINITY; uint32_t bi = kInvalid, bp = kInvalid;
        for (uint32_t ii = 0; ii < s.inst.size(); ++ii) {
            const Instance& e = s.inst[ii];
            Ray ob;
            ob.o = mul(e.minv, r.o);
            ob.d = mul3(e.minv, r.d);
            const Mesh& m = s.meshes[e.mesh];
            float lb = INFINITY; uint32_t lp = kInvalid; float lu = 0, lv = 0;
            for (uint32_t p = 0; p < m.idx.size() / 3; ++p) {
                float tt, u, v;
                if (m.intersect(ob, p, tt, u, v) && tt < lb) { lb = tt; lp = p; lu = u; lv = v; }
            }
            if (lp == kInvalid) continue;
            V4 P; V3 N;
            m.interpolate(lp, lu, lv, P, N);
            V4 wp = mul(e.m, P);
            V4 dd = wp - r.o;
            float dist = std::sqrt((dd.x * dd.x + dd.y * dd.y) + (dd.z * dd.z + dd.w * dd.w));
            if (dist < best) { best = dist; bi = ii; bp = lp; }
        }
        t[i] = best; inst[i] = bi; prim[i] = bp;
    });
}

void cpuref_camera_rays(const pico_camera* cam, uint32_t w, uint32_t h, uint32_t n, const uint32_t* px,
                        const uint32_t* py, const float* xi, float* dirs) {
    for (uint32_t i = 0; i < n; ++i) {
        Ray r = camera_ray(*cam, w, h, {xi[2 * i], xi[2 * i + 1]}, px[i], py[i]);
        dirs[3 * i] = r.d.x; dirs[3 * i + 1] = r.d.y; dirs[3 * i + 2] = r.d.z;
    }
}

void cpuref_primary_ids(const cpuref_scene* sc, const pico_camera* cam, uint32_t w, uint32_t h, float xx, float xy,
                        uint32_t* inst, uint32_t* prim) {
    parallel_for(h, 0, [&](uint32_t y) {
        for (uint32_t x = 0; x < w; ++x) {
            Ray r = camera_ray(*cam, w, h, {xx, xy}, x, y);
            Hit hh;
            bool ok = closest_hit(sc->s, r, hh);
            inst[y * w + x] = ok ? hh.inst : kInvalid;
            prim[y * w + x] = ok ? hh.prim : kInvalid;
        }
    });
}

uint32_t cpuref_pixel_seed(uint32_t frame_seed, uint32_t pixel) {
    return (uint32_t)mix64(((uint64_t)frame_seed << 32) ^ (uint64_t)pixel ^ 0x5851F42D4C957F2DULL);
}

// Per-pixel sample loop with the running mean / variance update and early-out,
// Source/Core/Scene.cpp:157-197; tiles in Util::Tiler order (Source/Util/Tiler.hpp:20-49)
// run on a pool of threads.
int cpuref_render(const cpuref_scene* sc, const pico_camera* cam, const pico_render_params* p, uint32_t rng_mode,
                  uint32_t threads, float* pixels, uint32_t* counts, float* variance, pico_stats* stats) {
    if (!sc || !cam || !p || !pixels || !counts || !variance) return PICO_ERR_ARG;
    const uint32_t W = p->width, H = p->height;
    const uint32_t tw = p->tile_w ? p->tile_w : 64, th = p->tile_h ? p->tile_h : 64;
    if (!(p->flags & PICO_RENDER_RESUME)) {
        std::memset(pixels, 0, sizeof(float) * 3 * W * H);
        std::memset(counts, 0, sizeof(uint32_t) * W * H);
        std::memset(variance, 0, sizeof(float) * 3 * W * H);
    }
    std::vector<std::pair<uint32_t, uint32_t>> tiles;
    for (uint32_t x = 0; x < W; x += tw)
        for (uint32_t y = 0; y < H; y += th) tiles.push_back({x, y});
    std::atomic<uint64_t> cam_rays{0}, ext_rays{0}, sh_rays{0}, nsamples{0};
    parallel_for((uint32_t)tiles.size(), threads, [&](uint32_t ti) {
        Counters c;
        uint64_t ns = 0;
        const uint32_t x0 = tiles[ti].first, y0 = tiles[ti].second;
        const uint32_t x1 = std::min(W, x0 + tw), y1 = std::min(H, y0 + th);
        for (uint32_t y = y0; y < y1; ++y)
            for (uint32_t x = x0; x < x1; ++x) {
                const uint32_t flat = y * W + x;
                if (counts[flat] >= p->spp) continue;
                Rng rng;
                if (rng_mode == CPUREF_RNG_REFERENCE) rng.seed_reference(cpuref_pixel_seed(p->frame_seed, flat));
                const uint32_t base = counts[flat];
                for (uint32_t i = 0; i < p->spp; ++i) {
                    if (rng_mode == CPUREF_RNG_PCG) rng.seed_pcg(p->frame_seed, p->view_id, base + i, flat);
                    const uint32_t prev = counts[flat];
                    V3 x3 = integrate(sc->s, *cam, W, H, x, y, p->max_depth, rng, c);
                    ++ns;
                    V3 pix = {pixels[3 * flat], pixels[3 * flat + 1], pixels[3 * flat + 2]};
                    V3 var = {variance[3 * flat], variance[3 * flat + 1], variance[3 * flat + 2]};
                    if (prev < 1) {
                        counts[flat] = 1;
                        var = {0, 0, 0};
                    } else {
                        const V3 np = pix + ((x3 - pix) * (1.0f / (float)prev));
                        var = ((var * (float)prev) + ((x3 - pix) * (x3 - np))) / (float)(prev + 1);
                        counts[flat] = prev + 1;
                        x3 = np;
                    }
                    pixels[3 * flat] = x3.x; pixels[3 * flat + 1] = x3.y; pixels[3 * flat + 2] = x3.z;
                    variance[3 * flat] = var.x; variance[3 * flat + 1] = var.y; variance[3 * flat + 2] = var.z;
                    if (prev > 4 && var.x <= p->max_variance && var.y <= p->max_variance && var.z <= p->max_variance) break;
                }
            }
        cam_rays += c.camera; ext_rays += c.ext; sh_rays += c.shadow; nsamples += ns;
    });
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->camera_rays = cam_rays; stats->extension_rays = ext_rays; stats->shadow_rays = sh_rays;
        stats->samples = nsamples;
    }
    return PICO_OK;
}

int cpuref_render_samples(const cpuref_scene* sc, const pico_camera* cam, const pico_render_params* p,
                          uint32_t rng_mode, uint32_t threads, uint32_t sample_base, float* out, uint32_t* rays_out) {
    const uint32_t W = p->width, H = p->height, N = W * H;
    parallel_for(H, threads, [&](uint32_t y) {
        for (uint32_t x = 0; x < W; ++x) {
            const uint32_t flat = y * W + x;
            Rng rng;
            if (rng_mode == CPUREF_RNG_REFERENCE) rng.seed_reference(cpuref_pixel_seed(p->frame_seed, flat));
            for (uint32_t i = 0; i < p->spp; ++i) {
                if (rng_mode == CPUREF_RNG_PCG) rng.seed_pcg(p->frame_seed, p->view_id, sample_base + i, flat);
                Counters c;
                V3 r = integrate(sc->s, *cam, W, H, x, y, p->max_depth, rng, c);
                const size_t o = ((size_t)i * N + flat) * 3;
                out[o] = r.x; out[o + 1] = r.y; out[o + 2] = r.z;
                if (rays_out) rays_out[(size_t)i * N + flat] = (uint32_t)(c.camera + c.ext + c.shadow);
            }
        }
    });
    return PICO_OK;
}

// Scene::generate_denoiser_inputs, Source/Core/Scene.cpp:248-294; the view
// matrix is glm::lookAt(eye, eye + dir, up) (right-handed).
void cpuref_gbuffer(const cpuref_scene* sc, const pico_camera* cam, uint32_t w, uint32_t h, float* normal,
                    float* position, float* albedo) {
    const Scene& s = sc->s;
    const V3 eye = {cam->position[0], cam->position[1], cam->position[2]};
    const V3 D = {cam->direction[0], cam->direction[1], cam->direction[2]};
    const V3 up = {cam->up[0], cam->up[1], cam->up[2]};
    const V3 f = normalize((eye + D) - eye);
    const V3 sv = normalize(cross(f, up));
    const V3 u = cross(sv, f);
    M4 view;
    view.c[0] = {sv.x, u.x, -f.x, 0.0f};
    view.c[1] = {sv.y, u.y, -f.y, 0.0f};
    view.c[2] = {sv.z, u.z, -f.z, 0.0f};
    view.c[3] = {-dot(sv, eye), -dot(u, eye), dot(f, eye), 1.0f};
    parallel_for(h, 0, [&](uint32_t y) {
        for (uint32_t x = 0; x < w; ++x) {
            const uint32_t flat = y * w + x;
            Ray r = camera_ray(*cam, w, h, {0.0f, 0.0f}, x, y);
            Hit hh;
            V3 n, pp, a;
            if (closest_hit(s, r, hh)) {
                n = hh.nrm;
                pp = xyz(mul(view, hh.pos));
                a = mdiff(s.mats[s.inst[hh.inst].material]);
            } else {
                n = -r.d;
                pp = r.d * r.len;
                a = sky_lookup(s, r.d);
            }
            normal[3 * flat] = n.x; normal[3 * flat + 1] = n.y; normal[3 * flat + 2] = n.z;
            position[3 * flat] = pp.x; position[3 * flat + 1] = pp.y; position[3 * flat + 2] = pp.z;
            albedo[3 * flat] = a.x; albedo[3 * flat + 1] = a.y; albedo[3 * flat + 2] = a.z;
        }
    });
}

// A-trous filter, Source/Util/Denoisers.cpp:18-126 (kernel weights at :30-34,
// including the 0.09563 entry at tap 17; tile-local white point and position
// range over the tapped set at :45-63; per-tap weights at :65-99).
void cpuref_atrous(const float* rgb, const float* normal, const float* position, const float* albedo, uint32_t w,
                   uint32_t h, uint32_t tw, uint32_t th, uint32_t levels, float* out) {
    static const float K[25] = {0.00292f, 0.01306f, 0.02154f, 0.01306f, 0.00292f, 0.01306f, 0.05855f, 0.09653f, 0.05855f,
                                0.01306f, 0.02154f, 0.09653f, 0.15915f, 0.09653f, 0.02154f, 0.01306f, 0.05855f, 0.09563f,
                                0.05855f, 0.01306f, 0.00292f, 0.01306f, 0.02154f, 0.01306f, 0.00292f};
    const size_t n = (size_t)w * h;
    std::vector<float> buf[2] = {std::vector<float>(rgb, rgb + 3 * n), std::vector<float>(3 * n)};
    auto ld = [](const float* b, size_t i) { return V3{b[3 * i], b[3 * i + 1], b[3 * i + 2]}; };
    std::vector<std::pair<uint32_t, uint32_t>> tiles;
    for (uint32_t x = 0; x < w; x += tw)
        for (uint32_t y = 0; y < h; y += th) tiles.push_back({x, y});
    for (uint32_t level = 0; level < levels; ++level) {
        const float* src = buf[level % 2].data();
        float* dst = buf[(level + 1) % 2].data();
        const int scale = (int)std::pow(2, level);
        parallel_for((uint32_t)tiles.size(), 0, [&](uint32_t ti) {
            const uint32_t x0 = tiles[ti].first, y0 = tiles[ti].second;
            const uint32_t x1 = std::min(w, x0 + tw), y1 = std::min(h, y0 + th);
            auto tap = [&](uint32_t x, uint32_t y, uint32_t u) {
                int ox = (int)(u % 5) - 2, oy = (int)(u / 5) - 2;
                int ix = std::clamp((int)x + ox * scale, 0, (int)w - 1);
                int iy = std::clamp((int)y + oy * scale, 0, (int)h - 1);
                return (size_t)iy * w + (size_t)ix;
            };
            float white = 0.0f, dmin = INFINITY, dmax = -INFINITY;
            for (uint32_t x = x0; x < x1; ++x)
                for (uint32_t y = y0; y < y1; ++y)
                    for (uint32_t u = 0; u < 25; ++u) {
                        const size_t fi = tap(x, y, u);
                        const V3 c = ld(src, fi);
                        const float lum = c.x * 0.2126f + c.y * 0.587f + c.z * 0.114f;
                        white = smax(white, lum);
                        const float len = length(ld(position, fi));
                        dmin = smin(dmin, len);
                        dmax = smax(dmax, len);
                    }
            const float range = dmax - dmin;
            for (uint32_t x = x0; x < x1; ++x)
                for (uint32_t y = y0; y < y1; ++y) {
                    const size_t ci = (size_t)y * w + x;
                    V3 res = {0, 0, 0};
                    float wsum = 0.0f;
                    for (uint32_t u = 0; u < 25; ++u) {
                        const size_t fi = tap(x, y, u);
                        V3 t = ld(normal, ci) - ld(normal, fi);
                        float dist = dot(t, t);
                        const float nw = smin(std::exp(-dist / 0.25f), 1.0f);
                        t = ld(position, ci) - ld(position, fi);
                        dist = dot(t, t);
                        const float pw = smin(std::exp(-dist / (range * range)), 1.0f);
                        t = ld(albedo, ci) - ld(albedo, fi);
                        dist = dot(t, t);
                        const float aw = smin(std::exp(-dist / (white * white)), 1.0f);
                        const float tw_ = K[u] * nw * pw * aw;
                        res = res + ld(src, fi) * tw_;
                        wsum += tw_;
                    }
                    const V3 o = res / wsum;
                    dst[3 * ci] = o.x; dst[3 * ci + 1] = o.y; dst[3 * ci + 2] = o.z;
                }
        });
    }
    std::memcpy(out, buf[levels % 2].data(), sizeof(float) * 3 * n);
}

// Util::reinhard_tone_mapping, Source/Util/ToneMappers.cpp:12-44.
void cpuref_tonemap(float* rgb, uint32_t w, uint32_t h) {
    const size_t n = (size_t)w * h;
    float white = 0.0f;
    for (size_t i = 0; i < n; ++i) white = smax(white, rgb[3 * i] * 0.2126f + rgb[3 * i + 1] * 0.587f + rgb[3 * i + 2] * 0.114f);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const float p = rgb[3 * i + c];
            const float num = p * (1.0f + (p / (white * white)));
            rgb[3 * i + c] = num / (1.0f + p);
        }
}

// Fixed-width edge-avoiding a-trous (Dammertz et al., HPG 2010): no
// counterpart in the reference, which uses tile-local widths instead
// (Source/Util/Denoisers.cpp:45-63).  B3-spline taps (1,4,6,4,1)/16 per axis
// at dilation 2^l, neighbours clamped to the image, tap weight
// h(u)h(v) exp(-(|dn|^2/sn^2 + (dd/d_c)^2/sd^2 + |da|^2/sa^2 + |dc|^2/(sc 2^-l)^2)),
// d = |position|, a width of 0 dropping its term.  Sums in FP64.
void cpuref_atrous_fixed(const float* rgb, const float* normal, const float* position, const float* albedo, uint32_t w,
                         uint32_t h, uint32_t levels, float sn, float sd, float sa, float sc, float* out) {
    const size_t n = (size_t)w * h;
    std::vector<float> a(rgb, rgb + 3 * n), b(3 * n), dist(n);
    for (size_t i = 0; i < n; ++i) {
        const double x = position[3 * i], y = position[3 * i + 1], z = position[3 * i + 2];
        dist[i] = (float)std::sqrt(x * x + y * y + z * z);
    }
    const double hb[5] = {1.0 / 16, 1.0 / 4, 3.0 / 8, 1.0 / 4, 1.0 / 16};
    auto inv2 = [](double s) { return s > 0.0 ? 1.0 / (s * s) : 0.0; };
    for (uint32_t l = 0; l < levels; ++l) {
        const int step = 1 << l;
        const double isn2 = inv2(sn), isd2 = inv2(sd), isa2 = inv2(sa), isc2 = inv2((double)sc / step);
        for (int y = 0; y < (int)h; ++y)
            for (int x = 0; x < (int)w; ++x) {
                const size_t i = (size_t)y * w + x;
                double acc[3] = {0, 0, 0}, ws = 0.0;
                const double dc2 = (double)dist[i] * dist[i];
                for (int v = 0; v < 5; ++v)
                    for (int u = 0; u < 5; ++u) {
                        const int xx = std::min(std::max(x + (u - 2) * step, 0), (int)w - 1);
                        const int yy = std::min(std::max(y + (v - 2) * step, 0), (int)h - 1);
                        const size_t j = (size_t)yy * w + xx;
                        double en = 0, ea = 0, ec = 0;
                        for (int k = 0; k < 3; ++k) {
                            const double dn = (double)normal[3 * i + k] - normal[3 * j + k];
                            const double da = (double)albedo[3 * i + k] - albedo[3 * j + k];
                            const double dcl = (double)a[3 * i + k] - a[3 * j + k];
                            en += dn * dn; ea += da * da; ec += dcl * dcl;
                        }
                        const double dd = (double)dist[i] - dist[j];
                        const double e = en * isn2 + (dc2 > 0 ? dd * dd / dc2 : 0.0) * isd2 + ea * isa2 + ec * isc2;
                        const double wt = hb[u] * hb[v] * std::exp(-e);
                        for (int k = 0; k < 3; ++k) acc[k] += wt * a[3 * j + k];
                        ws += wt;
                    }
                for (int k = 0; k < 3; ++k) b[3 * i + k] = (float)(acc[k] / ws);
            }
        std::swap(a, b);
    }
    std::memcpy(out, a.data(), 12 * n);
}

double cpuref_relmse(const float* img, const double* ref_sum, double ref_count, uint32_t n_pixels) {
    double acc = 0.0;
    for (size_t i = 0; i < (size_t)n_pixels * 3; ++i) {
        const double r = ref_sum[i] / ref_count;
        const double d = (double)img[i] - r;
        acc += d * d / (r * r + 1e-2);
    }
    return acc / ((double)n_pixels * 3.0);
}

void cpuref_dist_sample(uint32_t kind, uint32_t n, const float* xi, const float* v, const float* r, float* h) {
    for (uint32_t i = 0; i < n; ++i) {
        V2 x = {xi[2 * i], xi[2 * i + 1]};
        V3 V = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
        V3 H = kind == DIST_COS ? cos_sample(x) : (kind == DIST_BECKMANN ? beckmann_sample(x, V, r[i]) : ggx_sample(x, V, r[i]));
        h[3 * i] = H.x; h[3 * i + 1] = H.y; h[3 * i + 2] = H.z;
    }
}

void cpuref_dist_pdf(uint32_t kind, uint32_t n, const float* wo, const float* h, const float* r, float* pdf) {
    (void)wo;
    for (uint32_t i = 0; i < n; ++i) {
        V3 H = {h[3 * i], h[3 * i + 1], h[3 * i + 2]};
        pdf[i] = kind == DIST_COS ? cos_pdf(H) : (kind == DIST_BECKMANN ? beckmann_pdf(H, r[i]) : ggx_pdf(H, r[i]));
    }
}

void cpuref_tangent_frame(uint32_t n, const float* v, const float* nrm, float* rows9) {
    for (uint32_t i = 0; i < n; ++i) {
        Frame f = tangent_frame({v[3 * i], v[3 * i + 1], v[3 * i + 2]}, {nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]});
        float* o = rows9 + 9 * i;
        o[0] = f.t.x; o[1] = f.t.y; o[2] = f.t.z; o[3] = f.b.x; o[4] = f.b.y; o[5] = f.b.z; o[6] = f.n.x; o[7] = f.n.y; o[8] = f.n.z;
    }
}

void cpuref_xorshift(uint32_t seed, uint32_t n, uint32_t* out) {
    XorShift x(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = x.next();
}

void cpuref_hammersley(uint32_t n, const uint32_t* bits, float* out2) {
    for (uint32_t i = 0; i < n; ++i) { V2 v = hammersley(bits[i]); out2[2 * i] = v.x; out2[2 * i + 1] = v.y; }
}

void cpuref_pcg32(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* out) {
    Pcg32 p; p.seed(initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) out[i] = p.next();
}

void cpuref_pcg32_advance(uint64_t initstate, uint64_t initseq, uint64_t delta, uint32_t n, uint32_t* out) {
    Pcg32 p; p.seed(initstate, initseq); p.advance(delta);
    for (uint32_t i = 0; i < n; ++i) out[i] = p.next();
}

void cpuref_alias_build(uint32_t n, const float* weights, float* bw, uint32_t* ba) {
    std::vector<float> w(weights, weights + n), ow; std::vector<uint32_t> oa;
    alias_build(w, ow, oa);
    std::copy(ow.begin(), ow.end(), bw); std::copy(oa.begin(), oa.end(), ba);
}

uint32_t cpuref_alias_sample(uint32_t n, const float* bw, const uint32_t* ba, uint32_t seed, uint32_t draws, uint32_t* out) {
    std::vector<float> w(bw, bw + n); std::vector<uint32_t> a(ba, ba + n);
    Rng rng; rng.seed_reference(seed);
    for (uint32_t i = 0; i < draws; ++i) out[i] = alias_sample(w, a, rng);
    return draws;
}

void cpuref_triangle_areas(uint32_t n_tris, const float* p, const uint32_t* idx, float* areas) {
    for (uint32_t t = 0; t < n_tris; ++t) {
        auto P = [&](uint32_t k) { uint32_t v = idx[3 * t + k]; return V3{p[3 * v], p[3 * v + 1], p[3 * v + 2]}; };
        areas[t] = tri_area(P(0), P(1), P(2));
    }
}

/* BTDF samples: each from a ray whose IOR stack is the camera's 1.0, plus
 * the material's index when inside[i]; depth_after[i] is the stack size after. */
void cpuref_sky_sample(const pico_sky_image* img, const float* sky_rgb, uint32_t n, const float* dirs, float* out) {
    Scene s;
    s.sky = {sky_rgb[0], sky_rgb[1], sky_rgb[2]};
    if (img) {
        s.sky_w = img->width; s.sky_h = img->height; s.sky_faces = img->faces; s.sky_fmt = img->format;
        const size_t bytes = (size_t)img->width * img->height * img->faces * (img->format == PICO_TEXEL_RGBA32F ? 16 : 4);
        const uint8_t* p = static_cast<const uint8_t*>(img->texels);
        s.sky_tex.assign(p, p + bytes);
    }
    for (uint32_t i = 0; i < n; ++i) {
        const V3 c = sky_lookup(s, {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]});
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
}

void cpuref_btdf_sample(uint32_t bsdf, uint32_t mf, const pico_material* m, uint32_t n, const float* v, const float* nrm,
                        const uint8_t* inside, uint32_t rng_mode, uint64_t seed, float* l, float* p, float* energy,
                        uint32_t* depth_after) {
    Rng rng;
    if (rng_mode == CPUREF_RNG_REFERENCE) {
        rng.seed_reference((uint32_t)seed);
    } else {
        rng.mode = CPUREF_RNG_PCG;
        rng.hs.seed(seed, 1);
        rng.ms.seed(seed ^ 0xA0761D6478BD642FULL, 2);
    }
    for (uint32_t i = 0; i < n; ++i) {
        IorStack st;
        if (inside[i]) st.push(m->index_of_refraction);
        const V3 V = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
        Sample s = any_sample(bsdf, mf, *m, {nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]}, -V, rng, st);
        l[3 * i] = s.L.x; l[3 * i + 1] = s.L.y; l[3 * i + 2] = s.L.z;
        p[i] = s.P;
        energy[3 * i] = s.energy.x; energy[3 * i + 1] = s.energy.y; energy[3 * i + 2] = s.energy.z;
        depth_after[i] = (uint32_t)st.v.size();
    }
}

void cpuref_bsdf_sample(uint32_t bsdf, uint32_t mf, const pico_material* m, uint32_t n, const float* v, const float* nrm,
                        uint32_t rng_mode, uint64_t seed, float* l, float* p, float* energy) {
    Rng rng;
    if (rng_mode == CPUREF_RNG_REFERENCE) {
        rng.seed_reference((uint32_t)seed);
    } else {
        rng.mode = CPUREF_RNG_PCG;
        rng.hs.seed(seed, 1);
        rng.ms.seed(seed ^ 0xA0761D6478BD642FULL, 2);
    }
    for (uint32_t i = 0; i < n; ++i) {
        // incoming ray direction is -V
        V3 V = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
        Sample s = bsdf_sample(bsdf, mf, *m, {nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]}, -V, rng);
        l[3 * i] = s.L.x; l[3 * i + 1] = s.L.y; l[3 * i + 2] = s.L.z;
        p[i] = s.P;
        energy[3 * i] = s.energy.x; energy[3 * i + 1] = s.energy.y; energy[3 * i + 2] = s.energy.z;
    }
}

}  // extern "C"
