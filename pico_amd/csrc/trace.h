// Device-side building blocks of the path tracer: vector math, PCG32 streams,
// the two-level closest-hit traversal and the shading functions.
//
// Shading follows the reference's arithmetic order (glm's dot products are
// (x*x + y*y) + z*z, matrix-vector products go column by column) so that the
// device and the CPU restatement consume random numbers identically and agree
// per path to rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/pico_hip.h"
#include "device_types.h"

namespace pico {

#define PICO_PI_D 3.14159265358979323846

// ------------------------------------------------------------------ math --
struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ f3 reflect(f3 i, f3 n) { return i - (n * dot(n, i)) * 2.0f; }
// std::max / std::min semantics
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ bool bad(f3 a) {
    return isnan(a.x) || isnan(a.y) || isnan(a.z) || isinf(a.x) || isinf(a.y) || isinf(a.z);
}
__device__ __forceinline__ f3 ld3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ f3 xyz(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ float4 f4(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// --------------------------------------------------------------- random ---
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

// PCG32 output step (O'Neill 2014); `inc` is derived from the stream id.
__device__ __forceinline__ uint32_t pcg_next(uint64_t& state, uint64_t inc) {
    const uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}
__device__ __forceinline__ uint64_t pcg_seed(uint64_t initstate, uint64_t initseq) {
    const uint64_t inc = (initseq << 1u) | 1u;
    uint64_t st = 0;
    pcg_next(st, inc);
    st += initstate;
    pcg_next(st, inc);
    return st;
}

// Per-path streams: hs replaces the reference's xorshift/Hammersley generator,
// ms its mt19937_64 uniform float.  Stream ids 2p and 2p+1 of pixel p.
struct Rng {
    uint64_t hs, ms;
    uint64_t hinc, minc;
    __device__ void seed(uint32_t frame, uint32_t view, uint32_t sample, uint32_t pixel) {
        const uint64_t base = mix64(((uint64_t)frame << 32) | sample) + (uint64_t)view * 0xD1B54A32D192ED03ULL;
        hinc = ((2ull * pixel) << 1) | 1u;
        minc = ((2ull * pixel + 1ull) << 1) | 1u;
        hs = pcg_seed(base, 2ull * pixel);
        ms = pcg_seed(base ^ 0xA0761D6478BD642FULL, 2ull * pixel + 1ull);
    }
    __device__ void set_streams(uint32_t pixel) {
        hinc = ((2ull * pixel) << 1) | 1u;
        minc = ((2ull * pixel + 1ull) << 1) | 1u;
    }
    __device__ uint32_t next_u32() { return pcg_next(hs, hinc); }
    // Hammersley_Generator::next (Source/Core/RandUtils.cpp:23-27,87-90): one
    // draw gives (i / 2^32, radical_inverse(i)).
    __device__ float2 next2d() {
        const uint32_t i = next_u32();
        return make_float2((float)i * 0x1p-32f, (float)__builtin_bitreverse32(i) * 0x1p-32f);
    }
    __device__ float uniform() { return (float)(pcg_next(ms, minc) >> 8) * 0x1p-24f; }
};

// --------------------------------------------------------------- scene ----
struct DevScene {
    const DNode4* nodes;
    const DTri* tris;
    const DInstance* inst;
    const DXform* xf;
    const DMaterial* mats;
    const DLight* lights;
    const float* alias_w;
    const uint32_t* alias_i;
    const float* pos;
    const float* nrm;
    const uint32_t* idx;
    // per triangle (global index, mesh order) its shading record: the three
    // vertex positions then the three normals, 18 floats padded to 5 float4
    // (kShadeRec), so a shading point loads one record instead of three
    // indices and six vertex gathers
    const float4* tshade;
    uint32_t tlas_root;
    uint32_t n_lights;
    float sky[3];
    uint32_t microfacet;
    const uint8_t* sky_tex;  // sky image texels, or nullptr for the constant sky
    uint32_t sky_w, sky_h, sky_faces, sky_fmt;
    uint32_t use_sun;
    float sun_to[3];         // towards the sun: -sun_direction
    float sun_col[3];
};

struct Hit { float t, u, v; uint32_t inst, prim; };

struct TraceCounters { uint32_t nodes, tris, insts, spill_push, spill_pop; };

// TraceState::bsp of a ray in the TLAS (no stack depth is this large)
constexpr uint32_t kWorld = 0xffffffffu;

__device__ __forceinline__ float4 ldg4(const void* p) { return *reinterpret_cast<const float4*>(p); }

// Reciprocal for the slab test: v_rcp_f32 (1 ulp) refined by one
// Newton-Raphson step, within about 1 ulp of the IEEE quotient, in 3
// instructions instead of the 10 of an IEEE division.  The slab distances only
// cull boxes, and the far distance is widened by 2^-21 (4 ulps) relative, so
// the test stays conservative; hit distances never use it.
__device__ __forceinline__ float rcp_nr(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
}

// Reciprocal direction for the fused slab test lo * inv + (-o * inv).  A zero
// component would give inf * lo - inf * o = NaN there (the reference's
// (lo - o) * inv gives +-inf); it is replaced by +-1e-20 so the slab bounds
// stay finite and keep their sign (and the reciprocal's input is never a
// denormal).
__device__ __forceinline__ f3 slab_rcp(f3 d) {
    const float e = 1e-20f;
    return mk(rcp_nr(fabsf(d.x) > e ? d.x : copysignf(e, d.x)), rcp_nr(fabsf(d.y) > e ? d.y : copysignf(e, d.y)),
              rcp_nr(fabsf(d.z) > e ? d.z : copysignf(e, d.z)));
}

// Traversal stack of one lane: the first STACK entries in LDS (`lds`, stride
// `lds_stride`), the next `spill_cap` in device memory (`spill`, stride
// `spill_stride`, private to the lane); pushing beyond both sets `overflow`.
// The LDS part is addressed through an address-space-3 pointer, so that its
// accesses are LDS instructions: through a generic pointer (which a select
// between the LDS and the spill pointer makes) they become flat accesses,
// which also wait for every outstanding global load.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(1))) uint32_t global_u32;

template <int STACK>
struct TraceStack {
    lds_u32* lds;
    uint32_t lds_stride;
    global_u32* spill;
    uint32_t spill_stride;
    uint32_t spill_cap;
    uint32_t sp = 0;
    __device__ __forceinline__ TraceStack(uint32_t* lds_, uint32_t lds_stride_, uint32_t* spill_, uint32_t spill_stride_,
                                          uint32_t spill_cap_)
        : lds((lds_u32*)lds_), lds_stride(lds_stride_), spill((global_u32*)spill_), spill_stride(spill_stride_),
          spill_cap(spill_cap_) {}
    __device__ __forceinline__ void push(uint32_t v, uint32_t* overflow) {
        if (sp < STACK) {
            lds[sp * lds_stride] = v;
        } else if (sp < STACK + spill_cap) {
            spill[(sp - STACK) * spill_stride] = v;
        } else {
            *overflow = 1u;
            return;
        }
        ++sp;
    }
    // Pushes e0, ..., e(n-1) (n <= 3) in that order.  While all three slots
    // above the top are in LDS they are all written, whatever n is (the slots
    // above the new top are free): no branch per entry.
    __device__ __forceinline__ void push3(uint32_t n, uint32_t e0, uint32_t e1, uint32_t e2, uint32_t* overflow) {
        if (sp + 3 <= STACK) {
            lds[sp * lds_stride] = e0;
            lds[(sp + 1) * lds_stride] = e1;
            lds[(sp + 2) * lds_stride] = e2;
            sp += n;
        } else {
            if (n > 0) push(e0, overflow);
            if (n > 1) push(e1, overflow);
            if (n > 2) push(e2, overflow);
        }
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        if (sp < STACK) return lds[sp * lds_stride];
        return spill[(sp - STACK) * spill_stride];
    }
};

// The reference's order of std::min / std::max (NaN operands fall through as
// the reference's comparisons make them).

// Object-space ray (M^-1 o, mat3(M^-1) d) of an instance, transform_ray
// (Source/Core/AABB.cpp:43-52): M^-1 (o, 1) as (c0*x + c1*y) + (c2*z + c3),
// mat3(M^-1) d as (c0*x + c1*y) + c2*z.
__device__ __forceinline__ void object_ray(const DInstance* I, f3 o, f3 d, f3& ro, f3& rd) {
    const float4 m0 = ldg4(&I->inv[0][0]);
    const float4 m1 = ldg4(&I->inv[1][0]);
    const float4 m2 = ldg4(&I->inv[2][0]);
    ro = mk((m0.x * o.x + m0.y * o.y) + (m0.z * o.z + m0.w), (m1.x * o.x + m1.y * o.y) + (m1.z * o.z + m1.w),
            (m2.x * o.x + m2.y * o.y) + (m2.z * o.z + m2.w));
    rd = mk((m0.x * d.x + m0.y * d.y) + m0.z * d.z, (m1.x * d.x + m1.y * d.y) + m1.z * d.z,
            (m2.x * d.x + m2.y * d.y) + m2.z * d.z);
}

// Built-in implicit shapes, Source/Core/LowerLevelImplicitShapesBVH.cpp:21-88,
// in object space with the (not renormalised) object-space direction: the
// sphere's quadratic as the reference writes it, with t clamped to 0 for a
// ray starting inside, and the cube's slab test AABB::intersection_distance
// (Source/Core/AABB.cpp:108-147) with 1/d.  Returns the object-space
// parameter, or false.
__device__ __forceinline__ bool implicit_hit(uint32_t shape, f3 ro, f3 rd, float& tt) {
    if (shape == PICO_SHAPE_SPHERE) {
        const float b = dot(ro, rd);
        const float c = dot(ro, ro) - 0.5f * 0.5f;
        if (c > 0.0f && b > 0.0f) return false;
        const float discr = b * b - c;
        if (discr < 0.0f) return false;
        float t = -b - sqrtf(discr);
        if (t < 0.0f) t = 0.0f;
        tt = t;
        return true;
    }
    const f3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    const float t1 = (-0.5f - ro.x) * inv.x, t2 = (0.5f - ro.x) * inv.x;
    const float t3 = (-0.5f - ro.y) * inv.y, t4 = (0.5f - ro.y) * inv.y;
    const float t5 = (-0.5f - ro.z) * inv.z, t6 = (0.5f - ro.z) * inv.z;
    const float a1 = t1 > t2 ? t2 : t1, b1 = t1 > t2 ? t1 : t2;
    const float a2 = t3 > t4 ? t4 : t3, b2 = t3 > t4 ? t3 : t4;
    const float a3 = t5 > t6 ? t6 : t5, b3 = t5 > t6 ? t5 : t6;
    const float tmin = smax(smax(a1, a2), a3);
    const float tmax = smin(smin(b1, b2), b3);
    if (tmax < 0.0f || tmin > tmax) return false;
    if (tmin == INFINITY) return false;
    tt = tmin;
    return true;
}

// World hit point of an object-space point: M (p, 1) as (c0*x + c1*y) + (c2*z + c3).
__device__ __forceinline__ f3 to_world_point(const float* M, f3 P) {
    return mk((M[0] * P.x + M[4] * P.y) + (M[8] * P.z + M[12]), (M[1] * P.x + M[5] * P.y) + (M[9] * P.z + M[13]),
              (M[2] * P.x + M[6] * P.y) + (M[10] * P.z + M[14]));
}

// Inner nodes a lane may take in one traversal step, and triangles of a
// BLAS leaf it may test in one step.  The wave runs every stage of a step
// that any of its lanes takes, so the common one (inner nodes) is batched and
// the leaf loop is short.  Instanced 1080p frame, Mrays/s, with the leaf and
// instance stages after the inner nodes of the same step: inner 1 / 2 / 3
// with 1 triangle: 2628 / 2821 / 2790; inner 2 with 2 triangles: 2691.  With
// a lane whose node misses popping its next entry into the next inner slot:
// inner 2 / 3 / 4 / 5: 2994 / 3076 / 3045 / 3000.
constexpr int kInnerSteps = 3;
constexpr int kLeafTris = 1;

// Two-level closest-hit traversal over 4-wide nodes.  Each node tests its
// four quantised child boxes (planes origin + q 2^e, which contain the exact
// boxes) with the slab test, the plane distances taken as
// fma(q, 2^e inv, origin inv - o inv) per axis, continues with the nearest hit
// child and pushes the others farthest first; the stack depth at an instance
// entry is kept, and the world-space ray is restored when the stack is back
// at that depth.  t is the ray
// parameter of the world ray (directions are unit length, so t is the hit
// distance).  Hits at or beyond `tmax` are not reported (infinity for the
// reference's unbounded query).  Box tests widen the far distance by 2^-21
// relative so that rounding never culls a triangle whose bound the ray only
// grazes.
//
// Traversal state of one ray, advanced a few nodes, a leaf or an instance at a time
// by trace_step, so that a persistent kernel can hand a lane a new ray as
// soon as its current one is finished.
struct TraceState {
    f3 o, d;      // world ray
    f3 ro, rd;    // ray in the current space (world or the instance's object space)
    f3 inv, oi;   // slab-test reciprocals and -ro * inv
    f3 winv, woi; // the same of the world ray, restored on leaving an instance
    Hit h;
    uint32_t ref, cur;
    uint32_t bsp;  // stack depth at the instance entry (kWorld: in the TLAS)
};

template <int STACK>
__device__ __forceinline__ void trace_begin(const DevScene& s, TraceState& t, TraceStack<STACK>& st, f3 o, f3 d,
                                            float tmax = INFINITY) {
    t.o = o; t.d = d; t.ro = o; t.rd = d;
    t.inv = slab_rcp(d);
    t.oi = mk(-o.x * t.inv.x, -o.y * t.inv.y, -o.z * t.inv.z);
    t.winv = t.inv;
    t.woi = t.oi;
    t.h.t = tmax; t.h.u = 0.0f; t.h.v = 0.0f; t.h.inst = 0xffffffffu; t.h.prim = 0xffffffffu;
    t.ref = s.tlas_root;
    t.cur = 0xffffffffu;
    t.bsp = kWorld;
    st.sp = 0;
}

// Inner node t.ref: its four quantised child boxes against the ray; on a hit
// t.ref becomes the nearest child and the others are pushed farthest first
// (returns true), else nothing is left of it (returns false).
template <int STACK, bool COUNT>
__device__ __forceinline__ bool inner_node(const DevScene& s, TraceState& t, TraceStack<STACK>& st, TraceCounters* cnt,
                                           uint32_t* overflow) {
    // No FP contraction here (the library is built without it): the
    // object-space transform and Moller-Trumbore round exactly as the CPU
    // restatement does; only the box tests use explicit fmaf (conservative,
    // they only cull).
    const float kWide = 1.0000005f;
    const DNode4* n = s.nodes + t.ref;
    const float4 a = ldg4(&n->origin[0]);
    const uint4 qxy = *reinterpret_cast<const uint4*>(&n->qx[0]);
    const uint4 qzs = *reinterpret_cast<const uint4*>(&n->qz[0]);  // qz, scale_y, scale_z
    const uint4 r = *reinterpret_cast<const uint4*>(&n->ref[0]);
    if (COUNT) cnt->nodes++;
    const uint2 qz = make_uint2(qzs.x, qzs.y);
    float key[4];
    uint32_t cref[4] = {r.x, r.y, r.z, r.w};
    // 2^e inv per axis: a product by a power of two (exact, as ldexp)
    const float Ax = t.inv.x * a.w, Ay = t.inv.y * __uint_as_float(qzs.z), Az = t.inv.z * __uint_as_float(qzs.w);
    const float Bx = fmaf(a.x, t.inv.x, t.oi.x), By = fmaf(a.y, t.inv.y, t.oi.y), Bz = fmaf(a.z, t.inv.z, t.oi.z);
    // The plane distance fma(q, A, B) is monotonic in q (one rounding, A of
    // the reciprocal's sign and never 0), so on each axis the near plane is
    // the low one when the reciprocal is positive and the high one otherwise:
    // one select per axis and node of the four children's low / high bytes
    // replaces a min and a max per axis and child (same values; a NaN plane,
    // only possible with an infinite A, is skipped by max3 / min3, which can
    // only keep a box).
    const bool nx = t.inv.x < 0.0f, ny = t.inv.y < 0.0f, nz = t.inv.z < 0.0f;
    // the largest float below the hit distance found so far (always > 0)
    const float hp = __uint_as_float(__float_as_uint(t.h.t) - 1u);
    const uint32_t xn = nx ? qxy.y : qxy.x, xf = nx ? qxy.x : qxy.y;
    const uint32_t yn = ny ? qxy.w : qxy.z, yf = ny ? qxy.z : qxy.w;
    const uint32_t zn = nz ? qz.y : qz.x, zf = nz ? qz.x : qz.y;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int sh = 8 * i;
        const float x0 = fmaf((float)((xn >> sh) & 0xffu), Ax, Bx), x1 = fmaf((float)((xf >> sh) & 0xffu), Ax, Bx);
        const float y0 = fmaf((float)((yn >> sh) & 0xffu), Ay, By), y1 = fmaf((float)((yf >> sh) & 0xffu), Ay, By);
        const float z0 = fmaf((float)((zn >> sh) & 0xffu), Az, Bz), z1 = fmaf((float)((zf >> sh) & 0xffu), Az, Bz);
        const float tn = fmaxf(fmaxf(x0, y0), z0);
        const float tf = fminf(fminf(x1, y1), z1) * kWide;
        // tf >= 0, tn <= tf and tn < h as one comparison (h > 0, so tn < h
        // is tn <= hp); a NaN bound, only from a NaN plane, is skipped by
        // fmaxf / fminf, which can only keep a box
        const bool hit = cref[i] != kEmptyRef && fmaxf(tn, 0.0f) <= fminf(tf, hp);
        key[i] = hit ? tn : INFINITY;
    }
    // sort the four (distance, child) pairs, nearest first
#define PICO_CSWAP(i, j)                                                  \
    if (key[j] < key[i]) {                                                \
        const float tk = key[i]; key[i] = key[j]; key[j] = tk;            \
        const uint32_t tr = cref[i]; cref[i] = cref[j]; cref[j] = tr;     \
    }
    PICO_CSWAP(0, 1) PICO_CSWAP(2, 3) PICO_CSWAP(0, 2) PICO_CSWAP(1, 3) PICO_CSWAP(1, 2)
#undef PICO_CSWAP
    const int nh = (key[0] != INFINITY) + (key[1] != INFINITY) + (key[2] != INFINITY) + (key[3] != INFINITY);
    if (nh == 0) return false;
    if (COUNT && nh > 1) cnt->spill_push += (uint32_t)max(0, (int)st.sp + nh - 1 - max((int)st.sp, STACK));
    // the other hit children, farthest first: c3, c2, c1 of those that hit
    const uint32_t np = (uint32_t)nh - 1u;
    st.push3(np, np == 3 ? cref[3] : (np == 2 ? cref[2] : cref[1]), np == 3 ? cref[2] : cref[1], cref[1], overflow);
    t.ref = cref[0];
    return true;
}

// The triangles of BLAS leaf `leaf`, Moller-Trumbore with the reference's
// epsilons (Source/Core/LowerLevelMeshBVH.cpp:165-207), in the current
// (object) space.
// At most kLeafTris triangles per call: a leaf with more is left in t.ref
// with its remaining triangles (returns true), to go on in the next step.
template <bool COUNT>
__device__ __forceinline__ bool leaf_tris(const DevScene& s, TraceState& t, uint32_t leaf, TraceCounters* cnt) {
    const uint32_t first = leaf & kPrimMask, count = (leaf >> 27) & 15u;
    const uint32_t now = count < (uint32_t)kLeafTris ? count : (uint32_t)kLeafTris;
    for (uint32_t k = 0; k < now; ++k) {
        const DTri* tr = s.tris + first + k;
        const float4 p0 = ldg4(&tr->v0[0]);
        const float4 p1 = ldg4(&tr->e1[0]);
        const float4 p2 = ldg4(&tr->e2[0]);
        if (COUNT) cnt->tris++;
        const f3 e1 = mk(p1.x, p1.y, p1.z), e2 = mk(p2.x, p2.y, p2.z);
        const f3 hh = cross(t.rd, e2);
        const float det = dot(e1, hh);
        if (det > -0.0000001f && det < 0.0000001f) continue;
        const float f = 1.0f / det;
        const f3 sv = t.ro - mk(p0.x, p0.y, p0.z);
        const float u = f * dot(sv, hh);
        if (u < 0.0f || u > 1.0f) continue;
        const f3 q = cross(sv, e1);
        const float v = f * dot(t.rd, q);
        if (v < 0.0f || u + v > 1.0f) continue;
        const float tt = f * dot(e2, q);
        if (tt > 0.0000001f && tt < t.h.t) {
            t.h.t = tt; t.h.u = u; t.h.v = v; t.h.inst = t.cur; t.h.prim = __float_as_uint(p0.w);
        }
    }
    if (now == count) return false;
    t.ref = kLeafBit | ((count - now) << 27) | (first + now);
    return true;
}

// TLAS leaf t.ref (one instance): a built-in shape is intersected at once
// (returns false: nothing to descend into); a mesh instance records the stack
// depth of its entry, moves the ray into object space and makes the BLAS root
// the next item (returns true).
template <int STACK, bool COUNT>
__device__ __forceinline__ bool enter_instance(const DevScene& s, TraceState& t, TraceStack<STACK>& st,
                                               TraceCounters* cnt, uint32_t* overflow) {
    const uint32_t cur = t.ref & kPrimMask;
    const DInstance* I = s.inst + cur;
    const uint32_t root = I->blas_root;
    if (COUNT) cnt->insts++;
    f3 ro, rd;
    object_ray(I, t.o, t.d, ro, rd);
    if (root >= kImplicitRoot) {
        // a built-in shape: its hit, at the world distance |M p - o| the
        // reference compares (Source/Core/UpperLevelBVH.cpp:66-73); u keeps
        // the object-space parameter for the shading point
        float tt;
        if (implicit_hit(root & 0xffu, ro, rd, tt)) {
            const f3 w = to_world_point(s.xf[cur].m, ro + rd * tt);
            const f3 dd = w - t.o;
            const float dist = sqrtf((dd.x * dd.x + dd.y * dd.y) + (dd.z * dd.z + 0.0f));
            if (dist < t.h.t) {
                t.h.t = dist; t.h.u = tt; t.h.v = 0.0f; t.h.inst = cur; t.h.prim = 0;
            }
        }
        return false;
    }
    t.cur = cur;
    t.bsp = st.sp;
    t.ro = ro;
    t.rd = rd;
    t.inv = slab_rcp(t.rd);
    t.oi = mk(-t.ro.x * t.inv.x, -t.ro.y * t.inv.y, -t.ro.z * t.inv.z);
    t.ref = root;
    return true;
}

__device__ __forceinline__ void leave_instance(TraceState& t) {
    t.ro = t.o; t.rd = t.d;
    t.inv = t.winv;
    t.oi = t.woi;
    t.bsp = kWorld;
}

// One traversal step, in stages that a lane passes in order: up to
// kInnerSteps inner nodes (each continues with its nearest hit child, the
// others pushed farthest first; after a node that no child of hits, the next
// stack entry is popped into the next slot); then, when that reached a leaf
// (or the step began at one), a BLAS leaf's next triangles or a TLAS leaf's
// instance entry
// (the ray moves into its object space); then, when nothing is left to
// descend into, the next stack entry is popped.  The wave runs every stage in
// almost every step whichever lanes need it (the leaf and instance stages in
// 98 % and 85 % of the steps of the instanced 1080p frame), so a lane that
// reaches a leaf handles it in the same step instead of the next one: the
// per-ray node order is unchanged and a ray takes fewer steps (2821 Mrays/s
// against 2560 when each step took one kind of item).  Returns true once the
// stack is empty: t.h then holds the closest hit.
template <int STACK, bool COUNT>
__device__ __forceinline__ bool trace_step(const DevScene& s, TraceState& t, TraceStack<STACK>& st, TraceCounters* cnt,
                                           uint32_t* overflow) {
    bool done = false;  // nothing left below the current item: pop
    if (!(t.ref & kLeafBit)) {
#pragma unroll
        for (int k = 0; k < kInnerSteps; ++k) {
            if (!inner_node<STACK, COUNT>(s, t, st, cnt, overflow)) {
                if (k + 1 == kInnerSteps) { done = true; break; }
                // no child hit: the next stack entry takes the next slot of
                // this step (the pop below, done here)
                if (st.sp == t.bsp) leave_instance(t);
                if (st.sp == 0) return true;
                if (COUNT && st.sp > STACK) cnt->spill_pop++;
                t.ref = st.pop();
            }
            if (t.ref & kLeafBit) break;
        }
    }
    if (!done) {
        if (!(t.ref & kLeafBit)) return false;  // kInnerSteps inner nodes taken
        if (t.bsp != kWorld) {
            if (leaf_tris<COUNT>(s, t, t.ref, cnt)) return false;
        } else {
            if (enter_instance<STACK, COUNT>(s, t, st, cnt, overflow)) return false;
        }
    }
    // pop; back at the stack depth of its entry, the ray leaves the instance
    // (one pop per step: no sentinel entry marks the instance boundary)
    if (st.sp == t.bsp) leave_instance(t);
    if (st.sp == 0) return true;
    if (COUNT && st.sp > STACK) cnt->spill_pop++;
    t.ref = st.pop();
    return false;
}

// Closest hit of a world-space ray against the TLAS and the instanced
// BLASes: trace_begin, then trace_step until the stack is empty.
template <int STACK, bool COUNT>
__device__ __forceinline__ Hit trace_closest(const DevScene& s, f3 o, f3 d, TraceStack<STACK>& st, TraceCounters* cnt,
                                             uint32_t* overflow, float tmax = INFINITY) {
    TraceState t;
    trace_begin(s, t, st, o, d, tmax);
    while (!trace_step<STACK, COUNT>(s, t, st, cnt, overflow)) {
    }
    return t.h;
}

// ------------------------------------------------------------- shading ----
// Interpolated hit point in world space: interpolate_fragment
// (Source/Core/LowerLevelMeshBVH.cpp:134-163) then the instance transform
// (Source/Core/UpperLevelBVH.cpp:66-71): position M*p, normal
// normalize(mat3(M)*n).
struct Frag { f3 p; f3 n; };

// The shading point of a built-in shape from its object-space parameter:
// the sphere's normal normalize(p), the cube's the signed axis of p's largest
// component (Source/Core/LowerLevelImplicitShapesBVH.cpp:41-48, 80-88 and
// maximum_component_index, Source/Core/AABB.cpp:31-41), then the instance
// transform.
__device__ __forceinline__ Frag implicit_frag(const DevScene& s, const Hit& h, f3 o, f3 d) {
    const DInstance* I = s.inst + h.inst;
    f3 ro, rd;
    object_ray(I, o, d, ro, rd);
    const f3 P = ro + rd * h.u;
    f3 N;
    if ((I->blas_root & 0xffu) == PICO_SHAPE_SPHERE) {
        N = normalize(P);
    } else {
        const float ax = fabsf(P.x), ay = fabsf(P.y), az = fabsf(P.z);
        const int k = (ax >= ay && ax >= az) ? 0 : ((ay >= ax && ay >= az) ? 1 : 2);
        const float c = k == 0 ? P.x : (k == 1 ? P.y : P.z);
        const float sg = (float)((0.0f < c) - (c < 0.0f));  // glm::sign
        N = mk(k == 0 ? sg : 0.0f, k == 1 ? sg : 0.0f, k == 2 ? sg : 0.0f);
    }
    const float* M = s.xf[h.inst].m;
    Frag f;
    f.p = to_world_point(M, P);
    f.n = normalize(mk(M[0] * N.x + M[4] * N.y + M[8] * N.z, M[1] * N.x + M[5] * N.y + M[9] * N.z,
                       M[2] * N.x + M[6] * N.y + M[10] * N.z));
    return f;
}

__device__ __forceinline__ Frag make_frag(const DevScene& s, const Hit& h) {
    const DInstance& I = s.inst[h.inst];
    const uint32_t t = I.tri_offset + h.prim;
    const float4* r = s.tshade + (size_t)kShadeRec * t;
    const float4 a = r[0], b = r[1], c = r[2], d = r[3], e = r[4];
    const float w = 1.0f - h.v - h.u;
    const f3 P = ((w * mk(a.x, a.y, a.z)) + (h.u * mk(a.w, b.x, b.y))) + (h.v * mk(b.z, b.w, c.x));
    const f3 N = normalize(((w * mk(c.y, c.z, c.w)) + (h.u * mk(d.x, d.y, d.z))) + (h.v * mk(d.w, e.x, e.y)));
    const float* M = s.xf[h.inst].m;
    Frag f;
    // (c0*x + c1*y) + (c2*z + c3*1)
    f.p = mk((M[0] * P.x + M[4] * P.y) + (M[8] * P.z + M[12]), (M[1] * P.x + M[5] * P.y) + (M[9] * P.z + M[13]),
             (M[2] * P.x + M[6] * P.y) + (M[10] * P.z + M[14]));
    f.n = normalize(mk(M[0] * N.x + M[4] * N.y + M[8] * N.z, M[1] * N.x + M[5] * N.y + M[9] * N.z,
                       M[2] * N.x + M[6] * N.y + M[10] * N.z));
    return f;
}

// The shading point of any hit; (o, d) is the world ray that found it.
__device__ __forceinline__ Frag hit_frag(const DevScene& s, const Hit& h, f3 o, f3 d) {
    return s.inst[h.inst].blas_root >= kImplicitRoot ? implicit_frag(s, h, o, d) : make_frag(s, h);
}

// ImageCube::sample4 (Source/Core/Image.cpp:361-466) of the sky in direction
// d: the constant sky, or the nearest texel of the sky image.  A cube map
// takes the face of the dominant axis (ties go to z, then y) with
// resolve_cubemap_UV's face coordinates; a single image is equirectangular,
// u = atan2(z, x) / 2pi + 1/2, v = 1/2 - asin(y) / pi, with asin evaluated in
// double as the reference's call resolves.  Texel = uint(coordinate * size)
// mod size, so a coordinate of exactly 1 wraps to texel 0.
__device__ __forceinline__ f3 sky_radiance(const DevScene& s, f3 d) {
    if (!s.sky_tex) return mk(s.sky[0], s.sky[1], s.sky[2]);
    float u, v;
    uint32_t face = 0;
    if (s.sky_faces == 6) {
        const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
        float ma, a, b;
        if (az >= ax && az >= ay) {
            face = d.z < 0.0f ? 5u : 4u;
            ma = 0.5f / az;
            a = d.z < 0.0f ? -d.x : d.x;
            b = -d.y;
        } else if (ay >= ax) {
            face = d.y < 0.0f ? 3u : 2u;
            ma = 0.5f / ay;
            a = d.x;
            b = d.y < 0.0f ? -d.z : d.z;
        } else {
            face = d.x < 0.0f ? 1u : 0u;
            ma = 0.5f / ax;
            a = d.x < 0.0f ? d.z : -d.z;
            b = -d.y;
        }
        u = a * ma + 0.5f;
        v = b * ma + 0.5f;
    } else {
        const float inv2pi = (float)(1.0 / (3.14159265358979323846 * 2.0));
        const float invpi = (float)(1.0 / 3.14159265358979323846);
        u = atan2f(d.z, d.x) * inv2pi + 0.5f;
        v = 1.0f - ((float)asin((double)d.y) * invpi + 0.5f);
    }
    const uint32_t x = (uint32_t)(u * (float)s.sky_w) % s.sky_w;
    const uint32_t y = (uint32_t)(v * (float)s.sky_h) % s.sky_h;
    const size_t texel = ((size_t)face * s.sky_h + y) * s.sky_w + x;
    if (s.sky_fmt == PICO_TEXEL_RGBA32F) return ld3(reinterpret_cast<const float*>(s.sky_tex) + 4 * texel);
    const uchar4 c = reinterpret_cast<const uchar4*>(s.sky_tex)[texel];
    return mk(c.x / 255.0f, c.y / 255.0f, c.z / 255.0f);
}

// construct_world_to_tangent_transform, Source/Core/vectorUtils.cpp:9-35: rows (T, B, N).
struct Frame { f3 t, b, n; };
__device__ __forceinline__ Frame tangent_frame(f3 V, f3 N) {
    f3 t;
    if (fabsf(dot(V, N)) > 0.95f) {
        const f3 vv = fabsf(N.z) > 0.99f ? mk(1, 0, 0) : mk(0, 0, 1);
        t = normalize(cross(N, vv));
    } else {
        t = normalize(cross(V, N));
    }
    const f3 b = normalize(cross(t, N));
    return {t, b, N};
}
__device__ __forceinline__ f3 to_tangent(const Frame& f, f3 v) {
    return {f.t.x * v.x + f.t.y * v.y + f.t.z * v.z, f.b.x * v.x + f.b.y * v.y + f.b.z * v.z,
            f.n.x * v.x + f.n.y * v.y + f.n.z * v.z};
}
// Inverse of the frame matrix (adjugate / determinant), as the restatement does.
__device__ __forceinline__ f3 to_world(const Frame& f, f3 v) {
    const f3 c0 = cross(f.b, f.n), c1 = cross(f.n, f.t), c2 = cross(f.t, f.b);
    const float id = 1.0f / dot(f.t, c0);
    return {(c0.x * id) * v.x + (c1.x * id) * v.y + (c2.x * id) * v.z,
            (c0.y * id) * v.x + (c1.y * id) * v.y + (c2.y * id) * v.z,
            (c0.z * id) * v.x + (c1.z * id) * v.y + (c2.z * id) * v.z};
}

__device__ __forceinline__ float cos2t(f3 w) { return w.z * w.z; }
__device__ __forceinline__ float sin2t(f3 w) { return smax(0.0f, 1.0f - cos2t(w)); }
__device__ __forceinline__ float sint(f3 w) { return sqrtf(sin2t(w)); }
__device__ __forceinline__ float tan2t(f3 w) { return sin2t(w) / cos2t(w); }
__device__ __forceinline__ float clamp1(float x) { return fminf(fmaxf(x, -1.0f), 1.0f); }
__device__ __forceinline__ float cosp(f3 w) { const float s = sint(w); return (s == 0) ? 1.0f : clamp1(w.x / s); }
__device__ __forceinline__ float sinp(f3 w) { const float s = sint(w); return (s == 0) ? 0.0f : clamp1(w.y / s); }

// Cos_Weighted_Hemisphere_Distribution, Source/Render/Distributions.cpp:14-29.
__device__ __forceinline__ f3 cos_sample(float2 xi) {
    const float r = sqrtf(xi.x);
    const float phi = (float)((double)xi.y * PICO_PI_D * 2.0);
    return normalize(mk(r * cosf(phi), r * sinf(phi), sqrtf(smax(0.0f, 1.0f - xi.x))));
}
__device__ __forceinline__ float cos_pdf(f3 h) { return smax(0.0f, h.z / (float)PICO_PI_D); }

// Beckmann_All_Microfacet_Distribution, Source/Render/Distributions.cpp:31-78.
__device__ __forceinline__ f3 beckmann_sample(float2 xi, f3 V, float R) {
    const float alpha = R * R;
    float ls = logf(1.0f - xi.x);
    if (isinf(ls)) ls = 0.0f;
    const float tan2 = -alpha * alpha * ls;
    const float phi = (float)((double)(xi.y * 2.0f) * PICO_PI_D);
    const float ct = 1.0f / sqrtf(1.0f + tan2);
    const float st = sqrtf(smax(0.0f, 1.0f - ct * ct));
    f3 H = mk(st * cosf(phi), st * sinf(phi), ct);
    if (!(V.z * H.z > 0.0f)) H = -H;
    return normalize(H);
}
__device__ __forceinline__ float beckmann_pdf(f3 wh, float R) {
    const float r = smax(R, 1e-3f);
    const float x = logf(r);
    const float a = 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
    const float tan2 = tan2t(wh);
    float d;
    if (isinf(tan2)) {
        d = 0.0f;
    } else {
        const float cos4 = cos2t(wh) * cos2t(wh);
        const float c = cosp(wh), sn = sinp(wh);
        const float num = expf(-tan2 * ((c * c) / (a * a) + (sn * sn) / (a * a)));
        d = (float)((double)num / (PICO_PI_D * a * a * cos4));
    }
    d = fminf(fmaxf(d, 0.0f), 1.0f);
    return d * fabsf(wh.z);
}

// GGX (Walter et al. 2007) with alpha = R^2: D = a^2 / (pi cos^4 (a^2 + tan^2)^2),
// sampled by tan^2 = a^2 u / (1 - u); pdf_h = D |cos|.
__device__ __forceinline__ f3 ggx_sample(float2 xi, f3 V, float R) {
    const float a = R * R;
    const float tan2 = a * a * xi.x / (1.0f - xi.x);
    const float phi = (float)((double)(xi.y * 2.0f) * PICO_PI_D);
    const float ct = 1.0f / sqrtf(1.0f + tan2);
    const float st = sqrtf(smax(0.0f, 1.0f - ct * ct));
    f3 H = mk(st * cosf(phi), st * sinf(phi), ct);
    if (!(V.z * H.z > 0.0f)) H = -H;
    return normalize(H);
}
__device__ __forceinline__ float ggx_pdf(f3 wh, float R) {
    const float a = R * R;
    const float a2 = a * a;
    const float c2 = cos2t(wh);
    const float tan2 = tan2t(wh);
    if (isinf(tan2)) return 0.0f;
    const float den = a2 + tan2;
    const float D = a2 / ((float)PICO_PI_D * (c2 * c2) * (den * den));
    return D * fabsf(wh.z);
}

__device__ __forceinline__ f3 spec_sample(uint32_t mf, float2 xi, f3 V, float R) {
    return mf == 1u ? ggx_sample(xi, V, R) : beckmann_sample(xi, V, R);
}
__device__ __forceinline__ float spec_pdf(uint32_t mf, f3 H, float R) {
    return mf == 1u ? ggx_pdf(H, R) : beckmann_pdf(H, R);
}

struct BSample { f3 L; float P; f3 energy; };

// BSRDF::sample for the Diffuse, Specular, Dielectric, Light and Delta BSDFs,
// Source/Render/BSRDF.cpp:11-238.
__device__ __forceinline__ BSample bsdf_sample(uint32_t kind, uint32_t mf, const DMaterial& m, f3 N, f3 raydir, Rng& rng) {
    const f3 V = -raydir;
    BSample s;
    if (kind == PICO_BSDF_DELTA) {
        s.energy = ld3(m.specular);
        s.L = reflect(-V, N);
        s.P = 1.0f;
        return s;
    }
    const Frame f = tangent_frame(V, N);
    const f3 vt = normalize(to_tangent(f, V));
    if (kind == PICO_BSDF_DIFFUSE || kind == PICO_BSDF_LIGHT) {
        const float2 xi = rng.next2d();
        const f3 H = cos_sample(xi);
        s.P = cos_pdf(H);
        s.L = normalize(to_world(f, H));
        s.energy = kind == PICO_BSDF_LIGHT ? ld3(m.emissive) : ld3(m.diffuse);
        return s;
    }
    if (kind == PICO_BSDF_SPECULAR) {
        const float2 xi = rng.next2d();
        const f3 H = spec_sample(mf, xi, vt, m.roughness);
        const f3 L = normalize(reflect(-vt, H));
        s.L = normalize(to_world(f, L));
        s.P = spec_pdf(mf, H, m.roughness) / (2.0f * dot(vt, H));
        s.energy = ld3(m.specular);
        return s;
    }
    float2 xi = rng.next2d();
    if (xi.y >= m.reflectance) {
        const float2 x2 = rng.next2d();
        const f3 H = cos_sample(x2);
        s.L = normalize(to_world(f, H));
        s.P = cos_pdf(H);
        s.energy = ld3(m.diffuse);
    } else {
        xi = rng.next2d();
        const f3 H = spec_sample(mf, xi, vt, m.roughness);
        const f3 L = normalize(reflect(-vt, H));
        s.L = normalize(to_world(f, L));
        s.P = spec_pdf(mf, H, m.roughness) / (2.0f * dot(vt, H));
        s.energy = ld3(m.specular);
    }
    return s;
}

// The ray's index-of-refraction stack, Ray::push / pop /
// get_current_index_of_refraction / inside_geometry (Source/Core/AABB.cpp:54-79):
// `count` entries above the camera's 1.0 (the reference's size - 1), kept in
// a ring of kIorStack, so nesting up to kIorStack transmissive objects deep
// is exact.  A pop with nothing above the camera's entry leaves it in place
// (the reference asserts there).
constexpr uint32_t kIorStack = 4;  // the selects below assume 4
struct IorStack {
    float v[kIorStack];
    uint32_t count;
    // selects instead of a dynamic index keep the ring in registers
    __device__ void push(float x) {
        const uint32_t k = count % kIorStack;
        v[0] = k == 0 ? x : v[0];
        v[1] = k == 1 ? x : v[1];
        v[2] = k == 2 ? x : v[2];
        v[3] = k == 3 ? x : v[3];
        ++count;
    }
    __device__ void pop() { if (count) --count; }
    __device__ float current() const {
        if (!count) return 1.0f;
        const uint32_t k = (count - 1) % kIorStack;
        return k == 0 ? v[0] : (k == 1 ? v[1] : (k == 2 ? v[2] : v[3]));
    }
    __device__ bool inside() const { return count > 0; }
};

// Transparent_BTDF::sample, Source/Render/BSRDF.cpp:240-309: a microfacet
// normal from the specular distribution, the stack pushed on entering and
// popped on leaving, then glm::refract(view, H, 1 / IOR) as the reference
// calls it (NaN where the square root is of a negative, which the reference
// carries on to its NaN-pixel rule).
__device__ __forceinline__ BSample btdf_transparent(uint32_t mf, const DMaterial& m, f3 N, f3 raydir, Rng& rng,
                                                    IorStack& st) {
    const f3 V = -raydir;
    const Frame f = tangent_frame(V, N);
    const f3 vt = normalize(to_tangent(f, V));
    const bool enter = vt.z >= 0.0f;
    const float2 xi = rng.next2d();
    const f3 H = spec_sample(mf, xi, vt, m.roughness);
    const float eta = 1.0f / m.ior;
    if (enter) st.push(m.ior);
    else st.pop();
    const float d = dot(H, vt);
    const float k = 1.0f - eta * eta * (1.0f - d * d);
    const f3 L = (vt * eta - H * (eta * d + sqrtf(k))) * (k >= 0.0f ? 1.0f : 0.0f);
    BSample s;
    s.L = normalize(to_world(f, L));
    const float pdf = spec_pdf(mf, H, m.roughness);
    const f3 wh = normalize(vt + L * eta);
    const float den = dot(vt, wh) + eta * dot(L, wh);
    s.P = pdf * fabsf((eta * eta * dot(L, wh)) / (den * den));
    s.energy = ld3(m.diffuse);
    return s;
}

// Fresnel_BTDF::fresnel_factor, Source/Render/BSRDF.cpp:400-421 (dielectric
// Fresnel reflectance; eta_i is the material's index, eta_t the ray's).
__device__ __forceinline__ float fresnel_factor(float cos_i, float eta_i, float eta_t) {
    cos_i = cos_i < -1.0f ? -1.0f : (1.0f < cos_i ? 1.0f : cos_i);
    if (!(cos_i > 0.0f)) {
        const float t = eta_i;
        eta_i = eta_t;
        eta_t = t;
        cos_i = fabsf(cos_i);
    }
    const float sin_i = sqrtf(smax(0.0f, 1.0f - cos_i * cos_i));
    const float sin_t = eta_i / eta_t * sin_i;
    const float cos_t = sqrtf(smax(0.0f, 1.0f - sin_t * sin_t));
    const float rpa = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
    const float rpe = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
    return (rpa * rpa + rpe * rpe) / 2;
}

// Fresnel_BTDF::sample, Source/Render/BSRDF.cpp:335-362: one draw chooses
// reflection (the Specular BRDF) with the Fresnel reflectance, else
// transmission (the Transparent BTDF); the pdf is scaled by the choice.
__device__ __forceinline__ BSample btdf_fresnel(uint32_t mf, const DMaterial& m, f3 N, f3 raydir, Rng& rng,
                                                IorStack& st) {
    const f3 V = -raydir;
    const Frame f = tangent_frame(V, N);
    const f3 vt = normalize(to_tangent(f, V));
    const float2 xi = rng.next2d();
    const float F = fresnel_factor(vt.z, m.ior, st.current());
    BSample s;
    if (xi.x < F) {
        s = bsdf_sample(PICO_BSDF_SPECULAR, mf, m, N, raydir, rng);
        s.P *= F;
    } else {
        s = btdf_transparent(mf, m, N, raydir, rng, st);
        s.P *= 1.0f - F;
    }
    return s;
}

// BSRDF::pdf (Source/Render/BSRDF.cpp:42-45,87-91,151-155,229-232) at the
// half vector, as direct_lighting_pdf calls it (Source/Render/Integrators.cpp:13-22).
__device__ __forceinline__ float direct_pdf(uint32_t kind, uint32_t mf, const DMaterial& m, f3 N, f3 wi, f3 wo) {
    const Frame f = tangent_frame(wi, N);
    const f3 two = to_tangent(f, wo);
    f3 H = normalize(wi + wo);
    H = to_tangent(f, H);
    switch (kind) {
        case PICO_BSDF_DIFFUSE:
        case PICO_BSDF_LIGHT: return cos_pdf(H);
        case PICO_BSDF_SPECULAR: return spec_pdf(mf, H, m.roughness) / (2.0f * dot(two, H));
        case PICO_BSDF_DIELECTRIC: {
            const float r = m.reflectance;
            return (cos_pdf(H) * (1 - r)) + ((spec_pdf(mf, H, m.roughness) / (2.0f * dot(two, H))) * r);
        }
        default: return 0.0f;
    }
}

// Camera::generate_ray, Source/Core/Camera.cpp:11-28 (right = cross(normalize(D), up)).
struct DCam { f3 pos, dir, up, right; float aspect, far_plane; uint32_t w, h; };
__device__ __forceinline__ f3 camera_dir(const DCam& c, float2 rnd, uint32_t px, uint32_t py) {
    const float rx = rnd.x / 2.0f, ry = rnd.y / 2.0f;
    const f3 d = mk((((float)px + rx) / (float)c.w - 0.5f) * c.aspect, ((float)py + ry) / (float)c.h - 0.5f, 1.0f);
    return normalize(((d.z * c.dir) + (d.y * c.up)) + (d.x * c.right));
}

}  // namespace pico
