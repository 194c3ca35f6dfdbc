// Procedural scenes of the benchmark configurations (host code).
//
// The reference loads scenes from JSON/assimp (Source/Core/Scene.cpp:36-141);
// the configurations here are procedural and seeded, and are handed to the
// renderer (and to any checker) as one flattened pico_scene_desc.  Instance
// transforms are built as the reference builds them, T * R(quat) * S
// (Source/Core/Scene.cpp:359-361); materials are the reference's
// ConstantMetalnessRoughnessMaterial (Source/Render/BasicMaterials.cpp:28-39).
// Every light gets its own mesh: the reference regenerates a mesh's light
// sampling table once per light instance (Source/Core/Scene.cpp:405-408), so
// lights must not share a mesh.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/pico_hip.h"
#include "scene_store.h"

namespace {

struct Rng {  // PCG32 (O'Neill 2014)
    uint64_t state = 0, inc = 1;
    explicit Rng(uint64_t seed, uint64_t seq = 0x5eed) {
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
    float gauss() {
        float u1 = std::max(uniform(), 1e-7f), u2 = uniform();
        return std::sqrt(-2.0f * std::log(u1)) * std::cos(6.2831853f * u2);
    }
};

struct F3 { float x, y, z; };
static F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static F3 normalize(F3 a) { float l = std::sqrt(dot(a, a)); return l > 0 ? a * (1.0f / l) : a; }

struct Mesh {
    std::vector<F3> pos, nrm;
    std::vector<uint32_t> idx;
};

// Column-major 4x4.
struct Mat { float m[16]; };
static Mat identity() { Mat r{}; r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f; return r; }
static Mat mul(const Mat& a, const Mat& b) {
    Mat r{};
    for (int c = 0; c < 4; ++c)
        for (int rr = 0; rr < 4; ++rr) {
            float s = 0;
            for (int k = 0; k < 4; ++k) s += a.m[k * 4 + rr] * b.m[c * 4 + k];
            r.m[c * 4 + rr] = s;
        }
    return r;
}
static Mat translate(F3 t) { Mat r = identity(); r.m[12] = t.x; r.m[13] = t.y; r.m[14] = t.z; return r; }
static Mat scale(F3 s) { Mat r = identity(); r.m[0] = s.x; r.m[5] = s.y; r.m[10] = s.z; return r; }
// Rotation matrix of a unit quaternion (w, x, y, z), as glm::mat4_cast.
static Mat quat_mat(float w, float x, float y, float z) {
    Mat r = identity();
    r.m[0] = 1 - 2 * (y * y + z * z); r.m[1] = 2 * (x * y + w * z); r.m[2] = 2 * (x * z - w * y);
    r.m[4] = 2 * (x * y - w * z); r.m[5] = 1 - 2 * (x * x + z * z); r.m[6] = 2 * (y * z + w * x);
    r.m[8] = 2 * (x * z + w * y); r.m[9] = 2 * (y * z - w * x); r.m[10] = 1 - 2 * (x * x + y * y);
    return r;
}
static Mat rot_axis(F3 axis, float deg) {
    const float h = deg * 3.14159265358979f / 360.0f;
    F3 a = normalize(axis);
    const float s = std::sin(h);
    return quat_mat(std::cos(h), a.x * s, a.y * s, a.z * s);
}

static void smooth_normals(Mesh& m) {
    m.nrm.assign(m.pos.size(), F3{0, 0, 0});
    for (size_t t = 0; t < m.idx.size(); t += 3) {
        const F3 a = m.pos[m.idx[t]], b = m.pos[m.idx[t + 1]], c = m.pos[m.idx[t + 2]];
        const F3 n = cross(b - a, c - a);  // area weighted
        for (int k = 0; k < 3; ++k) m.nrm[m.idx[t + k]] = m.nrm[m.idx[t + k]] + n;
    }
    for (auto& n : m.nrm) n = normalize(n);
}

// Unit quad in the xz-plane (side 1, centred), shading normal n (+y or -y).
static Mesh quad(float ny) {
    Mesh m;
    m.pos = {{-0.5f, 0, -0.5f}, {0.5f, 0, -0.5f}, {0.5f, 0, 0.5f}, {-0.5f, 0, 0.5f}};
    m.nrm.assign(4, F3{0, ny, 0});
    m.idx = {0, 1, 2, 0, 2, 3};
    return m;
}

// Unit cube (side 1, centred) with flat per-face normals: 24 vertices, 12 triangles.
static Mesh cube() {
    Mesh m;
    const F3 axes[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int a = 0; a < 3; ++a)
        for (int sgn = -1; sgn <= 1; sgn += 2) {
            const F3 n = axes[a] * (float)sgn;
            const F3 u = axes[(a + 1) % 3], v = axes[(a + 2) % 3];
            const uint32_t b = (uint32_t)m.pos.size();
            const F3 c = n * 0.5f;
            m.pos.push_back(c - u * 0.5f - v * 0.5f);
            m.pos.push_back(c + u * 0.5f - v * 0.5f);
            m.pos.push_back(c + u * 0.5f + v * 0.5f);
            m.pos.push_back(c - u * 0.5f + v * 0.5f);
            for (int k = 0; k < 4; ++k) m.nrm.push_back(n);
            m.idx.insert(m.idx.end(), {b, b + 1, b + 2, b, b + 2, b + 3});
        }
    return m;
}

// Smooth random field: a few seeded sinusoids.
struct Noise {
    F3 k[6]; float ph[6];
    explicit Noise(Rng& r) {
        for (int i = 0; i < 6; ++i) {
            k[i] = F3{r.gauss(), r.gauss(), r.gauss()} * (1.5f + 2.0f * r.uniform());
            ph[i] = r.uniform(0.0f, 6.2831853f);
        }
    }
    float operator()(F3 p) const {
        float s = 0;
        for (int i = 0; i < 6; ++i) s += std::sin(dot(k[i], p) + ph[i]);
        return s / 6.0f;
    }
};

// Geodesic icosphere of frequency f (20 f^2 triangles), vertices shared along
// edges, radius perturbed by smooth noise of relative amplitude amp.
static Mesh icosphere(uint32_t f, float amp, Rng& rng) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    F3 v[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const uint32_t faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                                   {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                                   {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    Mesh m;
    Noise nz(rng);
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t>, uint32_t> key;
    // A grid point of face (A,B,C) with weights (a,b,c), a+b+c=f, is keyed by its
    // sorted (corner, weight) pairs so that faces sharing an edge share it.
    auto vid = [&](const uint32_t* fc, uint32_t wa, uint32_t wb, uint32_t wc) {
        std::pair<uint32_t, uint32_t> cw[3] = {{fc[0], wa}, {fc[1], wb}, {fc[2], wc}};
        std::sort(cw, cw + 3);
        // drop zero weights from the key so edge points match across faces
        uint32_t k0 = 0xffff, k1 = 0xffff, k2 = 0, k3 = 0;
        int n = 0;
        uint32_t ks[3][2];
        for (auto& p : cw)
            if (p.second) { ks[n][0] = p.first; ks[n][1] = p.second; ++n; }
        if (n == 1) { k0 = ks[0][0]; k2 = 0; }
        else if (n == 2) { k0 = ks[0][0]; k1 = ks[1][0]; k2 = ks[0][1]; }
        else { k0 = 1000 + (uint32_t)(&fc[0] - &faces[0][0]) / 3; k2 = wa; k3 = wb; }
        auto K = std::make_tuple(k0, k1, k2, k3);
        auto it = key.find(K);
        if (it != key.end()) return it->second;
        F3 p{0, 0, 0};
        for (int i = 0; i < n; ++i) p = p + v[ks[i][0]] * ((float)ks[i][1] / (float)f);
        p = normalize(p);
        p = p * (1.0f + amp * nz(p));
        uint32_t id = (uint32_t)m.pos.size();
        m.pos.push_back(p);
        key[K] = id;
        return id;
    };
    for (auto& fc : faces) {
        for (uint32_t i = 0; i < f; ++i)
            for (uint32_t j = 0; i + j < f; ++j) {
                // weights: A = f-i-j, B = i, C = j
                uint32_t p00 = vid(fc, f - i - j, i, j), p10 = vid(fc, f - i - j - 1, i + 1, j), p01 = vid(fc, f - i - j - 1, i, j + 1);
                m.idx.insert(m.idx.end(), {p00, p10, p01});
                if (i + j + 2 <= f) {
                    uint32_t p11 = vid(fc, f - i - j - 2, i + 1, j + 1);
                    m.idx.insert(m.idx.end(), {p10, p11, p01});
                }
            }
    }
    smooth_normals(m);
    return m;
}

// Torus (major radius 1, minor 0.35) on an nu x nv grid, displaced along the
// tube normal by smooth noise of relative amplitude amp.
static Mesh torus(uint32_t nu, uint32_t nv, float amp, Rng& rng) {
    Mesh m;
    Noise nz(rng);
    const float R = 1.0f, r = 0.35f;
    for (uint32_t i = 0; i < nu; ++i)
        for (uint32_t j = 0; j < nv; ++j) {
            const float u = 6.2831853f * (float)i / (float)nu, v = 6.2831853f * (float)j / (float)nv;
            const F3 c = {R * std::cos(u), 0.0f, R * std::sin(u)};
            const F3 n = {std::cos(v) * std::cos(u), std::sin(v), std::cos(v) * std::sin(u)};
            const F3 p = c + n * (r * (1.0f + amp * nz(c + n * r)));
            m.pos.push_back(p);
        }
    for (uint32_t i = 0; i < nu; ++i)
        for (uint32_t j = 0; j < nv; ++j) {
            const uint32_t a = i * nv + j, b = ((i + 1) % nu) * nv + j, c = ((i + 1) % nu) * nv + (j + 1) % nv,
                           d = i * nv + (j + 1) % nv;
            m.idx.insert(m.idx.end(), {a, b, c, a, c, d});
        }
    smooth_normals(m);
    return m;
}

// ConstantMetalnessRoughnessMaterial, Source/Render/BasicMaterials.cpp:28-39.
static pico_material metal_rough(F3 albedo, float metal, float rough, F3 emissive) {
    pico_material m{};
    const F3 d = albedo * (1.0f - 0.04f) * (1.0f - metal);
    m.diffuse[0] = d.x; m.diffuse[1] = d.y; m.diffuse[2] = d.z;
    const F3 f0 = F3{0.04f, 0.04f, 0.04f} + (albedo - F3{0.04f, 0.04f, 0.04f}) * metal;
    m.specular[0] = f0.x; m.specular[1] = f0.y; m.specular[2] = f0.z;
    m.roughness = rough;
    m.emissive[0] = emissive.x; m.emissive[1] = emissive.y; m.emissive[2] = emissive.z;
    return m;
}

}  // namespace

static uint32_t add_mesh(pico_scene* s, const Mesh& m) {
    static_assert(sizeof(F3) == 12, "packed float triples");
    return s->add_mesh(&m.pos[0].x, &m.nrm[0].x, m.pos.size(), m.idx.data(), m.idx.size());
}
static uint32_t add_instance(pico_scene* s, uint32_t mesh, const Mat& M, uint32_t mat, uint32_t bsdf) {
    return s->add_instance(mesh, M.m, mat, bsdf);
}

static void cornell_into(pico_scene* s, Rng& rng, bool jitter) {
    const uint32_t wall = add_mesh(s, quad(1.0f));
    const uint32_t box = add_mesh(s, cube());
    const uint32_t lightm = add_mesh(s, quad(-1.0f));
    auto rand_mat = [&]() {
        F3 a = {rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f)};
        return s->add_material(metal_rough(a, 0.0f, rng.uniform(0.1f, 0.9f), {0, 0, 0}));
    };
    // walls x=+-1, y=+-1, back z=+1 with inward normals; open front
    const Mat S = scale({2.0f, 1.0f, 2.0f});
    add_instance(s, wall, mul(translate({0, -1, 0}), S), rand_mat(), PICO_BSDF_DIELECTRIC);                          // floor
    add_instance(s, wall, mul(translate({0, 1, 0}), mul(rot_axis({1, 0, 0}, 180.0f), S)), rand_mat(), PICO_BSDF_DIELECTRIC);  // ceiling
    add_instance(s, wall, mul(translate({-1, 0, 0}), mul(rot_axis({0, 0, 1}, -90.0f), S)), rand_mat(), PICO_BSDF_DIELECTRIC); // left
    add_instance(s, wall, mul(translate({1, 0, 0}), mul(rot_axis({0, 0, 1}, 90.0f), S)), rand_mat(), PICO_BSDF_DIELECTRIC);   // right
    add_instance(s, wall, mul(translate({0, 0, 1}), mul(rot_axis({1, 0, 0}, -90.0f), S)), rand_mat(), PICO_BSDF_DIELECTRIC);  // back
    // two boxes rotated +-18 degrees about y
    const float jx = jitter ? rng.uniform(-0.15f, 0.15f) : 0.0f, jz = jitter ? rng.uniform(-0.15f, 0.15f) : 0.0f;
    const float r1 = jitter ? rng.uniform(-40.0f, 40.0f) : 18.0f, r2 = jitter ? rng.uniform(-40.0f, 40.0f) : -18.0f;
    add_instance(s, box, mul(translate({-0.35f + jx, -0.698f, 0.3f + jz}), mul(rot_axis({0, 1, 0}, r1), scale({0.6f, 0.6f, 0.6f}))),
                    rand_mat(), PICO_BSDF_DIELECTRIC);
    add_instance(s, box, mul(translate({0.35f - jx, -0.398f, -0.2f - jz}), mul(rot_axis({0, 1, 0}, r2), scale({0.6f, 1.2f, 0.6f}))),
                    rand_mat(), PICO_BSDF_DIELECTRIC);
    // quad light 0.5 x 0.5 at y = 0.99 facing down
    const float le = jitter ? rng.uniform(5.0f, 20.0f) : 10.0f;
    const uint32_t lm = s->add_material(metal_rough({0.8f, 0.8f, 0.8f}, 0.0f, 0.5f, {le, le, le}));
    const float lx = jitter ? rng.uniform(-0.4f, 0.4f) : 0.0f, lz = jitter ? rng.uniform(-0.4f, 0.4f) : 0.0f;
    add_instance(s, lightm, mul(translate({lx, 0.99f, lz}), scale({0.5f, 1.0f, 0.5f})), lm, PICO_BSDF_LIGHT);
    s->cam = pico_camera{{0.0f, 0.0f, -1.0f}, {0.0f, 0.0f, 1.0f}, {0.0f, 1.0f, 0.0f}, 1.0f, 10000.0f};
}

static void instanced_into(pico_scene* s, Rng& rng, uint32_t n_types, uint32_t n_inst, uint32_t tmin, uint32_t tmax,
                           float extent, uint32_t n_lights, float sky) {
    std::vector<uint32_t> meshes;
    for (uint32_t k = 0; k < n_types; ++k) {
        const double target = n_types > 1 ? (double)tmin * std::pow((double)tmax / tmin, (double)k / (n_types - 1)) : tmin;
        Mesh m;
        if (k % 2 == 0) {
            uint32_t f = std::max(1u, (uint32_t)std::lround(std::sqrt(target / 20.0)));
            m = icosphere(f, 0.06f, rng);
        } else {
            uint32_t nv = std::max(3u, (uint32_t)std::lround(std::sqrt(target / 2.0 / 2.5)));
            uint32_t nu = std::max(3u, (uint32_t)std::lround(target / 2.0 / nv));
            m = torus(nu, nv, 0.25f, rng);
        }
        meshes.push_back(add_mesh(s, m));
    }
    for (uint32_t i = 0; i < n_inst; ++i) {
        const uint32_t mesh = meshes[rng.next() % n_types];
        F3 p = {rng.uniform(0, extent), rng.uniform(0, extent), rng.uniform(0, extent)};
        float qw = rng.gauss(), qx = rng.gauss(), qy = rng.gauss(), qz = rng.gauss();
        const float ql = std::sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
        qw /= ql; qx /= ql; qy /= ql; qz /= ql;
        const float sc = rng.uniform(0.5f, 3.0f);
        const Mat M = mul(translate(p), mul(quat_mat(qw, qx, qy, qz), scale({sc, sc, sc})));
        F3 a = {rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f)};
        const uint32_t mat = s->add_material(metal_rough(a, 0.0f, rng.uniform(0.1f, 0.9f), {0, 0, 0}));
        add_instance(s, mesh, M, mat, PICO_BSDF_DIELECTRIC);
    }
    for (uint32_t l = 0; l < n_lights; ++l) {
        const uint32_t lmesh = add_mesh(s, quad(-1.0f));
        const float e = rng.uniform(5.0f, 20.0f);
        const uint32_t mat = s->add_material(metal_rough({0.8f, 0.8f, 0.8f}, 0.0f, 0.5f, {e, e, e}));
        const float side = 0.15f * extent;
        F3 p = {rng.uniform(0.2f * extent, 0.8f * extent), rng.uniform(0.85f * extent, extent), rng.uniform(0.2f * extent, 0.8f * extent)};
        add_instance(s, lmesh, mul(translate(p), scale({side, 1.0f, side})), mat, PICO_BSDF_LIGHT);
    }
    s->sky[0] = s->sky[1] = s->sky[2] = sky;
    const float c = 0.5f * extent;
    // outside the box, looking at its centre; the view is filled by the box
    s->cam = pico_camera{{c, c, -0.5f * extent}, {0.0f, 0.0f, 1.0f}, {0.0f, 1.0f, 0.0f}, 1.0f, 10000.0f};
}

extern "C" {

int pico_scene_cornell(uint32_t seed, uint32_t microfacet, pico_scene** out) {
    if (!out || microfacet > PICO_MICROFACET_GGX) return PICO_ERR_ARG;
    auto* s = new pico_scene();
    Rng rng(seed, 1);
    cornell_into(s, rng, false);
    s->microfacet = microfacet;
    *out = s;
    return PICO_OK;
}

int pico_scene_shapes(uint32_t seed, uint32_t microfacet, pico_scene** out) {
    if (!out || microfacet > PICO_MICROFACET_GGX) return PICO_ERR_ARG;
    auto* s = new pico_scene();
    Rng rng(seed, 4);
    const uint32_t wall = add_mesh(s, quad(1.0f));
    const uint32_t lightm = add_mesh(s, quad(-1.0f));
    const uint32_t sphere = s->add_shape(PICO_SHAPE_SPHERE);
    const uint32_t cubes = s->add_shape(PICO_SHAPE_CUBE);
    const uint32_t lamp_ball = s->add_shape(PICO_SHAPE_SPHERE);  // each light has its own shape entry
    const uint32_t lamp_box = s->add_shape(PICO_SHAPE_CUBE);
    auto rand_mat = [&]() {
        F3 a = {rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f), rng.uniform(0.2f, 0.8f)};
        return s->add_material(metal_rough(a, 0.0f, rng.uniform(0.1f, 0.9f), {0, 0, 0}));
    };
    const Mat S = scale({2.0f, 1.0f, 2.0f});
    add_instance(s, wall, mul(translate({0, -1, 0}), S), rand_mat(), PICO_BSDF_DIELECTRIC);
    add_instance(s, wall, mul(translate({0, 1, 0}), mul(rot_axis({1, 0, 0}, 180.0f), S)), rand_mat(), PICO_BSDF_DIELECTRIC);
    add_instance(s, wall, mul(translate({-1, 0, 0}), mul(rot_axis({0, 0, 1}, -90.0f), S)), rand_mat(), PICO_BSDF_DIELECTRIC);
    add_instance(s, wall, mul(translate({1, 0, 0}), mul(rot_axis({0, 0, 1}, 90.0f), S)), rand_mat(), PICO_BSDF_DIELECTRIC);
    add_instance(s, wall, mul(translate({0, 0, 1}), mul(rot_axis({1, 0, 0}, -90.0f), S)), rand_mat(), PICO_BSDF_DIELECTRIC);
    // Spheres keep unit scale: the reference solves the sphere's quadratic as
    // if the object-space direction were unit length, which holds only then.
    // A diffuse-and-specular sphere, a metal cube turned 30 degrees (just
    // above the floor: coplanar faces would tie), a glass sphere.
    add_instance(s, sphere, mul(translate({-0.45f, -0.5f, 0.4f}), rot_axis({0, 1, 0}, 25.0f)), rand_mat(), PICO_BSDF_DIELECTRIC);
    const uint32_t metal = s->add_material(metal_rough({0.9f, 0.8f, 0.6f}, 1.0f, 0.25f, {0, 0, 0}));
    add_instance(s, cubes, mul(translate({0.45f, -0.74f, 0.35f}), mul(rot_axis({0, 1, 0}, 30.0f), scale({0.5f, 0.5f, 0.5f}))),
                 metal, PICO_BSDF_SPECULAR);
    pico_material glass = metal_rough({0.95f, 0.95f, 0.95f}, 0.0f, 0.1f, {0, 0, 0});
    glass.diffuse[0] = glass.diffuse[1] = glass.diffuse[2] = 0.95f;
    glass.index_of_refraction = 1.5f;
    add_instance(s, sphere, translate({0.45f, 0.1f, -0.3f}), s->add_material(glass), PICO_BSDF_FRESNEL);
    // three lights: the ceiling quad, a sphere in a top corner, a small cube
    const uint32_t lm = s->add_material(metal_rough({0.8f, 0.8f, 0.8f}, 0.0f, 0.5f, {6, 6, 6}));
    add_instance(s, lightm, mul(translate({0, 0.99f, 0}), scale({0.5f, 1.0f, 0.5f})), lm, PICO_BSDF_LIGHT);
    const uint32_t lm2 = s->add_material(metal_rough({0.8f, 0.8f, 0.8f}, 0.0f, 0.5f, {3, 2.5f, 1.5f}));
    add_instance(s, lamp_ball, translate({-0.5f, 0.5f, 0.5f}), lm2, PICO_BSDF_LIGHT);
    const uint32_t lm3 = s->add_material(metal_rough({0.8f, 0.8f, 0.8f}, 0.0f, 0.5f, {4, 6, 8}));
    add_instance(s, lamp_box, mul(translate({-0.6f, 0.6f, -0.3f}), mul(rot_axis({1.0f, 0.7f, 0.2f}, 37.0f), scale({0.2f, 0.2f, 0.2f}))),
                 lm3, PICO_BSDF_LIGHT);
    s->cam = pico_camera{{0.0f, 0.0f, -2.6f}, {0.0f, 0.0f, 1.0f}, {0.0f, 1.0f, 0.0f}, 1.0f, 10000.0f};
    s->microfacet = microfacet;
    *out = s;
    return PICO_OK;
}

int pico_scene_cornell_glass(uint32_t seed, uint32_t microfacet, pico_scene** out) {
    if (!out || microfacet > PICO_MICROFACET_GGX) return PICO_ERR_ARG;
    auto* s = new pico_scene();
    Rng rng(seed, 1);
    cornell_into(s, rng, false);
    // instances 5 and 6 are the two boxes: a Fresnel glass block and a clear
    // Transparent one (water-like index)
    pico_material glass = metal_rough({0.95f, 0.95f, 0.95f}, 0.0f, 0.1f, {0, 0, 0});
    glass.diffuse[0] = glass.diffuse[1] = glass.diffuse[2] = 0.95f;  // transmitted energy
    glass.index_of_refraction = 1.5f;
    pico_material water = glass;
    water.roughness = 0.2f;
    water.index_of_refraction = 1.33f;
    s->imat[5] = s->add_material(glass);
    s->ibsdf[5] = PICO_BSDF_FRESNEL;
    s->imat[6] = s->add_material(water);
    s->ibsdf[6] = PICO_BSDF_TRANSPARENT;
    s->microfacet = microfacet;
    *out = s;
    return PICO_OK;
}

int pico_scene_instanced(uint32_t seed, uint32_t n_types, uint32_t n_inst, uint32_t tmin, uint32_t tmax, float extent,
                         uint32_t microfacet, pico_scene** out) {
    if (!out || n_types == 0 || n_inst == 0 || tmin < 2 || tmax < tmin || !(extent > 0) || microfacet > PICO_MICROFACET_GGX)
        return PICO_ERR_ARG;
    auto* s = new pico_scene();
    Rng rng(seed, 2);
    instanced_into(s, rng, n_types, n_inst, tmin, tmax, extent, 3, 0.1f);
    s->microfacet = microfacet;
    *out = s;
    return PICO_OK;
}

int pico_scene_training(uint32_t seed, uint32_t microfacet, pico_scene** out) {
    if (!out || microfacet > PICO_MICROFACET_GGX) return PICO_ERR_ARG;
    auto* s = new pico_scene();
    Rng rng(seed, 3);
    if (seed % 3 == 0) {
        cornell_into(s, rng, true);
    } else if (seed % 3 == 1) {
        instanced_into(s, rng, 4, 24 + rng.next() % 40, 200, 2000, 12.0f, 1 + rng.next() % 3, rng.uniform(0.0f, 0.3f));
        // camera in front of the cluster, looking roughly at its centre
        const float a = rng.uniform(-0.5f, 0.5f);
        s->cam.position[0] = 6.0f + 10.0f * std::sin(a);
        s->cam.position[2] = 6.0f - 10.0f * std::cos(a);
        F3 d = normalize(F3{6.0f - s->cam.position[0], 0.0f, 6.0f - s->cam.position[2]});
        s->cam.direction[0] = d.x; s->cam.direction[1] = d.y; s->cam.direction[2] = d.z;
    } else {
        // a field of the size of the benchmark scene (100^3 box, 2k-20k
        // triangle meshes, three quad lights, dim sky) seen from a random
        // point 60-110 units from its centre, looking near the centre
        instanced_into(s, rng, 8, 256 + rng.next() % 769, 2000, 20000, 100.0f, 3, rng.uniform(0.05f, 0.2f));
        F3 dir = normalize(F3{rng.gauss(), 0.5f * rng.gauss(), rng.gauss()});
        const float r = rng.uniform(60.0f, 110.0f);
        const F3 c = {50.0f, 50.0f, 50.0f};
        const F3 p = c + dir * r;
        const F3 t = c + F3{rng.uniform(-15.0f, 15.0f), rng.uniform(-15.0f, 15.0f), rng.uniform(-15.0f, 15.0f)};
        const F3 d = normalize(t - p);
        s->cam.position[0] = p.x; s->cam.position[1] = p.y; s->cam.position[2] = p.z;
        s->cam.direction[0] = d.x; s->cam.direction[1] = d.y; s->cam.direction[2] = d.z;
    }
    s->microfacet = microfacet;
    *out = s;
    return PICO_OK;
}

int pico_scene_desc_get(const pico_scene* s, pico_scene_desc* d) {
    if (!s || !d) return PICO_ERR_ARG;
    std::memset(d, 0, sizeof(*d));
    d->n_meshes = (uint32_t)s->vofs.size() - 1;
    d->mesh_vertex_offset = s->vofs.data();
    d->mesh_triangle_offset = s->tofs.data();
    d->positions = s->pos.data();
    d->normals = s->nrm.data();
    d->indices = s->idx.data();
    d->n_instances = (uint32_t)s->imesh.size();
    d->transforms = s->xf.data();
    d->instance_mesh = s->imesh.data();
    d->instance_material = s->imat.data();
    d->instance_bsdf = s->ibsdf.data();
    d->n_materials = (uint32_t)s->mats.size();
    d->materials = s->mats.data();
    d->n_lights = (uint32_t)s->lights.size();
    d->light_instances = s->lights.data();
    std::memcpy(d->sky_rgb, s->sky, sizeof(s->sky));
    d->microfacet = s->microfacet;
    d->mesh_shape = s->shape.data();
    d->use_sun = s->use_sun;
    std::memcpy(d->sun_direction, s->sun_dir, sizeof(s->sun_dir));
    std::memcpy(d->sun_colour, s->sun_col, sizeof(s->sun_col));
    d->sky_image = s->sky_img.width ? &s->sky_img : nullptr;
    return PICO_OK;
}

int pico_scene_set_sun(pico_scene* s, const float* direction, const float* colour) {
    if (!s || (!direction) != (!colour)) return PICO_ERR_ARG;
    s->use_sun = direction ? 1u : 0u;
    if (direction) {
        std::memcpy(s->sun_dir, direction, sizeof(s->sun_dir));
        std::memcpy(s->sun_col, colour, sizeof(s->sun_col));
    }
    return PICO_OK;
}

int pico_scene_set_sky_image(pico_scene* s, const pico_sky_image* img) {
    if (!s) return PICO_ERR_ARG;
    if (!img) {
        s->sky_img = pico_sky_image{};
        s->sky_texels.clear();
        return PICO_OK;
    }
    if ((img->faces != 1 && img->faces != 6) || img->format > PICO_TEXEL_RGBA32F || !img->width || !img->height ||
        !img->texels || (uint64_t)img->width * img->height * img->faces > (1ull << 26))
        return PICO_ERR_ARG;
    const size_t bytes = (size_t)img->width * img->height * img->faces * (img->format == PICO_TEXEL_RGBA32F ? 16 : 4);
    const uint8_t* p = static_cast<const uint8_t*>(img->texels);
    s->sky_texels.assign(p, p + bytes);
    s->sky_img = *img;
    s->sky_img.texels = s->sky_texels.data();
    return PICO_OK;
}

int pico_scene_default_camera(const pico_scene* s, float aspect, pico_camera* out) {
    if (!s || !out) return PICO_ERR_ARG;
    *out = s->cam;
    out->aspect = aspect;
    return PICO_OK;
}

void pico_scene_destroy(pico_scene* s) { delete s; }

}  // extern "C"
