// Device-side scene layout shared by the host builder and the HIP kernels.
//
// BVH nodes are binary "child-pair" nodes: one 64-byte record holds the boxes
// of both children and their references, so one node fetch (4 x 16 B,
// coalesced per lane) resolves a whole traversal step.  Child references:
//   bit 31 clear  -> internal node index
//   bit 31 set    -> leaf: bits 30..27 primitive count (1..15), bits 26..0 first
//                    primitive (BLAS: triangle record, TLAS: instance)
//   0xffffffff    -> empty child (its box is empty and never hit)
// The device traverses 4-wide nodes (DNode4) made by collapsing that binary
// tree: one 64-byte record per node with the four child boxes quantised to
// 8 bits per plane against the node's own origin and power-of-two scales
// (rounded outwards, so a child box only grows), and four child references.
// Triangles are stored in leaf order as (v0, v1 - v0, v2 - v0) with the
// mesh-local primitive id in v0.w, so Moller-Trumbore needs 48 bytes per test.
#pragma once
#include <cstdint>

namespace pico {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmptyRef = 0xffffffffu;
constexpr uint32_t kMaxLeafPrims = 15u;
constexpr uint32_t kPrimMask = (1u << 27) - 1u;

inline constexpr uint32_t make_leaf(uint32_t first, uint32_t count) { return kLeafBit | (count << 27) | first; }

// Shading record of a triangle: positions p0, p1, p2 then normals n0, n1, n2
// of its vertices (18 floats), padded to kShadeRec float4.
constexpr uint32_t kShadeRec = 5;

// BLAS root of an instance of a built-in implicit shape: kImplicitRoot | PICO_SHAPE_*
// (no nodes; the traversal intersects the shape directly).  Lights on such a
// shape have n_tris == 0 and the shape in tri_offset.
constexpr uint32_t kImplicitRoot = 0xffffff00u;

struct alignas(16) DNode {
    float c0x[2], c0y[2];   // child 0: lo.x hi.x lo.y hi.y
    float c1x[2], c1y[2];   // child 1: lo.x hi.x lo.y hi.y
    float c0z[2], c1z[2];   // lo.z hi.z of child 0, then child 1
    uint32_t ref[2];        // child references
    uint32_t pad[2];
};
static_assert(sizeof(DNode) == 64, "node is one 64-byte record");

struct alignas(16) DNode4 {
    float origin[3];        // union lower corner of the children
    float scale_x;          // 2^e of the x planes
    uint8_t qx[8];          // lo[4], hi[4]: child box = origin + q * 2^e
    uint8_t qy[8];
    uint8_t qz[8];
    float scale_y, scale_z; // 2^e of the y and z planes
    uint32_t ref[4];        // child references (kEmptyRef: no child)
};
static_assert(sizeof(DNode4) == 64, "wide node is one 64-byte record");

struct alignas(16) DTri {
    float v0[3]; uint32_t prim;
    float e1[3]; float pad1;
    float e2[3]; float pad2;
};
static_assert(sizeof(DTri) == 48, "triangle record is 48 bytes");

// Traversal part of an instance: rows of the inverse affine transform and the
// BLAS root reference.  64 bytes.
struct alignas(16) DInstance {
    float inv[3][4];        // object = inv * (x, y, z, 1)
    uint32_t blas_root;     // node index of the mesh's root child-pair node
    uint32_t tri_offset;    // first triangle of the mesh in the global index array
    uint32_t material;
    uint32_t bsdf;
};
static_assert(sizeof(DInstance) == 64, "instance record is 64 bytes");

// Shading part of an instance: the column-major 4x4 transform (glm layout).
struct alignas(16) DXform { float m[16]; };

struct alignas(16) DMaterial {
    float diffuse[3];  float reflectance;   // |spec| / (|spec| + |diff|)
    float specular[3]; float roughness;
    float emissive[3]; float ior;           // index of refraction (transmissive BSDFs)
};

struct alignas(16) DLight {
    uint32_t instance;
    uint32_t n_tris;
    uint32_t alias_offset;  // into the alias-table arrays
    uint32_t tri_offset;    // first triangle of the light mesh (global index array)
};

}  // namespace pico
