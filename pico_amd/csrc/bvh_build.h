// Host-side builder of the device scene (BLAS per mesh, TLAS over instances).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pico_hip.h"
#include "device_types.h"

namespace pico {

struct HostScene {
    std::vector<DNode> nodes;          // binary build: BLAS nodes of every mesh, then the TLAS
    std::vector<DNode4> nodes4;        // the 4-wide nodes the device traverses
    std::vector<DTri> tris;            // triangle records in BLAS leaf order
    std::vector<DInstance> instances;  // original instance order
    std::vector<DXform> xforms;
    std::vector<DMaterial> materials;
    std::vector<DLight> lights;
    std::vector<float> alias_weight;
    std::vector<uint32_t> alias_index;
    std::vector<float> positions;      // 3 x f32 per vertex (global)
    std::vector<float> normals;        // 3 x f32 per vertex (global)
    std::vector<uint32_t> indices;     // 3 x u32 per triangle, global vertex ids
    std::vector<float> tshade;         // 4 x kShadeRec f32 per triangle: its vertices' positions and normals
    std::vector<uint32_t> mesh_tri_offset;
    std::vector<float> mesh_box;       // 6 x f32 per mesh: object-space vertex bounds lo, hi
    std::vector<uint32_t> mesh_vrange; // 2 x u32 per mesh: first and end vertex (empty for the built-in shapes)
    std::vector<uint32_t> instance_mesh;
    uint32_t tlas_root = 0;            // DNode4 index
    uint32_t tlas_base = 0;            // first TLAS DNode4 (every BLAS node comes before it)
    uint32_t blas_bound = 0;           // most stack entries any BLAS traversal can hold
    uint32_t blas_nodes2 = 0;          // binary BLAS nodes (every mesh)
    uint32_t tlas_depth = 0;
    uint32_t blas_depth = 0;           // deepest BLAS (binary levels)
    uint32_t stack_bound = 0;          // most stack entries any 4-wide traversal can hold
    uint32_t n_instances = 0;
    float sky[3] = {0, 0, 0};
    uint32_t microfacet = 0;
    uint32_t use_sun = 0;
    float sun_to[3] = {0, 0, 0};       // -sun_direction
    float sun_col[3] = {0, 0, 0};
    pico_sky_image sky_image{};        // texels still the description's (width 0: none)
};

// Bytes of a sky image's texels.
size_t sky_bytes(const pico_sky_image& k);

// Binned SAH (bins per axis, all three axes) BVH2 over boxes; returns the
// child-pair root node index.  Leaves hold at most max_leaf primitives.
// `order` receives the primitive order of the leaves.
uint32_t build_bvh(const std::vector<float>& box_lo, const std::vector<float>& box_hi, uint32_t bins,
                   uint32_t max_leaf, uint32_t prim_base, std::vector<DNode>& nodes, std::vector<uint32_t>& order,
                   uint32_t& depth);

// Collapses the binary tree under child-pair node `root2` into 4-wide nodes
// appended to `out` (children chosen by opening the largest-area internal
// child until four); returns the DNode4 index of the root and, in `bound`,
// the most traversal-stack entries a ray can hold in this subtree.
uint32_t collapse_bvh4(const std::vector<DNode>& n2, uint32_t root2, std::vector<DNode4>& out, uint32_t& bound);

// Validates the description and builds everything; returns an error text or "".
std::string build_scene(const pico_scene_desc& d, uint32_t bins, HostScene& out);
// per mesh (first, end) vertex of its triangles; an empty range for a built-in shape
std::vector<uint32_t> mesh_vertex_ranges(const pico_scene_desc& d);
// the shading records of triangles [t0, t1) (global indices `idx`, global
// positions and normals) into out[4 * kShadeRec * t]
inline void shade_records(const uint32_t* idx, const float* pos, const float* nrm, size_t t0, size_t t1, float* out) {
    for (size_t t = t0; t < t1; ++t) {
        float* r = out + 4 * kShadeRec * t;
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) {
                r[3 * c + k] = pos[3 * (size_t)idx[3 * t + c] + k];
                r[9 + 3 * c + k] = nrm[3 * (size_t)idx[3 * t + c] + k];
            }
        r[18] = 0.0f;
        r[19] = 0.0f;
    }
}
// the largest range of such (first, end) pairs
inline uint32_t max_range(const std::vector<uint32_t>& r) {
    uint32_t m = 0;
    for (size_t k = 0; k + 1 < r.size(); k += 2) m = r[k + 1] - r[k] > m ? r[k + 1] - r[k] : m;
    return m;
}

// The parts of build_scene that do not touch the triangles: the structural
// checks of the description, and the materials, instances (inverses, no BLAS
// roots), transforms and light tables.  The device builder uses these.
std::string check_scene(const pico_scene_desc& d);
uint32_t mesh_shape(const pico_scene_desc& d, uint32_t mesh);  // PICO_SHAPE_*
std::string scene_tables(const pico_scene_desc& d, HostScene& out);

}  // namespace pico
