// Host-side scene builder: binned-SAH BVH2 per mesh (BLAS) and over instances
// (TLAS), emitted as 64-byte child-pair nodes, plus the light-sampling tables.
//
// Reference counterparts: BVHFactory/SAH_parition_Scheme (Source/Core/BVH.cpp:65-197)
// build the same kind of tree with 12 buckets on one axis; here 16 bins are
// evaluated on all three axes and leaves hold up to 4 triangles.  Closest-hit
// results do not depend on the topology except for exact ties.  Instance
// bounds follow AABB * mat4 (Source/Core/AABB.cpp:253-270); light tables
// follow LowerLevelMeshBVH::generate_sampling_data (Source/Core/LowerLevelMeshBVH.cpp:92-117)
// and AliasTable::build_table (Source/Util/AliasTable.cpp:12-87).
#include "bvh_build.h"
#include "scene_math.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <set>

namespace pico {

namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float* l, const float* h) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], l[k]); hi[k] = std::max(hi[k], h[k]); }
    }
    void grow_pt(const float* p) { grow(p, p); }
    float area() const {
        float d[3];
        for (int k = 0; k < 3; ++k) d[k] = std::max(0.0f, hi[k] - lo[k]);
        return 2.0f * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2]);
    }
    bool empty() const { return !(lo[0] <= hi[0]); }
};

struct Builder {
    const std::vector<float>& lo;
    const std::vector<float>& hi;
    std::vector<uint32_t>& order;
    std::vector<float> cen;
    uint32_t bins, max_leaf, prim_base;
    std::vector<DNode>& nodes;
    uint32_t max_depth = 0;

    Box range_box(uint32_t b, uint32_t e) const {
        Box bx;
        for (uint32_t i = b; i < e; ++i) bx.grow(&lo[3 * order[i]], &hi[3 * order[i]]);
        return bx;
    }

    // Splits [b, e) and returns the split position, or e for "make a leaf".
    uint32_t split(uint32_t b, uint32_t e, const Box& bounds) {
        const uint32_t n = e - b;
        if (n == 1) return e;
        Box cb;
        for (uint32_t i = b; i < e; ++i) cb.grow_pt(&cen[3 * order[i]]);
        float best_cost = INFINITY;
        int best_axis = -1;
        uint32_t best_bin = 0;
        std::vector<Box> bb(bins);
        std::vector<uint32_t> cnt(bins);
        std::vector<float> rarea(bins);
        std::vector<uint32_t> rcnt(bins);
        for (int axis = 0; axis < 3; ++axis) {
            const float ext = cb.hi[axis] - cb.lo[axis];
            if (!(ext > 0)) continue;
            const float sc = (float)bins / ext;
            std::fill(bb.begin(), bb.end(), Box());
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (uint32_t i = b; i < e; ++i) {
                const uint32_t p = order[i];
                uint32_t k = (uint32_t)std::min<float>((float)(bins - 1), (cen[3 * p + axis] - cb.lo[axis]) * sc);
                cnt[k]++;
                bb[k].grow(&lo[3 * p], &hi[3 * p]);
            }
            Box acc;
            uint32_t c = 0;
            for (int k = (int)bins - 1; k > 0; --k) {
                acc.grow(bb[k].lo, bb[k].hi);
                c += cnt[k];
                rarea[k] = acc.empty() ? 0.0f : acc.area();
                rcnt[k] = c;
            }
            acc = Box();
            c = 0;
            for (uint32_t k = 0; k + 1 < bins; ++k) {
                acc.grow(bb[k].lo, bb[k].hi);
                c += cnt[k];
                if (c == 0 || rcnt[k + 1] == 0) continue;
                const float cost = (float)c * acc.area() + (float)rcnt[k + 1] * rarea[k + 1];
                if (cost < best_cost) { best_cost = cost; best_axis = axis; best_bin = k; }
            }
        }
        const float parent = bounds.area();
        const float leaf_cost = (float)n * parent;
        const float split_cost = kSahTraversal * parent + best_cost;
        if (best_axis < 0) {
            // all centroids coincide: leaf if it fits, else split the index range
            return n <= max_leaf ? e : b + n / 2;
        }
        if (n <= max_leaf && leaf_cost <= split_cost) return e;
        const int axis = best_axis;
        const float sc = (float)bins / (cb.hi[axis] - cb.lo[axis]);
        auto mid = std::partition(order.begin() + b, order.begin() + e, [&](uint32_t p) {
            uint32_t k = (uint32_t)std::min<float>((float)(bins - 1), (cen[3 * p + axis] - cb.lo[axis]) * sc);
            return k <= best_bin;
        });
        uint32_t m = (uint32_t)(mid - order.begin());
        if (m == b || m == e) m = b + n / 2;
        return m;
    }

    // Returns a child reference for the range [b, e); emits nodes in DFS order.
    uint32_t build(uint32_t b, uint32_t e, uint32_t depth) {
        const Box bounds = range_box(b, e);
        const uint32_t s = split(b, e, bounds);
        if (s == e) return make_leaf(prim_base + b, e - b);
        return emit_internal(b, s, e, depth);
    }

    uint32_t emit_internal(uint32_t b, uint32_t s, uint32_t e, uint32_t depth) {
        max_depth = std::max(max_depth, depth + 1);
        const uint32_t id = (uint32_t)nodes.size();
        nodes.emplace_back();
        const Box l = range_box(b, s), r = range_box(s, e);
        const uint32_t rl = build(b, s, depth + 1);
        const uint32_t rr = build(s, e, depth + 1);
        DNode& nd = nodes[id];
        set_child(nd, 0, l, rl);
        set_child(nd, 1, r, rr);
        return id;
    }

    static void set_child(DNode& nd, int c, const Box& bx, uint32_t ref) {
        float* x = c == 0 ? nd.c0x : nd.c1x;
        float* y = c == 0 ? nd.c0y : nd.c1y;
        float* z = c == 0 ? nd.c0z : nd.c1z;
        x[0] = bx.lo[0]; x[1] = bx.hi[0];
        y[0] = bx.lo[1]; y[1] = bx.hi[1];
        z[0] = bx.lo[2]; z[1] = bx.hi[2];
        nd.ref[c] = ref;
    }
};

inline float fdot(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Triangle "area" of the reference's light sampler (the sine takes the angle
// scaled by 180/pi), Source/Core/LowerLevelMeshBVH.cpp:98-112, in its float/double mix.
float light_tri_weight(const float* a, const float* b, const float* c) {
    float ac[3], bc[3];
    for (int k = 0; k < 3; ++k) { ac[k] = a[k] - c[k]; bc[k] = b[k] - c[k]; }
    const float bc_len = std::sqrt(fdot(bc, bc));
    const float base = std::sqrt(fdot(ac, ac));
    float nac[3], nbc[3];
    const float ia = 1.0f / std::sqrt(fdot(ac, ac)), ib = 1.0f / std::sqrt(fdot(bc, bc));
    for (int k = 0; k < 3; ++k) { nac[k] = ac[k] * ia; nbc[k] = bc[k] * ib; }
    const float theta = (float)std::sin((double)std::acos(fdot(nac, nbc)) * (180.0f / M_PI));
    const float height = theta * bc_len;
    return (base * height) / 2.0f;
}

// Vose alias table as AliasTable::build_table builds it (leftovers alias 0, weight 1).
void alias_table(const std::vector<float>& w, std::vector<float>& bw, std::vector<uint32_t>& ba) {
    const uint32_t n = (uint32_t)w.size();
    bw.assign(n, 0.0f);
    ba.assign(n, 0u);
    const float total = (float)std::accumulate(w.begin(), w.end(), 0.0);
    for (uint32_t i = 0; i < n; ++i) bw[i] = w[i] / total;
    struct O { float p; uint32_t i; };
    std::vector<O> over, under;
    for (uint32_t i = 0; i < n; ++i) {
        const float ph = bw[i] * (float)n;
        (ph < 1.0f ? under : over).push_back({ph, i});
    }
    while (!over.empty() && !under.empty()) {
        const O ov = over.back(); over.pop_back();
        const O un = under.back(); under.pop_back();
        bw[un.i] = un.p;
        ba[un.i] = ov.i;
        const float excess = un.p + ov.p - 1.0f;
        (excess < 1.0f ? under : over).push_back({excess, ov.i});
    }
    for (auto& u : under) { ba[u.i] = 0; bw[u.i] = 1.0f; }
    for (auto& o : over) { ba[o.i] = 0; bw[o.i] = 1.0f; }
}

}  // namespace

namespace {
struct WideBuilder {
    const std::vector<DNode>& n2;
    std::vector<DNode4>& out;

    static void child_box(const DNode& n, int c, float lo[3], float hi[3]) {
        const float* x = c == 0 ? n.c0x : n.c1x;
        const float* y = c == 0 ? n.c0y : n.c1y;
        const float* z = c == 0 ? n.c0z : n.c1z;
        lo[0] = x[0]; hi[0] = x[1]; lo[1] = y[0]; hi[1] = y[1]; lo[2] = z[0]; hi[2] = z[1];
    }
    static float area(const float lo[3], const float hi[3]) {
        const float dx = std::max(0.0f, hi[0] - lo[0]), dy = std::max(0.0f, hi[1] - lo[1]), dz = std::max(0.0f, hi[2] - lo[2]);
        return dx * dy + dx * dz + dy * dz;
    }

    uint32_t collapse(uint32_t id2, uint32_t& bound) {
        struct Item { uint32_t ref; float lo[3], hi[3]; };
        Item ch[4];
        int nc = 0;
        auto add_children = [&](uint32_t id) {
            for (int c = 0; c < 2; ++c) {
                if (n2[id].ref[c] == kEmptyRef) continue;
                Item it;
                it.ref = n2[id].ref[c];
                child_box(n2[id], c, it.lo, it.hi);
                ch[nc++] = it;
            }
        };
        add_children(id2);
        while (nc < 4) {
            int best = -1;
            float ba = -1.0f;
            for (int i = 0; i < nc; ++i)
                if (!(ch[i].ref & kLeafBit) && area(ch[i].lo, ch[i].hi) > ba) { ba = area(ch[i].lo, ch[i].hi); best = i; }
            if (best < 0) break;
            const uint32_t open = ch[best].ref;
            ch[best] = ch[--nc];
            add_children(open);
        }
        const uint32_t me = (uint32_t)out.size();
        out.emplace_back();
        DNode4 nd;
        float lo[4][3], hi[4][3];
        for (int i = 0; i < nc; ++i)
            for (int k = 0; k < 3; ++k) { lo[i][k] = ch[i].lo[k]; hi[i][k] = ch[i].hi[k]; }
        quantise_children(lo, hi, nc, nd);
        uint32_t deepest = 0;
        for (int i = 0; i < nc; ++i) {
            if (ch[i].ref & kLeafBit) {
                nd.ref[i] = ch[i].ref;
            } else {
                uint32_t b = 0;
                nd.ref[i] = collapse(ch[i].ref, b);
                deepest = std::max(deepest, b);
            }
        }
        // entering any child leaves at most nc - 1 siblings on the stack
        bound = (nc > 0 ? (uint32_t)nc - 1 : 0) + deepest;
        out[me] = nd;
        return me;
    }
};
}  // namespace

uint32_t collapse_bvh4(const std::vector<DNode>& n2, uint32_t root2, std::vector<DNode4>& out, uint32_t& bound) {
    WideBuilder wb{n2, out};
    return wb.collapse(root2, bound);
}

uint32_t build_bvh(const std::vector<float>& box_lo, const std::vector<float>& box_hi, uint32_t bins, uint32_t max_leaf,
                   uint32_t prim_base, std::vector<DNode>& nodes, std::vector<uint32_t>& order, uint32_t& depth) {
    const uint32_t n = (uint32_t)(box_lo.size() / 3);
    order.resize(n);
    std::iota(order.begin(), order.end(), 0u);
    Builder b{box_lo, box_hi, order, {}, bins, max_leaf, prim_base, nodes};
    b.cen.resize(3 * n);
    for (uint32_t i = 0; i < 3 * n; ++i) b.cen[i] = 0.5f * (box_lo[i] + box_hi[i]);
    const Box all = b.range_box(0, n);
    uint32_t root;
    const uint32_t s = n ? b.split(0, n, all) : 0;
    if (n == 0 || s == n) {
        // a single leaf: wrap it in a node whose second child is empty
        root = (uint32_t)nodes.size();
        nodes.emplace_back();
        DNode& nd = nodes[root];
        Builder::set_child(nd, 0, all, n ? make_leaf(prim_base, n) : kEmptyRef);
        Builder::set_child(nd, 1, Box(), kEmptyRef);
        b.max_depth = 1;
    } else {
        root = b.emit_internal(0, s, n, 0);
    }
    depth = b.max_depth;
    return root;
}

std::string check_scene(const pico_scene_desc& d) {
    if (d.n_meshes == 0 || !d.mesh_vertex_offset || !d.mesh_triangle_offset || !d.positions || !d.normals || !d.indices)
        return "scene has no meshes";
    if (d.n_instances == 0 || !d.transforms || !d.instance_mesh || !d.instance_material || !d.instance_bsdf)
        return "scene has no instances";
    if (d.n_materials == 0 || !d.materials) return "scene has no materials";
    if (d.microfacet > PICO_MICROFACET_GGX) return "unknown microfacet distribution";
    if (d.n_lights && !d.light_instances) return "light list missing";
    if (d.mesh_vertex_offset[0] != 0 || d.mesh_triangle_offset[0] != 0) return "mesh offsets must start at 0";
    for (uint32_t m = 0; m < d.n_meshes; ++m)
        if (d.mesh_vertex_offset[m + 1] < d.mesh_vertex_offset[m] || d.mesh_triangle_offset[m + 1] < d.mesh_triangle_offset[m])
            return "mesh offsets must be non-decreasing";
    if (d.mesh_triangle_offset[d.n_meshes] > kPrimMask || d.n_instances > kPrimMask) return "too many triangles or instances";
    if (d.mesh_shape)
        for (uint32_t m = 0; m < d.n_meshes; ++m) {
            if (d.mesh_shape[m] > PICO_SHAPE_CUBE) return "unknown mesh shape";
            if (d.mesh_shape[m] != PICO_SHAPE_TRIANGLES &&
                (d.mesh_vertex_offset[m + 1] != d.mesh_vertex_offset[m] || d.mesh_triangle_offset[m + 1] != d.mesh_triangle_offset[m]))
                return "a built-in shape has no vertices or triangles";
        }
    if (d.use_sun)
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(d.sun_direction[c]) || !std::isfinite(d.sun_colour[c])) return "sun direction and colour must be finite";
    if (d.sky_image) {
        const pico_sky_image& k = *d.sky_image;
        if (k.faces != 1 && k.faces != 6) return "sky image needs 1 (equirectangular) or 6 (cube) faces";
        if (k.format > PICO_TEXEL_RGBA32F) return "unknown sky texel format";
        if (k.width == 0 || k.height == 0 || !k.texels) return "sky image is empty";
        if ((uint64_t)k.width * k.height * k.faces > (1ull << 26)) return "sky image larger than 2^26 texels";
    }
    return "";
}

size_t sky_bytes(const pico_sky_image& k) {
    return (size_t)k.width * k.height * k.faces * (k.format == PICO_TEXEL_RGBA32F ? 16 : 4);
}

uint32_t mesh_shape(const pico_scene_desc& d, uint32_t m) { return d.mesh_shape ? d.mesh_shape[m] : PICO_SHAPE_TRIANGLES; }

std::string scene_tables(const pico_scene_desc& d, HostScene& out) {
    // materials
    for (uint32_t i = 0; i < d.n_materials; ++i) {
        const pico_material& pm = d.materials[i];
        DMaterial m{};
        for (int c = 0; c < 3; ++c) { m.diffuse[c] = pm.diffuse[c]; m.specular[c] = pm.specular[c]; m.emissive[c] = pm.emissive[c]; }
        const float s = std::sqrt(fdot(pm.specular, pm.specular));
        m.reflectance = s / (s + std::sqrt(fdot(pm.diffuse, pm.diffuse)));
        m.roughness = pm.roughness;
        m.ior = pm.index_of_refraction;
        out.materials.push_back(m);
    }
    // instances (the BLAS roots are filled in by the builder)
    out.n_instances = d.n_instances;
    for (uint32_t i = 0; i < d.n_instances; ++i) {
        const uint32_t m = d.instance_mesh[i];
        if (m >= d.n_meshes) return "instance mesh out of range";
        if (d.instance_material[i] >= d.n_materials) return "instance material out of range";
        if (d.instance_bsdf[i] > PICO_BSDF_FRESNEL) return "unknown BSDF kind";
        if ((d.instance_bsdf[i] == PICO_BSDF_TRANSPARENT || d.instance_bsdf[i] == PICO_BSDF_FRESNEL) &&
            !(d.materials[d.instance_material[i]].index_of_refraction > 0.0f))
            return "a transmissive BSDF needs a material with an index of refraction";
        const float* M = &d.transforms[16 * (size_t)i];
        if (M[3] != 0.0f || M[7] != 0.0f || M[11] != 0.0f || M[15] != 1.0f) return "instance transform must be affine";
        DInstance di{};
        if (!affine_inverse_rows(M, di.inv)) return "instance transform is singular";
        di.blas_root = 0;
        di.tri_offset = d.mesh_triangle_offset[m];
        di.material = d.instance_material[i];
        di.bsdf = d.instance_bsdf[i];
        out.instances.push_back(di);
        out.instance_mesh.push_back(m);
        DXform xf;
        std::memcpy(xf.m, M, sizeof(xf.m));
        out.xforms.push_back(xf);
    }
    // lights: alias tables over the light meshes' triangles (mesh-local indices)
    std::set<uint32_t> light_meshes;
    for (uint32_t l = 0; l < d.n_lights; ++l) {
        const uint32_t ii = d.light_instances[l];
        if (ii >= d.n_instances) return "light instance out of range";
        if (d.instance_bsdf[ii] != PICO_BSDF_LIGHT) return "light instance must use the Light BSDF";
        const uint32_t m = d.instance_mesh[ii];
        if (!light_meshes.insert(m).second) return "lights must not share a mesh";
        const uint32_t t0 = d.mesh_triangle_offset[m], t1 = d.mesh_triangle_offset[m + 1];
        const uint32_t v0 = d.mesh_vertex_offset[m], nv = d.mesh_vertex_offset[m + 1] - v0;
        if (mesh_shape(d, m) != PICO_SHAPE_TRIANGLES) {
            // the shape's own sample_geometry: no table
            out.lights.push_back(DLight{ii, 0u, 0u, mesh_shape(d, m)});
            continue;
        }
        if (t1 == t0) return "light mesh has no triangles";
        std::vector<float> w;
        for (uint32_t t = t0; t < t1; ++t) {
            const uint32_t* li = &d.indices[3 * (size_t)t];
            if (li[0] >= nv || li[1] >= nv || li[2] >= nv) return "triangle index out of range";
            w.push_back(light_tri_weight(&d.positions[3 * (size_t)(v0 + li[0])], &d.positions[3 * (size_t)(v0 + li[1])],
                                         &d.positions[3 * (size_t)(v0 + li[2])]));
        }
        std::vector<float> bw;
        std::vector<uint32_t> ba;
        alias_table(w, bw, ba);
        DLight dl{ii, t1 - t0, (uint32_t)out.alias_weight.size(), t0};
        out.alias_weight.insert(out.alias_weight.end(), bw.begin(), bw.end());
        out.alias_index.insert(out.alias_index.end(), ba.begin(), ba.end());
        out.lights.push_back(dl);
    }
    std::memcpy(out.sky, d.sky_rgb, sizeof(out.sky));
    out.microfacet = d.microfacet;
    out.use_sun = d.use_sun ? 1u : 0u;
    for (int c = 0; c < 3; ++c) {
        out.sun_to[c] = -d.sun_direction[c];
        out.sun_col[c] = d.sun_colour[c];
    }
    if (d.sky_image) out.sky_image = *d.sky_image;
    return "";
}

std::vector<uint32_t> mesh_vertex_ranges(const pico_scene_desc& d) {
    std::vector<uint32_t> r(2 * (size_t)d.n_meshes);
    for (uint32_t m = 0; m < d.n_meshes; ++m) {
        const bool tri = mesh_shape(d, m) == PICO_SHAPE_TRIANGLES;
        r[2 * m] = d.mesh_vertex_offset[m];
        r[2 * m + 1] = tri ? d.mesh_vertex_offset[m + 1] : d.mesh_vertex_offset[m];
    }
    return r;
}

std::string build_scene(const pico_scene_desc& d, uint32_t bins, HostScene& out) {
    out = HostScene();
    std::string e = check_scene(d);
    if (!e.empty()) return e;
    const uint32_t nv = d.mesh_vertex_offset[d.n_meshes], nt = d.mesh_triangle_offset[d.n_meshes];
    out.positions.assign(d.positions, d.positions + 3 * (size_t)nv);
    out.normals.assign(d.normals, d.normals + 3 * (size_t)nv);
    out.indices.resize(3 * (size_t)nt);
    out.mesh_tri_offset.assign(d.mesh_triangle_offset, d.mesh_triangle_offset + d.n_meshes + 1);
    std::vector<uint32_t> mesh_root(d.n_meshes);
    std::vector<Box> mesh_box(d.n_meshes);
    out.mesh_vrange = mesh_vertex_ranges(d);
    for (uint32_t m = 0; m < d.n_meshes; ++m) {
        const uint32_t v0 = d.mesh_vertex_offset[m], v1 = d.mesh_vertex_offset[m + 1];
        const uint32_t t0 = d.mesh_triangle_offset[m], t1 = d.mesh_triangle_offset[m + 1];
        if (mesh_shape(d, m) != PICO_SHAPE_TRIANGLES) {
            // a built-in shape: bounds [-0.5, 0.5]^3 (its get_bounds), no BLAS
            for (int k = 0; k < 3; ++k) { mesh_box[m].lo[k] = -0.5f; mesh_box[m].hi[k] = 0.5f; }
            out.mesh_box.insert(out.mesh_box.end(), mesh_box[m].lo, mesh_box[m].lo + 3);
            out.mesh_box.insert(out.mesh_box.end(), mesh_box[m].hi, mesh_box[m].hi + 3);
            mesh_root[m] = kImplicitRoot | mesh_shape(d, m);
            continue;
        }
        std::vector<float> lo, hi;
        for (uint32_t t = t0; t < t1; ++t) {
            Box b;
            for (int k = 0; k < 3; ++k) {
                const uint32_t li = d.indices[3 * (size_t)t + k];
                if (li >= v1 - v0) return "triangle index out of range";
                out.indices[3 * (size_t)t + k] = v0 + li;
                b.grow_pt(&d.positions[3 * (size_t)(v0 + li)]);
            }
            lo.insert(lo.end(), b.lo, b.lo + 3);
            hi.insert(hi.end(), b.hi, b.hi + 3);
        }
        for (uint32_t v = v0; v < v1; ++v) mesh_box[m].grow_pt(&d.positions[3 * (size_t)v]);
        out.mesh_box.insert(out.mesh_box.end(), mesh_box[m].lo, mesh_box[m].lo + 3);
        out.mesh_box.insert(out.mesh_box.end(), mesh_box[m].hi, mesh_box[m].hi + 3);
        std::vector<uint32_t> order;
        uint32_t depth = 0;
        const uint32_t base = (uint32_t)out.tris.size();
        mesh_root[m] = build_bvh(lo, hi, bins, 4, base, out.nodes, order, depth);
        out.blas_depth = std::max(out.blas_depth, depth);
        for (uint32_t k = 0; k < order.size(); ++k) {
            const uint32_t t = t0 + order[k];
            const float* p0 = &out.positions[3 * (size_t)out.indices[3 * (size_t)t]];
            const float* p1 = &out.positions[3 * (size_t)out.indices[3 * (size_t)t + 1]];
            const float* p2 = &out.positions[3 * (size_t)out.indices[3 * (size_t)t + 2]];
            DTri tr{};
            for (int c = 0; c < 3; ++c) { tr.v0[c] = p0[c]; tr.e1[c] = p1[c] - p0[c]; tr.e2[c] = p2[c] - p0[c]; }
            tr.prim = order[k];
            out.tris.push_back(tr);
        }
    }
    out.blas_nodes2 = (uint32_t)out.nodes.size();
    out.tshade.resize(4 * kShadeRec * (size_t)nt);
    shade_records(out.indices.data(), out.positions.data(), out.normals.data(), 0, nt, out.tshade.data());
    e = scene_tables(d, out);
    if (!e.empty()) return e;
    // TLAS over the instance world boxes: a mesh instance's vertices through
    // its transform, a built-in shape's box corners
    std::vector<float> lo, hi;
    for (uint32_t i = 0; i < d.n_instances; ++i) {
        const uint32_t m = d.instance_mesh[i];
        const float* M = &d.transforms[16 * (size_t)i];
        Box wb;
        const uint32_t v0 = out.mesh_vrange[2 * m], v1 = out.mesh_vrange[2 * m + 1];
        if (v0 < v1) {
            for (uint32_t v = v0; v < v1; ++v) grow_world_box(M, &d.positions[3 * (size_t)v], wb.lo, wb.hi);
            widen_world_box(wb.lo, wb.hi);
        } else {
            instance_world_box(M, mesh_box[m].lo, mesh_box[m].hi, wb.lo, wb.hi);
        }
        lo.insert(lo.end(), wb.lo, wb.lo + 3);
        hi.insert(hi.end(), wb.hi, wb.hi + 3);
    }
    std::vector<uint32_t> order;
    uint32_t depth = 0;
    out.tlas_root = build_bvh(lo, hi, bins, 1, 0, out.nodes, order, depth);
    out.tlas_depth = depth;
    // TLAS leaves reference positions in `order`; remap them to instance ids.
    {
        const uint32_t first_tlas = out.tlas_root;
        for (size_t k = first_tlas; k < out.nodes.size(); ++k)
            for (int c = 0; c < 2; ++c) {
                uint32_t& r = out.nodes[k].ref[c];
                if (r != kEmptyRef && (r & kLeafBit)) r = make_leaf(order[r & kPrimMask], 1);
            }
    }
    // 4-wide nodes: every BLAS, then the TLAS
    {
        std::vector<uint32_t> root4(d.n_meshes);
        uint32_t blas_bound = 0;
        for (uint32_t m = 0; m < d.n_meshes; ++m) {
            if (mesh_root[m] >= kImplicitRoot) {
                root4[m] = mesh_root[m];
                continue;
            }
            uint32_t b = 0;
            root4[m] = collapse_bvh4(out.nodes, mesh_root[m], out.nodes4, b);
            blas_bound = std::max(blas_bound, b);
        }
        for (uint32_t i = 0; i < d.n_instances; ++i) out.instances[i].blas_root = root4[d.instance_mesh[i]];
        uint32_t tb = 0;
        out.tlas_base = (uint32_t)out.nodes4.size();
        out.blas_bound = blas_bound;
        out.tlas_root = collapse_bvh4(out.nodes, out.tlas_root, out.nodes4, tb);
        out.stack_bound = tb + blas_bound;  // TLAS entries, then BLAS entries
    }
    return "";
}

}  // namespace pico
