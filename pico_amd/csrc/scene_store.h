// The flattened scene behind the opaque pico_scene handle: procedural
// scenes (scene_gen.cpp) and scenes loaded from the reference's JSON format
// (scene_json.cpp) fill it, pico_scene_desc_get hands it out.
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pico_hip.h"

struct pico_scene {
    std::vector<uint32_t> vofs{0}, tofs{0};
    std::vector<float> pos, nrm, xf;
    std::vector<uint32_t> idx, imesh, imat, ibsdf, lights;
    std::vector<uint32_t> shape;  // per mesh: PICO_SHAPE_*
    std::vector<pico_material> mats;
    float sky[3] = {0, 0, 0};
    uint32_t microfacet = PICO_MICROFACET_GGX;
    uint32_t use_sun = 0;
    float sun_dir[3] = {0, -1, 0}, sun_col[3] = {1, 1, 1};  // Scene.cpp:692-712 defaults
    pico_sky_image sky_img{};       // texels point into sky_texels (width 0: constant sky)
    std::vector<uint8_t> sky_texels;
    pico_camera cam{};
    std::vector<std::pair<std::string, pico_camera>> cameras;  // named cameras of a loaded scene, by name

    // nv vertices (3 floats each), ni indices (mesh-local); returns the mesh id
    uint32_t add_mesh(const float* p, const float* n, size_t nv, const uint32_t* ix, size_t ni) {
        pos.insert(pos.end(), p, p + 3 * nv);
        nrm.insert(nrm.end(), n, n + 3 * nv);
        idx.insert(idx.end(), ix, ix + ni);
        vofs.push_back(vofs.back() + (uint32_t)nv);
        tofs.push_back(tofs.back() + (uint32_t)(ni / 3));
        shape.push_back(PICO_SHAPE_TRIANGLES);
        return (uint32_t)vofs.size() - 2;
    }
    // a built-in implicit shape (no vertices or triangles)
    uint32_t add_shape(uint32_t kind) {
        vofs.push_back(vofs.back());
        tofs.push_back(tofs.back());
        shape.push_back(kind);
        return (uint32_t)vofs.size() - 2;
    }
    uint32_t add_material(const pico_material& m) {
        mats.push_back(m);
        return (uint32_t)mats.size() - 1;
    }
    // M: column-major 4x4
    uint32_t add_instance(uint32_t mesh, const float* M, uint32_t mat, uint32_t bsdf) {
        xf.insert(xf.end(), M, M + 16);
        imesh.push_back(mesh);
        imat.push_back(mat);
        ibsdf.push_back(bsdf);
        const uint32_t id = (uint32_t)imesh.size() - 1;
        if (bsdf == PICO_BSDF_LIGHT) lights.push_back(id);
        return id;
    }
};
