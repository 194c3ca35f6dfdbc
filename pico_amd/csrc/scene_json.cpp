// Scenes in the reference's JSON format, read into the flattened scene
// (host code).  The reference's loader is Scene::Scene(ThreadPool&, path)
// (Source/Core/Scene.cpp:36-84) with one handler per section:
//   GLOBALS   process_globals    (Scene.cpp:639-720)
//   MESH      add_mesh           (Scene.cpp:297-318), meshes through assimp
//   MATERIALS add_material       (Scene.cpp:424-582)
//   INSTANCE  add_mesh_instance  (Scene.cpp:321-422)
//   CAMERA    add_camera         (Scene.cpp:584-636)
// Sections are read in that order and the entries of a section in the
// sorted order of their names (jsoncpp keeps object members in a std::map);
// the reference runs the entries of a section as parallel tasks, so its
// instance order is not fixed, and the order here is the sorted one.
//
// What is read:
//   - meshes from Wavefront OBJ files (the reference hands any format to
//     assimp with Triangulate | JoinIdenticalVertices | GenNormals and keeps
//     the file's first mesh): polygons become triangle fans, corners without
//     a normal get the flat normal of their triangle, and corners with equal
//     position and normal are joined in order of first use; the first mesh
//     ends where a later `o`, `g` or `usemtl` starts another one;
//   - Constant materials in both forms, metalness/roughness and
//     diffuse/specular/gloss, evaluated as Source/Render/BasicMaterials.cpp
//     evaluates them (a missing Metalness, uninitialised in the reference,
//     is 0 here);
//   - instances with T * R(normalised quaternion x, y, z, w) * S and the
//     BSRDFs Diffuse, Specular, Dielectric, Light, Delta, Transparent and
//     Fresnel (the microfacet lobes use the reference's Beckmann_All
//     distribution; the transmissive ones need a material with an
//     IndexOfRefraction);
//   - cameras (up (0, 1, 0), aspect 1, far plane 10 unless given; the
//     reference's rays ignore the near plane and the field of view).
// The built-in assets Sphere (radius 0.5) and Cube ([-0.5, 0.5]^3) become
// implicit-shape meshes.  GLOBALS give the sky (a cube map of six PNG faces
// or one equirectangular PNG / Radiance image, decoded as stb_image does) and
// the sun; without them the sky is black and there is no sun.  Refused with
// an error naming the entry: image materials (Metalic, Gloss), other image
// formats, and anything malformed.  A light instance
// whose mesh is already used by another light gets its own copy of the mesh
// (the device's light tables are per mesh).  Paths are resolved against the
// scene file's directory, without regard to case when the exact path is
// missing (as File_System_Mappings::resolve_path, Source/Core/FileMappings.cpp:37-61).
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "../../include/pico_hip.h"
#include "image_io.h"
#include "json_reader.h"
#include "scene_store.h"

namespace fs = std::filesystem;

namespace {

using JVal = pico_json::Value;

struct LoadError {
    std::string msg;
};

std::string read_file(const fs::path& p) {
    std::ifstream f(p, std::ios::binary | std::ios::ate);
    if (!f) throw LoadError{"cannot open " + p.string()};
    std::string s((size_t)f.tellg(), '\0');
    f.seekg(0);
    if (!f.read(s.data(), (std::streamsize)s.size())) throw LoadError{"cannot read " + p.string()};
    return s;
}

// The reference's path rule: relative to the scene's directory; when that
// file is missing, the file under the directory whose path matches without
// regard to case.
fs::path resolve(const fs::path& root, std::string rel) {
    std::replace(rel.begin(), rel.end(), '\\', '/');
    fs::path p = fs::path(rel).is_relative() ? (root / rel).lexically_normal() : fs::path(rel).lexically_normal();
    std::error_code ec;
    if (fs::exists(p, ec)) return p;
    auto lower = [](std::string s) {
        std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        return s;
    };
    const std::string want = lower(p.string());
    for (auto it = fs::recursive_directory_iterator(root, ec); !ec && it != fs::recursive_directory_iterator();
         it.increment(ec)) {
        if (it->is_regular_file(ec) && lower(it->path().lexically_normal().string()) == want) return it->path();
    }
    return p;
}

// ---- Wavefront OBJ --------------------------------------------------------
struct MeshData {
    std::vector<float> pos, nrm;
    std::vector<uint32_t> idx;
    std::string material;  // usemtl name ("" for the default material)
};

struct Corner { int64_t v, n; };

// Faces of one mesh, as corners (three per triangle) into the file's vertex
// and normal lists, and the material the mesh uses.
struct RawMesh {
    std::vector<Corner> corners;
    std::string material;
};

// An object of the file (`o`, or `g` with a new name): its meshes, one per
// run of faces under one material.
struct RawObject {
    std::string name;
    std::vector<RawMesh> meshes;
};

struct RawObj {
    std::vector<float> v, vn;
    std::vector<RawObject> objects;  // in the order the file creates them
    std::vector<std::string> mtllibs;
};

// Reads the file's vertices and faces.  Meshes are split as assimp's OBJ
// importer splits them (the reference's loader): `o name` continues the
// object of that name or starts one, `g name` with a name other than the
// active group starts an object, `usemtl` starts a new mesh in the current
// object once the current mesh has faces and the material changes, and faces
// before any `o` or `g` go to a default object.  With first_only the reading
// stops where the file's first mesh ends (a later `o`, `g` or `usemtl`).
RawObj parse_obj(const fs::path& path, bool first_only) {
    const std::string text = read_file(path);
    RawObj r;
    std::string material, active_group;
    int cur = -1;  // current object
    bool started = false;
    size_t line_no = 0;
    auto bad = [&](const std::string& what) {
        throw LoadError{path.string() + ":" + std::to_string(line_no) + ": " + what};
    };
    auto new_object = [&](const std::string& name) {
        r.objects.push_back({name, {RawMesh{{}, material}}});
        cur = (int)r.objects.size() - 1;
    };
    const char* p = text.data();
    const char* const end = p + text.size();
    auto blank = [](char c) { return c == ' ' || c == '\t' || c == '\r'; };
    while (p < end) {
        const char* eol = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        if (!eol) eol = end;
        ++line_no;
        const char* q = p;
        p = eol + 1;
        while (q < eol && blank(*q)) ++q;
        const char* t0 = q;
        while (q < eol && !blank(*q)) ++q;
        const std::string_view tag(t0, (size_t)(q - t0));
        if (tag.empty() || tag[0] == '#') continue;
        auto num = [&](float& x) {
            while (q < eol && blank(*q)) ++q;
            if (q < eol && *q == '+') ++q;
            const auto res = std::from_chars(q, eol, x);
            if (res.ec != std::errc()) return false;
            q = res.ptr;
            return true;
        };
        auto rest = [&]() {
            while (q < eol && blank(*q)) ++q;
            const char* e = eol;
            while (e > q && blank(e[-1])) --e;
            return std::string(q, (size_t)(e - q));
        };
        if (tag == "v" || tag == "vn") {
            float x = 0, y = 0, z = 0;
            if (!num(x) || !num(y) || !num(z)) bad(tag == "v" ? "malformed vertex" : "malformed normal");
            (tag == "v" ? r.v : r.vn).insert((tag == "v" ? r.v : r.vn).end(), {x, y, z});
        } else if (tag == "f") {
            if (cur < 0) new_object("defaultobject");
            std::vector<Corner>& corners = r.objects[cur].meshes.back().corners;
            Corner poly_first{}, prev{};
            int n_corner = 0;
            for (;;) {
                while (q < eol && blank(*q)) ++q;
                if (q >= eol) break;
                const char* tk = q;
                int64_t ref[3] = {0, 0, 0};
                for (int part = 0; part < 3; ++part) {
                    if (q < eol && !blank(*q) && *q != '/') {
                        if (*q == '+') ++q;
                        const auto res = std::from_chars(q, eol, ref[part]);
                        if (res.ec != std::errc() || ref[part] == 0) {
                            while (q < eol && !blank(*q)) ++q;
                            bad("malformed face index '" + std::string(tk, q) + "'");
                        }
                        q = res.ptr;
                    }
                    if (q < eol && *q == '/') ++q;
                    else break;
                }
                if (q < eol && !blank(*q)) {
                    while (q < eol && !blank(*q)) ++q;
                    bad("malformed face index '" + std::string(tk, q) + "'");
                }
                // 1-based, negative counts back from the last vertex read
                auto fix = [&](int64_t x, size_t count) -> int64_t {
                    if (x == 0) return -1;
                    const int64_t i = x > 0 ? x - 1 : (int64_t)count + x;
                    if (i < 0 || i >= (int64_t)count) bad("face index out of range '" + std::string(tk, q) + "'");
                    return i;
                };
                const Corner c{fix(ref[0], r.v.size() / 3), fix(ref[2], r.vn.size() / 3)};
                if (c.v < 0) bad("face corner without a vertex '" + std::string(tk, q) + "'");
                // a fan: (first, previous, this) from the third corner on
                if (n_corner == 0) poly_first = c;
                else if (n_corner >= 2) corners.insert(corners.end(), {poly_first, prev, c});
                prev = c;
                ++n_corner;
            }
            if (n_corner < 3) bad("face with fewer than 3 corners");
            started = true;
        } else if (tag == "o" || tag == "g" || tag == "usemtl") {
            if (first_only && started) break;  // the file's first mesh ends here
            const std::string name = rest();
            if (tag == "usemtl") {
                if (cur >= 0 && !r.objects[cur].meshes.back().corners.empty() && name != material)
                    r.objects[cur].meshes.push_back(RawMesh{{}, name});
                else if (cur >= 0)
                    r.objects[cur].meshes.back().material = name;
                material = name;
            } else if (tag == "o") {
                int found = -1;
                for (size_t i = 0; i < r.objects.size(); ++i)
                    if (r.objects[i].name == name) found = (int)i;
                if (found >= 0) cur = found;
                else new_object(name);
            } else if (name != active_group) {
                active_group = name;
                new_object(name);
            }
        } else if (tag == "mtllib") {
            r.mtllibs.push_back(rest());
        }
        // vt, s, l, p, ...: not needed
    }
    return r;
}

// One mesh of the file: flat normals for corners without one, then corners
// with equal position and normal joined in order of first use
// (Triangulate | JoinIdenticalVertices | GenNormals).
MeshData build_mesh(const RawObj& r, const RawMesh& raw) {
    const std::vector<float>& v = r.v;
    const std::vector<float>& vn = r.vn;
    const std::vector<Corner>& corners = raw.corners;
    MeshData m;
    m.material = raw.material;
    struct Key {
        uint32_t b[6];
        bool operator==(const Key& o) const { return std::memcmp(b, o.b, sizeof(b)) == 0; }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const {
            uint64_t h = 1469598103934665603ull;
            for (uint32_t x : k.b) { h ^= x; h *= 1099511628211ull; }
            return (size_t)h;
        }
    };
    std::unordered_map<Key, uint32_t, KeyHash> seen;
    seen.reserve(corners.size());
    for (size_t t = 0; t < corners.size(); t += 3) {
        const float* p0 = &v[3 * corners[t].v];
        const float* p1 = &v[3 * corners[t + 1].v];
        const float* p2 = &v[3 * corners[t + 2].v];
        float fn[3] = {0, 0, 0};
        {
            const float a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            const float b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            fn[0] = a[1] * b[2] - a[2] * b[1];
            fn[1] = a[2] * b[0] - a[0] * b[2];
            fn[2] = a[0] * b[1] - a[1] * b[0];
            const float l = std::sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
            if (l > 0.0f) for (float& x : fn) x /= l;
        }
        for (int c = 0; c < 3; ++c) {
            const Corner& cr = corners[t + c];
            const float* p = &v[3 * cr.v];
            const float* n = cr.n >= 0 ? &vn[3 * cr.n] : fn;
            Key k;
            std::memcpy(k.b, p, 12);
            std::memcpy(k.b + 3, n, 12);
            auto ins = seen.emplace(k, (uint32_t)(m.pos.size() / 3));
            if (ins.second) {
                m.pos.insert(m.pos.end(), p, p + 3);
                m.nrm.insert(m.nrm.end(), n, n + 3);
            }
            m.idx.push_back(ins.first->second);
        }
    }
    return m;
}

// The file's first mesh (Scene::add_mesh keeps mMeshes[0], Scene.cpp:297-318).
MeshData read_obj(const fs::path& path) {
    const RawObj r = parse_obj(path, true);
    for (const RawObject& o : r.objects)
        for (const RawMesh& m : o.meshes)
            if (!m.corners.empty()) return build_mesh(r, m);
    throw LoadError{path.string() + ": no faces"};
}

// ---- scene ----------------------------------------------------------------
float as_float(const JVal& v, const std::string& where) {
    if (v.kind != JVal::Num) throw LoadError{where + ": expected a number"};
    return v.as_float();
}

std::string as_string(const JVal* v, const std::string& where) {
    if (!v || v->kind != JVal::Str) throw LoadError{where + ": expected a string"};
    return v->str;
}

void vec(const JVal* v, int n, float* out, const std::string& where) {
    if (!v || v->kind != JVal::Arr || (int)v->arr.size() < n)
        throw LoadError{where + ": expected an array of " + std::to_string(n) + " numbers"};
    for (int k = 0; k < n; ++k) out[k] = as_float(v->arr[k], where);
}

// The sky of a GLOBALS entry (Scene.cpp:641-680): an array of six images,
// the faces +x, -x, +y, -y, +z, -z of a cube map read as 8-bit RGBA
// (stbi_load), or one equirectangular image read as float RGBA (stbi_loadf).
// The reference uses the last face's size for all six; faces of different
// sizes are refused here.
void load_sky(pico_scene& sc, const fs::path& root, const JVal& v, const std::string& where) {
    pico_sky_image img{};
    if (v.kind == JVal::Arr) {
        if (v.arr.size() < 6) throw LoadError{where + ": a cube map needs six faces"};
        for (int f = 0; f < 6; ++f) {
            const fs::path p = resolve(root, as_string(&v.arr[f], where + "[" + std::to_string(f) + "]"));
            uint32_t w = 0, h = 0;
            std::vector<uint8_t> px;
            const std::string e = pico_img::load_rgba8(p.string(), w, h, px);
            if (!e.empty()) throw LoadError{where + ": " + e};
            if (f && (w != img.width || h != img.height)) throw LoadError{where + ": the cube faces differ in size"};
            img.width = w;
            img.height = h;
            sc.sky_texels.insert(sc.sky_texels.end(), px.begin(), px.end());
        }
        img.faces = 6;
        img.format = PICO_TEXEL_RGBA8;
    } else if (v.kind == JVal::Str) {
        const fs::path p = resolve(root, v.str);
        std::vector<float> px;
        const std::string e = pico_img::load_rgba32f(p.string(), img.width, img.height, px);
        if (!e.empty()) throw LoadError{where + ": " + e};
        sc.sky_texels.resize(px.size() * sizeof(float));
        std::memcpy(sc.sky_texels.data(), px.data(), sc.sky_texels.size());
        img.faces = 1;
        img.format = PICO_TEXEL_RGBA32F;
    } else {
        throw LoadError{where + ": expected six image paths or one image path"};
    }
    img.texels = sc.sky_texels.data();
    sc.sky_img = img;
}

// column-major a * b
void mat_mul(const float* a, const float* b, float* r) {
    for (int c = 0; c < 4; ++c)
        for (int rr = 0; rr < 4; ++rr) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + rr] * b[c * 4 + k];
            r[c * 4 + rr] = s;
        }
}

void load_into(pico_scene& sc, const fs::path& file) {
    const fs::path root = file.parent_path().empty() ? fs::path(".") : file.parent_path();
    const std::string text = read_file(file);
    JVal doc;
    try {
        doc = pico_json::Reader(text, file.string()).parse();
    } catch (const pico_json::Error& e) {
        throw LoadError{e.msg};
    }
    if (doc.kind != JVal::Obj) throw LoadError{file.string() + ": the top level is not an object"};
    sc.microfacet = PICO_MICROFACET_BECKMANN_ALL;
    // GLOBALS: process_globals per entry (Scene.cpp:639-720); every entry
    // sets the sky and the sun again, so the last one in name order decides
    if (const JVal* g = doc.get("GLOBALS")) {
        if (g->kind != JVal::Obj) throw LoadError{"GLOBALS: expected an object"};
        for (const auto& [name, e] : g->obj) {
            const std::string where = "GLOBALS/" + name;
            if (e.kind != JVal::Obj) throw LoadError{where + ": expected an object"};
            sc.sky_img = pico_sky_image{};
            sc.sky_texels.clear();
            if (const JVal* sb = e.get("Skybox")) load_sky(sc, root, *sb, where + "/Skybox");
            // SunDirection is normalised as glm::normalize does, v * (1 / sqrt(dot(v, v)))
            float dir[3] = {0.0f, -1.0f, 0.0f}, col[3] = {1.0f, 1.0f, 1.0f};
            if (e.get("SunDirection")) {
                vec(e.get("SunDirection"), 3, dir, where + "/SunDirection");
                const float inv = 1.0f / std::sqrt((dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2]);
                for (float& x : dir) x = x * inv;
            }
            if (e.get("SunColour")) vec(e.get("SunColour"), 3, col, where + "/SunColour");
            sc.use_sun = e.get("SunDirection") && e.get("SunColour") ? 1u : 0u;
            std::memcpy(sc.sun_dir, dir, sizeof(dir));
            std::memcpy(sc.sun_col, col, sizeof(col));
        }
    }
    // MESH
    std::map<std::string, uint32_t> mesh_of;
    std::vector<MeshData> meshes;
    if (const JVal* ms = doc.get("MESH")) {
        if (ms->kind != JVal::Obj) throw LoadError{"MESH: expected an object"};
        for (const auto& [name, e] : ms->obj) {
            const std::string where = "MESH/" + name;
            if (name == "Sphere" || name == "Cube") throw LoadError{where + ": the name of a built-in shape"};
            const fs::path p = resolve(root, as_string(e.get("Path"), where + "/Path"));
            std::string ext = p.extension().string();
            std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return (char)std::tolower(c); });
            if (ext != ".obj") throw LoadError{where + ": only Wavefront .obj meshes are read (" + p.string() + ")"};
            mesh_of[name] = (uint32_t)meshes.size();
            meshes.push_back(read_obj(p));
        }
    }
    // MATERIALS
    std::map<std::string, uint32_t> mat_of;
    if (const JVal* mm = doc.get("MATERIALS")) {
        if (mm->kind != JVal::Obj) throw LoadError{"MATERIALS: expected an object"};
        for (const auto& [name, e] : mm->obj) {
            const std::string where = "MATERIALS/" + name;
            const std::string type = as_string(e.get("Type"), where + "/Type");
            if (type == "Metalic" || type == "Gloss")
                throw LoadError{where + ": image materials (" + type + ") are not supported"};
            if (type != "Constant") throw LoadError{where + ": unknown material type " + type};
            pico_material m{};
            float emissive[3] = {0, 0, 0};
            const bool transparent = e.get("IndexOfRefraction") != nullptr;  // its emission is 0 in the reference
            if (e.get("Emissive") && !transparent) vec(e.get("Emissive"), 3, emissive, where + "/Emissive");
            if (e.get("Albedo")) {
                float albedo[3];
                vec(e.get("Albedo"), 3, albedo, where + "/Albedo");
                const float metal = e.get("Metalness") ? as_float(*e.get("Metalness"), where + "/Metalness") : 0.0f;
                const float rough = e.get("Roughness") ? as_float(*e.get("Roughness"), where + "/Roughness") : 0.0f;
                for (int k = 0; k < 3; ++k) {
                    m.diffuse[k] = albedo[k] * (1.0f - 0.04f) * (1.0f - metal);
                    m.specular[k] = 0.04f + metal * (albedo[k] - 0.04f);  // glm::lerp(0.04, albedo, metalness)
                }
                m.roughness = rough;
            } else if (e.get("Diffuse") || e.get("Specular")) {
                if (e.get("Diffuse")) vec(e.get("Diffuse"), 3, m.diffuse, where + "/Diffuse");
                if (e.get("Specular")) vec(e.get("Specular"), 3, m.specular, where + "/Specular");
                const float gloss = e.get("Gloss") ? as_float(*e.get("Gloss"), where + "/Gloss") : 0.0f;
                m.roughness = 1.0f - gloss * gloss;
            } else {
                throw LoadError{where + ": a Constant material needs Albedo, Diffuse or Specular"};
            }
            for (int k = 0; k < 3; ++k) m.emissive[k] = emissive[k];
            if (transparent) {
                m.index_of_refraction = as_float(*e.get("IndexOfRefraction"), where + "/IndexOfRefraction");
                if (!(m.index_of_refraction > 0.0f)) throw LoadError{where + "/IndexOfRefraction: must be positive"};
            }
            mat_of[name] = sc.add_material(m);
        }
    }
    if (sc.mats.empty()) throw LoadError{"the scene has no materials"};
    // INSTANCE
    std::vector<uint32_t> mesh_id(meshes.size(), UINT32_MAX);
    std::vector<bool> light_uses(meshes.size(), false);
    uint32_t builtin_mesh[3] = {UINT32_MAX, UINT32_MAX, UINT32_MAX};
    bool builtin_light[3] = {false, false, false};
    auto add_mesh = [&](uint32_t m) {
        const MeshData& d = meshes[m];
        return sc.add_mesh(d.pos.data(), d.nrm.data(), d.pos.size() / 3, d.idx.data(), d.idx.size());
    };
    if (const JVal* is = doc.get("INSTANCE")) {
        if (is->kind != JVal::Obj) throw LoadError{"INSTANCE: expected an object"};
        for (const auto& [name, e] : is->obj) {
            const std::string where = "INSTANCE/" + name;
            const std::string asset = as_string(e.get("Asset"), where + "/Asset");
            const uint32_t builtin = asset == "Sphere" ? PICO_SHAPE_SPHERE : (asset == "Cube" ? PICO_SHAPE_CUBE : 0u);
            auto mi = mesh_of.find(asset);
            if (!builtin && mi == mesh_of.end()) throw LoadError{where + ": unknown asset " + asset};
            float position[3] = {0, 0, 0}, scale[3] = {1, 1, 1}, q[4] = {0, 0, 0, 1};  // q: x, y, z, w
            if (e.get("Position")) vec(e.get("Position"), 3, position, where + "/Position");
            if (e.get("Scale")) vec(e.get("Scale"), 3, scale, where + "/Scale");
            if (e.get("Rotation")) {
                vec(e.get("Rotation"), 4, q, where + "/Rotation");
                const float l = std::sqrt(q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);  // glm::normalize(quat)
                if (!(l > 0.0f)) throw LoadError{where + "/Rotation: zero quaternion"};
                for (float& x : q) x /= l;
            }
            uint32_t material = 0;
            if (e.get("Material")) {
                const std::string mn = as_string(e.get("Material"), where + "/Material");
                auto it = mat_of.find(mn);
                if (it == mat_of.end()) throw LoadError{where + ": unknown material " + mn};
                material = it->second;
            }
            const std::string kind = as_string(e.get("BSRDF"), where + "/BSRDF");
            uint32_t bsdf;
            if (kind == "Diffuse") bsdf = PICO_BSDF_DIFFUSE;
            else if (kind == "Specular") bsdf = PICO_BSDF_SPECULAR;
            else if (kind == "Dielectric") bsdf = PICO_BSDF_DIELECTRIC;
            else if (kind == "Light") bsdf = PICO_BSDF_LIGHT;
            else if (kind == "Delta") bsdf = PICO_BSDF_DELTA;
            else if (kind == "Transparent") bsdf = PICO_BSDF_TRANSPARENT;
            else if (kind == "Fresnel") bsdf = PICO_BSDF_FRESNEL;
            else throw LoadError{where + ": unknown BSRDF " + kind};
            if ((bsdf == PICO_BSDF_TRANSPARENT || bsdf == PICO_BSDF_FRESNEL) && !(sc.mats[material].index_of_refraction > 0.0f))
                throw LoadError{where + ": the " + kind + " BSRDF needs a material with an IndexOfRefraction"};
            // T * R * S (Scene.cpp:359-361); R = glm::mat4_cast(q)
            const float x = q[0], y = q[1], z = q[2], w = q[3];
            float T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, position[0], position[1], position[2], 1};
            float R[16] = {1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y), 0,
                           2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x), 0,
                           2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y), 0,
                           0, 0, 0, 1};
            float S[16] = {scale[0], 0, 0, 0, 0, scale[1], 0, 0, 0, 0, scale[2], 0, 0, 0, 0, 1};
            float TR[16], M[16];
            mat_mul(T, R, TR);
            mat_mul(TR, S, M);
            uint32_t id;
            if (builtin) {
                // the built-in shapes (Scene.cpp:49-56): one mesh entry each,
                // and its own entry for every further light on the shape
                if (bsdf == PICO_BSDF_LIGHT && builtin_light[builtin]) {
                    id = sc.add_shape(builtin);
                } else {
                    if (builtin_mesh[builtin] == UINT32_MAX) builtin_mesh[builtin] = sc.add_shape(builtin);
                    id = builtin_mesh[builtin];
                }
                if (bsdf == PICO_BSDF_LIGHT) builtin_light[builtin] = true;
            } else {
                const uint32_t m = mi->second;
                if (bsdf == PICO_BSDF_LIGHT && light_uses[m]) {
                    id = add_mesh(m);  // a second light on this mesh: its own copy
                } else {
                    if (mesh_id[m] == UINT32_MAX) mesh_id[m] = add_mesh(m);
                    id = mesh_id[m];
                }
                if (bsdf == PICO_BSDF_LIGHT) light_uses[m] = true;
            }
            sc.add_instance(id, M, material, bsdf);
        }
    }
    if (sc.imesh.empty()) throw LoadError{"the scene has no instances"};
    // CAMERA
    if (const JVal* cs = doc.get("CAMERA")) {
        if (cs->kind != JVal::Obj) throw LoadError{"CAMERA: expected an object"};
        for (const auto& [name, e] : cs->obj) {
            const std::string where = "CAMERA/" + name;
            pico_camera c{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, 1.0f, 10.0f};
            if (e.get("Position")) vec(e.get("Position"), 3, c.position, where + "/Position");
            if (e.get("Direction")) {
                vec(e.get("Direction"), 3, c.direction, where + "/Direction");
                const float l = std::sqrt(c.direction[0] * c.direction[0] + c.direction[1] * c.direction[1] +
                                          c.direction[2] * c.direction[2]);
                if (!(l > 0.0f)) throw LoadError{where + "/Direction: zero vector"};
                for (float& d : c.direction) d /= l;
            }
            if (e.get("Aspect")) c.aspect = as_float(*e.get("Aspect"), where + "/Aspect");
            if (e.get("FarPlane")) c.far_plane = as_float(*e.get("FarPlane"), where + "/FarPlane");
            sc.cameras.emplace_back(name, c);
        }
    }
    sc.cam = sc.cameras.empty() ? pico_camera{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, 1.0f, 10.0f} : sc.cameras.front().second;
}

// ---- a model file through the reference's import path ----------------------
// Scene::Scene(ThreadPool&, dir, aiScene*, options) (Source/Core/Scene.cpp:86-141)
// with parse_node (Scene.cpp:722-783) and add_material (Scene.cpp:785-971) on
// an OBJ file as assimp imports it (Source/main.cpp:53-59).  Every mesh of
// the file becomes an instance with the identity transform (OBJ nodes have
// none): a Light when its material emits, else a Dielectric with the
// Beckmann_All lobe; lights are listed in mesh order (the reference lists them
// in the order its loading tasks finish).  Materials come from the file's
// mtllib libraries as the constant branch of add_material reads them: Kd, Ks
// and Ke into ConstantDiffuseSpecularMaterial with "gloss" = the glTF
// roughness factor, which OBJ materials do not carry, so 1, and the material
// roughness 1 - 1^2 = 0 (Source/Render/BasicMaterials.cpp:40-46); OBJ sets no
// transparency factor, so no material is transparent.  A mesh without a
// material, or naming one no library defines, gets assimp's default material
// (diffuse 0.6).  The sky is black and there is no sun unless the options
// set them (pico_scene_set_sky_image, pico_scene_set_sun).
struct ObjMaterial {
    float kd[3] = {0.6f, 0.6f, 0.6f}, ks[3] = {0, 0, 0}, ke[3] = {0, 0, 0};
};

void read_mtl(const fs::path& path, std::map<std::string, ObjMaterial>& out) {
    const std::string text = read_file(path);
    std::istringstream in(text);
    std::string line, cur;
    size_t line_no = 0;
    while (std::getline(in, line)) {
        ++line_no;
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag) || tag[0] == '#') continue;
        if (tag == "newmtl") {
            std::getline(ls >> std::ws, cur);
            while (!cur.empty() && (cur.back() == '\r' || cur.back() == ' ' || cur.back() == '\t')) cur.pop_back();
            out[cur] = ObjMaterial{};
        } else if (tag == "Kd" || tag == "Ks" || tag == "Ke") {
            if (cur.empty() && !out.count(cur)) out[cur] = ObjMaterial{};
            ObjMaterial& m = out[cur];
            float* dst = tag == "Kd" ? m.kd : tag == "Ks" ? m.ks : m.ke;
            float c[3];
            if (!(ls >> c[0])) throw LoadError{path.string() + ":" + std::to_string(line_no) + ": malformed " + tag};
            // one value sets all three channels
            if (!(ls >> c[1] >> c[2])) c[1] = c[2] = c[0];
            std::copy(c, c + 3, dst);
        }
        // Ns, Ni, d, Tr, illum, map_*: no effect on the reference's constant materials
    }
}

void load_obj_scene(pico_scene& sc, const fs::path& file) {
    const RawObj r = parse_obj(file, false);
    const fs::path root = file.parent_path();
    std::map<std::string, ObjMaterial> lib;
    for (const std::string& m : r.mtllibs) read_mtl(resolve(root, m), lib);
    std::map<std::string, uint32_t> mat_id;
    for (const RawObject& o : r.objects)
        for (const RawMesh& raw : o.meshes) {
            if (raw.corners.empty()) continue;
            const MeshData m = build_mesh(r, raw);
            auto it = mat_id.find(raw.material);
            if (it == mat_id.end()) {
                const auto f = lib.find(raw.material);
                const ObjMaterial om = f == lib.end() ? ObjMaterial{} : f->second;
                pico_material pm{};
                std::copy(om.kd, om.kd + 3, pm.diffuse);
                std::copy(om.ks, om.ks + 3, pm.specular);
                pm.roughness = 1.0f - 1.0f * 1.0f;
                std::copy(om.ke, om.ke + 3, pm.emissive);
                it = mat_id.emplace(raw.material, sc.add_material(pm)).first;
            }
            const pico_material& pm = sc.mats[it->second];
            const bool light = pm.emissive[0] > 0.0f || pm.emissive[1] > 0.0f || pm.emissive[2] > 0.0f;
            const uint32_t id = sc.add_mesh(m.pos.data(), m.nrm.data(), m.pos.size() / 3, m.idx.data(), m.idx.size());
            const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
            sc.add_instance(id, I, it->second, light ? PICO_BSDF_LIGHT : PICO_BSDF_DIELECTRIC);
        }
    if (sc.imesh.empty()) throw LoadError{file.string() + ": no faces"};
    sc.microfacet = PICO_MICROFACET_BECKMANN_ALL;
    sc.cam = pico_camera{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, 1.0f, 10.0f};
}

void set_err(char* err, size_t n, const std::string& msg) {
    if (!err || n == 0) return;
    const size_t k = std::min(n - 1, msg.size());
    std::memcpy(err, msg.data(), k);
    err[k] = 0;
}

}  // namespace

extern "C" {

int pico_scene_load_json(const char* path, pico_scene** out, char* err, size_t err_len) {
    set_err(err, err_len, "");
    if (!path || !out) {
        set_err(err, err_len, "NULL argument");
        return PICO_ERR_ARG;
    }
    *out = nullptr;
    auto* s = new pico_scene();
    try {
        load_into(*s, fs::path(path));
    } catch (const LoadError& e) {
        delete s;
        set_err(err, err_len, e.msg);
        return PICO_ERR_ARG;
    } catch (const std::exception& e) {
        delete s;
        set_err(err, err_len, std::string("scene load failed: ") + e.what());
        return PICO_ERR_ARG;
    }
    *out = s;
    return PICO_OK;
}

int pico_scene_load_obj(const char* path, pico_scene** out, char* err, size_t err_len) {
    set_err(err, err_len, "");
    if (!path || !out) {
        set_err(err, err_len, "NULL argument");
        return PICO_ERR_ARG;
    }
    *out = nullptr;
    auto* s = new pico_scene();
    try {
        load_obj_scene(*s, fs::path(path));
    } catch (const LoadError& e) {
        delete s;
        set_err(err, err_len, e.msg);
        return PICO_ERR_ARG;
    } catch (const std::exception& e) {
        delete s;
        set_err(err, err_len, std::string("scene load failed: ") + e.what());
        return PICO_ERR_ARG;
    }
    *out = s;
    return PICO_OK;
}

int pico_scene_camera_count(const pico_scene* s, uint32_t* n) {
    if (!s || !n) return PICO_ERR_ARG;
    *n = (uint32_t)s->cameras.size();
    return PICO_OK;
}

int pico_scene_camera_get(const pico_scene* s, uint32_t i, char* name, size_t name_len, pico_camera* out) {
    if (!s || i >= s->cameras.size()) return PICO_ERR_ARG;
    set_err(name, name_len, s->cameras[i].first);
    if (out) *out = s->cameras[i].second;
    return PICO_OK;
}

}  // extern "C"
