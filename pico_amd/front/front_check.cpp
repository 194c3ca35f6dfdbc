// Drives the C++ front the way Pico's main() drives Scene::render_scene_to_file
// (Source/main.cpp:44-97): an OBJ model file through the import path and a
// JSON scene file loaded with its named camera (no
// device needed), one Cornell-box frame with denoising and tone mapping, the
// a-trous entry point on its own, a cancelled render, a Radiance file, and the
// scene file rendered with its camera.  Exit status 0 on success; without a
// visible GPU it reports the device error after the loading part and exits 0
// as well ("nodevice").
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "pico_front.hpp"

struct vec3 { float x, y, z; };

extern "C" int pico_scene_cornell(uint32_t seed, uint32_t microfacet, struct pico_scene** out);
extern "C" int pico_scene_desc_get(const struct pico_scene* s, pico_scene_desc* d);
extern "C" int pico_scene_default_camera(const struct pico_scene* s, float aspect, pico_camera* out);
extern "C" void pico_scene_destroy(struct pico_scene* s);

// A scene file with a named camera, as Pico's main() loads one (written to
// the working directory).
static void write_scene_file() {
    FILE* o = std::fopen("front_check_quad.obj", "w");
    FILE* j = std::fopen("front_check_scene.json", "w");
    if (!o || !j) return;
    std::fputs("v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nf 1 2 3 4\n", o);
    std::fputs("{ // comments and trailing commas, as jsoncpp reads them\n"
               " \"GLOBALS\": {\"S\": {\"SunDirection\": [0.3, -1, 0.2], \"SunColour\": [2, 2, 2]}},\n"
               " \"MESH\": {\"Quad\": {\"Path\": \"front_check_quad.obj\"}},\n"
               " \"MATERIALS\": {\"Grey\": {\"Type\": \"Constant\", \"Albedo\": [0.6, 0.6, 0.6], \"Roughness\": 0.5},},\n"
               " \"INSTANCE\": {\"floor\": {\"Asset\": \"Quad\", \"BSRDF\": \"Diffuse\", \"Material\": \"Grey\"},\n"
               "              \"ball\": {\"Asset\": \"Sphere\", \"BSRDF\": \"Dielectric\", \"Material\": \"Grey\",\n"
               "                       \"Position\": [0, 0.5, 0]}},\n"
               " \"CAMERA\": {\"main\": {\"Position\": [0, 1, -3], \"Direction\": [0, -0.3, 1], \"FarPlane\": 100}},\n"
               "}\n", j);
    std::fclose(o);
    std::fclose(j);
}

int main() {
    // a model file through the import path (no device needed): two meshes,
    // the emitting one a light
    {
        FILE* m = std::fopen("front_check_room.mtl", "w");
        FILE* o = std::fopen("front_check_room.obj", "w");
        if (!m || !o) return 23;
        std::fputs("newmtl Grey\nKd 0.5 0.5 0.5\nnewmtl Lamp\nKe 4 4 4\n", m);
        std::fputs("mtllib front_check_room.mtl\nv -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\n"
                   "o floor\nusemtl Grey\nf 1 2 3 4\no lamp\nusemtl Lamp\nf 4 3 2 1\n", o);
        std::fclose(m);
        std::fclose(o);
        pico_front::SceneFile model("front_check_room.obj");
        std::remove("front_check_room.mtl");
        std::remove("front_check_room.obj");
        if (model.status() != PICO_OK) { std::printf("model file %d %s\n", model.status(), model.error()); return 24; }
        if (model.desc().n_instances != 2 || model.desc().n_lights != 1 || model.get_camera("main")) return 25;
    }
    // the scene file loads without a device
    write_scene_file();
    pico_front::SceneFile sf("front_check_scene.json");
    std::remove("front_check_quad.obj");
    std::remove("front_check_scene.json");
    if (sf.status() != PICO_OK) { std::printf("scene file %d %s\n", sf.status(), sf.error()); return 11; }
    const pico_camera* c = sf.get_camera("main");
    if (!c || sf.get_camera("none") || sf.desc().n_instances != 2 || !sf.desc().use_sun) return 12;
    pico_front::Renderer r(0);
    if (r.status() != PICO_OK) {
        std::printf("nodevice status=%d\n", r.status());
        return 0;
    }
    pico_scene* s = nullptr;
    pico_scene_desc d;
    pico_camera cam;
    if (pico_scene_cornell(1, PICO_MICROFACET_GGX, &s) || pico_scene_desc_get(s, &d) ||
        pico_scene_default_camera(s, 1.0f, &cam))
        return 2;
    if (int rc = r.upload(d)) { std::printf("upload %d %s\n", rc, r.last_error().c_str()); return 3; }
    const uint32_t W = 64, H = 64;
    std::vector<vec3> pix(W * H), var(W * H);
    std::vector<uint32_t> cnt(W * H, 0);
    pico_front::RenderParams<vec3> p{5, 8, 0.0f, H, W, true, true, pix.data(), cnt.data(), var.data()};
    bool quit = false;
    if (int rc = r.render_scene_to_memory(cam, p, &quit)) { std::printf("render %d %s\n", rc, r.last_error().c_str()); return 4; }
    for (uint32_t i = 0; i < W * H; ++i) {
        if (cnt[i] == 0 || !std::isfinite(pix[i].x) || pix[i].x < 0.0f || pix[i].x > 1.0f) {
            std::printf("bad pixel %u: count %u value %g\n", i, cnt[i], pix[i].x);
            return 5;
        }
    }
    vec3* den = pico_front::atrous_denoise(r.context(), pix.data(), var.data(), var.data(), pix.data(), W, H);
    if (!den) return 6;
    delete[] den;
    // the learned filter with its fallback: no network, and a network whose
    // parameters are NaN, both give the a-trous filter's image
    {
        std::vector<vec3> nrm(W * H), pos(W * H), alb(W * H);
        if (pico_hip_gbuffer(r.context(), &cam, W, H, &nrm[0].x, &pos[0].x, &alb[0].x)) return 17;
        vec3* ref = pico_front::atrous_denoise(r.context(), pix.data(), nrm.data(), pos.data(), alb.data(), W, H, 16);
        int32_t fb = 0;
        vec3* a = pico_front::cnn_denoise(r.context(), (pico_cnn*)nullptr, pix.data(), nrm.data(), pos.data(), alb.data(),
                                          W, H, 16, &fb);
        pico_cnn* net = nullptr;
        if (!ref || !a || fb != 1 || pico_hip_cnn_create(r.context(), W, H, 1, &net)) return 18;
        std::vector<float> nan(pico_hip_cnn_num_params(), NAN);
        int32_t fb2 = 0;
        vec3* b = pico_hip_cnn_set_params(net, nan.data()) ? nullptr
                  : pico_front::cnn_denoise(r.context(), net, pix.data(), nrm.data(), pos.data(), alb.data(), W, H, 16, &fb2);
        if (!b || fb2 != 1) return 19;
        for (uint32_t i = 0; i < W * H; ++i) {
            const bool same_a = (a[i].x == ref[i].x || (std::isnan(a[i].x) && std::isnan(ref[i].x)));
            const bool same_b = (b[i].x == ref[i].x || (std::isnan(b[i].x) && std::isnan(ref[i].x)));
            if (!same_a || !same_b) { std::printf("fallback pixel %u: %g %g vs %g\n", i, a[i].x, b[i].x, ref[i].x); return 20; }
        }
        delete[] ref;
        delete[] a;
        delete[] b;
        pico_hip_cnn_destroy(net);
    }
    quit = true;  // a set flag stops the next render before it starts
    std::vector<uint32_t> cnt2(W * H, 0);
    pico_front::RenderParams<vec3> p2{5, 8, 0.0f, H, W, false, false, pix.data(), cnt2.data(), var.data()};
    if (r.render_scene_to_memory(cam, p2, &quit) != PICO_ERR_CANCELLED) return 7;
    // a cancelled render still tone-maps the (unchanged) frame, as the reference does
    for (uint32_t i = 0; i < W * H; ++i) pix[i] = {4.0f, 4.0f, 4.0f};
    pico_front::RenderParams<vec3> p2b{5, 8, 0.0f, H, W, true, false, pix.data(), cnt2.data(), var.data()};
    if (r.render_scene_to_memory(cam, p2b, &quit) != PICO_ERR_CANCELLED) return 21;
    for (uint32_t i = 0; i < W * H; ++i)
        if (!(pix[i].x > 0.0f && pix[i].x < 4.0f) || cnt2[i] != 0) { std::printf("cancelled frame pixel %u: %g\n", i, pix[i].x); return 22; }
    // render_scene_to_file: a Radiance file whose first line is the magic
    std::vector<uint32_t> cnt3(W * H, 0);
    pico_front::RenderParams<vec3> p3{5, 4, 0.0f, H, W, false, false, pix.data(), cnt3.data(), var.data()};
    const char* path = "front_check.hdr";
    if (int rc = r.render_scene_to_file(cam, p3, path)) { std::printf("file %d %s\n", rc, r.last_error().c_str()); return 8; }
    FILE* f = std::fopen(path, "rb");
    char magic[11] = {0};
    if (!f || std::fread(magic, 1, 10, f) != 10 || std::string(magic) != "#?RADIANCE") return 9;
    std::fclose(f);
    std::remove(path);
    pico_scene_destroy(s);
    // the scene file rendered with its camera
    {
        if (int rc = r.upload(sf)) { std::printf("upload file %d %s\n", rc, r.last_error().c_str()); return 13; }
        std::vector<uint32_t> cnt4(W * H, 0);
        pico_front::RenderParams<vec3> p4{5, 4, 0.0f, H, W, false, false, pix.data(), cnt4.data(), var.data()};
        bool go = false;
        if (int rc = r.render_scene_to_memory(*c, p4, &go)) { std::printf("render file %d %s\n", rc, r.last_error().c_str()); return 15; }
        double lit = 0.0;
        for (uint32_t i = 0; i < W * H; ++i) lit += pix[i].x;
        if (!(lit > 0.0)) return 16;  // the sun lights the floor
    }
    std::printf("ok\n");
    return 0;
}
