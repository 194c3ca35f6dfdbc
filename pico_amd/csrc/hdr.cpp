// Radiance HDR (RGBE) output of a frame: the file half of
// Scene::render_scene_to_file (Source/Core/Scene.cpp:236-246), which reverses
// the pixel order end to end and writes 3-channel float pixels as .hdr.
// Encoding (Greg Ward's shared-exponent RGBE): v = max(r, g, b); below 1e-32
// the pixel is (0,0,0,0); else v = m 2^e with m in [0.5, 1), the bytes are
// (r, g, b) * m * 256 / v and e + 128.  Scanlines are written flat (one RGBE
// quadruple per pixel), which every RGBE reader accepts.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/pico_hip.h"

extern "C" int pico_write_hdr(const char* path, uint32_t width, uint32_t height, const float* rgb, int reverse) {
    if (!path || !rgb || width == 0 || height == 0) return PICO_ERR_ARG;
    const size_t n = (size_t)width * height;
    std::vector<uint8_t> px(4 * n);
    for (size_t i = 0; i < n; ++i) {
        const float* p = rgb + 3 * (reverse ? n - 1 - i : i);
        const float v = std::fmax(p[0], std::fmax(p[1], p[2]));
        uint8_t* q = &px[4 * i];
        if (!(v >= 1e-32f)) {
            q[0] = q[1] = q[2] = q[3] = 0;
            continue;
        }
        int e = 0;
        const float m = std::frexp(v, &e);
        const float s = m * 256.0f / v;
        for (int k = 0; k < 3; ++k) q[k] = (uint8_t)(p[k] > 0.0f ? p[k] * s : 0.0f);
        q[3] = (uint8_t)(e + 128);
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return PICO_ERR_ARG;
    const int hdr = std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", height, width);
    const bool ok = hdr > 0 && std::fwrite(px.data(), 1, px.size(), f) == px.size();
    return (std::fclose(f) == 0 && ok) ? PICO_OK : PICO_ERR_ARG;
}
