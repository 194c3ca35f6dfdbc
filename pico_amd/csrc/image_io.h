// Image files for scene ingest (host code): the sky images a scene file
// names.  The reference reads them with stb_image (Source/Core/Scene.cpp:641-680):
// stbi_load(path, 4 channels) for the six faces of a cube map, stbi_loadf(path,
// 4 channels) for an equirectangular image.  These are the same conversions,
// for the formats read here: PNG (every colour type and bit depth, tRNS,
// Adam7) and Radiance RGBE (.hdr, flat or run-length scanlines).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pico_img {

// RGBA8 texels, as stbi_load(path, &w, &h, &comp, 4): PNG only (an RGBE file
// is refused).  Returns an empty string or the reason it failed.
std::string load_rgba8(const std::string& path, uint32_t& w, uint32_t& h, std::vector<uint8_t>& out);

// RGBA float texels, as stbi_loadf(path, &w, &h, &comp, 4): an RGBE file
// decoded to floats (alpha 1), or a PNG whose 8-bit colour is linearised
// with gamma 2.2 (alpha divided by 255).
std::string load_rgba32f(const std::string& path, uint32_t& w, uint32_t& h, std::vector<float>& out);

}  // namespace pico_img
