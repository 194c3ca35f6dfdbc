// PNG and Radiance RGBE decoding for scene ingest (host code); the
// conversions are stb_image's (Source/Core/Scene.cpp:641-680 loads sky
// images through stbi_load / stbi_loadf with 4 channels):
//   - PNG (ISO/IEC 15948): zlib stream (RFC 1950/1951) of the IDAT chunks,
//     per-row filters, Adam7 passes; 1/2/4-bit grey scaled to 0..255,
//     16-bit samples reduced to their high byte; tRNS gives palette alpha or
//     a colour key (compared at the file's bit depth, the key of a <= 8-bit
//     image taken from its low byte and scaled like the samples); grey is
//     replicated to RGB, missing alpha is 255.  CRCs and the Adler-32 are not
//     checked (as stb_image).
//   - RGBE (Greg Ward's Radiance format): "#?RADIANCE" or "#?RGBE", header
//     lines up to an empty line with FORMAT=32-bit_rle_rgbe, then "-Y h +X w";
//     scanlines flat, or run-length per channel when the width is 8..32767
//     and the scanline starts with 2, 2; texel = mantissa * 2^(e - 136)
//     (0 when e = 0), alpha 1.
//   - stbi_loadf of a PNG: colour powf(v / 255, 2.2), alpha v / 255.
#include "image_io.h"

#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>

namespace pico_img {
namespace {

bool read_bytes(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

// ------------------------------------------------------------- inflate ----

struct BitReader {
    const uint8_t* p;
    size_t n, pos = 0;
    uint64_t buf = 0;
    int cnt = 0;
    // at least k (<= 32) bits in buf; past the end of the data it reads
    // zeros (look-ahead), and overrun() tells whether any were consumed
    void fill(int k) {
        while (cnt < k) {
            const uint64_t b = pos < n ? p[pos] : 0;
            ++pos;
            buf |= b << cnt;
            cnt += 8;
        }
    }
    bool overrun() const { return pos * 8 - (size_t)cnt > n * 8; }
    uint32_t get(int k) {
        if (!k) return 0;
        fill(k);
        const uint32_t v = (uint32_t)(buf & ((1ull << k) - 1));
        buf >>= k;
        cnt -= k;
        return v;
    }
    void align() {
        const int r = cnt & 7;
        buf >>= r;
        cnt -= r;
    }
};

// Canonical Huffman code: a 9-bit first-level table, then the codes longer
// than 9 bits by length (canonical order).
struct Huffman {
    static constexpr int kFast = 9;
    uint16_t fast[1 << kFast];   // (length << 12) | symbol, 0: slow path
    uint16_t count[16];
    uint16_t first_code[16], first_index[16];
    uint16_t sym[320];
    bool build(const uint8_t* len, int n) {
        std::memset(count, 0, sizeof(count));
        std::memset(fast, 0, sizeof(fast));
        for (int i = 0; i < n; ++i) ++count[len[i]];
        count[0] = 0;
        int left = 1;
        for (int l = 1; l < 16; ++l) {
            left = 2 * left - count[l];
            if (left < 0) return false;  // over-subscribed
        }
        uint16_t offs[16];
        int code = 0, index = 0;
        for (int l = 1; l < 16; ++l) {
            first_code[l] = (uint16_t)code;
            first_index[l] = (uint16_t)index;
            offs[l] = (uint16_t)index;
            code = (code + count[l]) << 1;
            index += count[l];
        }
        for (int i = 0; i < n; ++i)
            if (len[i]) sym[offs[len[i]]++] = (uint16_t)i;
        // first-level table: codes are sent most significant bit first, the
        // reader delivers bits LSB first, so index by the reversed code
        for (int l = 1; l <= kFast; ++l)
            for (int k = 0; k < count[l]; ++k) {
                const int c = first_code[l] + k;
                int r = 0;
                for (int b = 0; b < l; ++b) r |= ((c >> b) & 1) << (l - 1 - b);
                for (int fill = r; fill < (1 << kFast); fill += 1 << l)
                    fast[fill] = (uint16_t)((l << 12) | sym[first_index[l] + k]);
            }
        return true;
    }
    int decode(BitReader& br) const {
        br.fill(16);
        const uint16_t e = fast[br.buf & ((1u << kFast) - 1)];
        if (e) {
            const int l = e >> 12;
            br.buf >>= l;
            br.cnt -= l;
            return e & 0xfff;
        }
        int code = 0;
        for (int l = 1; l < 16; ++l) {
            code |= (int)br.get(1);
            const int d = code - first_code[l];
            if (d >= 0 && d < count[l]) return sym[first_index[l] + d];
            code <<= 1;
        }
        return -1;
    }
};

const uint16_t kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                               35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193,
                                257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

std::string inflate_zlib(const uint8_t* data, size_t n, std::vector<uint8_t>& out, size_t expect) {
    if (n < 2) return "zlib stream too short";
    const uint8_t cmf = data[0], flg = data[1];
    if (((cmf << 8) | flg) % 31 != 0) return "bad zlib header";
    if (flg & 32) return "zlib preset dictionary";
    if ((cmf & 15) != 8) return "zlib method is not deflate";
    BitReader br{data + 2, n - 2};
    out.clear();
    out.reserve(expect);
    for (;;) {
        const uint32_t final = br.get(1);
        const uint32_t type = br.get(2);
        if (type == 0) {
            br.align();
            const uint32_t len = br.get(16), nlen = br.get(16);
            if ((len ^ 0xffff) != nlen) return "corrupt stored block";
            for (uint32_t i = 0; i < len; ++i) out.push_back((uint8_t)br.get(8));
            if (br.overrun()) return "stored block runs past the end";
        } else if (type == 1 || type == 2) {
            Huffman lit, dist;
            uint8_t lens[320];
            if (type == 1) {
                for (int i = 0; i < 144; ++i) lens[i] = 8;
                for (int i = 144; i < 256; ++i) lens[i] = 9;
                for (int i = 256; i < 280; ++i) lens[i] = 7;
                for (int i = 280; i < 288; ++i) lens[i] = 8;
                lit.build(lens, 288);
                for (int i = 0; i < 30; ++i) lens[i] = 5;
                dist.build(lens, 30);
            } else {
                const int hlit = (int)br.get(5) + 257, hdist = (int)br.get(5) + 1, hclen = (int)br.get(4) + 4;
                static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
                uint8_t cl[19] = {0};
                for (int i = 0; i < hclen; ++i) cl[order[i]] = (uint8_t)br.get(3);
                Huffman clh;
                if (!clh.build(cl, 19)) return "corrupt code lengths";
                int i = 0;
                while (i < hlit + hdist) {
                    const int s = clh.decode(br);
                    if (s < 0) return "corrupt code lengths";
                    if (s < 16) {
                        lens[i++] = (uint8_t)s;
                    } else {
                        int rep = 0;
                        uint8_t v = 0;
                        if (s == 16) {
                            if (i == 0) return "corrupt code lengths";
                            v = lens[i - 1];
                            rep = 3 + (int)br.get(2);
                        } else if (s == 17) {
                            rep = 3 + (int)br.get(3);
                        } else {
                            rep = 11 + (int)br.get(7);
                        }
                        if (i + rep > hlit + hdist) return "corrupt code lengths";
                        while (rep--) lens[i++] = v;
                    }
                }
                if (!lit.build(lens, hlit) || !dist.build(lens + hlit, hdist)) return "corrupt Huffman code";
            }
            for (;;) {
                const int s = lit.decode(br);
                if (s < 0 || br.overrun()) return "corrupt deflate data";
                if (s < 256) {
                    out.push_back((uint8_t)s);
                } else if (s == 256) {
                    break;
                } else {
                    const int li = s - 257;
                    if (li >= 29) return "corrupt length code";
                    const uint32_t len = kLenBase[li] + br.get(kLenExtra[li]);
                    const int ds = dist.decode(br);
                    if (ds < 0 || ds >= 30) return "corrupt distance code";
                    const uint32_t d = kDistBase[ds] + br.get(kDistExtra[ds]);
                    if (d > out.size()) return "distance before the start of the data";
                    const size_t from = out.size() - d;
                    for (uint32_t k = 0; k < len; ++k) {
                        const uint8_t v = out[from + k];
                        out.push_back(v);
                    }
                }
            }
        } else {
            return "bad deflate block type";
        }
        if (br.overrun()) return "deflate data ends early";
        if (final) break;
    }
    return "";
}

// ----------------------------------------------------------------- PNG ----

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

std::string decode_png(const std::vector<uint8_t>& f, uint32_t& w, uint32_t& h, std::vector<uint8_t>& rgba) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return "not a PNG file";
    size_t pos = 8;
    bool have_ihdr = false, ended = false;
    uint32_t depth = 0, color = 0, interlace = 0;
    uint8_t palette[256][4];
    uint32_t pal_len = 0;
    bool pal_alpha = false, has_key = false;
    uint16_t key[3] = {0, 0, 0};
    std::vector<uint8_t> idat;
    while (!ended) {
        if (pos + 8 > f.size()) return "PNG ends before IEND";
        const uint32_t len = be32(&f[pos]);
        const uint32_t type = be32(&f[pos + 4]);
        if (len > f.size() - pos - 8) return "PNG chunk runs past the end of the file";
        const uint8_t* d = &f[pos + 8];
        auto is = [&](const char* t) { return type == be32(reinterpret_cast<const uint8_t*>(t)); };
        if (is("IHDR")) {
            if (have_ihdr || len != 13) return "bad IHDR";
            have_ihdr = true;
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            color = d[9];
            interlace = d[12];
            if (d[10] || d[11] || interlace > 1) return "unknown compression, filter or interlace method";
            if (depth != 1 && depth != 2 && depth != 4 && depth != 8 && depth != 16) return "bad bit depth";
            if (color > 6 || (color != 3 && (color & 1)) || (color == 3 && depth == 16)) return "bad colour type";
            if ((color == 2 || color == 4 || color == 6) && depth < 8) return "bad bit depth for the colour type";
            if (!w || !h || w > (1u << 24) || h > (1u << 24) || (uint64_t)w * h > (1ull << 26)) return "image larger than 2^26 pixels";
        } else if (!have_ihdr) {
            return "the first chunk is not IHDR";
        } else if (is("PLTE")) {
            if (len > 768 || len % 3) return "bad PLTE";
            pal_len = len / 3;
            for (uint32_t i = 0; i < pal_len; ++i) {
                palette[i][0] = d[3 * i];
                palette[i][1] = d[3 * i + 1];
                palette[i][2] = d[3 * i + 2];
                palette[i][3] = 255;
            }
        } else if (is("tRNS")) {
            if (!idat.empty()) return "tRNS after IDAT";
            if (color == 3) {
                if (!pal_len || len > pal_len) return "bad tRNS";
                for (uint32_t i = 0; i < len; ++i) palette[i][3] = d[i];
                pal_alpha = true;
            } else {
                const uint32_t n = color == 2 ? 3 : 1;
                if (color & 4) return "tRNS with an alpha channel";
                if (len != 2 * n) return "bad tRNS";
                static const uint8_t scale8[17] = {0, 0xff, 0x55, 0, 0x11, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0};
                for (uint32_t k = 0; k < n; ++k) {
                    const uint16_t v = (uint16_t)((d[2 * k] << 8) | d[2 * k + 1]);
                    key[k] = depth == 16 ? v : (uint16_t)(uint8_t)((v & 255) * scale8[depth]);
                }
                has_key = true;
            }
        } else if (is("IDAT")) {
            if (color == 3 && !pal_len) return "IDAT before PLTE";
            idat.insert(idat.end(), d, d + len);
        } else if (is("IEND")) {
            ended = true;
        } else if (is("CgBI")) {
            return "Apple-optimised PNG is not supported";
        } else if (!(type & (1u << 29))) {
            return "unknown critical PNG chunk";
        }
        pos += 12 + (size_t)len;
    }
    if (!have_ihdr || idat.empty()) return "PNG has no image data";
    const uint32_t n = color == 3 ? 1 : (color & 2 ? 3 : 1) + (color & 4 ? 1 : 0);
    const uint32_t bpp = std::max<uint32_t>(1, n * depth / 8);  // filter unit in bytes
    // the passes: (x origin, y origin, x step, y step)
    static const uint32_t adam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                         {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    static const uint32_t whole[1][4] = {{0, 0, 1, 1}};
    const uint32_t (*passes)[4] = interlace ? adam7 : whole;
    const int n_passes = interlace ? 7 : 1;
    size_t expect = 0;
    for (int p = 0; p < n_passes; ++p) {
        const uint32_t pw = (w - passes[p][0] + passes[p][2] - 1) / passes[p][2];
        const uint32_t ph = (h - passes[p][1] + passes[p][3] - 1) / passes[p][3];
        if (w > passes[p][0] && h > passes[p][1]) expect += ((size_t)pw * n * depth + 7) / 8 * ph + ph;
    }
    std::vector<uint8_t> raw;
    std::string e = inflate_zlib(idat.data(), idat.size(), raw, expect);
    if (!e.empty()) return e;
    if (raw.size() < expect) return "PNG image data too short";
    // samples at the file's depth (grey of 1/2/4 bits scaled to 0..255)
    std::vector<uint16_t> smp((size_t)w * h * n);
    static const uint16_t gscale[5] = {0, 0xff, 0x55, 0, 0x11};
    size_t rp = 0;
    for (int p = 0; p < n_passes; ++p) {
        if (w <= passes[p][0] || h <= passes[p][1]) continue;
        const uint32_t pw = (w - passes[p][0] + passes[p][2] - 1) / passes[p][2];
        const uint32_t ph = (h - passes[p][1] + passes[p][3] - 1) / passes[p][3];
        const size_t rb = ((size_t)pw * n * depth + 7) / 8;
        std::vector<uint8_t> prev(rb, 0), cur(rb);
        for (uint32_t y = 0; y < ph; ++y) {
            const uint8_t ft = raw[rp++];
            const uint8_t* in = &raw[rp];
            rp += rb;
            for (size_t i = 0; i < rb; ++i) {
                const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
                int v = in[i];
                switch (ft) {
                    case 0: break;
                    case 1: v += a; break;
                    case 2: v += b; break;
                    case 3: v += (a + b) >> 1; break;
                    case 4: v += paeth(a, b, c); break;
                    default: return "bad PNG filter type";
                }
                cur[i] = (uint8_t)v;
            }
            const uint32_t yy = passes[p][1] + y * passes[p][3];
            for (uint32_t x = 0; x < pw; ++x) {
                const uint32_t xx = passes[p][0] + x * passes[p][2];
                uint16_t* o = &smp[((size_t)yy * w + xx) * n];
                for (uint32_t k = 0; k < n; ++k) {
                    const size_t s = (size_t)x * n + k;  // sample index in the row
                    uint16_t v;
                    if (depth == 16) {
                        v = (uint16_t)((cur[2 * s] << 8) | cur[2 * s + 1]);
                    } else if (depth == 8) {
                        v = cur[s];
                    } else {
                        const size_t bit = s * depth;
                        v = (uint16_t)((cur[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1));
                        if (color != 3) v = (uint16_t)(v * gscale[depth]);
                    }
                    o[k] = v;
                }
            }
            std::swap(prev, cur);
        }
    }
    // to RGBA8
    rgba.assign((size_t)w * h * 4, 0);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint16_t* s = &smp[i * n];
        uint8_t* o = &rgba[i * 4];
        auto b8 = [&](uint16_t v) { return (uint8_t)(depth == 16 ? v >> 8 : v); };
        if (color == 3) {
            const uint8_t* e8 = s[0] < pal_len ? palette[s[0]] : nullptr;
            o[0] = e8 ? e8[0] : 0;
            o[1] = e8 ? e8[1] : 0;
            o[2] = e8 ? e8[2] : 0;
            o[3] = e8 && pal_alpha ? e8[3] : 255;
            continue;
        }
        const bool rgb = color & 2;
        o[0] = b8(s[0]);
        o[1] = b8(rgb ? s[1] : s[0]);
        o[2] = b8(rgb ? s[2] : s[0]);
        if (color & 4) o[3] = b8(s[rgb ? 3 : 1]);
        else if (has_key && s[0] == key[0] && (!rgb || (s[1] == key[1] && s[2] == key[2]))) o[3] = 0;
        else o[3] = 255;
    }
    return "";
}

// ---------------------------------------------------------------- RGBE ----

void rgbe_texel(const uint8_t* e, float* o) {
    if (e[3]) {
        const float f = std::ldexp(1.0f, (int)e[3] - 136);
        o[0] = e[0] * f;
        o[1] = e[1] * f;
        o[2] = e[2] * f;
    } else {
        o[0] = o[1] = o[2] = 0.0f;
    }
    o[3] = 1.0f;
}

bool is_rgbe(const std::vector<uint8_t>& f) {
    auto starts = [&](const char* s) {
        const size_t k = std::strlen(s);
        return f.size() >= k && std::memcmp(f.data(), s, k) == 0;
    };
    return starts("#?RADIANCE\n") || starts("#?RGBE\n");
}

std::string decode_rgbe(const std::vector<uint8_t>& f, uint32_t& w, uint32_t& h, std::vector<float>& rgba) {
    size_t pos = 0;
    auto line = [&]() {
        std::string s;
        while (pos < f.size() && f[pos] != '\n') s.push_back((char)f[pos++]);
        if (pos < f.size()) ++pos;
        return s;
    };
    line();  // signature
    bool fmt = false;
    for (;;) {
        const std::string s = line();
        if (s.empty()) break;
        if (s == "FORMAT=32-bit_rle_rgbe") fmt = true;
        if (pos >= f.size()) return "RGBE header has no end";
    }
    if (!fmt) return "RGBE file is not FORMAT=32-bit_rle_rgbe";
    const std::string dims = line();
    long hh = 0, ww = 0;
    char* end = nullptr;
    if (dims.compare(0, 3, "-Y ") != 0) return "RGBE layout other than -Y h +X w";
    hh = std::strtol(dims.c_str() + 3, &end, 10);
    while (*end == ' ') ++end;
    if (std::strncmp(end, "+X ", 3) != 0) return "RGBE layout other than -Y h +X w";
    ww = std::strtol(end + 3, nullptr, 10);
    if (ww <= 0 || hh <= 0 || ww > (1 << 24) || hh > (1 << 24) || (uint64_t)ww * hh > (1ull << 26)) return "RGBE image empty or larger than 2^26 pixels";
    w = (uint32_t)ww;
    h = (uint32_t)hh;
    rgba.assign((size_t)w * h * 4, 0.0f);
    auto flat = [&](size_t from_texel) -> std::string {
        for (size_t t = from_texel; t < (size_t)w * h; ++t) {
            if (pos + 4 > f.size()) return "RGBE data ends early";
            rgbe_texel(&f[pos], &rgba[4 * t]);
            pos += 4;
        }
        return "";
    };
    if (w < 8 || w >= 32768) return flat(0);
    std::vector<uint8_t> scan((size_t)w * 4);
    for (uint32_t y = 0; y < h; ++y) {
        if (pos + 4 > f.size()) return "RGBE data ends early";
        if (f[pos] != 2 || f[pos + 1] != 2 || (f[pos + 2] & 0x80)) {
            // a scanline that is not run-length encoded: the image is flat
            if (y == 0) return flat(0);
            return "RGBE file mixes flat and run-length scanlines";
        }
        if ((((uint32_t)f[pos + 2] << 8) | f[pos + 3]) != w) return "RGBE scanline length differs from the width";
        pos += 4;
        for (int k = 0; k < 4; ++k) {
            uint32_t x = 0;
            while (x < w) {
                if (pos >= f.size()) return "RGBE data ends early";
                uint32_t count = f[pos++];
                if (count > 128) {
                    count -= 128;
                    if (count > w - x || pos >= f.size()) return "bad RGBE run";
                    const uint8_t v = f[pos++];
                    for (uint32_t z = 0; z < count; ++z) scan[4 * (x++) + k] = v;
                } else {
                    if (count == 0 || count > w - x || pos + count > f.size()) return "bad RGBE run";
                    for (uint32_t z = 0; z < count; ++z) scan[4 * (x++) + k] = f[pos++];
                }
            }
        }
        for (uint32_t x = 0; x < w; ++x) rgbe_texel(&scan[4 * x], &rgba[4 * ((size_t)y * w + x)]);
    }
    return "";
}

}  // namespace

std::string load_rgba8(const std::string& path, uint32_t& w, uint32_t& h, std::vector<uint8_t>& out) {
    std::vector<uint8_t> f;
    if (!read_bytes(path, f)) return "cannot read " + path;
    if (is_rgbe(f)) return path + ": a Radiance image where 8-bit texels are read (a cube map face)";
    const std::string e = decode_png(f, w, h, out);
    return e.empty() ? e : path + ": " + e;
}

std::string load_rgba32f(const std::string& path, uint32_t& w, uint32_t& h, std::vector<float>& out) {
    std::vector<uint8_t> f;
    if (!read_bytes(path, f)) return "cannot read " + path;
    if (is_rgbe(f)) {
        const std::string e = decode_rgbe(f, w, h, out);
        return e.empty() ? e : path + ": " + e;
    }
    std::vector<uint8_t> px;
    const std::string e = decode_png(f, w, h, px);
    if (!e.empty()) return path + ": " + e;
    out.resize(px.size());
    for (size_t i = 0; i < px.size(); ++i)
        out[i] = (i & 3) == 3 ? px[i] / 255.0f : std::pow(px[i] / 255.0f, 2.2f);
    return "";
}

}  // namespace pico_img
