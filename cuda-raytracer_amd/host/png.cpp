// png.cpp -- 8-bit RGB PNG writer (stored deflate blocks, CRC-32, Adler-32).
//
// The reference converts each channel with min(max(c, 0), 1) * 255 into an
// unsigned char (src/util/image.hpp:29-34, so NaN becomes 0) and writes the
// bytes with stb_image_write (src/util/image.cu:15-18). Only the byte
// conversion affects results; the encoder here is lossless and self-made.
#include "png.h"

#include <cstdio>
#include <fstream>
#include <stdexcept>

namespace koishi {

namespace {
uint32_t crc_table[256];
bool crc_ready = false;

void crc_init() {
    for (uint32_t n = 0; n < 256; ++n) {
        uint32_t c = n;
        for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
    crc_ready = true;
}

uint32_t crc32(const uint8_t *p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    if (!crc_ready) crc_init();
    for (size_t i = 0; i < n; ++i) c = crc_table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c;
}

void be32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back(uint8_t(x >> 24));
    v.push_back(uint8_t(x >> 16));
    v.push_back(uint8_t(x >> 8));
    v.push_back(uint8_t(x));
}

void chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data) {
    be32(out, uint32_t(data.size()));
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    be32(out, crc32(td.data(), td.size()) ^ 0xFFFFFFFFu);
}
}  // namespace

uint8_t to_byte(double c) {
    float f = float(c);
    float m = f > 0.f ? f : 0.f;  // max(c, 0) as `a > b ? a : b`: NaN -> 0
    m = m < 1.f ? m : 1.f;
    return uint8_t(m * 255.f);
}

std::vector<uint8_t> encode_png(uint32_t w, uint32_t h, const uint8_t *rgb) {
    std::vector<uint8_t> raw;
    raw.reserve(size_t(h) * (3 * w + 1));
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);  // filter: none
        raw.insert(raw.end(), rgb + size_t(y) * 3 * w, rgb + size_t(y + 1) * 3 * w);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    do {
        size_t n = std::min<size_t>(65535, raw.size() - pos);
        bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(uint8_t(n));
        z.push_back(uint8_t(n >> 8));
        z.push_back(uint8_t(~n));
        z.push_back(uint8_t(~n >> 8));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
        pos += n;
    } while (pos < raw.size());
    be32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    be32(ihdr, w);
    be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit, RGB, deflate, no filter set, no interlace
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    return out;
}

void write_png(const std::string &path, uint32_t w, uint32_t h, const uint8_t *rgb) {
    auto bytes = encode_png(w, h, rgb);
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::logic_error("cannot write '" + path + "'");
    f.write(reinterpret_cast<const char *>(bytes.data()), std::streamsize(bytes.size()));
}

}  // namespace koishi
