// png.h -- 8-bit RGB PNG output of the rendered image.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace koishi {

// Channel -> byte as Image<3>'s Component does: min(max(c, 0), 1) * 255,
// truncated (src/util/image.hpp:29-34).
uint8_t to_byte(double c);

std::vector<uint8_t> encode_png(uint32_t w, uint32_t h, const uint8_t *rgb);
void write_png(const std::string &path, uint32_t w, uint32_t h, const uint8_t *rgb);

}  // namespace koishi
