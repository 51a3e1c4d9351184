// Image decoding for texture maps named in scene files (the role of
// stbi_load(..., 4) in Texture::LoadTexture, Texture.cpp:12-24): every
// result is RGBA8, rows top to bottom, gray expanded to (g, g, g), missing
// alpha = 255.  Formats: PNG (8-bit gray / gray+alpha / RGB / RGBA / palette,
// non-interlaced) and binary PPM (P6, maxval 255).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lfh {

struct Image {
    int width = 0, height = 0;
    std::vector<uint8_t> rgba;   // width * height * 4
};

// Decodes an image file; false (and *err) if the file is missing, unreadable
// or in an unsupported format.
bool load_image(const std::string& path, Image* out, std::string* err);
// The same from bytes in memory.
bool decode_image(const uint8_t* data, size_t size, Image* out, std::string* err);
// zlib stream (RFC 1950) -> raw bytes (DEFLATE, RFC 1951).
bool zlib_inflate(const uint8_t* data, size_t size, std::vector<uint8_t>* out, std::string* err);

}  // namespace lfh
