// Image file output: the RGBA8 frames of krt_hdr_to_ldr written as PNG, the
// format the reference's capture path writes with stbi_write_png
// (bench.cu:226-233, source/capture.cpp snapshot: 4 channels, stride
// width x 4, rows in memory order).  Encoder: the PNG container
// (signature, IHDR, one IDAT, IEND, CRC-32 per chunk) around a zlib stream of
// filter-0 scanlines (zlib's compress2).
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "krt_internal.h"

namespace krt {

namespace {

void put_be32(std::vector<uint8_t> &out, uint32_t v) {
    out.push_back((uint8_t)(v >> 24));
    out.push_back((uint8_t)(v >> 16));
    out.push_back((uint8_t)(v >> 8));
    out.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t> &out, const char type[4], const uint8_t *data, size_t n) {
    put_be32(out, (uint32_t)n);
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    if (n) out.insert(out.end(), data, data + n);
    const uLong crc = crc32(0L, out.data() + at, (uInt)(4 + n));
    put_be32(out, (uint32_t)crc);
}

}  // namespace

void write_png(const std::string &path, const uint32_t *rgba, uint32_t width, uint32_t height) {
    KRT_REQUIRE(width > 0 && height > 0 && width <= (1u << 16) && height <= (1u << 16), KRT_ERR_INVALID,
                "image size out of range");
    // scanlines: a filter byte (0, none) then R, G, B, A of each pixel; the
    // packed uint32 is r | g << 8 | b << 16 | a << 24
    const size_t row = 1 + 4 * (size_t)width;
    std::vector<uint8_t> raw(row * height);
    for (uint32_t y = 0; y < height; y++) {
        uint8_t *dst = raw.data() + row * y;
        dst[0] = 0;
        for (uint32_t x = 0; x < width; x++) {
            const uint32_t p = rgba[(size_t)y * width + x];
            dst[1 + 4 * x] = (uint8_t)(p & 0xffu);
            dst[2 + 4 * x] = (uint8_t)((p >> 8) & 0xffu);
            dst[3 + 4 * x] = (uint8_t)((p >> 16) & 0xffu);
            dst[4 + 4 * x] = (uint8_t)((p >> 24) & 0xffu);
        }
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    KRT_REQUIRE(compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) == Z_OK, KRT_ERR_INVALID,
                "zlib compression failed");
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    uint8_t ihdr[13];
    const uint32_t wh[2] = {width, height};
    for (int k = 0; k < 2; k++)
        for (int b = 0; b < 4; b++) ihdr[4 * k + b] = (uint8_t)(wh[k] >> (24 - 8 * b));
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 6;   // colour type: RGBA
    ihdr[10] = 0;  // deflate
    ihdr[11] = 0;  // adaptive filtering (filter type 0 used on every row)
    ihdr[12] = 0;  // no interlace
    chunk(png, "IHDR", ihdr, sizeof(ihdr));
    chunk(png, "IDAT", z.data(), zlen);
    chunk(png, "IEND", nullptr, 0);
    FILE *f = std::fopen(path.c_str(), "wb");
    KRT_REQUIRE(f != nullptr, KRT_ERR_INVALID, "cannot open " + path + " for writing");
    const size_t wrote = std::fwrite(png.data(), 1, png.size(), f);
    const int closed = std::fclose(f);
    KRT_REQUIRE(wrote == png.size() && closed == 0, KRT_ERR_INVALID, "short write to " + path);
}

}  // namespace krt
