// Image file output: the RGBA8 frames of krt_hdr_to_ldr written as PNG, the
// format the reference's capture path writes with stbi_write_png
// (bench.cu:226-233, source/capture.cpp snapshot: 4 channels, stride
// width x 4, rows in memory order).  Encoder: the PNG container
// (signature, IHDR, one IDAT, IEND, CRC-32 per chunk) around a zlib stream of
// filter-0 scanlines (zlib's compress2).
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
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

// ------------------------------------------------------------------ decoder
// PNG images to RGBA8, for material textures (what the reference's texture
// loader takes from stbi_load with 4 channels, source/image.cpp:96-118):
// every colour type (grey, RGB, palette, grey + alpha, RGBA) at every bit
// depth, Adam7 interlacing, tRNS transparency; 16-bit samples keep their
// high byte and low-depth grey is scaled to 0..255.  Rows come in file order.
namespace {

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

struct PngInfo {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> plte, trns, idat;
};

PngInfo png_parse(const std::string &path, bool header_only) {
    FILE *f = std::fopen(path.c_str(), "rb");
    KRT_REQUIRE(f != nullptr, KRT_ERR_INVALID, "cannot open " + path);
    std::vector<uint8_t> buf;
    uint8_t tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) {
        buf.insert(buf.end(), tmp, tmp + n);
        if (header_only && buf.size() >= 33) break;
    }
    std::fclose(f);
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    KRT_REQUIRE(buf.size() >= 33 && std::memcmp(buf.data(), sig, 8) == 0, KRT_ERR_INVALID, path + ": not a PNG file");
    PngInfo I;
    size_t at = 8;
    bool ended = false;
    while (at + 12 <= buf.size()) {
        const uint32_t len = be32(&buf[at]);
        KRT_REQUIRE(at + 12 + (size_t)len <= buf.size(), KRT_ERR_INVALID, path + ": truncated chunk");
        const uint8_t *type = &buf[at + 4], *data = &buf[at + 8];
        const uLong crc = crc32(crc32(0L, nullptr, 0), type, (uInt)(4 + len));
        KRT_REQUIRE(crc == be32(&buf[at + 8 + len]), KRT_ERR_INVALID, path + ": chunk CRC mismatch");
        if (!std::memcmp(type, "IHDR", 4)) {
            KRT_REQUIRE(len == 13, KRT_ERR_INVALID, path + ": bad IHDR");
            I.w = be32(data);
            I.h = be32(data + 4);
            I.depth = data[8];
            I.ctype = data[9];
            KRT_REQUIRE(data[10] == 0 && data[11] == 0 && data[12] <= 1, KRT_ERR_INVALID,
                        path + ": unknown compression, filter or interlace method");
            I.interlace = data[12];
            if (header_only) return I;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            I.plte.assign(data, data + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            I.trns.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            I.idat.insert(I.idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            ended = true;
            break;
        } else {
            KRT_REQUIRE((type[0] & 0x20) != 0, KRT_ERR_INVALID, path + ": unknown critical chunk");
        }
        at += 12 + (size_t)len;
    }
    KRT_REQUIRE(ended && I.w > 0, KRT_ERR_INVALID, path + ": missing IHDR or IEND");
    return I;
}

int png_channels(int ctype) {
    switch (ctype) {
    case 0: return 1;
    case 2: return 3;
    case 3: return 1;
    case 4: return 2;
    case 6: return 4;
    default: return 0;
    }
}

uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (uint8_t)((pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c));
}

// Reverses the scanline filters of one (sub)image of w x h in place.
void unfilter(uint8_t *p, uint32_t w, uint32_t h, size_t rowbytes, int bpp, const std::string &path) {
    std::vector<uint8_t> zero(rowbytes, 0);
    const uint8_t *prev = zero.data();
    for (uint32_t y = 0; y < h; y++) {
        uint8_t *row = p + y * (rowbytes + 1);
        const int ft = row[0];
        uint8_t *x = row + 1;
        for (size_t i = 0; i < rowbytes; i++) {
            const int a = i >= (size_t)bpp ? x[i - bpp] : 0, b = prev[i], c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            switch (ft) {
            case 0: break;
            case 1: x[i] = (uint8_t)(x[i] + a); break;
            case 2: x[i] = (uint8_t)(x[i] + b); break;
            case 3: x[i] = (uint8_t)(x[i] + ((a + b) >> 1)); break;
            case 4: x[i] = (uint8_t)(x[i] + paeth(a, b, c)); break;
            default: KRT_REQUIRE(false, KRT_ERR_INVALID, path + ": bad filter type");
            }
        }
        prev = x;
        (void)w;
    }
}

}  // namespace

void read_png_size(const std::string &path, uint32_t &w, uint32_t &h) {
    const PngInfo I = png_parse(path, true);
    w = I.w;
    h = I.h;
}

void read_png(const std::string &path, std::vector<uint32_t> &rgba, uint32_t &w, uint32_t &h) {
    PngInfo I = png_parse(path, false);
    const int ch = png_channels(I.ctype), d = I.depth;
    KRT_REQUIRE(ch > 0, KRT_ERR_INVALID, path + ": unknown colour type");
    const bool ok_depth = (I.ctype == 0 && (d == 1 || d == 2 || d == 4 || d == 8 || d == 16)) ||
                          (I.ctype == 3 && (d == 1 || d == 2 || d == 4 || d == 8)) ||
                          ((I.ctype == 2 || I.ctype == 4 || I.ctype == 6) && (d == 8 || d == 16));
    KRT_REQUIRE(ok_depth, KRT_ERR_INVALID, path + ": bad bit depth for its colour type");
    KRT_REQUIRE(I.w <= 16384 && I.h <= 16384, KRT_ERR_INVALID, path + ": image larger than 16384 x 16384");
    KRT_REQUIRE(I.ctype != 3 || (!I.plte.empty() && I.plte.size() % 3 == 0), KRT_ERR_INVALID,
                path + ": palette image without PLTE");
    w = I.w;
    h = I.h;
    const int bits = ch * d, bpp = std::max(1, bits / 8);
    // passes: the whole image, or Adam7's seven
    static const int x0[7] = {0, 4, 0, 2, 0, 1, 0}, y0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const int dx[7] = {8, 8, 4, 4, 2, 2, 1}, dy[7] = {8, 8, 8, 4, 4, 2, 2};
    const int npass = I.interlace ? 7 : 1;
    uint32_t pw[7], ph[7];
    size_t total = 0;
    for (int p = 0; p < npass; p++) {
        pw[p] = I.interlace ? (I.w > (uint32_t)x0[p] ? (I.w - x0[p] + dx[p] - 1) / dx[p] : 0) : I.w;
        ph[p] = I.interlace ? (I.h > (uint32_t)y0[p] ? (I.h - y0[p] + dy[p] - 1) / dy[p] : 0) : I.h;
        if (pw[p] && ph[p]) total += (size_t)ph[p] * (1 + ((size_t)pw[p] * bits + 7) / 8);
    }
    std::vector<uint8_t> raw(total);
    uLongf got = (uLongf)total;
    KRT_REQUIRE(uncompress(raw.data(), &got, I.idat.data(), (uLong)I.idat.size()) == Z_OK && got == total,
                KRT_ERR_INVALID, path + ": bad image data");
    rgba.assign((size_t)I.w * I.h, 0u);
    const int maxv = (1 << (d == 16 ? 8 : d)) - 1;
    size_t off = 0;
    for (int p = 0; p < npass; p++) {
        if (!pw[p] || !ph[p]) continue;
        const size_t rowbytes = ((size_t)pw[p] * bits + 7) / 8;
        uint8_t *img = raw.data() + off;
        unfilter(img, pw[p], ph[p], rowbytes, bpp, path);
        for (uint32_t y = 0; y < ph[p]; y++) {
            const uint8_t *row = img + y * (rowbytes + 1) + 1;
            for (uint32_t x = 0; x < pw[p]; x++) {
                uint32_t s[4] = {0, 0, 0, 0}, raw16[4] = {0, 0, 0, 0};
                for (int c = 0; c < ch; c++) {
                    if (d == 16) {
                        const size_t at = ((size_t)x * ch + c) * 2;
                        raw16[c] = (uint32_t)row[at] << 8 | row[at + 1];
                        s[c] = row[at];
                    } else if (d == 8) {
                        s[c] = row[(size_t)x * ch + c];
                        raw16[c] = s[c];
                    } else {
                        const size_t bit = (size_t)x * d;
                        s[c] = (row[bit / 8] >> (8 - d - bit % 8)) & (uint32_t)maxv;
                        raw16[c] = s[c];
                    }
                }
                uint32_t r, g, b, a = 255;
                if (I.ctype == 3) {
                    KRT_REQUIRE(3 * (size_t)s[0] + 2 < I.plte.size(), KRT_ERR_INVALID, path + ": palette index out of range");
                    r = I.plte[3 * s[0]];
                    g = I.plte[3 * s[0] + 1];
                    b = I.plte[3 * s[0] + 2];
                    if (s[0] < I.trns.size()) a = I.trns[s[0]];
                } else if (I.ctype == 0 || I.ctype == 4) {
                    const uint32_t v = d < 8 ? s[0] * 255u / (uint32_t)maxv : s[0];
                    r = g = b = v;
                    if (I.ctype == 4) a = s[1];
                    else if (I.trns.size() >= 2 && raw16[0] == ((uint32_t)I.trns[0] << 8 | I.trns[1])) a = 0;
                } else {
                    r = s[0];
                    g = s[1];
                    b = s[2];
                    if (I.ctype == 6) a = s[3];
                    else if (I.trns.size() >= 6 && raw16[0] == ((uint32_t)I.trns[0] << 8 | I.trns[1]) &&
                             raw16[1] == ((uint32_t)I.trns[2] << 8 | I.trns[3]) &&
                             raw16[2] == ((uint32_t)I.trns[4] << 8 | I.trns[5]))
                        a = 0;
                }
                const size_t X = I.interlace ? x0[p] + (size_t)x * dx[p] : x;
                const size_t Y = I.interlace ? y0[p] + (size_t)y * dy[p] : y;
                rgba[Y * I.w + X] = r | g << 8 | b << 16 | a << 24;
            }
        }
        off += (size_t)ph[p] * (rowbytes + 1);
    }
}

}  // namespace krt
