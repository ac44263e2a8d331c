// Film output: the 8-bit sRGB conversion of film::write (cameras/film.cpp:114-153,
// utilities.cpp:33-38) with a PNG writer of our own (zlib deflate, RGBA8 as
// file_system::write_image writes it, file_system.cpp:17-24), and a raw FP32
// dump (PFM) of the film's pixel values for error measurements.
#include "image.hpp"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "host.hpp"

namespace rbw {
namespace image {

using host::rbw_error;

// utilities.cpp:33-38; the float overload of pow, as the reference's float call resolves
float gamma_correct(float v) {
    if (v <= 0.0031308f) return 12.92f * v;
    return 1.055f * std::pow(v, 1.0f / 2.4f) - 0.055f;
}

// film.cpp:118-124: clamp(255 * gamma + 0.5, 0, 255) in the order of glm::clamp
// (max, then min; a NaN passes through both), then the conversion to a byte.  A
// NaN converts through int32 as x86's truncation does (0x80000000), giving byte 0.
uint8_t to_byte(float v) {
    float x = 255.0f * gamma_correct(v) + 0.5f;
    x = (x < 0.0f) ? 0.0f : x;
    x = (255.0f < x) ? 255.0f : x;
    if (std::isnan(x)) return 0;
    return static_cast<uint8_t>(static_cast<int32_t>(x));
}

void film_to_rgba8(const float* rgb, uint32_t width, uint32_t height, const int32_t* bmin, const int32_t* bmax,
                   float scale, uint8_t* out) {
    const int64_t x0 = bmin ? bmin[0] : 0, y0 = bmin ? bmin[1] : 0;
    const int64_t x1 = bmax ? bmax[0] : int64_t(width), y1 = bmax ? bmax[1] : int64_t(height);
    const size_t n = size_t(width) * height;
    std::memset(out, 255, n * 4);
    for (size_t i = 0; i < n; i++) {
        const int64_t x = int64_t(i % width), y = int64_t(i / width);
        if (x < x0 || x >= x1 || y < y0 || y >= y1) continue;
        for (int c = 0; c < 3; c++) out[i * 4 + c] = to_byte(rgb[i * 3 + c] * scale);
        out[i * 4 + 3] = 255;
    }
}

namespace {

struct file_closer {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};
using file_ptr = std::unique_ptr<FILE, file_closer>;

file_ptr open_for_write(const std::string& path) {
    file_ptr f(std::fopen(path.c_str(), "wb"));
    if (!f) throw rbw_error(RBW_ERR_INVALID_ARGUMENT, "cannot open " + path + " for writing");
    return f;
}

void put(FILE* f, const void* p, size_t n, const std::string& path) {
    if (n && std::fwrite(p, 1, n, f) != n) throw rbw_error(RBW_ERR_DEVICE, "write to " + path + " failed");
}

void be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back(uint8_t(x >> 24));
    v.push_back(uint8_t(x >> 16));
    v.push_back(uint8_t(x >> 8));
    v.push_back(uint8_t(x));
}

// one PNG chunk: length, type, data, CRC-32 of type and data
void chunk(std::vector<uint8_t>& png, const char type[4], const uint8_t* data, size_t n) {
    be32(png, uint32_t(n));
    const size_t at = png.size();
    png.insert(png.end(), type, type + 4);
    png.insert(png.end(), data, data + n);
    uLong crc = crc32(0L, Z_NULL, 0);
    crc = crc32(crc, png.data() + at, uInt(n + 4));
    be32(png, uint32_t(crc));
}

}  // namespace

void write_png(const std::string& path, const uint8_t* rgba, uint32_t width, uint32_t height) {
    if (width == 0 || height == 0) throw rbw_error(RBW_ERR_INVALID_ARGUMENT, "empty image");
    // scanlines with filter type 0 (none), zlib-compressed into one IDAT chunk
    const size_t row = size_t(width) * 4;
    std::vector<uint8_t> raw((row + 1) * height);
    for (uint32_t y = 0; y < height; y++) {
        raw[y * (row + 1)] = 0;
        std::memcpy(&raw[y * (row + 1) + 1], rgba + y * row, row);
    }
    uLongf zlen = compressBound(uLong(raw.size()));
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK)
        throw rbw_error(RBW_ERR_DEVICE, "zlib compression failed");
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    be32(ihdr, width);
    be32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8 bits, RGBA, deflate, adaptive filtering, no interlace
    chunk(png, "IHDR", ihdr.data(), ihdr.size());
    chunk(png, "IDAT", z.data(), zlen);
    chunk(png, "IEND", nullptr, 0);
    file_ptr f = open_for_write(path);
    put(f.get(), png.data(), png.size(), path);
}

void write_pfm(const std::string& path, const float* rgb, uint32_t width, uint32_t height) {
    if (width == 0 || height == 0) throw rbw_error(RBW_ERR_INVALID_ARGUMENT, "empty image");
    file_ptr f = open_for_write(path);
    // colour PFM, little-endian (negative scale), rows from the bottom up
    char header[64];
    const int n = std::snprintf(header, sizeof(header), "PF\n%u %u\n-1.0\n", width, height);
    put(f.get(), header, size_t(n), path);
    for (uint32_t y = height; y-- > 0;) put(f.get(), rgb + size_t(y) * width * 3, size_t(width) * 3 * sizeof(float), path);
}

}  // namespace image
}  // namespace rbw
