// Film output formats (host code): 8-bit sRGB RGBA as film::write produces it,
// PNG and PFM files.
#pragma once

#include <cstdint>
#include <string>

namespace rbw {
namespace image {

float gamma_correct(float v);
uint8_t to_byte(float v);
// bmin/bmax: the film's pixel bound (NULL: the whole film); pixels outside stay 255
void film_to_rgba8(const float* rgb, uint32_t width, uint32_t height, const int32_t* bmin, const int32_t* bmax,
                   float scale, uint8_t* out_rgba);
void write_png(const std::string& path, const uint8_t* rgba, uint32_t width, uint32_t height);
void write_pfm(const std::string& path, const float* rgb, uint32_t width, uint32_t height);

}  // namespace image
}  // namespace rbw
