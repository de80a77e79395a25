// 8-bit greyscale PNG writer: signature, IHDR (colour type 0, bit depth 8), one IDAT holding the
// zlib stream of the filtered rows (filter type 0 on every row), IEND; chunk CRCs by zlib's crc32.
#include <zlib.h>

#include <cstdio>
#include <string>
#include <vector>

#include "halide_buffer_compat.h"
#include "npy.hpp"
#include "sarbp_host.h"

namespace {

void put_u32(std::vector<unsigned char> &b, uint32_t v) {
    b.push_back((unsigned char)(v >> 24));
    b.push_back((unsigned char)(v >> 16));
    b.push_back((unsigned char)(v >> 8));
    b.push_back((unsigned char)v);
}

void chunk(std::vector<unsigned char> &out, const char *type, const unsigned char *data, size_t n) {
    put_u32(out, (uint32_t)n);
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data, data + n);
    put_u32(out, (uint32_t)crc32(0L, out.data() + at, (uInt)(n + 4)));
}

}  // namespace

extern "C" int sarbp_png_write_gray8(const char *path, const uint8_t *pixels, int32_t width,
                                     int32_t height, int64_t row_stride) {
    if (!path || !pixels) {
        sarbp_host::set_error("sarbp_png_write_gray8: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    if (width < 1 || height < 1 || row_stride < width) {
        sarbp_host::set_error("sarbp_png_write_gray8: bad image size or row stride");
        return halide_error_code_param_too_small;
    }
    std::vector<unsigned char> raw((size_t)height * ((size_t)width + 1));
    for (int32_t y = 0; y < height; ++y) {
        unsigned char *dst = raw.data() + (size_t)y * ((size_t)width + 1);
        dst[0] = 0;  // filter type None
        const uint8_t *src = pixels + (size_t)y * (size_t)row_stride;
        for (int32_t x = 0; x < width; ++x) dst[1 + x] = src[x];
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) {
        sarbp_host::set_error("sarbp_png_write_gray8: deflate failed");
        return halide_error_code_generic_error;
    }
    std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    put_u32(ihdr, (uint32_t)width);
    put_u32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 0, 0, 0, 0});  // depth 8, greyscale, deflate, adaptive, no interlace
    chunk(out, "IHDR", ihdr.data(), ihdr.size());
    chunk(out, "IDAT", z.data(), zlen);
    chunk(out, "IEND", nullptr, 0);
    FILE *f = std::fopen(path, "wb");
    if (!f) {
        sarbp_host::set_error(std::string("sarbp_png_write_gray8: cannot open ") + path);
        return halide_error_code_generic_error;
    }
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (std::fclose(f) != 0 || !ok) {
        sarbp_host::set_error(std::string("sarbp_png_write_gray8: write failed: ") + path);
        return halide_error_code_generic_error;
    }
    return 0;
}
