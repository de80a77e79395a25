// Reader of NumPy .npy files (format versions 1.0, 2.0 and 3.0): the on-disk format of the
// reference's platform directories (read there with cnpy, PlatformData.cpp:4,26).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace sarbp_host {

struct NpyArray {
    char kind = 0;              // 'f' float, 'c' complex, 'i' signed int, 'u' unsigned int, 'b' bool
    size_t word_size = 0;       // bytes per element
    std::vector<size_t> shape;  // empty for a 0-d array
    std::vector<uint8_t> bytes;

    size_t num_vals() const {
        size_t n = 1;
        for (size_t s : shape) n *= s;
        return n;
    }
    template <typename T>
    const T *data() const {
        return reinterpret_cast<const T *>(bytes.data());
    }
};

// Throws std::runtime_error naming the file on a missing file, a malformed header, a big-endian or
// Fortran-ordered array, or a short data section.
NpyArray npy_load(const std::string &path);

bool file_exists(const std::string &path);

// Message returned by sarbp_host_last_error() (per thread).
void set_error(const std::string &msg);

}  // namespace sarbp_host
