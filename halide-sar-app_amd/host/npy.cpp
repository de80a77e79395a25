// NumPy .npy reader. Layout of the format: the 6-byte magic "\x93NUMPY", a major and a minor
// version byte, the header length (u16 little-endian for version 1, u32 for versions 2 and 3), then
// an ASCII Python dict literal with keys 'descr', 'fortran_order' and 'shape', padded with spaces
// and a newline, then the raw array data.
#include "npy.hpp"

#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace sarbp_host {

namespace {

[[noreturn]] void fail(const std::string &path, const std::string &what) {
    throw std::runtime_error("npy " + path + ": " + what);
}

// The value text that follows 'key': in the header dict.
std::string dict_value(const std::string &hdr, const std::string &key, const std::string &path) {
    const std::string k = "'" + key + "'";
    size_t at = hdr.find(k);
    if (at == std::string::npos) fail(path, "header has no " + k);
    at = hdr.find(':', at + k.size());
    if (at == std::string::npos) fail(path, "malformed header at " + k);
    ++at;
    while (at < hdr.size() && hdr[at] == ' ') ++at;
    return hdr.substr(at);
}

}  // namespace

bool file_exists(const std::string &path) {
    if (FILE *f = std::fopen(path.c_str(), "rb")) {
        std::fclose(f);
        return true;
    }
    return false;
}

NpyArray npy_load(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) fail(path, "cannot open");
    struct Closer {
        FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{f};

    unsigned char pre[8];
    if (std::fread(pre, 1, 8, f) != 8 || std::memcmp(pre, "\x93NUMPY", 6) != 0)
        fail(path, "not an npy file");
    const int major = pre[6];
    size_t hlen = 0;
    if (major == 1) {
        unsigned char b[2];
        if (std::fread(b, 1, 2, f) != 2) fail(path, "truncated header");
        hlen = (size_t)b[0] | ((size_t)b[1] << 8);
    } else if (major == 2 || major == 3) {
        unsigned char b[4];
        if (std::fread(b, 1, 4, f) != 4) fail(path, "truncated header");
        hlen = (size_t)b[0] | ((size_t)b[1] << 8) | ((size_t)b[2] << 16) | ((size_t)b[3] << 24);
    } else {
        fail(path, "unsupported format version " + std::to_string(major));
    }
    std::string hdr(hlen, '\0');
    if (hlen && std::fread(&hdr[0], 1, hlen, f) != hlen) fail(path, "truncated header");

    NpyArray a;
    // 'descr': '<f4' (byte order, kind, size)
    std::string d = dict_value(hdr, "descr", path);
    if (d.empty() || (d[0] != '\'' && d[0] != '"')) fail(path, "descr is not a simple type string");
    const size_t end = d.find(d[0], 1);
    if (end == std::string::npos || end < 3) fail(path, "malformed descr");
    const std::string descr = d.substr(1, end - 1);
    const char order = descr[0];
    if (order == '>') fail(path, "big-endian data is not supported");
    if (order != '<' && order != '|' && order != '=') fail(path, "unsupported descr " + descr);
    a.kind = descr[1];
    if (std::strchr("fciub", a.kind) == nullptr) fail(path, "unsupported element type " + descr);
    a.word_size = (size_t)std::strtoul(descr.c_str() + 2, nullptr, 10);
    if (a.word_size == 0) fail(path, "unsupported element size " + descr);

    const std::string fo = dict_value(hdr, "fortran_order", path);
    const bool fortran = fo.compare(0, 4, "True") == 0;

    const std::string sh = dict_value(hdr, "shape", path);
    if (sh.empty() || sh[0] != '(') fail(path, "malformed shape");
    const size_t close = sh.find(')');
    if (close == std::string::npos) fail(path, "malformed shape");
    const std::string dims = sh.substr(1, close - 1);
    size_t pos = 0;
    while (pos < dims.size()) {
        while (pos < dims.size() && (dims[pos] == ' ' || dims[pos] == ',')) ++pos;
        if (pos >= dims.size()) break;
        char *e = nullptr;
        const unsigned long long v = std::strtoull(dims.c_str() + pos, &e, 10);
        if (e == dims.c_str() + pos) fail(path, "malformed shape");
        a.shape.push_back((size_t)v);
        pos = (size_t)(e - dims.c_str());
    }
    if (fortran && a.shape.size() > 1) fail(path, "Fortran-ordered arrays are not supported");

    const size_t nbytes = a.num_vals() * a.word_size;
    a.bytes.resize(nbytes);
    if (nbytes && std::fread(a.bytes.data(), 1, nbytes, f) != nbytes) fail(path, "truncated data");
    return a;
}

}  // namespace sarbp_host
