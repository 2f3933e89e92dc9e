// OpenEXR writer for raw-XYZ films (counterpart of write_exr,
// src/exr_output.cpp:62-120): converts the XYZ film to linear sRGB, writes an
// uncompressed scanline EXR (FLOAT or HALF channels B, G, R) with string
// metadata, and publishes it atomically through a sibling temporary file and
// rename(2).  Format per the OpenEXR file-layout documentation: magic
// 20000630, version 2, attribute header, scanline offset table, then one
// block per scanline (y, byte count, channels in alphabetical order).
#include "rp_internal.hpp"

#include <cerrno>
#include <cmath>
#include <cstring>
#include <fcntl.h>
#include <string>
#include <unistd.h>
#include <vector>

namespace rp {
namespace {

struct Bytes {
    std::vector<uint8_t> b;
    void u8(uint8_t v) { b.push_back(v); }
    void u32(uint32_t v) {
        for (int i = 0; i < 4; ++i) b.push_back(uint8_t(v >> (8 * i)));
    }
    void i32(int32_t v) { u32(uint32_t(v)); }
    void u64(uint64_t v) {
        for (int i = 0; i < 8; ++i) b.push_back(uint8_t(v >> (8 * i)));
    }
    void f32(float v) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        u32(u);
    }
    void str(const std::string &s) {
        b.insert(b.end(), s.begin(), s.end());
        b.push_back(0);
    }
    void attr(const std::string &name, const std::string &type, const Bytes &value) {
        str(name);
        str(type);
        u32(uint32_t(value.b.size()));
        b.insert(b.end(), value.b.begin(), value.b.end());
    }
};

// IEEE binary16 with round-to-nearest-even (values already range-checked).
uint16_t to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    int32_t exp = int32_t((x >> 23) & 0xff) - 127 + 15;
    uint32_t mant = x & 0x7fffffu;
    if (((x >> 23) & 0xff) == 0xff) return uint16_t(sign | 0x7c00u | (mant ? 0x200u : 0));
    if (exp >= 31) return uint16_t(sign | 0x7c00u);
    if (exp <= 0) {
        if (exp < -10) return uint16_t(sign);
        mant |= 0x800000u;
        uint32_t shift = uint32_t(14 - exp);
        uint32_t half = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1), mid = 1u << (shift - 1);
        if (rem > mid || (rem == mid && (half & 1))) ++half;
        return uint16_t(sign | half);
    }
    uint32_t half = sign | (uint32_t(exp) << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (half & 1))) ++half;
    return uint16_t(half);
}

} // namespace

std::string write_exr_file(const char *path, uint32_t width, uint32_t height, const float *xyz, double exposure,
                           bool half, const char *const *keys, const char *const *values, uint32_t nmeta) {
    if (!path || !*path) return "invalid EXR output path";
    if (width == 0 || height == 0 || !xyz) return "invalid EXR film dimensions";
    if (!std::isfinite(exposure)) return "cannot write a film with nonfinite exposure";
    const size_t npx = size_t(width) * height;
    std::vector<float> rgb(3 * npx);
    for (size_t i = 0; i < npx; ++i) {
        double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
            return "cannot write a film with nonfinite XYZ values";
        // XYZ -> linear sRGB (D65), src/render.cpp:2282-2296
        double r = exposure * (3.240444291419 * x - 1.537133830545 * y - 0.498529890988 * z);
        double g = exposure * (-0.969266752560 * x + 1.876012242980 * y + 0.041556048487 * z);
        double b = exposure * (0.055643519672 * x - 0.204026238798 * y + 1.057226873771 * z);
        for (double v : {r, g, b}) {
            if (!std::isfinite(v)) return "cannot write a film with nonfinite linear RGB values";
            if (std::fabs(v) > 3.4028234663852886e38) return "EXR FLOAT output would overflow";
            if (half && std::fabs(v) > 65504.0) return "EXR HALF output would overflow";
        }
        rgb[3 * i] = float(r), rgb[3 * i + 1] = float(g), rgb[3 * i + 2] = float(b);
    }
    const uint32_t pixel_type = half ? 1u : 2u, bpc = half ? 2u : 4u;
    // Attribute names are 1..255 bytes; names above 31 bytes need the
    // long-names flag (bit 10 of the version word).  Attributes the writer
    // owns, or that give the file another layout, are reserved.
    static const char *const reserved[] = {"channels",   "compression",        "dataWindow",        "displayWindow",
                                           "lineOrder",  "pixelAspectRatio",   "screenWindowCenter", "screenWindowWidth",
                                           "tiles",      "type",               "name",              "version",
                                           "chunkCount", "maxSamplesPerPixel"};
    bool long_names = false;
    for (uint32_t m = 0; m < nmeta; ++m) {
        if (!keys || !values || !keys[m] || !values[m]) return "invalid EXR metadata entry";
        size_t len = std::strlen(keys[m]);
        if (len == 0 || len > 255) return "EXR metadata names must be 1 to 255 bytes";
        long_names |= len > 31;
        for (const char *r : reserved)
            if (std::strcmp(keys[m], r) == 0) return std::string("EXR metadata name '") + r + "' is reserved";
        for (uint32_t k = 0; k < m; ++k)
            if (std::strcmp(keys[k], keys[m]) == 0) return "duplicate EXR metadata name";
    }
    Bytes h;
    h.u32(20000630u);
    h.u32(2u | (long_names ? 0x400u : 0u));
    Bytes ch;
    for (const char *c : {"B", "G", "R"}) {
        ch.str(c);
        ch.u32(pixel_type);
        ch.u8(0), ch.u8(0), ch.u8(0), ch.u8(0); // pLinear + reserved
        ch.i32(1), ch.i32(1);                    // sampling
    }
    ch.u8(0);
    h.attr("channels", "chlist", ch);
    Bytes comp;
    comp.u8(0); // NO_COMPRESSION
    h.attr("compression", "compression", comp);
    Bytes win;
    win.i32(0), win.i32(0), win.i32(int32_t(width - 1)), win.i32(int32_t(height - 1));
    h.attr("dataWindow", "box2i", win);
    h.attr("displayWindow", "box2i", win);
    Bytes lo;
    lo.u8(0); // INCREASING_Y
    h.attr("lineOrder", "lineOrder", lo);
    Bytes par;
    par.f32(1.0f);
    h.attr("pixelAspectRatio", "float", par);
    Bytes swc;
    swc.f32(0.0f), swc.f32(0.0f);
    h.attr("screenWindowCenter", "v2f", swc);
    Bytes sww;
    sww.f32(1.0f);
    h.attr("screenWindowWidth", "float", sww);
    for (uint32_t m = 0; m < nmeta; ++m) {
        Bytes v;
        v.b.assign(values[m], values[m] + std::strlen(values[m]));
        h.attr(keys[m], "string", v);
    }
    h.u8(0); // end of header
    const uint64_t line_bytes = uint64_t(width) * 3 * bpc;
    uint64_t offset = h.b.size() + 8ull * height;
    for (uint32_t y = 0; y < height; ++y) {
        h.u64(offset);
        offset += 8 + line_bytes;
    }
    for (uint32_t y = 0; y < height; ++y) {
        h.i32(int32_t(y));
        h.u32(uint32_t(line_bytes));
        for (int c : {2, 1, 0}) { // B, G, R
            for (uint32_t x = 0; x < width; ++x) {
                float v = rgb[3 * (size_t(y) * width + x) + size_t(c)];
                if (half) {
                    uint16_t hv = to_half(v);
                    h.u8(uint8_t(hv)), h.u8(uint8_t(hv >> 8));
                } else {
                    h.f32(v);
                }
            }
        }
    }
    std::string target(path);
    auto slash = target.find_last_of('/');
    std::string dir = slash == std::string::npos ? "." : target.substr(0, slash);
    std::string tmpl = dir + "/.rypath-exr.XXXXXX";
    int fd = mkstemp(tmpl.data());
    if (fd < 0) return std::string("could not create sibling EXR temporary file: ") + std::strerror(errno);
    size_t done = 0;
    while (done < h.b.size()) {
        ssize_t w = ::write(fd, h.b.data() + done, h.b.size() - done);
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) {
            int e = errno;
            ::close(fd);
            ::unlink(tmpl.c_str());
            return std::string("could not write EXR temporary file: ") + std::strerror(e);
        }
        done += size_t(w);
    }
    if (::close(fd) != 0) {
        ::unlink(tmpl.c_str());
        return std::string("could not close EXR temporary file: ") + std::strerror(errno);
    }
    if (std::rename(tmpl.c_str(), path) != 0) {
        int e = errno;
        ::unlink(tmpl.c_str());
        return std::string("could not publish EXR image: ") + std::strerror(e);
    }
    return {};
}

} // namespace rp

extern "C" int rp_write_exr(const char *path, uint32_t width, uint32_t height, const float *xyz, double exposure,
                            uint32_t half, const char *const *keys, const char *const *values, uint32_t count) {
    std::string e = rp::write_exr_file(path, width, height, xyz, exposure, half != 0, keys, values, count);
    if (!e.empty()) {
        rp::set_error("rp_write_exr: " + e);
        return RP_FAILURE;
    }
    return RP_SUCCESS;
}
