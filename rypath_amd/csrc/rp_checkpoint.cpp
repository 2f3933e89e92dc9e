// Film checkpoints (include/rp_io.h): counterpart of write_checkpoint /
// read_checkpoint (src/checkpoint.hpp, src/checkpoint.cpp): a running film
// with its sample count, sampler settings, scene identity and provenance
// segments, published atomically (sibling temporary, fsync, rename, fsync of
// the directory) and verified on read by size, SHA-256 trailer and content
// checks.  The layout is this library's own (format 1), documented in
// include/rp_io.h.
#include "rp_internal.hpp"
#include "rp_sha256.hpp"
#include "../../include/rp_io.h"

#include <cerrno>
#include <cmath>
#include <cstring>
#include <fcntl.h>
#include <memory>
#include <stdexcept>
#include <sys/stat.h>
#include <unistd.h>

namespace rp {
namespace {

constexpr char kMagic[8] = {'R', 'P', 'A', 'F', 'I', 'L', 'M', '\n'};
constexpr uint32_t kFormat = 1;
constexpr uint64_t kMaxPixels = uint64_t(1) << 28; // 2^28 pixels = 12 GiB of planes
constexpr uint32_t kMaxSegments = 1024;
constexpr uint32_t kMaxName = 4096;

struct Writer {
    std::vector<uint8_t> b;
    void raw(const void *p, size_t n) {
        const uint8_t *c = static_cast<const uint8_t *>(p);
        b.insert(b.end(), c, c + n);
    }
    void u32(uint32_t v) {
        for (int i = 0; i < 4; ++i) b.push_back(uint8_t(v >> (8 * i)));
    }
    void u64(uint64_t v) {
        for (int i = 0; i < 8; ++i) b.push_back(uint8_t(v >> (8 * i)));
    }
};

struct Reader {
    const uint8_t *p;
    size_t n;
    void need(size_t k) const {
        if (k > n) throw std::runtime_error("truncated checkpoint");
    }
    uint32_t u32() {
        need(4);
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) v |= uint32_t(p[i]) << (8 * i);
        p += 4, n -= 4;
        return v;
    }
    uint64_t u64() {
        need(8);
        uint64_t v = 0;
        for (int i = 0; i < 8; ++i) v |= uint64_t(p[i]) << (8 * i);
        p += 8, n -= 8;
        return v;
    }
    void raw(void *out, size_t k) {
        need(k);
        std::memcpy(out, p, k);
        p += k, n -= k;
    }
};

std::string sys_error(const char *what) { return std::string(what) + ": " + std::strerror(errno); }

bool write_all(int fd, const uint8_t *p, size_t n) {
    while (n) {
        ssize_t w = ::write(fd, p, n);
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) {
            if (w == 0) errno = EIO;
            return false;
        }
        p += w, n -= size_t(w);
    }
    return true;
}

bool fsync_retry(int fd) {
    int r;
    do r = ::fsync(fd);
    while (r < 0 && errno == EINTR);
    return r == 0;
}

} // namespace

std::string validate_film(const rp_film &f) {
    const uint64_t npx = uint64_t(f.width) * f.height;
    if (npx == 0 || npx > kMaxPixels) return "invalid film dimensions";
    if (f.planes.size() != 6 * npx) return "film planes do not match its dimensions";
    if (f.randomization > RP_RANDOMIZE_OWEN) return "invalid film sampler randomization";
    if (f.segments.size() > kMaxSegments) return "film exceeds 1024 provenance segments";
    uint64_t end = 0;
    for (const auto &s : f.segments) {
        if (s.begin != end || s.end <= s.begin) return "film provenance segments are not contiguous";
        if (s.device.size() > kMaxName) return "film provenance name exceeds 4096 bytes";
        end = s.end;
    }
    if (end != f.samples) return "film provenance does not cover its samples";
    for (size_t i = 0; i < f.planes.size(); ++i) {
        double v = f.planes[i];
        if (!std::isfinite(v)) return "film holds nonfinite sums";
        if (f.samples == 0 && v != 0) return "empty film holds nonzero sums";
    }
    for (uint64_t c = 0; c < 3; ++c)
        for (uint64_t px = 0; px < npx; ++px)
            if (!std::isfinite(f.planes[2 * c * npx + px] + f.planes[(2 * c + 1) * npx + px]))
                return "film holds nonfinite sums";
    return {};
}

std::string write_checkpoint_file(const rp_film &f, const char *path) {
    if (!path || !*path) return "invalid checkpoint path";
    if (std::string e = validate_film(f); !e.empty()) return e;
    Writer w;
    w.raw(kMagic, 8);
    w.u32(kFormat);
    w.u32(f.width);
    w.u32(f.height);
    w.u32(f.randomization);
    w.u64(f.seed);
    w.u64(f.samples);
    w.raw(f.identity.data(), 32);
    w.u32(uint32_t(f.segments.size()));
    for (const auto &s : f.segments) {
        w.u64(s.begin);
        w.u64(s.end);
        w.u32(uint32_t(s.device.size()));
        w.raw(s.device.data(), s.device.size());
    }
    Sha256 hash;
    hash.update(w.b.data(), w.b.size());
    // planes are little-endian IEEE binary64 on every supported host
    static_assert(sizeof(double) == 8);
    const uint8_t *payload = reinterpret_cast<const uint8_t *>(f.planes.data());
    const size_t payload_bytes = f.planes.size() * sizeof(double);
    hash.update(payload, payload_bytes);
    const auto digest = hash.finish();

    std::string target(path);
    auto slash = target.find_last_of('/');
    std::string dir = slash == std::string::npos ? "." : (slash == 0 ? "/" : target.substr(0, slash));
    std::string tmp = dir + "/.rypath-checkpoint.XXXXXX";
    int fd = mkstemp(tmp.data());
    if (fd < 0) return sys_error("checkpoint temporary creation failed");
    auto cleanup = [&](const std::string &msg) {
        int e = errno;
        if (fd >= 0) ::close(fd);
        ::unlink(tmp.c_str());
        errno = e;
        return msg;
    };
    if (!write_all(fd, w.b.data(), w.b.size()) || !write_all(fd, payload, payload_bytes) ||
        !write_all(fd, digest.data(), digest.size()))
        return cleanup(sys_error("checkpoint write failed"));
    if (!fsync_retry(fd)) return cleanup(sys_error("checkpoint file sync failed"));
    int c = ::close(fd);
    fd = -1;
    if (c != 0) return cleanup(sys_error("checkpoint close failed"));
    int r;
    do r = std::rename(tmp.c_str(), path);
    while (r < 0 && errno == EINTR);
    if (r != 0) return cleanup(sys_error("checkpoint publication failed"));
    int dfd;
    do dfd = ::open(dir.c_str(), O_RDONLY | O_DIRECTORY);
    while (dfd < 0 && errno == EINTR);
    if (dfd < 0) return sys_error("checkpoint published but durability is uncertain");
    bool ok = fsync_retry(dfd);
    ::close(dfd);
    if (!ok) return sys_error("checkpoint published but durability is uncertain");
    return {};
}

std::unique_ptr<rp_film> read_checkpoint_file(const char *path) {
    if (!path || !*path) throw std::runtime_error("invalid checkpoint path");
    int fd;
    do fd = ::open(path, O_RDONLY);
    while (fd < 0 && errno == EINTR);
    if (fd < 0) throw std::runtime_error(sys_error("checkpoint open failed"));
    struct stat st {};
    if (fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {
        ::close(fd);
        throw std::runtime_error("checkpoint must be a regular file");
    }
    std::vector<uint8_t> data(size_t(st.st_size));
    size_t got = 0;
    while (got < data.size()) {
        ssize_t r = ::read(fd, data.data() + got, data.size() - got);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) break;
        got += size_t(r);
    }
    ::close(fd);
    if (got != data.size()) throw std::runtime_error("checkpoint read failed");
    if (data.size() < 8 + 32 || std::memcmp(data.data(), kMagic, 8) != 0)
        throw std::runtime_error("not a film checkpoint (bad magic)");
    Sha256 hash;
    hash.update(data.data(), data.size() - 32);
    if (hash.finish() != *reinterpret_cast<const std::array<uint8_t, 32> *>(data.data() + data.size() - 32))
        throw std::runtime_error("checkpoint checksum mismatch");
    Reader r{data.data() + 8, data.size() - 8 - 32};
    if (r.u32() != kFormat) throw std::runtime_error("unsupported checkpoint format (expected format 1)");
    auto f = std::make_unique<rp_film>();
    f->width = r.u32();
    f->height = r.u32();
    f->randomization = r.u32();
    f->seed = r.u64();
    f->samples = r.u64();
    r.raw(f->identity.data(), 32);
    uint32_t nseg = r.u32();
    if (nseg > kMaxSegments) throw std::runtime_error("checkpoint exceeds 1024 provenance segments");
    for (uint32_t i = 0; i < nseg; ++i) {
        rp_film::Segment s;
        s.begin = r.u64();
        s.end = r.u64();
        uint32_t len = r.u32();
        if (len > kMaxName) throw std::runtime_error("checkpoint provenance name exceeds 4096 bytes");
        s.device.resize(len);
        r.raw(s.device.data(), len);
        f->segments.push_back(std::move(s));
    }
    const uint64_t npx = uint64_t(f->width) * f->height;
    if (npx == 0 || npx > kMaxPixels) throw std::runtime_error("invalid checkpoint film dimensions");
    if (r.n != 6 * npx * sizeof(double)) throw std::runtime_error("checkpoint payload size mismatch");
    f->planes.resize(6 * npx);
    r.raw(f->planes.data(), r.n);
    if (std::string e = validate_film(*f); !e.empty()) throw std::runtime_error("invalid checkpoint: " + e);
    return f;
}

std::array<uint8_t, 32> scene_identity(const rp_scene_desc &d) {
    Sha256 h;
    auto put = [&](const void *p, size_t n) {
        if (n) h.update(p, n);
    };
    auto u32 = [&](uint32_t v) { put(&v, 4); };
    const char tag[] = "rypath_amd scene v1";
    put(tag, sizeof tag);
    u32(d.vertex_count), u32(d.triangle_count), u32(d.object_count), u32(d.material_count);
    u32(d.emission_count), u32(d.medium_count), u32(d.spot_count), u32(d.spectra_count);
    put(d.positions, 3 * size_t(d.vertex_count) * sizeof(float));
    put(d.indices, 3 * size_t(d.triangle_count) * sizeof(uint32_t));
    put(d.tri_object, size_t(d.triangle_count) * sizeof(uint32_t));
    put(d.objects, size_t(d.object_count) * sizeof(rp_object));
    put(d.materials, size_t(d.material_count) * sizeof(rp_material));
    put(d.emissions, size_t(d.emission_count) * sizeof(rp_emission));
    put(d.media, size_t(d.medium_count) * sizeof(rp_medium));
    put(d.spots, size_t(d.spot_count) * sizeof(rp_spot_light));
    put(d.spectra, size_t(d.spectra_count) * sizeof(float));
    u32(d.has_environment);
    put(&d.environment, sizeof d.environment);
    put(&d.camera, sizeof d.camera);
    u32(d.max_depth), u32(d.hero_wavelengths);
    return h.finish();
}

} // namespace rp

extern "C" {

int rp_film_create(uint32_t width, uint32_t height, rp_film **out) {
    if (!out) {
        rp::set_error("rp_film_create: null argument");
        return RP_INVALID;
    }
    *out = nullptr;
    const uint64_t npx = uint64_t(width) * height;
    if (npx == 0 || npx > rp::kMaxPixels) {
        rp::set_error("rp_film_create: invalid film dimensions");
        return RP_INVALID;
    }
    try {
        auto f = std::make_unique<rp_film>();
        f->width = width;
        f->height = height;
        f->planes.assign(6 * npx, 0.0);
        *out = f.release();
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        rp::set_error(std::string("rp_film_create: ") + e.what());
        return RP_FAILURE;
    }
}

int rp_film_get_info(const rp_film *f, rp_film_info *info) {
    if (!f || !info) {
        rp::set_error("rp_film_get_info: null argument");
        return RP_INVALID;
    }
    info->width = f->width;
    info->height = f->height;
    info->samples = f->samples;
    info->seed = f->seed;
    info->randomization = f->randomization;
    info->segments = uint32_t(f->segments.size());
    std::memcpy(info->scene_identity, f->identity.data(), 32);
    return RP_SUCCESS;
}

int rp_film_resolve(const rp_film *f, float *xyz) {
    if (!f || !xyz) {
        rp::set_error("rp_film_resolve: null argument");
        return RP_INVALID;
    }
    if (f->samples == 0) {
        rp::set_error("rp_film_resolve: the film holds no samples");
        return RP_INVALID;
    }
    const size_t npx = size_t(f->width) * f->height;
    for (size_t px = 0; px < npx; ++px)
        for (size_t c = 0; c < 3; ++c)
            xyz[3 * px + c] = float((f->planes[2 * c * npx + px] + f->planes[(2 * c + 1) * npx + px]) / double(f->samples));
    return RP_SUCCESS;
}

int rp_checkpoint_write(const rp_film *f, const char *path) {
    if (!f) {
        rp::set_error("rp_checkpoint_write: null argument");
        return RP_INVALID;
    }
    std::string e = rp::write_checkpoint_file(*f, path);
    if (!e.empty()) {
        rp::set_error("rp_checkpoint_write: " + e);
        return RP_FAILURE;
    }
    return RP_SUCCESS;
}

int rp_checkpoint_read(const char *path, rp_film **out) {
    if (!path || !out) {
        rp::set_error("rp_checkpoint_read: null argument");
        return RP_INVALID;
    }
    *out = nullptr;
    try {
        *out = rp::read_checkpoint_file(path).release();
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        rp::set_error(std::string("rp_checkpoint_read: ") + e.what());
        return RP_FAILURE;
    }
}

void rp_film_destroy(rp_film *f) { delete f; }

} // extern "C"
