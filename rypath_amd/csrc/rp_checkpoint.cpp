// Film checkpoints (include/rp_io.h): counterpart of write_checkpoint /
// read_checkpoint (src/checkpoint.hpp, src/checkpoint.cpp).  Files are the
// reference's RYPCHK format 3 (src/checkpoint.cpp:17-122, contract words of
// src/checkpoint_contract.h:14): a 44-byte header, the render settings, film
// flags, scene identity and provenance segments as metadata, 216 bytes per
// pixel, and the SHA-256 of all preceding bytes.  They are published
// atomically (sibling temporary, fsync, rename, fsync of the directory) and
// verified on read by size, SHA-256 and content checks.  Format 1 (this
// library's earlier "RPAFILM" layout) is still read.
#include "rp_internal.hpp"
#include "rp_sha256.hpp"
#include "../../include/rp_io.h"

#include <bit>
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

// format 3 (the reference's current contract)
constexpr char kMagic3[8] = {'R', 'Y', 'P', 'C', 'H', 'K', '\r', '\n'};
constexpr uint32_t kContract[5] = {3, 12, 29, 11, 2}; // format, identity, renderer, sampling, accumulation
constexpr size_t kHeader3 = 44, kPixel3 = 216, kFixedMeta3 = 118;
constexpr uint64_t kMaxPayload3 = uint64_t(4) << 30, kMaxMeta3 = uint64_t(1) << 20;
// format 1 (read only)
constexpr char kMagic1[8] = {'R', 'P', 'A', 'F', 'I', 'L', 'M', '\n'};

constexpr uint64_t kMaxPixels = uint64_t(1) << 28; // 2^28 pixels = 12 GiB of planes
constexpr uint32_t kMaxSegments = 1024;
constexpr uint32_t kMaxName = 4096;

struct Writer {
    std::vector<uint8_t> b;
    void raw(const void *p, size_t n) {
        const uint8_t *c = static_cast<const uint8_t *>(p);
        b.insert(b.end(), c, c + n);
    }
    void uint(uint64_t v, int n) {
        for (int i = 0; i < n; ++i) b.push_back(uint8_t(v >> (8 * i)));
    }
    void u32(uint32_t v) { uint(v, 4); }
    void u64(uint64_t v) { uint(v, 8); }
    void f64(double v) { u64(std::bit_cast<uint64_t>(v)); }
    void f32(float v) { u32(std::bit_cast<uint32_t>(v)); }
    void str(const std::string &s) {
        u32(uint32_t(s.size()));
        raw(s.data(), s.size());
    }
};

struct Reader {
    const uint8_t *p;
    size_t n;
    void need(size_t k) const {
        if (k > n) throw std::runtime_error("truncated checkpoint metadata");
    }
    uint64_t uint(int k) {
        need(size_t(k));
        uint64_t v = 0;
        for (int i = 0; i < k; ++i) v |= uint64_t(p[i]) << (8 * i);
        p += k, n -= size_t(k);
        return v;
    }
    uint32_t u32() { return uint32_t(uint(4)); }
    uint64_t u64() { return uint(8); }
    double f64() { return std::bit_cast<double>(u64()); }
    float f32() { return std::bit_cast<float>(u32()); }
    void raw(void *out, size_t k) {
        need(k);
        std::memcpy(out, p, k);
        p += k, n -= k;
    }
    std::string str() {
        const uint32_t len = u32();
        if (len > kMaxName || len > n) throw std::runtime_error("invalid checkpoint string length");
        std::string v(reinterpret_cast<const char *>(p), len);
        p += len, n -= len;
        return v;
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

// The render settings a format-3 file records for the film: the job's
// sample budget and batch (the film's own count when the job set none or the
// film has grown past it).
uint64_t job_samples(const rp_film &f) { return std::max<uint64_t>(std::max(f.job_samples, f.samples), 1); }
uint64_t job_batch(const rp_film &f) { return f.batch_size ? f.batch_size : std::max<uint64_t>(f.samples, 1); }

// Format 3 metadata (src/checkpoint.cpp:85-107).
std::vector<uint8_t> metadata3(const rp_film &f) {
    Writer w;
    w.u32(f.width);
    w.u32(f.height);
    w.u64(job_samples(f));
    w.u64(job_batch(f));
    w.u64(f.seed);
    w.u32(f.sampler);
    w.u32(f.randomization);
    w.u32(uint32_t(f.material_seed));
    w.u32(f.model);
    w.f64(f.exposure);
    w.uint(f.packet ? 1 : 0, 1);
    w.uint(0, 1); // compatibility (clamped) output: never accumulated here
    w.f64(f.max_component);
    w.u32(0);     // box reconstruction filter,
    w.f32(0.5f);  // radius 0.5 x 0.5 (one pixel, unit weight),
    w.f32(0.5f);
    w.f32(0.5f);  // default sigma
    w.raw(f.identity.data(), 32);
    w.u32(uint32_t(f.segments.size()));
    for (const auto &s : f.segments) {
        w.u64(s.begin);
        w.u64(s.end);
        w.str(s.build);
        w.str(s.device);
        w.str(s.architecture);
        w.str(s.os);
    }
    return w.b;
}

// Format 3 pixel records (src/checkpoint.cpp:108-117): sum[3], correction[3],
// count, six zero splat and compatibility triples, reconstruction weight.
void pixels3(const rp_film &f, uint64_t first, uint64_t count, Writer &w) {
    const uint64_t npx = uint64_t(f.width) * f.height;
    const double weight = f.packet ? double(f.samples) : 0.0;
    for (uint64_t px = first; px < first + count; ++px) {
        for (int c = 0; c < 3; ++c) w.f64(f.planes[2 * c * npx + px]);
        for (int c = 0; c < 3; ++c) w.f64(f.planes[(2 * c + 1) * npx + px]);
        w.u64(f.samples);
        for (int k = 0; k < 18; ++k) w.f64(0.0);
        w.f64(weight);
        w.f64(0.0);
    }
}

std::string publish(const char *path, const std::vector<uint8_t> &head, const std::vector<uint8_t> &meta,
                    const rp_film &f) {
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
    Sha256 hash;
    auto emit = [&](const std::vector<uint8_t> &b) {
        hash.update(b.data(), b.size());
        return write_all(fd, b.data(), b.size());
    };
    if (!emit(head) || !emit(meta)) return cleanup(sys_error("checkpoint write failed"));
    const uint64_t npx = uint64_t(f.width) * f.height;
    for (uint64_t first = 0; first < npx; first += 1024) {
        Writer w;
        w.b.reserve(kPixel3 * 1024);
        pixels3(f, first, std::min<uint64_t>(1024, npx - first), w);
        if (!emit(w.b)) return cleanup(sys_error("checkpoint write failed"));
    }
    const auto digest = hash.finish();
    if (!write_all(fd, digest.data(), digest.size())) return cleanup(sys_error("checkpoint checksum write failed"));
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

std::vector<uint8_t> read_file(const char *path) {
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
    return data;
}

// Format 3 (src/checkpoint.cpp:319-455), then the conditions under which its
// film carries over into this library's unit-weight box-filter film.
std::unique_ptr<rp_film> parse3(const std::vector<uint8_t> &data) {
    if (data.size() < kHeader3) throw std::runtime_error("invalid checkpoint header");
    Reader h{data.data() + 8, kHeader3 - 8};
    if (h.u32() != kContract[0]) throw std::runtime_error("unsupported checkpoint format version (expected format 3)");
    for (int i = 1; i < 5; ++i)
        if (h.u32() != kContract[i]) throw std::runtime_error("unsupported checkpoint contract version");
    const uint64_t meta_size = h.u64(), payload = h.u64();
    if (meta_size > kMaxMeta3 || payload > kMaxPayload3 || payload % kPixel3 != 0 ||
        uint64_t(data.size()) != kHeader3 + meta_size + payload + 32)
        throw std::runtime_error("invalid checkpoint length or size limit");
    Sha256 hash;
    hash.update(data.data(), data.size() - 32);
    if (hash.finish() != *reinterpret_cast<const std::array<uint8_t, 32> *>(data.data() + data.size() - 32))
        throw std::runtime_error("checkpoint checksum mismatch");
    Reader m{data.data() + kHeader3, size_t(meta_size)};
    auto f = std::make_unique<rp_film>();
    f->format = 3;
    f->width = m.u32();
    f->height = m.u32();
    f->job_samples = m.u64();
    f->batch_size = m.u64();
    f->seed = m.u64();
    f->sampler = m.u32();
    f->randomization = m.u32();
    f->material_seed = int32_t(m.u32());
    f->model = m.u32();
    f->exposure = m.f64();
    const uint64_t packet = m.uint(1), compatibility = m.uint(1);
    if (packet > 1 || compatibility > 1) throw std::runtime_error("invalid checkpoint film flags");
    f->packet = packet != 0;
    f->max_component = m.f64();
    const uint32_t filter = m.u32();
    const float rx = m.f32(), ry = m.f32(), sigma = m.f32();
    m.raw(f->identity.data(), 32);
    const uint32_t nseg = m.u32();
    if (nseg > kMaxSegments) throw std::runtime_error("checkpoint exceeds 1024 provenance segments");
    for (uint32_t i = 0; i < nseg; ++i) {
        rp_film::Segment s;
        s.begin = m.u64();
        s.end = m.u64();
        s.build = m.str();
        s.device = m.str();
        s.architecture = m.str();
        s.os = m.str();
        f->segments.push_back(std::move(s));
    }
    if (m.n != 0) throw std::runtime_error("trailing checkpoint metadata");
    // the reference's own validation (validate_render_settings, validate_checkpoint)
    const uint64_t npx = uint64_t(f->width) * f->height;
    if (npx == 0 || npx > 0xffffffffull) throw std::runtime_error("invalid checkpoint film dimensions");
    if (f->job_samples == 0 || f->batch_size == 0)
        throw std::runtime_error("render sample count and batch size must be positive");
    if (f->randomization > 3) throw std::runtime_error("invalid checkpoint sampler randomization");
    if (npx != payload / kPixel3) throw std::runtime_error("checkpoint film payload size mismatch");
    if ((f->model != 0 && f->model != 1) || !std::isfinite(f->exposure) || f->exposure < 0 ||
        (f->model == 0 && (f->exposure != 1 || f->packet)))
        throw std::runtime_error("invalid checkpoint film model or exposure");
    if (!(f->max_component > 0) || filter > 2 || !(rx > 0 && ry > 0) || !std::isfinite(rx) || !std::isfinite(ry) ||
        !std::isfinite(sigma) || (filter == 2 && !(sigma > 0)))
        throw std::runtime_error("invalid checkpoint film filter or clamp");
    // what this library's film can hold
    if (npx > kMaxPixels) throw std::runtime_error("checkpoint film exceeds 2^28 pixels");
    if (f->sampler != RP_SAMPLER_INDEPENDENT && f->sampler != RP_SAMPLER_SOBOL)
        throw std::runtime_error("checkpoint sampler (Halton or ZSobol) is not rendered by this library");
    if (filter != 0 || rx != 0.5f || ry != 0.5f)
        throw std::runtime_error("checkpoint film uses a reconstruction filter other than the one-pixel box");
    if (compatibility) throw std::runtime_error("checkpoint film selects clamped compatibility output");
    f->planes.assign(6 * npx, 0.0);
    Reader r{data.data() + kHeader3 + meta_size, size_t(payload)};
    for (uint64_t px = 0; px < npx; ++px) {
        double sum[3], corr[3], extra[18];
        for (double &v : sum) v = r.f64();
        for (double &v : corr) v = r.f64();
        const uint64_t count = r.u64();
        for (double &v : extra) v = r.f64();
        const double wsum = r.f64(), wcorr = r.f64();
        if (px == 0) f->samples = count;
        if (count != f->samples || count > f->job_samples) throw std::runtime_error("invalid checkpoint sample counts");
        for (int c = 0; c < 3; ++c) {
            if (!std::isfinite(sum[c]) || !std::isfinite(corr[c]) || !std::isfinite(sum[c] + corr[c]) ||
                (count == 0 && (sum[c] != 0 || corr[c] != 0)))
                throw std::runtime_error("invalid checkpoint accumulation");
            f->planes[2 * c * npx + px] = sum[c];
            f->planes[(2 * c + 1) * npx + px] = corr[c];
        }
        for (double v : extra)
            if (v != 0) throw std::runtime_error("checkpoint film holds splat or compatibility sums");
        const double expect = f->packet ? double(count) : 0.0;
        if (!std::isfinite(wsum) || !std::isfinite(wcorr) || wsum + wcorr != expect || wcorr != 0)
            throw std::runtime_error(f->packet ? "checkpoint reconstruction weights are not one per sample"
                                               : "invalid empty-film reconstruction weights");
    }
    return f;
}

std::unique_ptr<rp_film> parse1(const std::vector<uint8_t> &data) {
    Sha256 hash;
    hash.update(data.data(), data.size() - 32);
    if (hash.finish() != *reinterpret_cast<const std::array<uint8_t, 32> *>(data.data() + data.size() - 32))
        throw std::runtime_error("checkpoint checksum mismatch");
    Reader r{data.data() + 8, data.size() - 8 - 32};
    if (r.u32() != 1) throw std::runtime_error("unsupported checkpoint format (expected format 3 or 1)");
    auto f = std::make_unique<rp_film>();
    f->format = 1;
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
        s.device = r.str();
        f->segments.push_back(std::move(s));
    }
    const uint64_t npx = uint64_t(f->width) * f->height;
    if (npx == 0 || npx > kMaxPixels) throw std::runtime_error("invalid checkpoint film dimensions");
    if (r.n != 6 * npx * sizeof(double)) throw std::runtime_error("checkpoint payload size mismatch");
    f->planes.resize(6 * npx);
    r.raw(f->planes.data(), r.n);
    return f;
}

} // namespace

std::string validate_film(const rp_film &f) {
    const uint64_t npx = uint64_t(f.width) * f.height;
    if (npx == 0 || npx > kMaxPixels) return "invalid film dimensions";
    if (f.planes.size() != 6 * npx) return "film planes do not match its dimensions";
    if (f.randomization > RP_RANDOMIZE_OWEN) return "invalid film sampler randomization";
    if (f.sampler != RP_SAMPLER_INDEPENDENT && f.sampler != RP_SAMPLER_SOBOL) return "invalid film sampler";
    if (f.segments.size() > kMaxSegments) return "film exceeds 1024 provenance segments";
    uint64_t end = 0;
    for (const auto &s : f.segments) {
        if (s.begin != end || s.end <= s.begin) return "film provenance segments are not contiguous";
        for (const std::string *t : {&s.build, &s.device, &s.architecture, &s.os})
            if (t->size() > kMaxName) return "film provenance string exceeds 4096 bytes";
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
    const uint64_t npx = uint64_t(f.width) * f.height;
    if (npx * kPixel3 > kMaxPayload3) return "film payload exceeds 4 GiB";
    const std::vector<uint8_t> meta = metadata3(f);
    if (meta.size() > kMaxMeta3) return "checkpoint metadata exceeds 1 MiB";
    Writer head;
    head.raw(kMagic3, 8);
    for (uint32_t v : kContract) head.u32(v);
    head.u64(meta.size());
    head.u64(npx * kPixel3);
    return publish(path, head.b, meta, f);
}

std::unique_ptr<rp_film> read_checkpoint_file(const char *path) {
    if (!path || !*path) throw std::runtime_error("invalid checkpoint path");
    const std::vector<uint8_t> data = read_file(path);
    std::unique_ptr<rp_film> f;
    if (data.size() >= 8 + 32 && std::memcmp(data.data(), kMagic3, 8) == 0) f = parse3(data);
    else if (data.size() >= 8 + 32 && std::memcmp(data.data(), kMagic1, 8) == 0) f = parse1(data);
    else throw std::runtime_error("invalid checkpoint header");
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
    info->sampler_kind = f->sampler;
    info->model = f->model;
    info->job_samples = f->job_samples;
    info->batch_size = f->batch_size;
    info->packet = f->packet ? 1u : 0u;
    info->format = f->format;
    return RP_SUCCESS;
}

int rp_film_set_job(rp_film *f, uint64_t job_samples, uint64_t batch_size) {
    if (!f) {
        rp::set_error("rp_film_set_job: null argument");
        return RP_INVALID;
    }
    if (job_samples == 0 || batch_size == 0 || job_samples < f->samples) {
        rp::set_error("rp_film_set_job: the job needs positive samples and batch size, and at least the "
                      "film's samples");
        return RP_INVALID;
    }
    f->job_samples = job_samples;
    f->batch_size = batch_size;
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
