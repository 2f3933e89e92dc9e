// `rypath` command line (counterpart of src/main.cpp:429-530 --render-demo,
// :532-800 --render-pbrt, :885-899 --check-pbrt): renders a PBRT-v4 scene file
// or the reference's demo scene (scalar demo model, Philox sampler) on the
// GPU and publishes a linear-sRGB OpenEXR image.
// Exit status: 0 success, 1 failure (bad arguments, invalid scene, render or
// write error), 2 unavailable (no HIP device), as src/main.cpp:215-230.
#include "../../include/rp_api.h"
#include "../../include/rp_io.h"
#include "../../include/rp_synthetic.h"

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::string last_error() {
    std::vector<char> buf(4096);
    rp_last_error(buf.data(), buf.size());
    return buf.data();
}

int fail(const std::string &m) {
    std::fprintf(stderr, "rypath: %s\n", m.c_str());
    return 1;
}

int status_exit(int st, const char *where) {
    std::fprintf(stderr, "rypath: %s: %s\n", where, last_error().c_str());
    return st == RP_NO_DEVICE || st == RP_NOT_BUILT ? 2 : 1;
}

bool parse_u64(const char *s, uint64_t &out) {
    if (!s || !*s || *s == '-') return false;
    char *end = nullptr;
    errno = 0;
    unsigned long long v = std::strtoull(s, &end, 10);
    if (errno || *end) return false;
    out = v;
    return true;
}

void usage() {
    std::puts("usage:\n"
              "  rypath --render-pbrt SCENE.pbrt --output IMAGE.exr [--samples N] [--seed S]\n"
              "         [--mode megakernel|wavefront|sorted (default: wavefront)] [--hero-wavelengths uniform|visible]\n"
              "         [--exposure E] [--half] [--timings] [--batch-size N]\n"
              "         [--checkpoint PATH] [--resume PATH] [--checkpoint-interval SECONDS (default: 60)]\n"
              "  rypath --check-pbrt SCENE.pbrt\n"
              "  rypath --render-demo --output IMAGE.exr [--width W] [--height H] [--samples N] [--seed S]\n"
              "  rypath --device-info | --version");
}

struct Options {
    std::string scene, output, checkpoint, resume;
    uint64_t samples = 0, seed = 0, width = 256, height = 192, batch = 0;
    bool have_seed = false, have_mode = false, half = false, timings = false;
    double exposure = 1, interval = 60;
    uint32_t mode = RP_MODE_WAVEFRONT; // the fastest mode on MI355X; films are identical in every mode
    int hero = -1;
};

int parse_options(int argc, char **argv, int first, Options &o) {
    for (int i = first; i < argc; ++i) {
        std::string k = argv[i];
        if (k == "--half") {
            o.half = true;
            continue;
        }
        if (k == "--timings") {
            o.timings = true;
            continue;
        }
        if (i + 1 >= argc || std::strncmp(argv[i + 1], "--", 2) == 0) return fail(k + " requires a value");
        const char *v = argv[++i];
        if (k == "--output") o.output = v;
        else if (k == "--checkpoint") o.checkpoint = v;
        else if (k == "--resume") o.resume = v;
        else if (k == "--checkpoint-interval") {
            char *end = nullptr;
            o.interval = std::strtod(v, &end);
            if (*end || !(o.interval >= 0)) return fail("--checkpoint-interval must be a nonnegative number of seconds");
        } else if (k == "--batch-size") {
            if (!parse_u64(v, o.batch) || o.batch == 0 || o.batch > 0xffffffffull)
                return fail("--batch-size must be a positive 32-bit unsigned integer");
        } else if (k == "--samples") {
            if (!parse_u64(v, o.samples) || o.samples == 0 || o.samples > 0xffffffffull)
                return fail("--samples must be a positive 32-bit unsigned integer");
        } else if (k == "--seed") {
            uint64_t s;
            if (!parse_u64(v, s)) return fail("--seed must be an unsigned integer");
            o.seed = s;
            o.have_seed = true;
        } else if (k == "--width") {
            if (!parse_u64(v, o.width) || o.width == 0 || o.width > 16384) return fail("--width must be 1..16384");
        } else if (k == "--height") {
            if (!parse_u64(v, o.height) || o.height == 0 || o.height > 16384) return fail("--height must be 1..16384");
        } else if (k == "--exposure") {
            char *end = nullptr;
            o.exposure = std::strtod(v, &end);
            if (*end || !(o.exposure > 0)) return fail("--exposure must be a positive number");
        } else if (k == "--mode") {
            std::string m = v;
            if (m == "megakernel") o.mode = RP_MODE_MEGAKERNEL;
            else if (m == "wavefront") o.mode = RP_MODE_WAVEFRONT;
            else if (m == "sorted") o.mode = RP_MODE_WAVEFRONT_SORTED;
            else return fail("--mode requires megakernel, wavefront or sorted");
            o.have_mode = true;
        } else if (k == "--hero-wavelengths") {
            std::string m = v;
            if (m == "uniform") o.hero = 0;
            else if (m == "visible") o.hero = 1;
            else return fail("--hero-wavelengths requires uniform or visible");
        } else {
            return fail("unknown option " + k);
        }
    }
    return 0;
}

bool ends_with(const std::string &s, const char *suffix) {
    size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

int render(const rp_scene_desc &desc, const Options &o, uint32_t spp, uint64_t seed, uint32_t randomization,
           const std::string &source, uint32_t frame_flags = 0) {
    auto t0 = std::chrono::steady_clock::now();
    // The film accumulates batches of samples; a resumed or checkpointed
    // render equals the uninterrupted one bit for bit.  A resumed film must
    // match the scene's size and the sampler before anything renders
    // (src/render_job.cpp:125-140); its scene identity is checked once the
    // scene is built.
    rp_film *film = nullptr;
    int st = !o.resume.empty() ? rp_checkpoint_read(o.resume.c_str(), &film)
                               : rp_film_create(desc.camera.width, desc.camera.height, &film);
    if (st != RP_SUCCESS) return status_exit(st, o.resume.empty() ? "film" : "resume");
    rp_film_info fi{};
    rp_film_get_info(film, &fi);
    std::string mismatch;
    if (fi.width != desc.camera.width || fi.height != desc.camera.height)
        mismatch = "the checkpoint's film is " + std::to_string(fi.width) + "x" + std::to_string(fi.height) +
                   ", the scene renders " + std::to_string(desc.camera.width) + "x" +
                   std::to_string(desc.camera.height);
    else if (fi.samples > spp)
        mismatch = "the checkpoint holds more samples than requested";
    else if (fi.samples && (fi.seed != seed || fi.randomization != randomization))
        mismatch = "the checkpoint was rendered with another seed or randomization";
    else if (fi.samples && fi.sampler_kind != ((frame_flags & RP_FRAME_INDEPENDENT_SAMPLER) ? uint32_t(RP_SAMPLER_INDEPENDENT)
                                                                                             : uint32_t(RP_SAMPLER_SOBOL)))
        mismatch = "the checkpoint holds samples of another sampler";
    if (!mismatch.empty()) {
        rp_film_destroy(film);
        return fail("checkpoint does not match the scene or sampler: " + mismatch);
    }
    if (rp_device_count() <= 0) {
        rp_film_destroy(film);
        std::fprintf(stderr, "rypath: no HIP device is available\n");
        return 2;
    }
    rp_scene *scene = nullptr;
    st = rp_scene_create(&desc, nullptr, &scene);
    if (st != RP_SUCCESS) {
        rp_film_destroy(film);
        return status_exit(st, "scene build");
    }
    rp_build_info info{};
    rp_scene_info(scene, &info);
    if ((st = rp_film_check(scene, film, seed, randomization)) != RP_SUCCESS) {
        rp_film_destroy(film);
        rp_scene_destroy(scene);
        return status_exit(st, "resume");
    }
    const uint64_t batch = o.batch ? o.batch : (o.checkpoint.empty() ? spp : std::max<uint64_t>(1, spp / 16));
    rp_film_set_job(film, spp, batch); // recorded in checkpoints; spp >= the film's samples was checked above
    rp_frame_stats total{};
    auto last_write = std::chrono::steady_clock::now();
    while (fi.samples < spp) {
        rp_frame_config cfg{};
        cfg.mode = o.mode;
        cfg.spp = uint32_t(std::min<uint64_t>(batch, spp - fi.samples));
        cfg.sample_begin = fi.samples;
        cfg.seed = seed;
        cfg.randomization = randomization;
        cfg.flags = frame_flags;
        rp_frame_stats stats{};
        st = rp_film_accumulate(scene, film, &cfg, nullptr, &stats);
        if (st != RP_SUCCESS) {
            rp_film_destroy(film);
            rp_scene_destroy(scene);
            return status_exit(st, "render");
        }
        total.seconds += stats.seconds;
        total.paths += stats.paths;
        total.extension_rays += stats.extension_rays;
        total.shadow_rays += stats.shadow_rays;
        total.failed_paths += stats.failed_paths;
        rp_film_get_info(film, &fi);
        const bool last = fi.samples == spp;
        const double since = std::chrono::duration<double>(std::chrono::steady_clock::now() - last_write).count();
        if (!o.checkpoint.empty() && (last || since >= o.interval)) {
            if ((st = rp_checkpoint_write(film, o.checkpoint.c_str())) != RP_SUCCESS) {
                rp_film_destroy(film);
                rp_scene_destroy(scene);
                return status_exit(st, "checkpoint");
            }
            last_write = std::chrono::steady_clock::now();
        }
    }
    rp_scene_destroy(scene);
    rp_film_get_info(film, &fi);
    std::vector<float> xyz(size_t(fi.width) * fi.height * 3);
    st = rp_film_resolve(film, xyz.data());
    rp_film_destroy(film);
    if (st != RP_SUCCESS) return status_exit(st, "resolve");
    if (total.failed_paths) std::fprintf(stderr, "rypath: %llu paths failed\n", (unsigned long long)total.failed_paths);
    std::string spp_s = std::to_string(spp), seed_s = std::to_string(seed), ident;
    for (uint8_t b : fi.scene_identity) {
        char h[3];
        std::snprintf(h, sizeof h, "%02x", b);
        ident += h;
    }
    const char *keys[] = {"rypath.source", "rypath.samples", "rypath.seed", "rypath.checkpointIdentity"};
    const char *vals[] = {source.c_str(), spp_s.c_str(), seed_s.c_str(), ident.c_str()};
    st = rp_write_exr(o.output.c_str(), fi.width, fi.height, xyz.data(), o.exposure, o.half ? 1 : 0,
                      keys, vals, 4);
    if (st != RP_SUCCESS) return status_exit(st, "write");
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (o.timings)
        std::printf("bvh_build_s=%.4f frame_s=%.4f paths=%llu mpaths_per_s=%.2f extension_rays=%llu "
                    "shadow_rays=%llu wall_s=%.3f\n",
                    info.build_seconds, total.seconds, (unsigned long long)total.paths,
                    total.seconds > 0 ? total.paths / total.seconds * 1e-6 : 0.0,
                    (unsigned long long)total.extension_rays, (unsigned long long)total.shadow_rays, wall);
    return 0;
}

int render_pbrt(int argc, char **argv, bool check_only) {
    if (argc < 3 || std::strncmp(argv[2], "--", 2) == 0) return fail(std::string(argv[1]) + " requires a scene path");
    Options o;
    o.scene = argv[2];
    if (int e = parse_options(argc, argv, 3, o)) return e;
    rp_pbrt *p = nullptr;
    int st = rp_pbrt_load(o.scene.c_str(), &p);
    if (st != RP_SUCCESS) return status_exit(st, "scene");
    rp_pbrt_info info{};
    rp_pbrt_get_info(p, &info);
    rp_scene_desc desc = *rp_pbrt_desc(p);
    if (check_only) {
        std::printf("%s: %ux%u, %u spp, maxdepth %u, %u triangles, %u materials, %u emitters, %u media\n",
                    o.scene.c_str(), info.width, info.height, info.spp, info.max_depth, desc.triangle_count,
                    desc.material_count, desc.emission_count, desc.medium_count);
        rp_pbrt_destroy(p);
        return 0;
    }
    if (o.output.empty() && info.output[0]) o.output = info.output;
    if (o.output.empty()) {
        rp_pbrt_destroy(p);
        return fail("--render-pbrt requires --output IMAGE.exr");
    }
    if (!ends_with(o.output, ".exr")) {
        rp_pbrt_destroy(p);
        return fail("--render-pbrt output must have the .exr extension");
    }
    if (o.hero >= 0) desc.hero_wavelengths = uint32_t(o.hero);
    int rc = render(desc, o, o.samples ? uint32_t(o.samples) : info.spp, o.have_seed ? o.seed : info.seed,
                    info.randomization, o.scene);
    rp_pbrt_destroy(p);
    return rc;
}

// The reference's demo (src/main.cpp:429-530): demo-v1 scene, scalar demo
// model, Philox4x32-10 independent sampler, 256 x 192, 128 samples, seed 1.
int render_demo(int argc, char **argv) {
    Options o;
    if (int e = parse_options(argc, argv, 2, o)) return e;
    if (o.output.empty()) return fail("--render-demo requires --output PATH");
    if (!ends_with(o.output, ".exr")) return fail("--render-demo output must have the .exr extension");
    if (o.hero >= 0) return fail("--hero-wavelengths requires effective hero transport");
    if (o.have_mode && o.mode != RP_MODE_MEGAKERNEL) return fail("the demo's scalar transport renders as a megakernel");
    o.mode = RP_MODE_MEGAKERNEL;
    rp_synthetic_options so{};
    so.width = uint32_t(o.width);
    so.height = uint32_t(o.height);
    rp_synthetic *syn = nullptr;
    int st = rp_synthetic_create(RP_SYNTH_DEMO, 0, &so, &syn);
    if (st != RP_SUCCESS) return status_exit(st, "demo scene");
    rp_scene_desc desc = *rp_synthetic_desc(syn);
    std::printf("Scene: demo-v1\nSampler: philox4x32-10-v1\nResolution: %ux%u\n", desc.camera.width,
                desc.camera.height);
    int rc = render(desc, o, o.samples ? uint32_t(o.samples) : 128, o.have_seed ? o.seed : 1, RP_RANDOMIZE_NONE,
                    "demo-v1", RP_FRAME_INDEPENDENT_SAMPLER);
    rp_synthetic_destroy(syn);
    return rc;
}

} // namespace

int main(int argc, char **argv) {
    if (argc == 1 || (argc == 2 && std::strcmp(argv[1], "--help") == 0)) {
        usage();
        return argc == 1 ? 1 : 0;
    }
    std::string cmd = argv[1];
    if (cmd == "--version") {
        std::printf("rypath-amd (C ABI %u, gfx950)\n", rp_api_version());
        return 0;
    }
    if (cmd == "--device-info") {
        int n = rp_device_count();
        std::printf("hip devices: %d\n", n);
        return n > 0 ? 0 : 2;
    }
    if (cmd == "--render-pbrt") return render_pbrt(argc, argv, false);
    if (cmd == "--check-pbrt") return render_pbrt(argc, argv, true);
    if (cmd == "--render-demo") return render_demo(argc, argv);
    usage();
    return 1;
}
