// HIP kernels and the C ABI (include/rp_api.h) of the gfx950 path tracer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rp_internal.hpp"
#include "rp_path.h"
#include "rp_wavefront.h"

namespace rp {

#define HIP_CHECK(expr)                                                                                       \
    do {                                                                                                       \
        hipError_t err_ = (expr);                                                                              \
        if (err_ != hipSuccess)                                                                                \
            throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(err_));                     \
    } while (0)

template <typename T> struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count <= n && p) return;
        release();
        if (count == 0) count = 1;
        HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T)));
        n = count;
    }
    void upload(const T *src, size_t count) {
        alloc(count);
        if (count) HIP_CHECK(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
    }
    void upload(const std::vector<T> &v) { upload(v.data(), v.size()); }
};

} // namespace rp

using namespace rp;

struct rp_scene {
    HostScene host;
    Bvh4 bvh;
    rp_build_info info{};
    int device = 0;
    DevBuf<float4> tris;
    DevBuf<QNode> nodes;
    DevBuf<rp_object> objects;
    DevBuf<rp_material> materials;
    DevBuf<rp_emission> emissions;
    DevBuf<rp_medium> media;
    DevBuf<rp_spot_light> spots;
    DevBuf<float> spectra;
    DevBuf<rpd::LightEntry> lights;
    DevBuf<float4> cie;
    DevBuf<uint32_t> sobol_dir;
    DevBuf<uint64_t> sobol_intervals, sobol_inverse;
    rpd::Scene dev{};
    // frame scratch
    DevBuf<float> path_xyz;
    DevBuf<double> film;
    DevBuf<rp_path_record> records;
    DevBuf<unsigned long long> counters;
    DevBuf<rp_ray> rays;
    DevBuf<rp_hit> hits;
    DevBuf<uint8_t> occluded;
    hipStream_t stream = nullptr;
    rpd::WavefrontBuffers wf;
    ~rp_scene() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

constexpr int kBlock = 128;

template <int B> struct LdsStack {
    static __device__ __forceinline__ void init(rpd::Stack &st, uint32_t *n, float *t) {
        st.lds_node = n + threadIdx.x;
        st.lds_t = t + threadIdx.x;
        st.stride = B;
        st.sp = 0;
    }
};

__device__ __forceinline__ void count_rays(unsigned long long *counters, uint32_t ext, uint32_t sh, uint32_t failed) {
    // wave-level reduction, one atomic per wave per counter
    unsigned long long e = ext, s = sh, f = failed;
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_down(e, off);
        s += __shfl_down(s, off);
        f += __shfl_down(f, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&counters[0], e);
        atomicAdd(&counters[1], s);
        atomicAdd(&counters[2], f);
    }
}

__global__ void __launch_bounds__(kBlock) k_records(rpd::Scene S, rpd::PathParams P, uint32_t pixel_begin,
                                                     uint32_t pixel_count, uint64_t sample_begin, uint64_t n,
                                                     rp_path_record *out) {
    __shared__ uint32_t s_node[rpd::kShortStack * kBlock];
    __shared__ float s_t[rpd::kShortStack * kBlock];
    rpd::Stack st;
    LdsStack<kBlock>::init(st, s_node, s_t);
    uint64_t lane = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (lane >= n) return;
    uint32_t pixel = pixel_begin + uint32_t(lane % pixel_count);
    uint64_t sample = sample_begin + lane / pixel_count;
    rpd::PathResult r;
    rpd::trace_path(S, P, st, pixel, sample, r);
    rp_path_record rec;
    rec.pixel = pixel;
    rec.status = r.status;
    rec.sample = sample;
    rec.sobol_index = r.index;
    for (int k = 0; k < 4; ++k) {
        rec.wavelengths[k] = r.wl.lambda[k];
        rec.wavelength_pdfs[k] = r.wl.pdf[k];
        rec.radiance[k] = r.L[k];
    }
    float xyz[3] = {0, 0, 0};
    if (r.status == RP_PATH_SUCCESS) rpd::sensor_xyz(S, r.wl, r.L, xyz);
    rec.xyz[0] = xyz[0];
    rec.xyz[1] = xyz[1];
    rec.xyz[2] = xyz[2];
    rec.primary_triangle = r.primary;
    rec.vertices = r.vertices;
    rec.camera_words[0] = r.w0;
    rec.camera_words[1] = r.w1;
    rec.pad = 0;
    out[lane] = rec;
}

// Megakernel: one thread = one complete path; per-path XYZ written planar
// (x plane, y plane, z plane) for the film reduction.
__global__ void __launch_bounds__(kBlock) k_mega(rpd::Scene S, rpd::PathParams P, uint32_t npx, uint64_t sample_begin,
                                                  uint64_t n, float *xyz, unsigned long long *counters) {
    __shared__ uint32_t s_node[rpd::kShortStack * kBlock];
    __shared__ float s_t[rpd::kShortStack * kBlock];
    rpd::Stack st;
    LdsStack<kBlock>::init(st, s_node, s_t);
    uint64_t lane = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    uint32_t ext = 0, sh = 0, failed = 0;
    if (lane < n) {
        uint32_t pixel = uint32_t(lane % npx);
        uint64_t sample = sample_begin + lane / npx;
        rpd::PathResult r;
        rpd::trace_path(S, P, st, pixel, sample, r);
        float v[3] = {0, 0, 0};
        if (r.status == RP_PATH_SUCCESS) rpd::sensor_xyz(S, r.wl, r.L, v);
        else failed = 1;
        xyz[lane] = v[0];
        xyz[n + lane] = v[1];
        xyz[2 * n + lane] = v[2];
        ext = r.ext_rays;
        sh = r.shadow_rays;
    }
    count_rays(counters, ext, sh, failed);
}

// Atomic-free film: each pixel adds its samples in increasing sample order
// (Neumaier-compensated double sums, src/render.cpp:2326-2407).
__global__ void k_film_add(uint32_t npx, uint32_t samples, uint64_t n, const float *xyz, double *film) {
    uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= npx) return;
    double sum[3], corr[3];
    for (int c = 0; c < 3; ++c) {
        sum[c] = film[(2 * c) * size_t(npx) + pixel];
        corr[c] = film[(2 * c + 1) * size_t(npx) + pixel];
    }
    for (uint32_t s = 0; s < samples; ++s) {
        uint64_t lane = uint64_t(s) * npx + pixel;
        for (int c = 0; c < 3; ++c) {
            double x = double(xyz[uint64_t(c) * n + lane]);
            double t = sum[c] + x;
            if (fabs(sum[c]) >= fabs(x)) corr[c] += (sum[c] - t) + x;
            else corr[c] += (x - t) + sum[c];
            sum[c] = t;
        }
    }
    for (int c = 0; c < 3; ++c) {
        film[(2 * c) * size_t(npx) + pixel] = sum[c];
        film[(2 * c + 1) * size_t(npx) + pixel] = corr[c];
    }
}

__global__ void k_zero_counters(unsigned long long *c, uint32_t n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0;
}

__global__ void k_film_zero(uint64_t n, double *film) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) film[i] = 0.0;
}

__global__ void k_film_finish(uint32_t npx, uint32_t spp, const double *film, float *out) {
    uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= npx) return;
    for (int c = 0; c < 3; ++c)
        out[3 * size_t(pixel) + c] =
            float((film[(2 * c) * size_t(npx) + pixel] + film[(2 * c + 1) * size_t(npx) + pixel]) / double(spp));
}

__global__ void __launch_bounds__(kBlock) k_trace_closest(rpd::Scene S, const rp_ray *rays, uint64_t n, rp_hit *hits) {
    __shared__ uint32_t s_node[rpd::kShortStack * kBlock];
    __shared__ float s_t[rpd::kShortStack * kBlock];
    rpd::Stack st;
    LdsStack<kBlock>::init(st, s_node, s_t);
    uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    rp_ray r = rays[i];
    rpd::Hit h;
    bool hit = rpd::traverse<false>(S, st, {r.origin[0], r.origin[1], r.origin[2]},
                                    {r.direction[0], r.direction[1], r.direction[2]}, r.tmax, h);
    rp_hit o;
    o.triangle = hit ? h.id : RP_NONE;
    o.t = hit ? h.t : __builtin_inff();
    o.b1 = hit ? h.b1 : 0.0f;
    o.b2 = hit ? h.b2 : 0.0f;
    hits[i] = o;
}

__global__ void __launch_bounds__(kBlock) k_trace_any(rpd::Scene S, const rp_ray *rays, uint64_t n, uint8_t *occ) {
    __shared__ uint32_t s_node[rpd::kShortStack * kBlock];
    __shared__ float s_t[rpd::kShortStack * kBlock];
    rpd::Stack st;
    LdsStack<kBlock>::init(st, s_node, s_t);
    uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    rp_ray r = rays[i];
    rpd::Hit h;
    bool hit = rpd::traverse<true>(S, st, {r.origin[0], r.origin[1], r.origin[2]},
                                   {r.direction[0], r.direction[1], r.direction[2]}, r.tmax, h);
    occ[i] = hit ? 1 : 0;
}

int fail(const char *where, const std::exception &e) {
    set_error(std::string(where) + ": " + e.what());
    return RP_FAILURE;
}

void upload_scene(rp_scene *s) {
    const rp_scene_desc &d = s->host.desc;
    HIP_CHECK(hipSetDevice(s->device));
    if (!s->stream) HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    // leaf-order triangles: (v0, scene id), (v1, object), (v2, light)
    std::vector<uint32_t> tri_light;
    float radius = scene_bounding_radius(d);
    std::vector<LightEntry> lights = build_light_table(d, radius, tri_light);
    std::vector<uint32_t> leaf_pos(d.triangle_count);
    for (uint32_t p = 0; p < d.triangle_count; ++p) leaf_pos[s->bvh.tri_order[p]] = p;
    std::vector<float4> tris(3 * size_t(d.triangle_count));
    for (uint32_t p = 0; p < d.triangle_count; ++p) {
        uint32_t t = s->bvh.tri_order[p];
        const float *v[3];
        for (int k = 0; k < 3; ++k) v[k] = d.positions + 3 * size_t(d.indices[3 * t + k]);
        uint32_t w[3] = {t, d.tri_object[t], tri_light[t]};
        for (int k = 0; k < 3; ++k) {
            float wf;
            std::memcpy(&wf, &w[k], 4);
            tris[3 * size_t(p) + k] = make_float4(v[k][0], v[k][1], v[k][2], wf);
        }
    }
    std::vector<rpd::LightEntry> dl(lights.size());
    for (size_t i = 0; i < lights.size(); ++i) {
        dl[i].kind = lights[i].kind;
        dl[i].index = lights[i].kind == 0 ? leaf_pos[lights[i].index] : lights[i].index;
        dl[i].pmf = lights[i].pmf;
        dl[i].cdf = lights[i].cdf;
        dl[i].inv_area = lights[i].inv_area;
        dl[i].pad[0] = dl[i].pad[1] = dl[i].pad[2] = 0;
    }
    s->tris.upload(tris);
    s->nodes.upload(s->bvh.nodes);
    s->objects.upload(s->host.objects);
    s->materials.upload(s->host.materials);
    s->emissions.upload(s->host.emissions);
    s->media.upload(s->host.media);
    s->spots.upload(s->host.spots);
    s->spectra.upload(s->host.spectra);
    s->lights.upload(dl);
    std::vector<float4> cie(471);
    const auto &ct = cie_xyz();
    for (int i = 0; i < 471; ++i) cie[i] = make_float4(ct[i][0], ct[i][1], ct[i][2], 0.0f);
    s->cie.upload(cie);
    const SobolData &sd = sobol_data();
    std::vector<uint32_t> dirT(52 * 1024);
    for (uint32_t dim = 0; dim < 1024; ++dim)
        for (uint32_t c = 0; c < 52; ++c) dirT[c * 1024 + dim] = sd.directions[dim * 52 + c];
    s->sobol_dir.upload(dirT);
    s->sobol_intervals.upload(sd.intervals);
    s->sobol_inverse.upload(sd.inverse);
    s->counters.alloc(8);

    rpd::Scene &S = s->dev;
    S.tris = s->tris.p;
    S.nodes = reinterpret_cast<const uint4 *>(s->nodes.p);
    S.node_count = uint32_t(s->bvh.nodes.size());
    S.tri_count = d.triangle_count;
    S.objects = s->objects.p;
    S.materials = s->materials.p;
    S.emissions = s->emissions.p;
    S.media = s->media.p;
    S.spots = s->spots.p;
    S.spectra = s->spectra.p;
    S.lights = s->lights.p;
    S.light_count = uint32_t(dl.size());
    S.has_environment = d.has_environment;
    S.cie = s->cie.p;
    S.sobol_dir = s->sobol_dir.p;
    S.sobol_intervals = s->sobol_intervals.p;
    S.sobol_inverse = s->sobol_inverse.p;
    S.camera = d.camera;
    S.environment = d.environment;
    S.max_depth = d.max_depth;
    S.hero_mode = d.hero_wavelengths;
    uint32_t v = std::max(std::max(d.camera.width, d.camera.height), 1u) - 1, m = 0;
    while (v) { ++m; v >>= 1; }
    S.m = m;
    S.has_interfaces = 0;
    for (const auto &mt : s->host.materials)
        if (mt.kind == RP_MATERIAL_INTERFACE) S.has_interfaces = 1;
    S.has_media = d.medium_count ? 1 : 0;
    S.env_pdf = (d.has_environment && !dl.empty()) ? dl.back().pmf * rpd::kInv4Pi : 0.0f;
    s->info.emitter_count = 0;
    for (auto &l : dl)
        if (l.kind == 0) ++s->info.emitter_count;
    s->info.light_count = uint32_t(dl.size());
    s->info.device_bytes = s->tris.n * sizeof(float4) + s->nodes.n * sizeof(QNode) + s->spectra.n * 4 +
                           s->sobol_dir.n * 4 + s->lights.n * sizeof(rpd::LightEntry);
}

rpd::PathParams path_params(uint64_t seed, uint32_t randomization, uint32_t emergency, uint32_t boundary) {
    rpd::PathParams P;
    P.seed = seed;
    P.mode = randomization;
    P.emergency = emergency ? emergency : 1024u;
    P.boundary = boundary ? boundary : 4096u;
    return P;
}

// Renders the frame into the scene's film buffer; returns seconds (device events).
double render_frame_impl(rp_scene *s, const rp_frame_config &cfg, float *xyz_device, rp_frame_stats *stats) {
    const rp_camera &cam = s->host.desc.camera;
    const uint32_t npx = cam.width * cam.height;
    const uint32_t spp = cfg.spp;
    if (spp == 0) throw std::runtime_error("spp must be positive");
    if (cfg.mode > RP_MODE_WAVEFRONT_SORTED) throw std::runtime_error("unknown mode");
    if (cfg.randomization > RP_RANDOMIZE_OWEN) throw std::runtime_error("unknown randomization");
    HIP_CHECK(hipSetDevice(s->device));
    rpd::PathParams P = path_params(cfg.seed, cfg.randomization, 0, 0);
    const uint64_t max_chunk_paths = uint64_t(1) << 26;
    uint32_t chunk = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(spp, max_chunk_paths / npx)));
    s->film.alloc(6 * size_t(npx));
    hipLaunchKernelGGL(k_film_zero, dim3(uint32_t((6 * uint64_t(npx) + 255) / 256)), dim3(256), 0, s->stream,
                       6 * uint64_t(npx), s->film.p);
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s->stream, s->counters.p, 8u);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, s->stream));
    double trace_seconds = 0;
    uint64_t launches = 0;
    for (uint32_t s0 = 0; s0 < spp; s0 += chunk) {
        uint32_t ns = std::min(chunk, spp - s0);
        uint64_t n = uint64_t(ns) * npx;
        s->path_xyz.alloc(3 * n);
        if (cfg.mode == RP_MODE_MEGAKERNEL) {
            hipLaunchKernelGGL(k_mega, dim3(uint32_t((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream, s->dev, P,
                               npx, cfg.sample_begin + s0, n, s->path_xyz.p, s->counters.p);
        } else {
            trace_seconds += rpd::wavefront_render(s->dev, P, s->wf, cfg, npx, cfg.sample_begin + s0, n, s->path_xyz.p,
                                                   s->counters.p, s->stream, launches);
        }
        HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_film_add, dim3((npx + 255) / 256), dim3(256), 0, s->stream, npx, ns, n, s->path_xyz.p,
                           s->film.p);
    }
    hipLaunchKernelGGL(k_film_finish, dim3((npx + 255) / 256), dim3(256), 0, s->stream, npx, spp, s->film.p, xyz_device);
    HIP_CHECK(hipEventRecord(e1, s->stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_CHECK(hipGetLastError());
    if (stats) {
        unsigned long long c[8];
        HIP_CHECK(hipMemcpy(c, s->counters.p, sizeof c, hipMemcpyDeviceToHost));
        stats->seconds = ms * 1e-3;
        stats->paths = uint64_t(npx) * spp;
        stats->extension_rays = c[0];
        stats->shadow_rays = c[1];
        stats->failed_paths = c[2];
        stats->trace_seconds = trace_seconds;
        stats->extend_launches = launches;
        stats->node_fetches = c[3];
        stats->triangle_tests = c[4];
    }
    return ms * 1e-3;
}

} // namespace

extern "C" {

uint32_t rp_api_version(void) { return RP_API_VERSION; }

int rp_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

size_t rp_last_error(char *buffer, size_t capacity) {
    const std::string &e = last_error();
    if (buffer && capacity) {
        size_t k = std::min(capacity - 1, e.size());
        std::memcpy(buffer, e.data(), k);
        buffer[k] = 0;
    }
    return e.size();
}

int rp_scene_create(const rp_scene_desc *desc, const rp_build_options *options, rp_scene **out) {
    if (!desc || !out) {
        set_error("rp_scene_create: null argument");
        return RP_INVALID;
    }
    *out = nullptr;
    std::string err = validate_desc(*desc);
    if (!err.empty()) {
        set_error("rp_scene_create: " + err);
        return RP_INVALID;
    }
    if (rp_device_count() <= 0) {
        set_error("rp_scene_create: no HIP device");
        return RP_NO_DEVICE;
    }
    rp_build_options o{};
    if (options) o = *options;
    try {
        auto s = std::make_unique<rp_scene>();
        s->device = int(o.device);
        s->host.copy_from(*desc);
        s->bvh = build_bvh4(s->host.desc.positions, s->host.desc.indices, s->host.desc.triangle_count, o.leaf_size,
                            o.sah_bins, o.threads);
        if (s->bvh.max_stack > uint32_t(rpd::kMaxStack))
            throw std::runtime_error("BVH needs a traversal stack of " + std::to_string(s->bvh.max_stack) +
                                     " entries; the kernel supports " + std::to_string(rpd::kMaxStack));
        auto t0 = std::chrono::steady_clock::now();
        upload_scene(s.get());
        HIP_CHECK(hipDeviceSynchronize());
        s->info.build_seconds = s->bvh.seconds;
        s->info.upload_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        s->info.node_count = s->bvh.nodes.size();
        s->info.leaf_count = s->bvh.leaf_count;
        s->info.max_depth = s->bvh.depth;
        s->info.max_stack = s->bvh.max_stack;
        *out = s.release();
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_scene_create", e);
    }
}

int rp_scene_info(const rp_scene *scene, rp_build_info *info) {
    if (!scene || !info) {
        set_error("rp_scene_info: null argument");
        return RP_INVALID;
    }
    *info = scene->info;
    return RP_SUCCESS;
}

void rp_scene_destroy(rp_scene *scene) { delete scene; }

int rp_sample(rp_scene *s, const rp_request *rq, rp_path_record *out, uint64_t capacity) {
    if (!s || !rq || !out) {
        set_error("rp_sample: null argument");
        return RP_INVALID;
    }
    const uint32_t npx = s->host.desc.camera.width * s->host.desc.camera.height;
    if (rq->pixel_count == 0 || rq->sample_count == 0 || uint64_t(rq->pixel_begin) + rq->pixel_count > npx) {
        set_error("rp_sample: pixel range outside the film or empty request");
        return RP_INVALID;
    }
    if (rq->randomization > RP_RANDOMIZE_OWEN) {
        set_error("rp_sample: unknown randomization");
        return RP_INVALID;
    }
    uint64_t n = uint64_t(rq->pixel_count) * rq->sample_count;
    if (capacity < n) {
        set_error("rp_sample: output capacity too small");
        return RP_INVALID;
    }
    try {
        HIP_CHECK(hipSetDevice(s->device));
        s->records.alloc(n);
        rpd::PathParams P = path_params(rq->seed, rq->randomization, rq->emergency_limit, rq->boundary_limit);
        hipLaunchKernelGGL(k_records, dim3(uint32_t((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream, s->dev, P,
                           rq->pixel_begin, rq->pixel_count, rq->sample_begin, n, s->records.p);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(out, s->records.p, n * sizeof(rp_path_record), hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_sample", e);
    }
}

int rp_render_frame_device(rp_scene *s, const rp_frame_config *cfg, float *xyz_device, rp_frame_stats *stats) {
    if (!s || !cfg || !xyz_device) {
        set_error("rp_render_frame_device: null argument");
        return RP_INVALID;
    }
    try {
        render_frame_impl(s, *cfg, xyz_device, stats);
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_render_frame", e);
    }
}

int rp_render_frame(rp_scene *s, const rp_frame_config *cfg, float *xyz_host, rp_frame_stats *stats) {
    if (!s || !cfg) {
        set_error("rp_render_frame: null argument");
        return RP_INVALID;
    }
    try {
        HIP_CHECK(hipSetDevice(s->device));
        const uint32_t npx = s->host.desc.camera.width * s->host.desc.camera.height;
        DevBuf<float> out;
        out.alloc(3 * size_t(npx));
        render_frame_impl(s, *cfg, out.p, stats);
        if (xyz_host) HIP_CHECK(hipMemcpy(xyz_host, out.p, 3 * size_t(npx) * sizeof(float), hipMemcpyDeviceToHost));
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_render_frame", e);
    }
}

static int trace_common(rp_scene *s, const rp_ray *rays, uint64_t count, void *out, bool any, double *seconds) {
    if (!s || (!rays && count) || (!out && count)) {
        set_error("rp_trace: null argument");
        return RP_INVALID;
    }
    if (count == 0) return RP_SUCCESS;
    try {
        HIP_CHECK(hipSetDevice(s->device));
        s->rays.upload(rays, count);
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipEventRecord(e0, s->stream));
        dim3 grid(uint32_t((count + kBlock - 1) / kBlock));
        if (any) {
            s->occluded.alloc(count);
            hipLaunchKernelGGL(k_trace_any, grid, dim3(kBlock), 0, s->stream, s->dev, s->rays.p, count, s->occluded.p);
        } else {
            s->hits.alloc(count);
            hipLaunchKernelGGL(k_trace_closest, grid, dim3(kBlock), 0, s->stream, s->dev, s->rays.p, count, s->hits.p);
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipEventRecord(e1, s->stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (seconds) *seconds = ms * 1e-3;
        if (any) HIP_CHECK(hipMemcpy(out, s->occluded.p, count, hipMemcpyDeviceToHost));
        else HIP_CHECK(hipMemcpy(out, s->hits.p, count * sizeof(rp_hit), hipMemcpyDeviceToHost));
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_trace", e);
    }
}

int rp_trace_closest(rp_scene *s, const rp_ray *rays, uint64_t count, rp_hit *hits, double *seconds) {
    return trace_common(s, rays, count, hits, false, seconds);
}
int rp_trace_any(rp_scene *s, const rp_ray *rays, uint64_t count, uint8_t *occluded, double *seconds) {
    return trace_common(s, rays, count, occluded, true, seconds);
}

} // extern "C"

// Entry points whose implementation lives in later translation units are
// declared in include/rp_api.h; the ones below are not available yet.
extern "C" {
int rp_batch_create(const rp_batch_item *, uint32_t, rp_batch **out) {
    if (out) *out = nullptr;
    set_error("rp_batch_create: not available in this build");
    return RP_NOT_BUILT;
}
int rp_batch_render(rp_batch *, const rp_batch_config *, float **, rp_batch_stats *) {
    set_error("rp_batch_render: not available in this build");
    return RP_NOT_BUILT;
}
void rp_batch_destroy(rp_batch *) {}
int rp_autotune(rp_batch *, rp_batch_config *, double *) {
    set_error("rp_autotune: not available in this build");
    return RP_NOT_BUILT;
}
int rp_set_params(rp_scene *, const rp_param *, uint32_t, const float *) {
    set_error("rp_set_params: not available in this build");
    return RP_NOT_BUILT;
}
int rp_backward(rp_scene *, const rp_frame_config *, const rp_param *, uint32_t, const float *, float *,
                rp_frame_stats *) {
    set_error("rp_backward: not available in this build");
    return RP_NOT_BUILT;
}
}
