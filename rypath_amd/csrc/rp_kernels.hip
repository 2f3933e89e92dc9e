// HIP kernels and the C ABI (include/rp_api.h) of the gfx950 path tracer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rp_io.h"
#include "rp_internal.hpp"
#include "rp_path.h"
#include "rp_wavefront.h"
#include "rp_batch.h"
#include "rp_adjoint.h"
#include "rp_scalar.h"

static_assert(sizeof(rp_frame_config) == 48 && sizeof(rp_frame_stats) == 144 && sizeof(rp_batch_config) == 32 &&
                  sizeof(rp_request) == 72 && sizeof(rp_path_record) == 112,
              "C ABI struct sizes");

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

struct FrameSession;

} // namespace rp

using namespace rp;

struct rp_scene {
    HostScene host;
    uint64_t serial = 0;         // unique per scene in the process
    uint64_t params_version = 0; // bumped by rp_set_params
    Bvh4 bvh;
    rp_build_info info{};
    int device = 0;
    DevBuf<float4> tris;
    DevBuf<uint32_t> tri_material; // material index per triangle, BVH leaf order (sort keys)
    DevBuf<QNode> nodes;
    DevBuf<rp_object> objects;
    DevBuf<rp_material> materials;
    DevBuf<rp_emission> emissions;
    DevBuf<rp_medium> media;
    DevBuf<rp_spot_light> spots;
    DevBuf<float> spectra;
    DevBuf<rpd::LightEntry> lights;
    DevBuf<rpd::LightNode> light_nodes;
    DevBuf<float4> cie;
    DevBuf<uint32_t> sobol_dir;
    DevBuf<uint64_t> sobol_intervals, sobol_inverse;
    rpd::Scene dev{};
    // frame scratch
    DevBuf<float> path_xyz;
    DevBuf<double> film;
    DevBuf<rp_path_record> records;
    DevBuf<rp_render_sample> scalar_records;
    DevBuf<unsigned long long> counters;
    DevBuf<rp_ray> rays;
    DevBuf<rp_hit> hits;
    DevBuf<uint8_t> occluded;
    hipStream_t stream = nullptr;
    // Serialises device work on this scene between API calls and the
    // asynchronous frame worker.
    std::mutex exec;
    std::shared_ptr<rp::FrameSession> frames;
    std::array<uint8_t, 32> identity{}; // scene_identity(desc), computed on first use
    bool has_identity = false;
    std::string device_name, device_arch; // provenance of film segments
    ~rp_scene();
};

namespace {

constexpr int kBlock = 128;

template <int B> struct LdsStack {
    static __device__ __forceinline__ void init(rpd::Stack &st, uint32_t *n, float *t) {
        st.init(n, t, B);
    }
};

using rpd::kCounterStride;
__device__ __forceinline__ void count_rays(unsigned long long *counters, uint32_t ext, uint32_t sh, uint32_t failed) {
    // wave-level reduction, one atomic per wave per counter
    unsigned long long e = ext, s = sh, f = failed;
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_down(e, off);
        s += __shfl_down(s, off);
        f += __shfl_down(f, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (e) atomicAdd(&counters[0], e);
        if (s) atomicAdd(&counters[kCounterStride], s);
        if (f) atomicAdd(&counters[2 * kCounterStride], f);
    }
}

// Per-path records; B = the request's workgroup width (the records do not
// depend on it).
template <int B>
__global__ void __launch_bounds__(B) RP_PATH_ATTR k_records(rpd::Scene S, rpd::PathParams P, uint32_t pixel_begin,
                                               uint32_t pixel_count, uint64_t sample_begin, uint64_t n,
                                               rp_path_record *out) {
    __shared__ uint32_t s_node[rpd::kShortStack * B];
    __shared__ float s_t[rpd::kShortStack * B];
    RP_STACK(st);
    LdsStack<B>::init(st, s_node, s_t);
    uint64_t lane = uint64_t(blockIdx.x) * B + threadIdx.x;
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
    rec.filter_weight = 1.0f; // box pixel filter: every sample has unit weight
    rec.splat_begin = 0;      // no light-tracing splats on this path
    rec.splat_count = 0;
    out[lane] = rec;
}

// Megakernel: one thread = one complete path; per-path XYZ written planar
// (x plane, y plane, z plane) for the film reduction.
__global__ void __launch_bounds__(kBlock) RP_PATH_ATTR k_mega(rpd::Scene S, rpd::PathParams P, uint32_t npx, uint64_t sample_begin,
                                                  uint64_t n, float *xyz, unsigned long long *counters) {
    __shared__ uint32_t s_node[rpd::kShortStack * kBlock];
    __shared__ float s_t[rpd::kShortStack * kBlock];
    RP_STACK(st);
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

// Scalar-transport megakernel: one thread = one scalar path; per-path film
// weight planar as in k_mega.
__global__ void __launch_bounds__(kBlock) RP_PATH_ATTR k_scalar_mega(rpd::Scene S, rpd::PathParams P, uint32_t npx,
                                                         uint64_t sample_begin, uint64_t n, uint32_t independent,
                                                         float *xyz, unsigned long long *counters) {
    __shared__ uint32_t s_node[rpd::kShortStack * kBlock];
    __shared__ float s_t[rpd::kShortStack * kBlock];
    RP_STACK(st);
    LdsStack<kBlock>::init(st, s_node, s_t);
    uint64_t lane = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    uint32_t ext = 0, sh = 0, failed = 0;
    if (lane < n) {
        rpd::ScalarResult r;
        rpd::trace_scalar<true>(S, P, st, uint32_t(lane % npx), sample_begin + lane / npx, independent != 0, r);
        float v[3];
        rpd::scalar_xyz(S, r, v);
        failed = r.status != RP_PATH_SUCCESS;
        xyz[lane] = v[0];
        xyz[n + lane] = v[1];
        xyz[2 * n + lane] = v[2];
        ext = r.ext_rays;
        sh = r.shadow_rays;
    }
    count_rays(counters, ext, sh, failed);
}

// Scalar RenderSample records; B = the request's workgroup width.
template <int B>
__global__ void __launch_bounds__(B) k_scalar_records(rpd::Scene S, rpd::PathParams P, uint32_t pixel_begin,
                                                      uint32_t pixel_count, uint64_t sample_begin, uint64_t n,
                                                      uint32_t independent, rp_render_sample *out) {
    __shared__ uint32_t s_node[rpd::kShortStack * B];
    __shared__ float s_t[rpd::kShortStack * B];
    RP_STACK(st);
    LdsStack<B>::init(st, s_node, s_t);
    uint64_t lane = uint64_t(blockIdx.x) * B + threadIdx.x;
    if (lane >= n) return;
    const uint32_t pixel = pixel_begin + uint32_t(lane % pixel_count);
    const uint64_t sample = sample_begin + lane / pixel_count;
    rpd::ScalarResult r;
    rpd::trace_scalar<true>(S, P, st, pixel, sample, independent != 0, r);
    rp_render_sample rec;
    rec.pixel = pixel;
    rec.status = r.status;
    rec.sample = sample;
    rec.wavelength = r.wavelength;
    rec.radiance = r.radiance;
    rec.scatterings = r.scatterings;
    for (int k = 0; k < 4; ++k) rec.random_words[k] = r.words[k];
    rpd::scalar_xyz(S, r, rec.xyz);
    out[lane] = rec;
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

// Statistics counters live one per 256-B line (counter k at
// c[k * kCounterStride]): atomics on one line are serialised.
__global__ void k_zero_counters(unsigned long long *c, uint32_t n) {
    if (threadIdx.x < n) c[threadIdx.x * kCounterStride] = 0;
}
// Copies counters [0, n) of a strided counter block to the host.
void read_counters(const unsigned long long *dev, uint32_t n, unsigned long long *out) {
    std::vector<unsigned long long> tmp(size_t(n) * kCounterStride);
    HIP_CHECK(hipMemcpy(tmp.data(), dev, tmp.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < n; ++k) out[k] = tmp[size_t(k) * kCounterStride];
}

__global__ void k_film_zero(uint64_t n, double *film) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) film[i] = 0.0;
}

__global__ void k_film_finish(uint32_t npx, uint64_t samples, const double *film, float *out) {
    uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= npx) return;
    for (int c = 0; c < 3; ++c)
        out[3 * size_t(pixel) + c] =
            float((film[(2 * c) * size_t(npx) + pixel] + film[(2 * c + 1) * size_t(npx) + pixel]) / double(samples));
}

__global__ void __launch_bounds__(kBlock) k_trace_closest(rpd::Scene S, const rp_ray *rays, uint64_t n, rp_hit *hits) {
    __shared__ uint32_t s_node[rpd::kShortStack * kBlock];
    __shared__ float s_t[rpd::kShortStack * kBlock];
    RP_STACK(st);
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
    RP_STACK(st);
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
    LightTables lt = build_light_tables(d, radius, tri_light);
    const std::vector<LightEntry> &lights = lt.lights;
    std::vector<uint32_t> leaf_pos(d.triangle_count);
    for (uint32_t p = 0; p < d.triangle_count; ++p) leaf_pos[s->bvh.tri_order[p]] = p;
    std::vector<float4> tris(3 * size_t(d.triangle_count));
    std::vector<uint32_t> tri_material(std::max(d.triangle_count, 1u), 0u);
    for (uint32_t p = 0; p < d.triangle_count; ++p) {
        uint32_t t = s->bvh.tri_order[p];
        tri_material[p] = d.objects[d.tri_object[t]].material;
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
        const LightEntry &h = lights[i];
        rpd::LightEntry &e = dl[i];
        e = rpd::LightEntry{};
        e.kind = h.kind;
        // area lights address their (first) triangle by its leaf position
        e.index = (h.kind == 0 || h.kind == 3) ? leaf_pos[h.index] : h.index;
        e.leaf = h.leaf;
        e.rect = h.rect;
        e.pmf = h.pmf;
        e.threshold = h.threshold;
        e.alias = h.alias;
        e.inv_area = h.inv_area;
        for (int k = 0; k < 3; ++k) {
            e.p00[k] = h.p00[k];
            e.ex[k] = h.ex[k];
            e.ey[k] = h.ey[k];
            e.ux[k] = h.ux[k];
            e.uy[k] = h.uy[k];
            e.uz[k] = h.uz[k];
        }
        e.lx = h.lx;
        e.ly = h.ly;
    }
    std::vector<rpd::LightNode> dn(lt.nodes.size());
    for (size_t i = 0; i < dn.size(); ++i) {
        const LightNodeHost &h = lt.nodes[i];
        rpd::LightNode &e = dn[i];
        // bounding sphere of the box in binary32: centre (lo + hi) / 2,
        // squared radius |hi - lo|^2 / 4 (the expressions the oracle evaluates)
        const float dx = h.high[0] - h.low[0], dy = h.high[1] - h.low[1], dz = h.high[2] - h.low[2];
        for (int k = 0; k < 3; ++k) {
            e.center[k] = (h.low[k] + h.high[k]) * 0.5f;
            e.dir[k] = h.dir[k];
        }
        e.r2 = (dx * dx + dy * dy + dz * dz) * 0.25f;
        e.radius = std::sqrt(e.r2);
        e.cos_outer = h.cos_outer;
        e.sin_outer = std::sqrt(std::max(0.0f, 1.0f - h.cos_outer * h.cos_outer));
        e.cos_emission = h.cos_emission;
        e.power = h.power;
        e.child = h.child;
        e.parent = h.parent;
        e.light = h.light;
        e.flags = h.flags;
    }
    s->tris.upload(tris);
    s->tri_material.upload(tri_material);
    s->nodes.upload(s->bvh.nodes);
    s->objects.upload(s->host.objects);
    s->materials.upload(s->host.materials);
    s->emissions.upload(s->host.emissions);
    s->media.upload(s->host.media);
    s->spots.upload(s->host.spots);
    s->spectra.upload(s->host.spectra);
    s->lights.upload(dl);
    s->light_nodes.upload(dn);
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
    s->counters.alloc(8 * kCounterStride);

    rpd::Scene &S = s->dev;
    S.tris = s->tris.p;
    S.tri_material = s->tri_material.p;
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
    S.light_nodes = s->light_nodes.p;
    S.light_node_count = uint32_t(dn.size());
    S.light_sampler = d.light_sampler;
    S.env_light = lt.env_light;
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
    S.transport = d.transport;
    S.env_pdf = lt.env_light != RP_NONE ? lt.env_pmf * rpd::kInv4Pi : 0.0f;
    s->info.emitter_count = 0;
    for (auto &l : dl)
        if (l.kind == 0 || l.kind == 3) ++s->info.emitter_count;
    s->info.light_count = uint32_t(dl.size());
    s->info.device_bytes = s->tris.n * sizeof(float4) + s->tri_material.n * 4 + s->nodes.n * sizeof(QNode) +
                           s->spectra.n * 4 +
                           s->sobol_dir.n * 4 + s->lights.n * sizeof(rpd::LightEntry) +
                           s->light_nodes.n * sizeof(rpd::LightNode);
}

rpd::PathParams path_params(uint64_t seed, uint32_t randomization, uint32_t emergency, uint32_t boundary,
                            uint32_t frame_flags = 0) {
    rpd::PathParams P;
    P.seed = seed;
    P.mode = randomization;
    P.emergency = emergency ? emergency : 1024u;
    P.boundary = boundary ? boundary : 4096u;
    P.roulette = (frame_flags & RP_FRAME_NO_ROULETTE) ? 0u : 1u;
    return P;
}

// Wavefront buffers shared by every scene on a device (≈ 500 B per path in
// flight, up to 2^26 paths per chunk): one arena per device instead of one
// per scene keeps many resident scenes from multiplying that footprint.
// Frames of different scenes on one device take turns on it.  The arenas
// live until the process ends.
// The arena's per-lane radiance and status arrays hold the primal of the
// last wavefront render on the device; `primal` says which one, so that a
// path-replay adjoint right after a wavefront forward of the same frame
// (same scene, parameters and frame configuration) can skip its own primal.
struct PrimalTag {
    bool valid = false;
    uint64_t scene = 0, params_version = 0, seed = 0, sample_begin = 0, paths = 0;
    uint32_t randomization = 0, spp = 0, roulette_off = 0, npx = 0;
    uint64_t rays[2] = {0, 0};
    uint64_t failed = 0; // paths whose status is not success
    bool same(const rp_scene *s, const rp_frame_config &c, uint32_t px) const {
        return valid && scene == s->serial && params_version == s->params_version && seed == c.seed &&
               sample_begin == c.sample_begin && spp == c.spp && randomization == c.randomization &&
               roulette_off == ((c.flags & RP_FRAME_NO_ROULETTE) ? 1u : 0u) && npx == px &&
               paths == uint64_t(px) * c.spp;
    }
    void set(const rp_scene *s, const rp_frame_config &c, uint32_t px, const uint64_t r[2]) {
        valid = true;
        scene = s->serial;
        params_version = s->params_version;
        seed = c.seed;
        sample_begin = c.sample_begin;
        spp = c.spp;
        randomization = c.randomization;
        roulette_off = (c.flags & RP_FRAME_NO_ROULETTE) ? 1u : 0u;
        npx = px;
        paths = uint64_t(px) * c.spp;
        rays[0] = r[0];
        rays[1] = r[1];
        failed = 0;
    }
};
struct FrameArena {
    std::mutex m;
    rpd::FrameBuffers fb;
    PrimalTag primal;
    rpd::AdjBuffers adj; // path-replay scratch
};
FrameArena &frame_arena(int device) {
    static std::mutex guard;
    static std::vector<FrameArena *> *arenas = new std::vector<FrameArena *>();
    std::lock_guard<std::mutex> g(guard);
    if (device < 0) throw std::runtime_error("invalid device");
    if (size_t(device) >= arenas->size()) arenas->resize(size_t(device) + 1, nullptr);
    if (!(*arenas)[size_t(device)]) (*arenas)[size_t(device)] = new FrameArena();
    return *(*arenas)[size_t(device)];
}

// Continuation of a running film: the compensated sums (6 planes: sum and
// correction per XYZ channel) after `samples` samples per pixel.  Adding the
// next samples in order continues the same Neumaier recurrence, so a resumed
// render equals an uninterrupted one bit for bit.
struct FilmCarry {
    std::vector<double> *planes; // in: state before the frame; out: state after
    uint64_t samples;            // samples per pixel already in `planes`
};

// Renders the frame into the scene's film buffer; returns seconds (device events).
double render_frame_impl(rp_scene *s, const rp_frame_config &cfg, float *xyz_device, rp_frame_stats *stats,
                         FilmCarry *carry = nullptr) {
    const rp_camera &cam = s->host.desc.camera;
    const uint32_t npx = cam.width * cam.height;
    const uint32_t spp = cfg.spp;
    if (spp == 0) throw std::runtime_error("spp must be positive");
    if (cfg.mode > RP_MODE_WAVEFRONT_SORTED) throw std::runtime_error("unknown mode");
    if (cfg.randomization > RP_RANDOMIZE_OWEN) throw std::runtime_error("unknown randomization");
    const bool scalar = s->host.desc.transport != RP_TRANSPORT_HERO;
    if (scalar && cfg.mode != RP_MODE_MEGAKERNEL)
        throw std::runtime_error("scalar transport renders in RP_MODE_MEGAKERNEL");
    if (!scalar && (cfg.flags & RP_FRAME_INDEPENDENT_SAMPLER))
        throw std::runtime_error("the independent sampler is for scalar transport");
    const uint32_t independent = (cfg.flags & RP_FRAME_INDEPENDENT_SAMPLER) ? 1u : 0u;
    HIP_CHECK(hipSetDevice(s->device));
    rpd::PathParams P = path_params(cfg.seed, cfg.randomization, 0, 0, cfg.flags);
    const uint64_t max_chunk_paths = uint64_t(1) << 26;
    uint32_t chunk = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(spp, max_chunk_paths / npx)));
    s->film.alloc(6 * size_t(npx));
    if (carry && carry->samples) {
        if (carry->planes->size() != 6 * size_t(npx)) throw std::runtime_error("film size does not match the scene");
        HIP_CHECK(hipMemcpyAsync(s->film.p, carry->planes->data(), 6 * size_t(npx) * sizeof(double),
                                 hipMemcpyHostToDevice, s->stream));
    } else {
        hipLaunchKernelGGL(k_film_zero, dim3(uint32_t((6 * uint64_t(npx) + 255) / 256)), dim3(256), 0, s->stream,
                           6 * uint64_t(npx), s->film.p);
    }
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s->stream, s->counters.p, 8u);
    rpd::EventPair ev; // destroyed on every exit, exceptions included
    hipEvent_t e0 = ev.a, e1 = ev.b;
    HIP_CHECK(hipEventRecord(e0, s->stream));
    double trace_seconds = 0, shadow[2] = {0, 0};
    uint64_t launches = 0, wf_rays[2] = {0, 0}, fused[2] = {0, 0};
    // the device's wavefront arena is held until this frame's last kernel has
    // finished (the event synchronisation below happens before the unlock)
    std::unique_lock<std::mutex> arena_lock;
    FrameArena *arena = nullptr;
    if (cfg.mode != RP_MODE_MEGAKERNEL) {
        arena = &frame_arena(s->device);
        arena_lock = std::unique_lock<std::mutex>(arena->m);
    }
    for (uint32_t s0 = 0; s0 < spp; s0 += chunk) {
        uint32_t ns = std::min(chunk, spp - s0);
        uint64_t n = uint64_t(ns) * npx;
        s->path_xyz.alloc(3 * n);
        if (scalar) {
            hipLaunchKernelGGL(k_scalar_mega, dim3(uint32_t((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream,
                               s->dev, P, npx, cfg.sample_begin + s0, n, independent, s->path_xyz.p, s->counters.p);
        } else if (cfg.mode == RP_MODE_MEGAKERNEL) {
            hipLaunchKernelGGL(k_mega, dim3(uint32_t((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream, s->dev, P,
                               npx, cfg.sample_begin + s0, n, s->path_xyz.p, s->counters.p);
        } else {
            arena->primal.valid = false;
            trace_seconds += rpd::wavefront_render(s->dev, P, arena->fb, cfg, npx, cfg.sample_begin + s0, n,
                                                   s->path_xyz.p, s->counters.p, s->stream, launches, wf_rays,
                                                   s0 == 0, shadow, fused);
        }
        HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_film_add, dim3((npx + 255) / 256), dim3(256), 0, s->stream, npx, ns, n, s->path_xyz.p,
                           s->film.p);
    }
    // one chunk: the arena now holds this frame's primal for an adjoint
    if (arena && chunk >= spp) arena->primal.set(s, cfg, npx, wf_rays);
    const uint64_t total = (carry ? carry->samples : 0) + spp;
    if (xyz_device)
        hipLaunchKernelGGL(k_film_finish, dim3((npx + 255) / 256), dim3(256), 0, s->stream, npx, total, s->film.p,
                           xyz_device);
    HIP_CHECK(hipEventRecord(e1, s->stream));
    if (carry) {
        carry->planes->resize(6 * size_t(npx));
        HIP_CHECK(hipMemcpyAsync(carry->planes->data(), s->film.p, 6 * size_t(npx) * sizeof(double),
                                 hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
        carry->samples = total;
    }
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    HIP_CHECK(hipGetLastError());
    unsigned long long c[8] = {};
    if (stats || (arena && arena->primal.valid)) read_counters(s->counters.p, 8, c);
    if (arena && arena->primal.valid) arena->primal.failed = c[2]; // reported again when an adjoint reuses it
    if (stats) {
        uint64_t tail[2] = {0, 0};
        if (arena) arena->fb.tail_rays(tail);
        stats->seconds = ms * 1e-3;
        stats->paths = uint64_t(npx) * spp;
        stats->extension_rays = c[0] + wf_rays[0] + tail[0];
        stats->shadow_rays = c[1] + wf_rays[1] + tail[1];
        stats->extend_rays = wf_rays[0];
        stats->failed_paths = c[2];
        stats->trace_seconds = trace_seconds;
        stats->extend_launches = launches;
        stats->node_fetches = c[3];
        stats->triangle_tests = c[4];
        stats->shadow_node_fetches = c[5];
        stats->shadow_triangle_tests = c[6];
        stats->stack_spills = c[7];
        stats->shadow_seconds = shadow[0];
        stats->shadow_launches = uint64_t(shadow[1]);
        stats->shadow_kernel_rays = wf_rays[1];
        stats->fused_launches = fused[0];
        stats->fused_shadow_rays = fused[1];
    }
    return ms * 1e-3;
}

// SHA-256 identity of the scene's current description; the caller holds
// s->exec (rp_set_params rewrites the hashed materials, spectra and media
// under the same lock and drops the cached value).
std::array<uint8_t, 32> identity_locked(rp_scene *s) {
    if (!s->has_identity) {
        s->identity = scene_identity(s->host.desc);
        s->has_identity = true;
    }
    return s->identity;
}

// Whether film f can continue scene s with (seed, randomization): the
// resume contract of src/render_job.cpp:125-140 (size, sampler, identity).
// The caller holds s->exec.
// sampler: the rp_sampler_kind about to render, or UINT32_MAX to leave it unchecked
bool film_compatible(rp_scene *s, const rp_film *f, uint64_t seed, uint32_t randomization, const char *who,
                     uint32_t sampler = UINT32_MAX) {
    const rp_camera &cam = s->host.desc.camera;
    if (f->width != cam.width || f->height != cam.height) {
        set_error(std::string(who) + ": the film is " + std::to_string(f->width) + "x" + std::to_string(f->height) +
                  ", the scene's camera " + std::to_string(cam.width) + "x" + std::to_string(cam.height));
        return false;
    }
    if (f->samples == 0) return true;
    if (seed != f->seed || randomization != f->randomization) {
        set_error(std::string(who) + ": seed and randomization must match the film's");
        return false;
    }
    if (sampler != UINT32_MAX && sampler != f->sampler) {
        set_error(std::string(who) + ": the film holds samples of another sampler");
        return false;
    }
    if (identity_locked(s) != f->identity) {
        set_error(std::string(who) + ": the film belongs to another scene (identity mismatch)");
        return false;
    }
    return true;
}

} // namespace

// ------------------------------------------------------------- asynchronous frames
// FrameSession::submit / Frame::poll / Frame::wait (src/frame_session.hpp:145-212):
// one worker thread per scene renders submitted frames in order on the
// scene's stream; a retained frame holds one of two slots until released,
// and frames stay valid (film included) after the scene is destroyed.
namespace rp {

struct FrameJob {
    rp_frame_config cfg{};
    uint32_t width = 0, height = 0;
    DevBuf<float> film;
    std::mutex m;
    std::condition_variable cv;
    uint32_t state = 0; // 0 pending, 1 complete, 2 failed
    rp_frame_stats stats{};
    std::string error;
    std::shared_ptr<FrameSession> session;
};

struct FrameSession {
    rp_scene *scene;
    std::mutex m;
    std::condition_variable cv;
    std::deque<std::shared_ptr<FrameJob>> queue;
    uint32_t retained = 0;
    bool stopping = false;
    std::thread worker;

    explicit FrameSession(rp_scene *s) : scene(s) {
        worker = std::thread([this] { run(); });
    }
    void run() {
        for (;;) {
            std::shared_ptr<FrameJob> job;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return stopping || !queue.empty(); });
                if (queue.empty()) return;
                job = queue.front();
                queue.pop_front();
            }
            uint32_t state = 1;
            std::string error;
            rp_frame_stats st{};
            try {
                std::lock_guard<std::mutex> g(scene->exec);
                render_frame_impl(scene, job->cfg, job->film.p, &st);
            } catch (const std::exception &e) {
                state = 2;
                error = e.what();
            }
            {
                std::lock_guard<std::mutex> lk(job->m);
                job->stats = st;
                job->error = error;
                job->state = state;
            }
            job->cv.notify_all();
        }
    }
    // Renders every queued frame, then stops the worker.
    void shutdown() {
        {
            std::lock_guard<std::mutex> lk(m);
            stopping = true;
        }
        cv.notify_all();
        if (worker.joinable()) worker.join();
        scene = nullptr;
    }
};

} // namespace rp

rp_scene::~rp_scene() {
    if (frames) frames->shutdown();
    if (stream) (void)hipStreamDestroy(stream);
}

struct rp_frame {
    std::shared_ptr<rp::FrameJob> job;
};

// ---------------------------------------------------- allocation-free frames
// rp_render_frame_device_noalloc keeps the frame's statistics counters, its
// compensated film planes and the per-path XYZ in `scratch`, a file-scope
// device array (one copy per GPU, held until the process ends), instead of
// the scene's device buffers.  In host code `scratch` names a host-side
// stand-in: kernels get the device address that hipGetSymbolAddress returns
// on the current device, looked up once per device.
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
constexpr size_t kNoAllocCounterBytes = 8 * kCounterStride * sizeof(unsigned long long); // 2 KiB
constexpr size_t kNoAllocFilmBytesPerPixel = 6 * sizeof(double);
constexpr size_t kNoAllocXyzBytesPerPath = 3 * sizeof(float);
constexpr int kNoAllocMaxDevices = 64;

} // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct NoAllocSlot {
    std::mutex m;                // serialises this process's calls on the device
    unsigned char *ptr = nullptr; // device address of `scratch` on the device
};
NoAllocSlot g_noalloc[kNoAllocMaxDevices];

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// The caller holds the slot's lock with `device` current.
unsigned char *noalloc_scratch(NoAllocSlot &slot) {
    if (!slot.ptr) {
        void *p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned a null scratch address");
        slot.ptr = static_cast<unsigned char *>(p);
    }
    return slot.ptr;
}

// RenderRequest validation (src/metal_render.mm:1930-1946, src/sampling.hpp:29-31):
// raster size, sampler kind and randomization, the pixel and sample ranges
// (and the fixed sample budget when given), and the workgroup width.
bool request_valid(const rp_scene *s, const rp_request &rq, const char *who, bool scalar = false) {
    const rp_camera &cam = s->host.desc.camera;
    const uint32_t npx = cam.width * cam.height;
    auto bad = [&](const std::string &why) {
        set_error(std::string(who) + ": " + why);
        return false;
    };
    if ((rq.width || rq.height) && (rq.width != cam.width || rq.height != cam.height))
        return bad("request raster " + std::to_string(rq.width) + "x" + std::to_string(rq.height) +
                   " differs from the scene camera's " + std::to_string(cam.width) + "x" + std::to_string(cam.height));
    if (scalar ? (s->host.desc.transport == RP_TRANSPORT_HERO) : (s->host.desc.transport != RP_TRANSPORT_HERO))
        return bad(scalar ? "the scene uses hero transport (rp_sample)" : "the scene uses scalar transport (rp_sample_scalar)");
    if (rq.sampler_kind != RP_SAMPLER_SOBOL && !(scalar && rq.sampler_kind == RP_SAMPLER_INDEPENDENT))
        return bad(scalar ? "sampler kind not supported by this backend (Sobol or independent)"
                          : "sampler kind not supported by this backend (Sobol only)");
    if (rq.randomization > RP_RANDOMIZE_OWEN) return bad("unknown randomization");
    if (rq.pixel_count == 0 || rq.sample_count == 0 || uint64_t(rq.pixel_begin) + rq.pixel_count > npx)
        return bad("pixel range outside the film or empty request");
    if (rq.sample_begin + rq.sample_count < rq.sample_begin) return bad("sample range overflows");
    if (rq.total_samples &&
        (rq.sample_begin >= rq.total_samples || rq.sample_count > rq.total_samples - rq.sample_begin))
        return bad("the fixed total sample budget does not contain the sample range");
    if (rq.threadgroup_width != 0 && rq.threadgroup_width != 64 && rq.threadgroup_width != 128 &&
        rq.threadgroup_width != 256)
        return bad("threadgroup width exceeds the kernel's limit (0, 64, 128 or 256)");
    return true;
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
        static std::atomic<uint64_t> serials{0};
        s->serial = ++serials;
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
    if (!request_valid(s, *rq, "rp_sample")) return RP_INVALID;
    uint64_t n = uint64_t(rq->pixel_count) * rq->sample_count;
    if (capacity < n) {
        set_error("rp_sample: output capacity too small");
        return RP_INVALID;
    }
    try {
        std::lock_guard<std::mutex> lock(s->exec);
        HIP_CHECK(hipSetDevice(s->device));
        s->records.alloc(n);
        rpd::PathParams P = path_params(rq->seed, rq->randomization, rq->emergency_limit, rq->boundary_limit);
        const uint32_t B = rq->threadgroup_width ? rq->threadgroup_width : uint32_t(kBlock);
        const dim3 grid(uint32_t((n + B - 1) / B));
        if (B == 64)
            hipLaunchKernelGGL(k_records<64>, grid, dim3(64), 0, s->stream, s->dev, P, rq->pixel_begin,
                               rq->pixel_count, rq->sample_begin, n, s->records.p);
        else if (B == 256)
            hipLaunchKernelGGL(k_records<256>, grid, dim3(256), 0, s->stream, s->dev, P, rq->pixel_begin,
                               rq->pixel_count, rq->sample_begin, n, s->records.p);
        else
            hipLaunchKernelGGL(k_records<128>, grid, dim3(128), 0, s->stream, s->dev, P, rq->pixel_begin,
                               rq->pixel_count, rq->sample_begin, n, s->records.p);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(out, s->records.p, n * sizeof(rp_path_record), hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_sample", e);
    }
}

int rp_sample_scalar(rp_scene *s, const rp_request *rq, rp_render_sample *out, uint64_t capacity) {
    if (!s || !rq || !out) {
        set_error("rp_sample_scalar: null argument");
        return RP_INVALID;
    }
    if (!request_valid(s, *rq, "rp_sample_scalar", true)) return RP_INVALID;
    const uint64_t n = uint64_t(rq->pixel_count) * rq->sample_count;
    if (capacity < n) {
        set_error("rp_sample_scalar: output capacity too small");
        return RP_INVALID;
    }
    try {
        std::lock_guard<std::mutex> lock(s->exec);
        HIP_CHECK(hipSetDevice(s->device));
        s->scalar_records.alloc(n);
        rpd::PathParams P = path_params(rq->seed, rq->randomization, rq->emergency_limit, rq->boundary_limit);
        const uint32_t B = rq->threadgroup_width ? rq->threadgroup_width : uint32_t(kBlock);
        const uint32_t independent = rq->sampler_kind == RP_SAMPLER_INDEPENDENT ? 1u : 0u;
        const dim3 grid(uint32_t((n + B - 1) / B));
        if (B == 64)
            hipLaunchKernelGGL(k_scalar_records<64>, grid, dim3(64), 0, s->stream, s->dev, P, rq->pixel_begin,
                               rq->pixel_count, rq->sample_begin, n, independent, s->scalar_records.p);
        else if (B == 256)
            hipLaunchKernelGGL(k_scalar_records<256>, grid, dim3(256), 0, s->stream, s->dev, P, rq->pixel_begin,
                               rq->pixel_count, rq->sample_begin, n, independent, s->scalar_records.p);
        else
            hipLaunchKernelGGL(k_scalar_records<128>, grid, dim3(128), 0, s->stream, s->dev, P, rq->pixel_begin,
                               rq->pixel_count, rq->sample_begin, n, independent, s->scalar_records.p);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(out, s->scalar_records.p, n * sizeof(rp_render_sample), hipMemcpyDeviceToHost,
                                 s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_sample_scalar", e);
    }
}

int rp_render_frame_device(rp_scene *s, const rp_frame_config *cfg, float *xyz_device, rp_frame_stats *stats) {
    if (!s || !cfg || !xyz_device) {
        set_error("rp_render_frame_device: null argument");
        return RP_INVALID;
    }
    try {
        std::lock_guard<std::mutex> lock(s->exec);
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
        std::lock_guard<std::mutex> lock(s->exec);
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

int rp_render_frame_device_noalloc(rp_scene *s, const rp_frame_config *cfg, float *xyz_device,
                                   rp_frame_stats *stats) {
    static const char *who = "rp_render_frame_device_noalloc";
    if (!s || !cfg || !xyz_device) {
        set_error(std::string(who) + ": null argument");
        return RP_INVALID;
    }
    if (cfg->mode != RP_MODE_MEGAKERNEL) {
        set_error(std::string(who) + ": only RP_MODE_MEGAKERNEL runs without allocation");
        return RP_INVALID;
    }
    if (cfg->spp == 0 || cfg->randomization > RP_RANDOMIZE_OWEN) {
        set_error(std::string(who) + ": spp must be positive and randomization known");
        return RP_INVALID;
    }
    if (s->host.desc.transport == RP_TRANSPORT_HERO && (cfg->flags & RP_FRAME_INDEPENDENT_SAMPLER)) {
        set_error(std::string(who) + ": the independent sampler is for scalar transport");
        return RP_INVALID;
    }
    const rp_camera &cam = s->host.desc.camera;
    const uint64_t npx = uint64_t(cam.width) * cam.height;
    const size_t film_off = kNoAllocCounterBytes;
    const size_t xyz_off = align256(film_off + kNoAllocFilmBytesPerPixel * npx);
    if (npx == 0 || xyz_off + kNoAllocXyzBytesPerPath * npx > kNoAllocScratchBytes) {
        set_error(std::string(who) + ": a " + std::to_string(cam.width) + "x" + std::to_string(cam.height) +
                  " film needs " + std::to_string(xyz_off + kNoAllocXyzBytesPerPath * npx) +
                  " bytes of scratch; at most " + std::to_string(kNoAllocScratchBytes) +
                  " (60 bytes per pixel plus 2 KiB)");
        return RP_INVALID;
    }
    if (s->device < 0 || s->device >= kNoAllocMaxDevices) {
        set_error(std::string(who) + ": device index beyond " + std::to_string(kNoAllocMaxDevices - 1));
        return RP_INVALID;
    }
    try {
        std::lock_guard<std::mutex> lock(s->exec);
        NoAllocSlot &slot = g_noalloc[s->device];
        std::lock_guard<std::mutex> slot_lock(slot.m);
        HIP_CHECK(hipSetDevice(s->device));
        unsigned char *base = noalloc_scratch(slot);
        auto *counters = reinterpret_cast<unsigned long long *>(base);
        auto *film = reinterpret_cast<double *>(base + film_off);
        auto *xyz = reinterpret_cast<float *>(base + xyz_off);
        const auto t0 = std::chrono::steady_clock::now();
        // samples per chunk: what the per-path XYZ region holds (at least one,
        // by the size check), at most 2^26 paths
        const uint64_t room = (kNoAllocScratchBytes - xyz_off) / (kNoAllocXyzBytesPerPath * npx);
        const uint32_t chunk =
            uint32_t(std::max<uint64_t>(1, std::min<uint64_t>({cfg->spp, room, (uint64_t(1) << 26) / npx})));
        rpd::PathParams P = path_params(cfg->seed, cfg->randomization, 0, 0, cfg->flags);
        // the previous call's contents stay in the array: clear what starts at zero
        hipLaunchKernelGGL(k_film_zero, dim3(uint32_t((6 * npx + 255) / 256)), dim3(256), 0, s->stream, 6 * npx, film);
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s->stream, counters, 8u);
        for (uint32_t s0 = 0; s0 < cfg->spp; s0 += chunk) {
            const uint32_t ns = std::min(chunk, cfg->spp - s0);
            const uint64_t n = uint64_t(ns) * npx;
            if (s->host.desc.transport != RP_TRANSPORT_HERO)
                hipLaunchKernelGGL(k_scalar_mega, dim3(uint32_t((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream,
                                   s->dev, P, uint32_t(npx), cfg->sample_begin + s0, n,
                                   (cfg->flags & RP_FRAME_INDEPENDENT_SAMPLER) ? 1u : 0u, xyz, counters);
            else
                hipLaunchKernelGGL(k_mega, dim3(uint32_t((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s->stream,
                                   s->dev, P, uint32_t(npx), cfg->sample_begin + s0, n, xyz, counters);
            hipLaunchKernelGGL(k_film_add, dim3(uint32_t((npx + 255) / 256)), dim3(256), 0, s->stream, uint32_t(npx), ns,
                               n, xyz, film);
        }
        hipLaunchKernelGGL(k_film_finish, dim3(uint32_t((npx + 255) / 256)), dim3(256), 0, s->stream, uint32_t(npx),
                           uint64_t(cfg->spp), film, xyz_device);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(s->stream));
        if (stats) {
            unsigned long long c[3 * kCounterStride];
            HIP_CHECK(hipMemcpy(c, counters, sizeof c, hipMemcpyDeviceToHost));
            *stats = rp_frame_stats{};
            stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            stats->paths = npx * cfg->spp;
            stats->extension_rays = c[0];
            stats->shadow_rays = c[kCounterStride];
            stats->failed_paths = c[2 * kCounterStride];
        }
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail(who, e);
    }
}

int rp_scene_identity(rp_scene *s, uint8_t identity[32]) {
    if (!s || !identity) {
        set_error("rp_scene_identity: null argument");
        return RP_INVALID;
    }
    std::lock_guard<std::mutex> lock(s->exec);
    const auto id = identity_locked(s);
    std::memcpy(identity, id.data(), 32);
    return RP_SUCCESS;
}

int rp_film_check(rp_scene *s, const rp_film *f, uint64_t seed, uint32_t randomization) {
    if (!s || !f) {
        set_error("rp_film_check: null argument");
        return RP_INVALID;
    }
    std::lock_guard<std::mutex> lock(s->exec);
    return film_compatible(s, f, seed, randomization, "rp_film_check") ? RP_SUCCESS : RP_INVALID;
}

int rp_film_accumulate(rp_scene *s, rp_film *f, const rp_frame_config *cfg, float *xyz_host, rp_frame_stats *stats) {
    if (!s || !f || !cfg) {
        set_error("rp_film_accumulate: null argument");
        return RP_INVALID;
    }
    const rp_camera &cam = s->host.desc.camera;
    try {
        std::lock_guard<std::mutex> lock(s->exec);
        const uint32_t sampler =
            (cfg->flags & RP_FRAME_INDEPENDENT_SAMPLER) ? uint32_t(RP_SAMPLER_INDEPENDENT) : uint32_t(RP_SAMPLER_SOBOL);
        if (!film_compatible(s, f, cfg->seed, cfg->randomization, "rp_film_accumulate", sampler)) return RP_INVALID;
        if (cfg->sample_begin != f->samples) {
            set_error("rp_film_accumulate: sample_begin must equal the film's sample count " +
                      std::to_string(f->samples));
            return RP_INVALID;
        }
        const std::array<uint8_t, 32> id = identity_locked(s);
        HIP_CHECK(hipSetDevice(s->device));
        if (s->device_name.empty()) {
            hipDeviceProp_t prop{};
            HIP_CHECK(hipGetDeviceProperties(&prop, s->device));
            s->device_name = prop.name;
            s->device_arch = prop.gcnArchName;
        }
        const uint32_t npx = cam.width * cam.height;
        DevBuf<float> out;
        if (xyz_host) out.alloc(3 * size_t(npx));
        std::vector<double> planes = f->planes;
        FilmCarry carry{&planes, f->samples};
        render_frame_impl(s, *cfg, xyz_host ? out.p : nullptr, stats, &carry);
        if (xyz_host) HIP_CHECK(hipMemcpy(xyz_host, out.p, 3 * size_t(npx) * sizeof(float), hipMemcpyDeviceToHost));
        // commit only after the whole frame succeeded (all-or-nothing, as FilmBatch::commit)
        const uint64_t begin = f->samples;
        f->planes.swap(planes);
        f->samples = carry.samples;
        f->seed = cfg->seed;
        f->randomization = cfg->randomization;
        f->sampler = sampler;
        f->model = s->host.desc.transport == RP_TRANSPORT_SCALAR_DEMO ? 0u : 1u;
        f->packet = s->host.desc.transport == RP_TRANSPORT_HERO;
        f->identity = id;
        const rp_film::Segment seg{begin, f->samples, "rypath_amd C ABI " + std::to_string(RP_API_VERSION),
                                   s->device_name, s->device_arch, "Linux"};
        rp_film::Segment *last = f->segments.empty() ? nullptr : &f->segments.back();
        if (last && last->end == begin && last->build == seg.build && last->device == seg.device &&
            last->architecture == seg.architecture && last->os == seg.os)
            last->end = f->samples;
        else
            f->segments.push_back(seg);
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_film_accumulate", e);
    }
}

int rp_frame_submit(rp_scene *s, const rp_frame_config *cfg, rp_frame **out) {
    if (!s || !cfg || !out) {
        set_error("rp_frame_submit: null argument");
        return RP_INVALID;
    }
    *out = nullptr;
    if (cfg->spp == 0 || cfg->mode > RP_MODE_WAVEFRONT_SORTED || cfg->randomization > RP_RANDOMIZE_OWEN) {
        set_error("rp_frame_submit: spp must be positive and mode/randomization known");
        return RP_INVALID;
    }
    try {
        if (!s->frames) s->frames = std::make_shared<FrameSession>(s);
        FrameSession &fs = *s->frames;
        {
            std::lock_guard<std::mutex> lk(fs.m);
            if (fs.retained >= 2) {
                set_error("rp_frame_submit: both frame slots are retained; release a frame first");
                return RP_BUSY;
            }
        }
        auto job = std::make_shared<FrameJob>();
        job->cfg = *cfg;
        job->width = s->host.desc.camera.width;
        job->height = s->host.desc.camera.height;
        HIP_CHECK(hipSetDevice(s->device));
        job->film.alloc(3 * size_t(job->width) * job->height);
        job->session = s->frames;
        {
            std::lock_guard<std::mutex> lk(fs.m);
            ++fs.retained;
            fs.queue.push_back(job);
        }
        fs.cv.notify_all();
        *out = new rp_frame{job};
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_frame_submit", e);
    }
}

int rp_frame_poll(const rp_frame *f, uint32_t *done) {
    if (!f || !done) {
        set_error("rp_frame_poll: null argument");
        return RP_INVALID;
    }
    std::lock_guard<std::mutex> lk(f->job->m);
    *done = f->job->state != 0;
    return f->job->state == 2 ? RP_FAILURE : RP_SUCCESS;
}

int rp_frame_wait(rp_frame *f, float *xyz_host, rp_frame_stats *stats) {
    if (!f) {
        set_error("rp_frame_wait: null argument");
        return RP_INVALID;
    }
    FrameJob &job = *f->job;
    {
        std::unique_lock<std::mutex> lk(job.m);
        job.cv.wait(lk, [&] { return job.state != 0; });
        if (job.state == 2) {
            set_error("rp_frame_wait: " + job.error);
            return RP_FAILURE;
        }
    }
    try {
        if (xyz_host)
            HIP_CHECK(hipMemcpy(xyz_host, job.film.p, 3 * size_t(job.width) * job.height * sizeof(float),
                                hipMemcpyDeviceToHost));
        if (stats) *stats = job.stats;
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_frame_wait", e);
    }
}

void rp_frame_release(rp_frame *f) {
    if (!f) return;
    {
        // the worker may still be rendering this frame: wait, then free the slot
        std::unique_lock<std::mutex> lk(f->job->m);
        f->job->cv.wait(lk, [&] { return f->job->state != 0; });
    }
    if (auto fs = f->job->session) {
        std::lock_guard<std::mutex> lk(fs->m);
        --fs->retained;
    }
    f->job->session.reset();
    delete f;
}

static int trace_common(rp_scene *s, const rp_ray *rays, uint64_t count, void *out, bool any, double *seconds) {
    if (!s || (!rays && count) || (!out && count)) {
        set_error("rp_trace: null argument");
        return RP_INVALID;
    }
    if (count == 0) return RP_SUCCESS;
    try {
        std::lock_guard<std::mutex> lock(s->exec);
        HIP_CHECK(hipSetDevice(s->device));
        s->rays.upload(rays, count);
        rpd::EventPair ev; // destroyed on every exit, exceptions included
        hipEvent_t e0 = ev.a, e1 = ev.b;
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

namespace {
// Read-bandwidth probe: grid-stride float4 loads over the buffer, four in
// flight per iteration, `passes` times; the sum keeps the loads alive.
__global__ void __launch_bounds__(256) k_probe_read(const float4 *buf, uint64_t n4, uint32_t passes, float *out) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float acc = 0.0f;
    for (uint32_t p = 0; p < passes; ++p) {
        uint64_t i = t;
        for (; i + 3 * stride < n4; i += 4 * stride) {
            const float4 a = buf[i], b = buf[i + stride], c = buf[i + 2 * stride], d = buf[i + 3 * stride];
            acc += ((a.x + a.y) + (a.z + a.w)) + ((b.x + b.y) + (b.z + b.w)) + ((c.x + c.y) + (c.z + c.w)) +
                   ((d.x + d.y) + (d.z + d.w));
        }
        for (; i < n4; i += stride) {
            const float4 a = buf[i];
            acc += (a.x + a.y) + (a.z + a.w);
        }
    }
    out[t] = acc;
}
} // namespace

extern "C" {

int rp_probe_read_bandwidth(int device, uint64_t bytes, uint32_t passes, double *gbytes_per_second) {
    if (!gbytes_per_second || bytes < 4096 || bytes % 16 || passes == 0) {
        set_error("rp_probe_read_bandwidth: need bytes >= 4096, a multiple of 16, passes >= 1 and an output");
        return RP_INVALID;
    }
    try {
        HIP_CHECK(hipSetDevice(device));
        hipStream_t st = nullptr;
        HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        std::unique_ptr<ihipStream_t, void (*)(hipStream_t)> stream_guard(st, [](hipStream_t x) {
            (void)hipStreamDestroy(x);
        });
        DevBuf<float4> buf;
        buf.alloc(bytes / 16);
        HIP_CHECK(hipMemsetAsync(buf.p, 0, bytes, st));
        int cus = 0;
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        const dim3 grid(uint32_t(cus > 0 ? cus : 256) * 16), block(256);
        DevBuf<float> out;
        out.alloc(size_t(grid.x) * block.x);
        hipLaunchKernelGGL(k_probe_read, grid, block, 0, st, buf.p, bytes / 16, 1u, out.p); // warm-up
        rpd::EventPair ev; // destroyed on every exit, exceptions included
        hipEvent_t e0 = ev.a, e1 = ev.b;
        HIP_CHECK(hipEventRecord(e0, st));
        hipLaunchKernelGGL(k_probe_read, grid, block, 0, st, buf.p, bytes / 16, passes, out.p);
        HIP_CHECK(hipEventRecord(e1, st));
        HIP_CHECK(hipEventSynchronize(e1));
        HIP_CHECK(hipGetLastError());
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        *gbytes_per_second = double(bytes) * passes / (double(ms) * 1e-3) / 1e9;
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_probe_read_bandwidth", e);
    }
}

} // extern "C"

// ============================================================== batches
struct rp_batch {
    std::vector<rp_batch_item> items;
    std::vector<uint32_t> order;      // issue order (largest estimated cost first)
    std::vector<rpd::BatchScene> host_scenes;
    std::vector<uint64_t> offsets;    // per item offset (floats) into xyz
    uint64_t total = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    DevBuf<rpd::BatchScene> scenes;
    DevBuf<float> xyz;
    DevBuf<double> film;
    DevBuf<float> film_out;
    DevBuf<unsigned long long> work;
    // wavefront pool
    rpd::FrameBuffers wf;
    DevBuf<uint32_t> sid0, sid1, sh_sid, ov;
    DevBuf<uint4> e_id;
    bool any_interfaces = false; // some scene of the batch has null interfaces
    DevBuf<rpd::PoolCounters> counters;
    rpd::PoolCounters *host_counters = nullptr;
    uint32_t capacity = 0;
    hipGraphExec_t graph = nullptr;
    uint64_t graph_key = ~0ull;
    ~rp_batch() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (host_counters) (void)hipHostFree(host_counters);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int cu_count(int device) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) n = 256;
    return n;
}

template <int BS> void launch_persistent(rp_batch *b, dim3 grid) {
    hipLaunchKernelGGL(rpd::k_batch_persistent<BS>, grid, dim3(BS), 0, b->stream, b->scenes.p,
                       uint32_t(b->host_scenes.size()), b->work.p, b->total);
}

rpd::PoolView pool_view(rp_batch *b) {
    rpd::PoolView V{};
    V.S[0] = b->wf.state[0];
    V.S[1] = b->wf.state[1];
    V.sid[0] = b->sid0.p;
    V.sid[1] = b->sid1.p;
    V.F = b->wf.f;
    V.sh_sid = b->sh_sid.p;
    V.e_id = b->e_id.p;
    V.ov = b->ov.p;
    V.scenes = b->scenes.p;
    V.nscenes = uint32_t(b->host_scenes.size());
    V.capacity = b->capacity;
    V.c = b->counters.p;
    return V;
}

// `grid` is a fixed grid-stride grid; the extend kernel and the scene-uniform
// shade and shadow kernels take one thread per entry and a grid that covers
// the capacity (blocks past the live count return at once).
template <int BS>
void pool_iteration(rp_batch *b, const rpd::PoolView &V, uint32_t p, uint32_t bits, dim3 grid) {
    hipStream_t st = b->stream;
    const dim3 cover((V.capacity + BS - 1) / BS);
    hipLaunchKernelGGL(rpd::k_pool_refill, grid, dim3(256), 0, st, V, p);
    hipLaunchKernelGGL(rpd::k_pool_commit, dim3(1), dim3(256), 0, st, V, p, bits);
    hipLaunchKernelGGL(rpd::k_pool_extend<BS>, cover, dim3(BS), 0, st, V, p);
    if (bits) {
        hipLaunchKernelGGL(rpd::k_pool_sort_hist, grid, dim3(256), 0, st, V, p, bits);
        hipLaunchKernelGGL(rpd::k_sort_scan, dim3(1), dim3(64), 0, st, bits, V.F.hist);
        hipLaunchKernelGGL(rpd::k_pool_sort_scatter, grid, dim3(256), 0, st, V, p, bits);
    }
    hipLaunchKernelGGL(rpd::k_pool_shade<BS>, cover, dim3(BS), 0, st, V, p, bits ? 1u : 0u);
    hipLaunchKernelGGL(rpd::k_pool_shade_mixed, grid, dim3(256), 0, st, V, p);
    if (b->any_interfaces) {
        hipLaunchKernelGGL((rpd::k_pool_shadow<BS, true>), cover, dim3(BS), 0, st, V, p);
        hipLaunchKernelGGL(rpd::k_pool_shadow_mixed<true>, grid, dim3(256), 0, st, V, p);
    } else {
        hipLaunchKernelGGL((rpd::k_pool_shadow<BS, false>), cover, dim3(BS), 0, st, V, p);
        hipLaunchKernelGGL(rpd::k_pool_shadow_mixed<false>, grid, dim3(256), 0, st, V, p);
    }
    hipLaunchKernelGGL(rpd::k_pool_retire, grid, dim3(256), 0, st, V);
}

template <int BS> void pool_block(rp_batch *b, const rpd::PoolView &V, uint32_t bits, dim3 grid, int pairs) {
    for (int k = 0; k < pairs; ++k) {
        pool_iteration<BS>(b, V, 0, bits, grid);
        pool_iteration<BS>(b, V, 1, bits, grid);
    }
}

void run_pool(rp_batch *b, const rp_batch_config &cfg, uint32_t &graph_launches) {
    uint32_t cap = cfg.capacity ? cfg.capacity : (1u << 24);
    if (cap & (cap - 1)) throw std::runtime_error("wavefront capacity must be a power of two");
    if (cap > (1u << 28)) throw std::runtime_error("wavefront capacity too large");
    const uint32_t bs = cfg.group_size ? cfg.group_size : 128;
    if (bs != 64 && bs != 128 && bs != 256) throw std::runtime_error("group_size must be 64, 128 or 256");
    const uint32_t bits = cfg.sort_bits;
    if (bits > 8) throw std::runtime_error("sort_bits must be in [0, 8]");
    if (cap != b->capacity) {
        b->wf.ensure(cap);
        b->sid0.alloc(cap), b->sid1.alloc(cap), b->sh_sid.alloc(cap), b->e_id.alloc(cap), b->ov.alloc(cap);
        b->counters.alloc(1);
        b->capacity = cap;
        if (b->graph) {
            (void)hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
        }
    }
    if (!b->host_counters)
        HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&b->host_counters), sizeof(rpd::PoolCounters),
                                hipHostMallocDefault));
    rpd::PoolView V = pool_view(b);
    const dim3 grid(uint32_t(cu_count(b->device)) * 8);
    hipLaunchKernelGGL(rpd::k_pool_init, dim3(1), dim3(64), 0, b->stream, V, b->total);
    const int pairs = 4; // 8 bounce iterations per submission
    // live paths at or below which the drained pool finishes in k_pool_tail
    rp_frame_config tcfg{};
    tcfg.tail_paths = cfg.tail_paths;
    const uint32_t tail = (cfg.flags & RP_BATCH_NO_TAIL) ? 0u : rpd::tail_limit(tcfg);
    auto submit = [&] {
        if (bs == 64) pool_block<64>(b, V, bits, grid, pairs);
        else if (bs == 128) pool_block<128>(b, V, bits, grid, pairs);
        else pool_block<256>(b, V, bits, grid, pairs);
    };
    const uint64_t key = (uint64_t(cap) << 20) | (uint64_t(bs) << 8) | bits;
    if (cfg.use_graph && (b->graph == nullptr || b->graph_key != key)) {
        if (b->graph) {
            (void)hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
        }
        hipGraph_t g;
        HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        submit();
        HIP_CHECK(hipStreamEndCapture(b->stream, &g));
        HIP_CHECK(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        b->graph_key = key;
    }
    graph_launches = 0;
    const auto t_start = std::chrono::steady_clock::now();
    for (uint64_t guard = 0;; ++guard) {
        if (cfg.use_graph) {
            HIP_CHECK(hipGraphLaunch(b->graph, b->stream));
            ++graph_launches;
        } else {
            submit();
        }
        HIP_CHECK(hipMemcpyAsync(b->host_counters, b->counters.p, sizeof(rpd::PoolCounters), hipMemcpyDeviceToHost,
                                 b->stream));
        HIP_CHECK(hipStreamSynchronize(b->stream));
        if (b->host_counters->done >= b->total) break;
        // every path issued and few live (set 0 after an even number of
        // iterations): the latency tail finishes them
        const uint32_t live = b->host_counters->qn[0];
        if (b->host_counters->next >= b->total && live > 0 && live <= tail) {
            const dim3 tg((live + rpd::kWfBlock - 1) / rpd::kWfBlock);
            if (b->any_interfaces)
                hipLaunchKernelGGL(rpd::k_pool_tail<true>, tg, dim3(rpd::kWfBlock), 0, b->stream, V, 0u, live);
            else
                hipLaunchKernelGGL(rpd::k_pool_tail<false>, tg, dim3(rpd::kWfBlock), 0, b->stream, V, 0u, live);
            HIP_CHECK(hipGetLastError());
            // the final ray counts, read after the caller's synchronisation
            HIP_CHECK(hipMemcpyAsync(b->host_counters, b->counters.p, sizeof(rpd::PoolCounters),
                                     hipMemcpyDeviceToHost, b->stream));
            break;
        }
        const double waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        if (guard > 200000 || waited > 120.0)
            throw std::runtime_error("wavefront pool did not drain: done " + std::to_string(b->host_counters->done) +
                                     " of " + std::to_string(b->total) + " after " +
                                     std::to_string(guard + 1) + " submissions");
    }
}

double batch_render_impl(rp_batch *b, const rp_batch_config &cfg, float **xyz_host, rp_batch_stats *stats) {
    HIP_CHECK(hipSetDevice(b->device));
    if (cfg.randomization > RP_RANDOMIZE_OWEN) throw std::runtime_error("unknown randomization");
    if (cfg.scheduler > RP_SCHED_WAVEFRONT) throw std::runtime_error("unknown scheduler");
    for (auto &bs : b->host_scenes) bs.P.mode = cfg.randomization;
    b->scenes.upload(b->host_scenes);
    rpd::EventPair ev; // destroyed on every exit, exceptions included
    hipEvent_t e0 = ev.a, e1 = ev.b;
    HIP_CHECK(hipEventRecord(e0, b->stream));
    uint32_t launches = 0;
    unsigned long long rays[3] = {0, 0, 0}, mixed = 0;
    if (cfg.scheduler == RP_SCHED_PERSISTENT) {
        b->work.alloc(4 * kCounterStride);
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, b->stream, b->work.p, 4u);
        const uint32_t bs = cfg.group_size ? cfg.group_size : 128;
        const dim3 grid(uint32_t(cu_count(b->device)) * (bs == 64 ? 16 : bs == 128 ? 8 : 4));
        if (bs == 64) launch_persistent<64>(b, grid);
        else if (bs == 128) launch_persistent<128>(b, grid);
        else if (bs == 256) launch_persistent<256>(b, grid);
        else throw std::runtime_error("group_size must be 64, 128 or 256");
        HIP_CHECK(hipGetLastError());
        launches = 1;
    } else {
        run_pool(b, cfg, launches);
    }
    // per-scene films (atomic-free, samples in increasing order)
    for (size_t i = 0; i < b->host_scenes.size(); ++i) {
        const rpd::BatchScene &bs = b->host_scenes[i];
        const uint32_t npx = bs.npx, spp = uint32_t(bs.paths / npx);
        hipLaunchKernelGGL(k_film_zero, dim3(uint32_t((6 * uint64_t(npx) + 255) / 256)), dim3(256), 0, b->stream,
                           6 * uint64_t(npx), b->film.p);
        hipLaunchKernelGGL(k_film_add, dim3((npx + 255) / 256), dim3(256), 0, b->stream, npx, spp, bs.paths, bs.xyz,
                           b->film.p);
        hipLaunchKernelGGL(k_film_finish, dim3((npx + 255) / 256), dim3(256), 0, b->stream, npx, spp, b->film.p,
                           b->film_out.p);
        uint32_t item = b->order[i];
        if (xyz_host && xyz_host[item])
            HIP_CHECK(hipMemcpyAsync(xyz_host[item], b->film_out.p, 3 * size_t(npx) * sizeof(float),
                                     hipMemcpyDeviceToHost, b->stream));
    }
    HIP_CHECK(hipEventRecord(e1, b->stream));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipGetLastError());
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (cfg.scheduler == RP_SCHED_PERSISTENT) {
        unsigned long long w[4];
        read_counters(b->work.p, 4, w);
        rays[0] = w[1], rays[1] = w[2], rays[2] = w[3];
    } else {
        rays[0] = b->host_counters->rays[0], rays[1] = b->host_counters->rays[1], rays[2] = b->host_counters->rays[2];
        mixed = b->host_counters->mixed;
    }
    if (stats) {
        stats->seconds = ms * 1e-3;
        stats->paths = b->total;
        stats->extension_rays = rays[0];
        stats->shadow_rays = rays[1];
        stats->failed_paths = rays[2];
        stats->graph_launches = launches;
        stats->pad = 0;
        stats->mixed_scene_lanes = mixed;
    }
    return ms * 1e-3;
}

} // namespace

extern "C" {

int rp_batch_create(const rp_batch_item *items, uint32_t count, rp_batch **out) {
    if (!out || (!items && count) || count == 0) {
        set_error("rp_batch_create: null argument or empty batch");
        return RP_INVALID;
    }
    *out = nullptr;
    try {
        auto b = std::make_unique<rp_batch>();
        b->items.assign(items, items + count);
        b->device = items[0].scene ? items[0].scene->device : 0;
        std::vector<double> cost(count);
        uint32_t max_npx = 0;
        for (uint32_t i = 0; i < count; ++i) {
            const rp_batch_item &it = items[i];
            if (!it.scene || it.spp == 0) throw std::runtime_error("batch item without scene or spp");
            if (it.scene->host.desc.transport != RP_TRANSPORT_HERO)
                throw std::runtime_error("the batch pool renders hero-transport scenes");
            if (it.scene->device != b->device) throw std::runtime_error("batch scenes on different devices");
            const rp_camera &cam = it.scene->host.desc.camera;
            uint64_t paths = uint64_t(cam.width) * cam.height * it.spp;
            cost[i] = double(paths) * std::log2(2.0 + double(it.scene->host.desc.triangle_count));
            max_npx = std::max(max_npx, cam.width * cam.height);
        }
        b->order.resize(count);
        for (uint32_t i = 0; i < count; ++i) b->order[i] = i;
        std::stable_sort(b->order.begin(), b->order.end(), [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
        uint64_t floats = 0;
        for (uint32_t k = 0; k < count; ++k) {
            const rp_batch_item &it = items[b->order[k]];
            const rp_camera &cam = it.scene->host.desc.camera;
            rpd::BatchScene bs{};
            bs.S = it.scene->dev;
            bs.P = path_params(it.seed, RP_RANDOMIZE_FASTOWEN, 0, 0);
            bs.npx = cam.width * cam.height;
            bs.paths = uint64_t(bs.npx) * it.spp;
            bs.path_begin = b->total;
            bs.sample_begin = 0;
            b->total += bs.paths;
            b->offsets.push_back(floats);
            floats += 3 * bs.paths;
            b->any_interfaces = b->any_interfaces || bs.S.has_interfaces;
            b->host_scenes.push_back(bs);
        }
        HIP_CHECK(hipSetDevice(b->device));
        HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->xyz.alloc(floats);
        for (uint32_t k = 0; k < count; ++k) b->host_scenes[k].xyz = b->xyz.p + b->offsets[k];
        b->film.alloc(6 * size_t(max_npx));
        b->film_out.alloc(3 * size_t(max_npx));
        *out = b.release();
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_batch_create", e);
    }
}

int rp_batch_render(rp_batch *b, const rp_batch_config *cfg, float **xyz_host, rp_batch_stats *stats) {
    if (!b || !cfg) {
        set_error("rp_batch_render: null argument");
        return RP_INVALID;
    }
    try {
        batch_render_impl(b, *cfg, xyz_host, stats);
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_batch_render", e);
    }
}

void rp_batch_destroy(rp_batch *b) { delete b; }

// Coordinate-descent autotuner over scheduler, workgroup size, wavefront
// capacity and sort-key width; every candidate renders the whole batch.
int rp_autotune(rp_batch *b, rp_batch_config *best, double *best_seconds) {
    if (!b || !best) {
        set_error("rp_autotune: null argument");
        return RP_INVALID;
    }
    try {
        const uint32_t randomization = best->randomization, tail_paths = best->tail_paths, flags = best->flags;
        double best_t = 1e300;
        rp_batch_config best_cfg{};
        // Each candidate renders twice and the second render is timed: the
        // first one allocates the pool for a new capacity and captures its graph.
        auto trial = [&](rp_batch_config c) {
            c.randomization = randomization;
            c.tail_paths = tail_paths;
            c.flags = flags;
            (void)batch_render_impl(b, c, nullptr, nullptr);
            const double t = batch_render_impl(b, c, nullptr, nullptr);
            if (t < best_t) {
                best_t = t;
                best_cfg = c;
            }
            return t;
        };
        for (uint32_t g : {64u, 128u, 256u}) trial({0, g, 0, 0, 0, RP_SCHED_PERSISTENT});
        // wavefront pool: capacity x workgroup size at sort width 0, then the
        // sort width on the fastest of those
        rp_batch_config wf{};
        double wf_t = 1e300;
        for (uint32_t cap : {1u << 22, 1u << 24, 1u << 25})
            for (uint32_t g : {128u, 256u}) {
                const rp_batch_config c{cap, g, 0, 1, 0, RP_SCHED_WAVEFRONT};
                const double t = trial(c);
                if (t < wf_t) {
                    wf_t = t;
                    wf = c;
                }
            }
        for (uint32_t bits : {2u, 4u}) {
            rp_batch_config c = wf;
            c.sort_bits = bits;
            trial(c);
        }
        *best = best_cfg;
        best->randomization = randomization;
        if (best_seconds) *best_seconds = best_t;
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_autotune", e);
    }
}

} // extern "C"

// ============================================================== adjoint
namespace {

// Writes one parameter value into the owned scene description.
void set_param(rp_scene *s, const rp_param &p, float v) {
    HostScene &h = s->host;
    if (p.kind == RP_PARAM_REFLECTANCE_BIN || p.kind == RP_PARAM_SIGMA_S_BIN) {
        const rp_spectrum *sp = nullptr;
        if (p.kind == RP_PARAM_REFLECTANCE_BIN) {
            if (p.target >= h.materials.size() || h.materials[p.target].kind != RP_MATERIAL_DIFFUSE)
                throw std::runtime_error("reflectance parameter needs a diffuse material");
            sp = &h.materials[p.target].reflectance;
        } else {
            if (p.target >= h.media.size()) throw std::runtime_error("sigma_s parameter: medium out of range");
            sp = &h.media[p.target].sigma_s;
        }
        if (sp->kind != RP_SPECTRUM_TABLE || sp->scale != 1.0f)
            throw std::runtime_error("binned parameters need an unscaled 471-sample table spectrum");
        if (p.bins == 0 || p.bins > 471 || p.bin >= p.bins) throw std::runtime_error("bad parameter bin");
        for (uint32_t i = 0; i < 471; ++i)
            if (i * p.bins / 471u == p.bin) h.spectra[sp->table + i] = v;
    } else if (p.kind == RP_PARAM_ROUGHNESS) {
        if (p.target >= h.materials.size() || h.materials[p.target].kind == RP_MATERIAL_DIFFUSE)
            throw std::runtime_error("roughness parameter needs a conductor or dielectric");
        if (!(v >= 0)) throw std::runtime_error("roughness must be nonnegative");
        h.materials[p.target].roughness_u = v;
        h.materials[p.target].roughness_v = v;
    } else {
        throw std::runtime_error("unknown parameter kind");
    }
}

struct GradBuffers {
    DevBuf<int32_t> refl_base, rough, sigma_base;
    DevBuf<uint32_t> refl_bins, sigma_bins;
};

rpd::GradTables grad_tables(rp_scene *s, const rp_param *params, uint32_t count, GradBuffers &gb) {
    const HostScene &h = s->host;
    std::vector<int32_t> refl_base(h.materials.size() + 1, -1), rough(h.materials.size() + 1, -1),
        sigma_base(h.media.size() + 1, -1);
    std::vector<uint32_t> refl_bins(h.materials.size() + 1, 0), sigma_bins(h.media.size() + 1, 0);
    uint32_t sigma_medium = RP_NONE;
    for (uint32_t i = 0; i < count; ++i) {
        const rp_param &p = params[i];
        if (p.kind == RP_PARAM_ROUGHNESS) {
            if (p.target >= h.materials.size() || h.materials[p.target].kind != RP_MATERIAL_CONDUCTOR)
                throw std::runtime_error("roughness gradients are supported for conductors");
            rough[p.target] = int32_t(i);
            continue;
        }
        if (p.kind != RP_PARAM_REFLECTANCE_BIN && p.kind != RP_PARAM_SIGMA_S_BIN)
            throw std::runtime_error("unknown parameter kind");
        if (p.bins == 0 || p.bins > 471 || p.bin >= p.bins) throw std::runtime_error("bad parameter bin");
        std::vector<int32_t> &base = p.kind == RP_PARAM_REFLECTANCE_BIN ? refl_base : sigma_base;
        std::vector<uint32_t> &bins = p.kind == RP_PARAM_REFLECTANCE_BIN ? refl_bins : sigma_bins;
        size_t limit = p.kind == RP_PARAM_REFLECTANCE_BIN ? h.materials.size() : h.media.size();
        if (p.target >= limit) throw std::runtime_error("parameter target out of range");
        // bins of one spectrum must be consecutive parameters in bin order
        if (p.bin == 0) {
            base[p.target] = int32_t(i);
            bins[p.target] = p.bins;
        } else if (base[p.target] < 0 || bins[p.target] != p.bins || base[p.target] + int32_t(p.bin) != int32_t(i)) {
            throw std::runtime_error("binned parameters of one spectrum must be consecutive, bin 0 first");
        }
        if (p.kind == RP_PARAM_SIGMA_S_BIN) {
            if (sigma_medium != RP_NONE && sigma_medium != p.target)
                throw std::runtime_error("sigma_s gradients are supported for one medium");
            sigma_medium = p.target;
        }
        if (p.kind == RP_PARAM_REFLECTANCE_BIN && h.materials[p.target].kind != RP_MATERIAL_DIFFUSE)
            throw std::runtime_error("reflectance gradients need a diffuse material");
    }
    gb.refl_base.upload(refl_base);
    gb.rough.upload(rough);
    gb.sigma_base.upload(sigma_base);
    gb.refl_bins.upload(refl_bins);
    gb.sigma_bins.upload(sigma_bins);
    rpd::GradTables G{};
    G.refl_base = gb.refl_base.p;
    G.refl_bins = gb.refl_bins.p;
    G.rough = gb.rough.p;
    G.sigma_base = gb.sigma_base.p;
    G.sigma_bins = gb.sigma_bins.p;
    G.nparams = count;
    G.sigma_medium = sigma_medium;
    return G;
}

} // namespace

extern "C" {

int rp_set_params(rp_scene *s, const rp_param *params, uint32_t count, const float *values) {
    if (!s || (count && (!params || !values))) {
        set_error("rp_set_params: null argument");
        return RP_INVALID;
    }
    try {
        std::lock_guard<std::mutex> lock(s->exec);
        ++s->params_version;
        s->has_identity = false; // the identity hashes the values written below
        for (uint32_t i = 0; i < count; ++i) set_param(s, params[i], values[i]);
        HIP_CHECK(hipSetDevice(s->device));
        s->spectra.upload(s->host.spectra);
        s->materials.upload(s->host.materials);
        s->media.upload(s->host.media);
        HIP_CHECK(hipDeviceSynchronize());
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail("rp_set_params", e);
    }
}

} // extern "C"

namespace {
// dl_on_device: dl_dxyz is device memory on the scene's device (read in
// place, after a device synchronisation); else host memory (uploaded).
int backward_impl(const char *name, rp_scene *s, const rp_frame_config *cfg, const rp_param *params,
                  uint32_t count, const float *dl_dxyz, bool dl_on_device, float *grad, rp_frame_stats *stats) {
    if (!s || !cfg || (count && !params) || !dl_dxyz || !grad) {
        set_error(std::string(name) + ": null argument");
        return RP_INVALID;
    }
    if (count > rpd::kMaxParams) {
        set_error(std::string(name) + ": at most 256 parameters");
        return RP_INVALID;
    }
    if (s->host.desc.transport != RP_TRANSPORT_HERO) {
        set_error(std::string(name) + ": the adjoint differentiates hero-transport scenes");
        return RP_INVALID;
    }
    try {
        std::lock_guard<std::mutex> lock(s->exec);
        HIP_CHECK(hipSetDevice(s->device));
        if (dl_on_device) HIP_CHECK(hipDeviceSynchronize());
        if (cfg->spp == 0) throw std::runtime_error("spp must be positive");
        GradBuffers gb;
        rpd::GradTables G = grad_tables(s, params, count, gb);
        const rp_camera &cam = s->host.desc.camera;
        const uint32_t npx = cam.width * cam.height;
        DevBuf<float> dl, g;
        if (!dl_on_device) dl.upload(dl_dxyz, 3 * size_t(npx));
        const float *dl_dev = dl_on_device ? dl_dxyz : dl.p;
        g.alloc(std::max<uint32_t>(count, 1));
        HIP_CHECK(hipMemsetAsync(g.p, 0, std::max<uint32_t>(count, 1) * sizeof(float), s->stream));
        rpd::PathParams P = path_params(cfg->seed, cfg->randomization, 0, 0, cfg->flags);
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s->stream, s->counters.p, 8u);
        rpd::EventPair ev; // destroyed on every exit, exceptions included
        hipEvent_t e0 = ev.a, e1 = ev.b;
        HIP_CHECK(hipEventRecord(e0, s->stream));
        const uint64_t total = uint64_t(npx) * cfg->spp;
        const uint64_t per_chunk = std::max<uint64_t>(1, (uint64_t(1) << 26) / npx); // whole samples per launch
        // primal: the frame wavefront over the same paths, which leaves each
        // path's radiance and status in the device arena's per-lane arrays;
        // the arena is held until the replay of the last chunk has finished
        // (the event synchronisation below happens before the unlock)
        // A wavefront frame of the same scene, parameters and configuration
        // rendered just before (one chunk) left exactly this primal in the
        // arena: then only the replay runs.
        FrameArena &arena = frame_arena(s->device);
        std::unique_lock<std::mutex> arena_lock(arena.m);
        rp_frame_config wcfg = *cfg;
        wcfg.mode = RP_MODE_WAVEFRONT;
        uint64_t launches = 0, wf_rays[2] = {0, 0};
        const bool one_chunk = per_chunk >= cfg->spp;
        const bool reuse = one_chunk && arena.primal.same(s, *cfg, npx);
        if (reuse) {
            wf_rays[0] = arena.primal.rays[0];
            wf_rays[1] = arena.primal.rays[1];
        }
        for (uint64_t s0 = 0; s0 < cfg->spp; s0 += per_chunk) {
            uint64_t ns = std::min<uint64_t>(per_chunk, cfg->spp - s0);
            uint64_t n = ns * npx;
            if (!reuse) {
                arena.primal.valid = false;
                s->path_xyz.alloc(3 * n);
                (void)rpd::wavefront_render(s->dev, P, arena.fb, wcfg, npx, cfg->sample_begin + s0, n,
                                            s->path_xyz.p, s->counters.p, s->stream, launches, wf_rays, s0 == 0);
                if (one_chunk) arena.primal.set(s, *cfg, npx, wf_rays);
            }
            rpd::adjoint_replay(s->dev, P, G, arena.fb, arena.adj, npx, cfg->sample_begin + s0, n, cfg->spp, dl_dev, g.p,
                                rpd::tail_limit(*cfg), s->stream);
        }
        HIP_CHECK(hipEventRecord(e1, s->stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        HIP_CHECK(hipMemcpy(grad, g.p, count * sizeof(float), hipMemcpyDeviceToHost));
        unsigned long long c[8];
        read_counters(s->counters.p, 8, c);
        if (reuse) c[2] = arena.primal.failed; // the counters were cleared; the primal's frame counted them
        else if (one_chunk) arena.primal.failed = c[2];
        if (stats) {
            uint64_t tail[2];
            arena.fb.tail_rays(tail); // the primal's (the replay's tail does not count)
            std::memset(stats, 0, sizeof *stats);
            stats->seconds = ms * 1e-3;
            stats->paths = total;
            stats->extension_rays = wf_rays[0] + tail[0];
            stats->shadow_rays = wf_rays[1] + tail[1];
            stats->extend_rays = wf_rays[0];
            stats->failed_paths = c[2];
            stats->extend_launches = launches; // 0: the primal of the preceding wavefront frame was reused
        }
        return RP_SUCCESS;
    } catch (const std::exception &e) {
        return fail(name, e);
    }
}
} // namespace

extern "C" {

int rp_backward(rp_scene *s, const rp_frame_config *cfg, const rp_param *params, uint32_t count,
                const float *dl_dxyz, float *grad, rp_frame_stats *stats) {
    return backward_impl("rp_backward", s, cfg, params, count, dl_dxyz, false, grad, stats);
}
int rp_backward_device(rp_scene *s, const rp_frame_config *cfg, const rp_param *params, uint32_t count,
                       const float *dl_dxyz_device, float *grad, rp_frame_stats *stats) {
    return backward_impl("rp_backward_device", s, cfg, params, count, dl_dxyz_device, true, grad, stats);
}

} // extern "C"
