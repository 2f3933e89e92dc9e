/*
 * rp_api.h -- C ABI of the MI355X (gfx950) wavefront spectral path tracer.
 *
 * Plain C: opaque handles, plain pointers and sizes, integer status codes,
 * no C++ exceptions across the boundary.  Every synchronous call on one
 * handle must be externally serialised (one host thread per handle at a time).
 *
 * Reference interfaces replaced (rlefko/rypath):
 *   rp_scene_desc          RenderScene                      src/render.hpp:248-306
 *   rp_scene_create        build_metal_renderer             src/metal_render.hpp:73-93
 *   rp_sample              MetalRenderScene::sample         src/metal_render.hpp:56-57
 *   rp_request             RenderRequest                    src/metal_render.hpp:14-32
 *   rp_path_record         PacketSample (+ diagnostics)     src/render.hpp:307-316
 *   rp_render_frame        FrameSession::submit/Frame::wait src/frame_session.hpp:145-200
 *   rp_status              GeometryStatus                   src/metal_geometry.hpp:22
 *   rp_path_status         render_wire status codes         src/render_wire.h:20-26
 * Entry points with no counterpart in the reference (new patterns):
 *   rp_trace_closest / rp_trace_any  (traversal probes, parity + roofline)
 *   rp_batch_*                       (uneven multi-scene persistent work pool)
 *   rp_autotune                      (queue capacity / workgroup / sort-key width)
 *   rp_backward                      (path-replay adjoint)
 */
#ifndef RP_API_H
#define RP_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RP_API_VERSION 1u
#define RP_NONE 0xffffffffu
#define RP_SPECTRUM_SAMPLES 471u /* 360..830 nm at 1 nm, nearest-nm lookup */

/* ---- status codes (GeometryStatus, src/metal_geometry.hpp:22) ---------- */
typedef enum rp_status {
    RP_SUCCESS = 0,
    RP_NOT_BUILT = 1,   /* library built without a device backend */
    RP_NO_DEVICE = 2,   /* no HIP device visible */
    RP_FAILURE = 3,     /* invalid input or device error; see rp_last_error */
    RP_INVALID = 4,     /* invalid argument (null pointer, bad size, bad enum) */
    RP_BUSY = 5         /* rp_frame_submit: both frame slots retained (FrameSubmitResult::busy,
                           src/frame_session.hpp:176-181) */
} rp_status;

/* ---- per-path status (render_wire::Status, src/render_wire.h:20-26) ---- */
typedef enum rp_path_status {
    RP_PATH_SUCCESS = 0,
    RP_PATH_EMERGENCY_DEPTH = 1,
    RP_PATH_INVALID_ARITHMETIC = 2,
    RP_PATH_INVALID_HIT = 3,
    RP_PATH_BOUNDARY_EXHAUSTION = 4
} rp_path_status;

/* ---- spectra ------------------------------------------------------------ */
typedef enum rp_spectrum_kind {
    RP_SPECTRUM_CONSTANT = 0, /* c[0] */
    RP_SPECTRUM_SIGMOID = 1,  /* PBRT RGB sigmoid s(x), x = fma(fma(c0,l,c1),l,c2), l in nm */
    RP_SPECTRUM_TABLE = 2,    /* 471 samples at `table` (offset in floats into spectra pool) */
    RP_SPECTRUM_CAUCHY = 3,   /* c0 + c1 / (l_um * l_um) */
    RP_SPECTRUM_LINEAR = 4,   /* linear from c0 at 360 nm to c1 at 830 nm */
    RP_SPECTRUM_KNOTS8 = 5    /* piecewise linear through 8 values at `table` (offset into the pool) at the
                                 reference's knots 360, 420, 470, 520, 570, 620, 700, 830 nm
                                 (SpectralMaterial::reflectance, src/render.hpp:14,108) */
} rp_spectrum_kind;

typedef struct rp_spectrum {
    uint32_t kind;
    float c[3];
    float scale;     /* multiplies the value (sigmoid: unbounded scale; others: 1) */
    uint32_t table;  /* float offset into rp_scene_desc.spectra (TABLE kind) */
} rp_spectrum;

/* ---- materials (SpectralMaterial, src/render.hpp:108-129) -------------- */
typedef enum rp_material_kind {
    RP_MATERIAL_DIFFUSE = 0,
    RP_MATERIAL_CONDUCTOR = 1,
    RP_MATERIAL_DIELECTRIC = 2,
    RP_MATERIAL_INTERFACE = 3 /* null interface: medium boundary, no scattering */
} rp_material_kind;

typedef struct rp_material {
    uint32_t kind;
    uint32_t remap_roughness; /* nonzero: alpha = sqrt(roughness) (PBRT) */
    float roughness_u, roughness_v;
    rp_spectrum reflectance; /* diffuse */
    rp_spectrum eta;         /* conductor / dielectric */
    rp_spectrum k;           /* conductor */
} rp_material;

/* ---- objects: material + area emission + medium interface per object --- */
typedef struct rp_object {
    uint32_t material;
    uint32_t emission;       /* index into emissions or RP_NONE */
    uint32_t medium_inside;  /* RP_NONE = vacuum */
    uint32_t medium_outside;
} rp_object;

typedef struct rp_emission {
    rp_spectrum radiance; /* emitted radiance L(lambda) */
    uint32_t two_sided;   /* zero: emits on the geometric-normal side only */
    uint32_t pad[3];
} rp_emission;

typedef struct rp_spot_light {
    float position[3];
    float direction[3];   /* unit cone axis */
    float cos_falloff_start, cos_falloff_end;
    rp_spectrum intensity;
} rp_spot_light;

typedef struct rp_medium {
    rp_spectrum sigma_a, sigma_s; /* 1/unit */
    float g;                      /* Henyey-Greenstein asymmetry, clamped to [-0.99, 0.99] */
    uint32_t pad[3];
} rp_medium;

/* RasterCamera (src/render.hpp:153-158): d = corner + x*dx + y*dy, x,y in [0,1) */
typedef struct rp_camera {
    float origin[3];
    float corner[3];
    float dx[3];
    float dy[3];
    uint32_t width, height;
    uint32_t medium; /* RP_NONE = vacuum */
    uint32_t pad;
} rp_camera;

typedef struct rp_scene_desc {
    uint32_t api_version;
    uint32_t vertex_count, triangle_count, object_count;
    const float *positions;      /* 3 * vertex_count */
    const uint32_t *indices;     /* 3 * triangle_count */
    const uint32_t *tri_object;  /* triangle_count */
    const rp_object *objects;
    uint32_t material_count, emission_count, medium_count, spot_count;
    const rp_material *materials;
    const rp_emission *emissions;
    const rp_medium *media;
    const rp_spot_light *spots;
    const float *spectra;        /* pool of 471-sample tables */
    uint32_t spectra_count;      /* floats in the pool */
    uint32_t has_environment;
    rp_spectrum environment;     /* constant-direction environment radiance */
    rp_camera camera;
    uint32_t max_depth;          /* PBRT maxdepth: scattering events */
    uint32_t hero_wavelengths;   /* 0 = uniform (1/470), 1 = visible (PBRT fit) */
    uint32_t light_sampler;      /* rp_light_sampler (RenderScene::light_sampler, src/render.hpp:269) */
    uint32_t transport;          /* rp_transport (RenderScene::spectral_transport, src/render.hpp:270) */
} rp_scene_desc;

/* Spectral transport (SpectralTransport, src/render.hpp:213).  HERO: four
 * hero wavelengths with spectral balance MIS (the default).  SCALAR: one
 * wavelength per path, uniform on 360-830 nm, the reference's scalar
 * estimator (shaders/scalar_transport.h:446-899) for diffuse surfaces:
 * light sampling with power-heuristic MIS (beta = 2), cosine-hemisphere
 * continuation, survival 3/4 from the third scattering on; film weight
 * 470 CIE(round lambda) L.  SCALAR_DEMO: the reference's demo model
 * (docs/rendering.md:41-60): the environment radiance as the only light,
 * sampled over the uniform sphere, no depth limit (emergency only), film
 * weight L CIE(lambda) 470 / Y (linear CIE interpolation, Y = 106.857).
 * Scalar scenes render with rp_sample_scalar and with rp_render_frame* in
 * RP_MODE_MEGAKERNEL. */
typedef enum rp_transport {
    RP_TRANSPORT_HERO = 0, RP_TRANSPORT_SCALAR = 1, RP_TRANSPORT_SCALAR_DEMO = 2
} rp_transport;

/* LightSamplerKind (src/render.hpp:208): how next-event estimation picks a light.
 * BVH: PBRT's light bounding-volume hierarchy, importance by position and
 * normal of the receiving point (the reference's default); POWER: an alias
 * table proportional to emitted power; UNIFORM: every light alike.  Quad
 * emitters (two consecutive triangles (a,b,c), (a,c,d) of one emissive object
 * forming a rectangle) are one light, sampled over their spherical rectangle. */
typedef enum rp_light_sampler {
    RP_LIGHT_SAMPLER_BVH = 0, RP_LIGHT_SAMPLER_POWER = 1, RP_LIGHT_SAMPLER_UNIFORM = 2
} rp_light_sampler;

typedef struct rp_build_options {
    uint32_t leaf_size;       /* max triangles per BVH leaf (1..8), 0 = default 2 */
    uint32_t sah_bins;        /* 0 = default 32 */
    uint32_t threads;         /* host build threads, 0 = hardware concurrency */
    uint32_t device;          /* HIP device ordinal */
} rp_build_options;

typedef struct rp_build_info {
    double build_seconds, upload_seconds;
    uint64_t node_count, leaf_count, device_bytes;
    uint32_t max_depth, max_stack; /* BVH depth and worst-case traversal stack */
    uint32_t emitter_count, light_count;
} rp_build_info;

/* ---- sampler ------------------------------------------------------------ */
typedef enum rp_randomization {
    RP_RANDOMIZE_NONE = 0, RP_RANDOMIZE_PERMUTE = 1, RP_RANDOMIZE_FASTOWEN = 2, RP_RANDOMIZE_OWEN = 3
} rp_randomization;

/* RenderSampler::kind (SamplerKind, src/sampling.hpp:14).  This backend
 * renders with the Sobol sampler; the other kinds are refused (RP_INVALID). */
typedef enum rp_sampler_kind {
    RP_SAMPLER_INDEPENDENT = 0, RP_SAMPLER_SOBOL = 1, RP_SAMPLER_HALTON = 2, RP_SAMPLER_ZSOBOL = 3
} rp_sampler_kind;

/* RenderRequest (src/metal_render.hpp:14-32), field for field.  72 bytes. */
typedef struct rp_request {
    uint32_t width, height;   /* raster size: 0 = the scene camera's, else it must equal the camera's */
    uint64_t seed;            /* sampler seed (PBRT signed 32-bit seed in the low word) */
    int32_t material_seed;    /* PBRT's stochastic-material seed: carried for the caller; no material
                                 of this backend is stochastic, so it does not change the result */
    uint32_t pixel_begin, pixel_count;
    uint32_t sample_count;
    uint64_t sample_begin;
    uint64_t total_samples;   /* fixed sample budget: 0 = none, else it must contain
                                 [sample_begin, sample_begin + sample_count) */
    uint32_t threadgroup_width; /* 0 = the device's choice (128); else 64, 128 or 256 (the
                                   records do not depend on it) */
    uint32_t emergency_limit; /* 0 = default 1024 */
    uint32_t sampler_kind;    /* rp_sampler_kind: RP_SAMPLER_SOBOL */
    uint32_t randomization;   /* rp_randomization (RenderSampler::randomization) */
    uint32_t boundary_limit;  /* 0 = default 4096 */
    uint32_t pad;
} rp_request;

/* PacketSample (src/render.hpp:307-316) plus parity diagnostics. 112 bytes. */
typedef struct rp_path_record {
    uint32_t pixel, status;
    uint64_t sample;
    uint64_t sobol_index;
    float wavelengths[4];
    float wavelength_pdfs[4];
    float radiance[4];
    float xyz[3];
    uint32_t primary_triangle; /* scene triangle id of the camera ray's hit, RP_NONE = miss */
    uint32_t vertices;          /* path vertices including the camera */
    uint32_t camera_words[2];   /* Sobol words of dimensions 0 and 1 */
    float filter_weight;        /* PacketSample::filter_weight: 1 (box pixel filter) */
    uint32_t splat_begin, splat_count; /* PacketSample splat range: 0, 0 (no light splats) */
} rp_path_record;

/* ---- frames ------------------------------------------------------------- */
typedef enum rp_mode {
    RP_MODE_MEGAKERNEL = 0,      /* one thread = one complete path (plain) */
    RP_MODE_WAVEFRONT = 1,       /* queues, unsorted */
    RP_MODE_WAVEFRONT_SORTED = 2 /* queues, shade queue sorted by material key */
} rp_mode;

typedef struct rp_frame_config {
    uint32_t mode;             /* rp_mode */
    uint32_t spp;
    uint64_t sample_begin;
    uint64_t seed;
    uint32_t randomization;
    uint32_t capacity;         /* wavefront paths in flight, 0 = auto */
    uint32_t group_size;       /* threads per workgroup, 0 = auto */
    uint32_t sort_bits;        /* material sort-key width (1..8), 0 = default 4 */
    uint32_t flags;            /* RP_FRAME_* bits below */
    uint32_t tail_paths;       /* wavefront: once at most this many paths are live, one kernel
                                  finishes them, a thread per path; 0 = auto (the lanes the
                                  tail kernel holds resident: 512 per CU, 131072 on MI355X) */
} rp_frame_config;

#define RP_FRAME_COUNT_TRAFFIC 1u /* replay-count BVH node/triangle fetches of the closest-hit queries */
/* Throughput roulette off (forward and rp_backward): every path runs to
 * max_depth or escape.  For derivative checks by finite differences, where
 * the roulette decisions that flip with the parameters dominate the noise;
 * rendering keeps the reference's roulette (the default). */
#define RP_FRAME_NO_ROULETTE 2u
/* Wavefront: every bounce through the queue kernels (no latency tail). */
#define RP_FRAME_NO_TAIL 4u
/* Scalar transport only: the independent Philox4x32-10 sampler (the demo's
 * and PBRT's "independent"; counter (pixel, sample low, sample high, block),
 * key = the seed's two words) instead of Sobol. */
#define RP_FRAME_INDEPENDENT_SAMPLER 8u

typedef struct rp_frame_stats {
    double seconds;                 /* device time of the frame (events) */
    uint64_t paths;                 /* W*H*spp */
    uint64_t extension_rays, shadow_rays;
    uint64_t failed_paths;          /* status != success */
    double trace_seconds;           /* wavefront: device time in the launches that trace extension rays
                                       (closest hit; fused launches also trace shadow rays) */
    uint64_t extend_launches;       /* wavefront: those launches */
    uint64_t node_fetches, triangle_tests; /* with RP_FRAME_COUNT_TRAFFIC: closest-hit traversal totals */
    uint64_t extend_rays;           /* wavefront: extension rays traced by the closest-hit kernels
                                       (the rest of extension_rays by the latency tail) */
    /* with RP_FRAME_COUNT_TRAFFIC: any-hit (shadow) traversal totals, and the
     * traversal-stack pushes of both kernels that overflowed the LDS short
     * stack into per-lane scratch */
    uint64_t shadow_node_fetches, shadow_triangle_tests, stack_spills;
    double shadow_seconds;          /* wavefront: device time in the shadow-only (any-hit) launches */
    uint64_t shadow_launches;       /* wavefront: shadow-only launches */
    uint64_t shadow_kernel_rays;    /* wavefront: shadow rays traced by the bounce kernels, shadow-only
                                       and fused (the rest of shadow_rays by the latency tail) */
    /* wavefront, scenes without null interfaces: launches that trace a bounce's
     * extension rays together with the previous bounce's shadow rays (counted
     * in extend_launches and trace_seconds), and the shadow rays they traced */
    uint64_t fused_launches, fused_shadow_rays;
} rp_frame_stats;

/* ---- traversal probe ray/hit records ------------------------------------ */
typedef struct rp_ray {
    float origin[3]; float tmax;
    float direction[3]; uint32_t pad;
} rp_ray;

typedef struct rp_hit {
    uint32_t triangle; /* scene triangle id, RP_NONE = miss */
    float t, b1, b2;   /* barycentrics of vertices 1 and 2 */
} rp_hit;

typedef struct rp_scene rp_scene;
typedef struct rp_batch rp_batch;

/* Library version and device availability. */
uint32_t rp_api_version(void);
int rp_device_count(void);
/* Last error of the calling thread's most recent failing call (NUL-terminated). */
size_t rp_last_error(char *buffer, size_t capacity);

/* Build: deep-copies all scene data, builds the BVH on the host, uploads. */
int rp_scene_create(const rp_scene_desc *desc, const rp_build_options *options, rp_scene **out);
int rp_scene_info(const rp_scene *scene, rp_build_info *info);
void rp_scene_destroy(rp_scene *scene);

/* RenderSample (src/render.hpp:292-301): one scalar path. */
typedef struct rp_render_sample {
    uint32_t pixel;
    uint32_t status;             /* rp_path_status */
    uint64_t sample;
    float wavelength;            /* nm */
    float radiance;              /* L(lambda) along the path */
    uint32_t scatterings;
    uint32_t random_words[4];    /* the camera / wavelength words (Philox block 0 or Sobol dims 0-3) */
    float xyz[3];                /* the film weight of the path (diagnostic) */
} rp_render_sample;

/* Per-path records for request->pixel_count * request->sample_count paths,
 * lane = pixel_offset + pixel_count * sample_offset (src/render.metal:240-242).
 * `out` is host memory with room for that many records. */
int rp_sample(rp_scene *scene, const rp_request *request, rp_path_record *out, uint64_t capacity);
/* Scalar-transport scenes: RenderSample records (MetalRenderScene::sample's
 * `samples`, src/metal_render.hpp:35), lane = pixel_offset + pixel_count *
 * sample_offset.  request->sampler_kind is RP_SAMPLER_SOBOL or
 * RP_SAMPLER_INDEPENDENT.  Hero scenes return RP_INVALID (use rp_sample). */
int rp_sample_scalar(rp_scene *scene, const rp_request *request, rp_render_sample *out, uint64_t capacity);

/* One frame of W*H*spp paths.  Film: per-pixel mean XYZ (float32, W*H*3) of the
 * samples [sample_begin, sample_begin+spp), reduced in increasing sample order
 * with Neumaier-compensated double sums.  xyz_host may be NULL. */
int rp_render_frame(rp_scene *scene, const rp_frame_config *config, float *xyz_host,
                    rp_frame_stats *stats);
/* Same, with the film written to caller-owned device memory (W*H*3 floats). */
int rp_render_frame_device(rp_scene *scene, const rp_frame_config *config, float *xyz_device,
                           rp_frame_stats *stats);
/* Allocation-free frame.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same film as
 * rp_render_frame_device in RP_MODE_MEGAKERNEL, bit for bit; other modes
 * return RP_INVALID.  Its counters, film sums and per-path XYZ live in one
 * static 128 MiB device array per GPU, held until the process ends: the
 * film must fit in it (60 bytes per pixel plus 2 KiB, so at most 2,236,928
 * pixels; 1920x1080 fits), else RP_INVALID and nothing runs.  The call makes
 * no hipMalloc or hipFree.  Every call on a GPU shares that GPU's array and
 * clears what it needs; this library serialises its calls on one GPU within
 * the process, and calls that overlapped on one GPU would race.  Blocks
 * until the film is written; stats->seconds is host wall-clock time. */
int rp_render_frame_device_noalloc(rp_scene *scene, const rp_frame_config *config, float *xyz_device,
                                   rp_frame_stats *stats);

/* Closest-hit / any-hit traversal of caller rays (host buffers). any-hit writes
 * 1 (occluded) or 0 per ray into `occluded`. */
int rp_trace_closest(rp_scene *scene, const rp_ray *rays, uint64_t count, rp_hit *hits,
                     double *seconds);
int rp_trace_any(rp_scene *scene, const rp_ray *rays, uint64_t count, uint8_t *occluded,
                 double *seconds);

/* Device read bandwidth (no reference counterpart; used to state roofline
 * fractions against measured peaks): streams a `bytes` buffer `passes` times
 * with 16-byte loads over all CUs and returns GB/s (1e9 B/s).  A buffer far
 * above the L2 and Infinity Cache measures HBM; a few MiB read many times
 * measures L2 (every XCD caches the eighth of the buffer its CUs read). */
int rp_probe_read_bandwidth(int device, uint64_t bytes, uint32_t passes, double *gbytes_per_second);

/* ---- batched work pool (uneven scenes, one persistent launch) ----------- */
typedef struct rp_batch_item {
    rp_scene *scene;
    uint32_t spp;
    uint32_t pad;
    uint64_t seed;
} rp_batch_item;

typedef enum rp_scheduler {
    RP_SCHED_PERSISTENT = 0, /* one persistent megakernel launch pulling paths from a device work counter */
    RP_SCHED_WAVEFRONT = 1   /* device-resident wavefront pool: refill/extend/shade/shadow/retire */
} rp_scheduler;

typedef struct rp_batch_config {
    uint32_t capacity;    /* wavefront paths in flight, 0 = default */
    uint32_t group_size;  /* threads per workgroup (64, 128, 256), 0 = default */
    uint32_t sort_bits;   /* material sort-key width (0 = no sort) */
    uint32_t use_graph;   /* nonzero: the wavefront bounce loop is replayed as one captured HIP graph */
    uint32_t randomization;
    uint32_t scheduler;   /* rp_scheduler */
    uint32_t tail_paths;  /* wavefront pool: once every path is issued and at most this many are
                             live, one kernel finishes them (a thread per path); 0 = auto
                             (512 per CU) */
    uint32_t flags;       /* RP_BATCH_* bits below */
} rp_batch_config;

#define RP_BATCH_NO_TAIL 1u /* wavefront pool: bounce iterations until the pool is empty */

typedef struct rp_batch_stats {
    double seconds;
    uint64_t paths;
    uint64_t extension_rays, shadow_rays;
    uint64_t failed_paths;
    uint32_t graph_launches, pad;
    /* wavefront pool: path events and shadow rays of waves that held more
     * than one scene, handled by the per-lane-scene kernels */
    uint64_t mixed_scene_lanes;
} rp_batch_stats;

int rp_batch_create(const rp_batch_item *items, uint32_t count, rp_batch **out);
/* Renders every item; xyz_host[i] (may be NULL) receives item i's film. */
int rp_batch_render(rp_batch *batch, const rp_batch_config *config, float **xyz_host,
                    rp_batch_stats *stats);
void rp_batch_destroy(rp_batch *batch);

/* Autotuner: times candidate (capacity, group_size, sort_bits) triples on the
 * batch and returns the fastest in `best`. */
int rp_autotune(rp_batch *batch, rp_batch_config *best, double *best_seconds);

/* ---- differentiable rendering ------------------------------------------- */
typedef enum rp_param_kind {
    RP_PARAM_REFLECTANCE_BIN = 0, /* material `target`, bin `bin` of `bins` over 360..830 nm */
    RP_PARAM_ROUGHNESS = 1,       /* material `target` (u and v roughness) */
    RP_PARAM_SIGMA_S_BIN = 2      /* medium `target`, bin `bin` of `bins` */
} rp_param_kind;

typedef struct rp_param {
    uint32_t kind, target, bin, bins;
} rp_param;

/* Replace parameter values (same layout as rp_backward's gradient). */
int rp_set_params(rp_scene *scene, const rp_param *params, uint32_t count, const float *values);
/* Path-replay adjoint: dL/dparams for the frame defined by `config`, given
 * dL/dXYZ per pixel (W*H*3 floats, host).  grad receives `count` floats,
 * reproducible to float summation order.  The primal (each path's radiance) comes
 * from a frame-wavefront pass over the same paths; when the last wavefront
 * frame on the device was this scene with the same parameters, seed,
 * randomization, sample range and roulette setting in one chunk, its primal
 * is reused and stats->extend_launches is 0. */
int rp_backward(rp_scene *scene, const rp_frame_config *config, const rp_param *params,
                uint32_t count, const float *dl_dxyz, float *grad, rp_frame_stats *stats);
/* Same, with dL/dXYZ read in place from device memory on the scene's device
 * (a PyTorch tensor's data_ptr()); the call synchronises the device first, so
 * work that produced it on any stream is complete.  grad is host memory. */
int rp_backward_device(rp_scene *scene, const rp_frame_config *config, const rp_param *params,
                       uint32_t count, const float *dl_dxyz_device, float *grad, rp_frame_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* RP_API_H */
