/*
 * rp_io.h -- scene and image I/O of the host API (C ABI).
 *
 *   rp_pbrt_load   PBRT-v4 scene-file subset -> rp_scene_desc
 *                  (counterpart of pbrt::check + pbrt::translate,
 *                  src/pbrt.hpp:105-241, src/pbrt_render.hpp)
 *   rp_write_exr   raw-XYZ film -> linear sRGB OpenEXR, atomic publish
 *                  (counterpart of write_exr, src/exr_output.cpp:62-120)
 *   rp_frame_*     asynchronous frame submission with two retained slots
 *                  (counterpart of FrameSession::submit / Frame::poll / wait,
 *                  src/frame_session.hpp:145-212)
 *   rp_film_*, rp_checkpoint_*
 *                  running films, checkpoint and resume (counterpart of
 *                  Film + write_checkpoint / read_checkpoint,
 *                  src/render.hpp:317-362, src/checkpoint.hpp:8-24)
 */
#ifndef RP_IO_H
#define RP_IO_H

#include "rp_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rp_pbrt rp_pbrt;

typedef struct rp_pbrt_info {
    uint32_t width, height;
    uint32_t spp;            /* Sampler "pixelsamples" (default 16) */
    uint32_t randomization;  /* rp_randomization from Sampler "randomization" */
    uint64_t seed;           /* Sampler "seed" */
    uint32_t max_depth;      /* Integrator "maxdepth" (default 5) */
    uint32_t volumetric;     /* Integrator "volpath" */
    char output[256];        /* Film "filename" (may be empty) */
} rp_pbrt_info;

/* Parses and translates a scene file.  Unsupported directives, shapes,
 * materials or lights are errors (rp_last_error names them). */
int rp_pbrt_load(const char *path, rp_pbrt **out);
int rp_pbrt_load_string(const char *text, const char *base_dir, rp_pbrt **out);
const rp_scene_desc *rp_pbrt_desc(const rp_pbrt *scene);
int rp_pbrt_get_info(const rp_pbrt *scene, rp_pbrt_info *info);
void rp_pbrt_destroy(rp_pbrt *scene);

/* Writes the film (W*H*3 mean XYZ) as an uncompressed scanline EXR with
 * channels B, G, R = exposure * M_sRGB * XYZ, FLOAT (half == 0) or HALF,
 * plus `count` string attributes.  The file is written to a sibling
 * temporary and renamed into place. */
int rp_write_exr(const char *path, uint32_t width, uint32_t height, const float *xyz, double exposure,
                 uint32_t half, const char *const *keys, const char *const *values, uint32_t count);

/* Running films and checkpoints (counterpart of the progressive Film with
 * write_checkpoint / read_checkpoint, src/render.hpp:317-362,
 * src/checkpoint.hpp:8-24).  A film holds the compensated per-pixel sums of
 * the samples [0, samples) of one scene, seed and randomization.
 * rp_film_accumulate renders the next config->spp samples; config->sample_begin
 * must equal the film's sample count, and after the first accumulation the
 * seed, sampler, randomization and scene identity must match the film's.  A render
 * resumed from a checkpoint equals the uninterrupted render bit for bit.
 *
 * Checkpoint files are the reference's RYPCHK format 3 (src/checkpoint.cpp:
 * 17-122, src/checkpoint_contract.h:14), little endian:
 *   header (44 B): "RYPCHK\r\n", u32 contract words format 3, identity 12,
 *     renderer 29, sampling 11, accumulation 2, u64 metadata bytes,
 *     u64 payload bytes (216 per pixel);
 *   metadata: u32 width, u32 height, u64 job samples, u64 batch size,
 *     u64 seed, u32 sampler kind (0 independent, 1 Sobol), u32 randomization,
 *     i32 material seed, u32 film model (0 demo, 1 PBRT), f64 exposure,
 *     u8 packet, u8 compatibility output, f64 max component, u32 filter kind,
 *     f32 x radius, f32 y radius, f32 sigma, u8[32] scene identity,
 *     u32 segment count, segments { u64 begin, u64 end, 4 x (u32 n, n bytes):
 *     build, device, architecture, os };
 *   pixels: f64 sum[3], f64 correction[3], u64 count, f64[18] splat and
 *     compatibility sums, f64 weight sum, f64 weight correction;
 *   SHA-256 of all preceding bytes.
 * A film written here uses the one-pixel box filter (radius 0.5, unit
 * weight), no splats and no clamped compatibility sums; hero films are
 * packet films (weight sum = sample count), scalar films are not.  Reading a
 * format-3 file refuses what this film cannot hold: another filter,
 * splat or compatibility sums, compatibility output, Halton or ZSobol
 * samples, more than 2^28 pixels.  The scene identity is this library's
 * SHA-256 of its scene description (rp_scene_identity), so a film written by
 * the reference reads here but binds to no scene of this library, and the
 * reference likewise refuses to continue a film written here.  The earlier
 * format 1 of this library ("RPAFILM\n") is still read.
 * Files are published through a sibling temporary, fsync and rename. */
typedef struct rp_film rp_film;
typedef struct rp_film_info {
    uint32_t width, height;
    uint64_t samples, seed;
    uint32_t randomization, segments;
    uint8_t scene_identity[32]; /* SHA-256 of the scene description (zero until bound) */
    uint32_t sampler_kind;      /* rp_sampler_kind of the samples held */
    uint32_t model;             /* 0 demo (scalar demo transport), 1 PBRT */
    uint64_t job_samples, batch_size; /* the job settings a checkpoint records (0 = unset) */
    uint32_t packet;            /* unit-weight reconstruction sums (hero transport) */
    uint32_t format;            /* checkpoint format the film was read from, 0 if created */
} rp_film_info;
int rp_film_create(uint32_t width, uint32_t height, rp_film **out);
/* Whether `film` can continue rendering `scene` with this seed and
 * randomization: its size must equal the scene camera's, and a film that
 * holds samples must carry the same seed, randomization and scene identity
 * (the checks of read_checkpoint's compatibility contract,
 * src/checkpoint.cpp:300-360).  RP_SUCCESS, or RP_INVALID with the reason in
 * rp_last_error.  rp_film_accumulate applies the same checks. */
int rp_film_check(rp_scene *scene, const rp_film *film, uint64_t seed, uint32_t randomization);
int rp_film_accumulate(rp_scene *scene, rp_film *film, const rp_frame_config *config, float *xyz_host,
                       rp_frame_stats *stats);
int rp_film_resolve(const rp_film *film, float *xyz_host);
int rp_film_get_info(const rp_film *film, rp_film_info *info);
/* Records the render job's total samples and batch size, which a checkpoint
 * stores as the reference's RenderSettings.samples and batch_size
 * (src/checkpoint.cpp:87-89); unset, or once the film holds more samples,
 * the film's own sample count is written as the job's.  job_samples must be
 * at least the film's samples. */
int rp_film_set_job(rp_film *film, uint64_t job_samples, uint64_t batch_size);
int rp_checkpoint_write(const rp_film *film, const char *path);
int rp_checkpoint_read(const char *path, rp_film **out);
void rp_film_destroy(rp_film *film);
/* SHA-256 identity of the scene's description (the value films bind to). */
int rp_scene_identity(rp_scene *scene, uint8_t identity[32]);
/* SHA-256 (FIPS 180-4) of a byte range. */
void rp_sha256(const void *data, size_t size, uint8_t digest[32]);

/* Asynchronous frames: submit enqueues a whole frame on the scene's stream
 * and returns immediately; poll reports completion; wait blocks and copies
 * the film (W*H*3 floats) to the host when `xyz_host` is not NULL.  A scene
 * retains at most two in-flight frames (slots 0 and 1). */
typedef struct rp_frame rp_frame;
int rp_frame_submit(rp_scene *scene, const rp_frame_config *config, rp_frame **out);
int rp_frame_poll(const rp_frame *frame, uint32_t *done);
int rp_frame_wait(rp_frame *frame, float *xyz_host, rp_frame_stats *stats);
void rp_frame_release(rp_frame *frame);

#ifdef __cplusplus
}
#endif
#endif /* RP_IO_H */
