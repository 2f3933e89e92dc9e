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
