/*
 * rp_synthetic.h -- seeded synthetic scene generators (C ABI).
 *
 * Counterpart of the reference's built-in scene construction
 * (make_demo_scene, src/render.cpp:2060-2106) and of the RenderScene the PBRT
 * translator produces (src/pbrt_render.hpp translate), for the benchmark
 * workloads: Cornell box, fBm heightfield with icospheres, uneven batch
 * scenes, volumetric displaced icosphere, and the inverse-rendering Cornell
 * box with a fog volume.  The generated rp_scene_desc and every array it
 * points to are owned by the rp_synthetic handle.
 */
#ifndef RP_SYNTHETIC_H
#define RP_SYNTHETIC_H

#include "rp_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rp_synthetic_kind {
    RP_SYNTH_CORNELL = 0,     /* 555-unit Cornell box, 36 tris, blackbody 6500 K x 17 */
    RP_SYNTH_TERRAIN = 1,     /* fBm heightfield + icospheres, 40/30/30 materials, 16 quads + env */
    RP_SYNTH_BATCH = 2,       /* terrain-like scene sized to target_triangles */
    RP_SYNTH_VOLUME = 3,      /* displaced icosphere, dielectric 1.33 boundary, HG medium, spot + env */
    RP_SYNTH_INVERSE = 4,     /* Cornell box, conductor boxes, fog box (null interface) */
    RP_SYNTH_DEMO = 5         /* reference demo: floor, two walls, three boxes, unit environment, scalar demo model */
} rp_synthetic_kind;

typedef struct rp_synthetic_options {
    uint32_t width, height;      /* 0 = kind default */
    uint32_t max_depth;          /* 0 = kind default */
    uint32_t grid;               /* terrain vertices per side, 0 = 1024 */
    uint32_t spheres;            /* terrain icospheres, 0 = 64 */
    uint32_t sphere_subdivision; /* 0 = 4 */
    uint32_t target_triangles;   /* RP_SYNTH_BATCH */
    uint32_t volume_subdivision; /* RP_SYNTH_VOLUME, 0 = 6 */
    float hg_g;                  /* RP_SYNTH_VOLUME */
    uint32_t bins;               /* RP_SYNTH_INVERSE reflectance bins, 0 = 16 */
    uint32_t pad;
} rp_synthetic_options;

typedef struct rp_synthetic rp_synthetic;

int rp_synthetic_create(uint32_t kind, uint64_t seed, const rp_synthetic_options *options, rp_synthetic **out);
const rp_scene_desc *rp_synthetic_desc(const rp_synthetic *scene);
void rp_synthetic_destroy(rp_synthetic *scene);

#ifdef __cplusplus
}
#endif
#endif /* RP_SYNTHETIC_H */
