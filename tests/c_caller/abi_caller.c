/*
 * A plain C11 caller of the C ABI (include/rp_api.h, include/rp_synthetic.h):
 * builds the synthetic Cornell box, checks a few request validations, takes
 * per-path records through rp_sample, renders a frame through
 * rp_render_frame, and releases every handle.  It prints one line of
 * results that tests/test_c_caller.py compares with the same calls made
 * through ctypes.  Exit status: 0 success, 1 an unexpected status, 2 no
 * device (rp_scene_create returned RP_NO_DEVICE).
 *
 * Build: cc -std=c11 -Wall -Wextra -pedantic -I include abi_caller.c \
 *           -L rypath_amd -lrypath_amd -Wl,-rpath,<dir of librypath_amd.so>
 */
#include "rp_api.h"
#include "rp_synthetic.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int expect(int got, int want, const char *what) {
    if (got == want) return 1;
    char msg[512];
    rp_last_error(msg, sizeof msg);
    fprintf(stderr, "%s: status %d, expected %d (%s)\n", what, got, want, msg);
    return 0;
}

int main(int argc, char **argv) {
    const uint32_t width = argc > 1 ? (uint32_t)atoi(argv[1]) : 32;
    const uint32_t spp = argc > 2 ? (uint32_t)atoi(argv[2]) : 4;
    if (rp_api_version() != RP_API_VERSION) {
        fprintf(stderr, "API version mismatch\n");
        return 1;
    }
    rp_synthetic_options so;
    memset(&so, 0, sizeof so);
    so.width = width;
    so.height = width;
    rp_synthetic *synth = NULL;
    if (!expect(rp_synthetic_create(RP_SYNTH_CORNELL, 1, &so, &synth), RP_SUCCESS, "rp_synthetic_create")) return 1;
    const rp_scene_desc *desc = rp_synthetic_desc(synth);

    rp_build_options bo;
    memset(&bo, 0, sizeof bo);
    rp_scene *scene = NULL;
    int rc = rp_scene_create(desc, &bo, &scene);
    if (rc == RP_NO_DEVICE) {
        rp_synthetic_destroy(synth);
        return 2;
    }
    if (!expect(rc, RP_SUCCESS, "rp_scene_create")) return 1;
    /* the scene deep-copies the description: the generator can go now */
    rp_synthetic_destroy(synth);

    const uint32_t npx = width * width;
    rp_request rq;
    memset(&rq, 0, sizeof rq);
    rq.width = width;
    rq.height = width;
    rq.seed = 1;
    rq.sampler_kind = RP_SAMPLER_SOBOL;
    rq.randomization = RP_RANDOMIZE_FASTOWEN;
    rq.pixel_begin = 0;
    rq.pixel_count = npx;
    rq.sample_begin = 0;
    rq.sample_count = spp;
    rq.total_samples = spp;

    rp_path_record *rec = calloc((size_t)npx * spp, sizeof *rec);
    if (!rec) return 1;
    int ok = 1;
    /* documented refusals */
    rp_request bad = rq;
    bad.width = width + 1;
    ok &= expect(rp_sample(scene, &bad, rec, (uint64_t)npx * spp), RP_INVALID, "raster size mismatch");
    bad = rq;
    bad.sampler_kind = RP_SAMPLER_HALTON;
    ok &= expect(rp_sample(scene, &bad, rec, (uint64_t)npx * spp), RP_INVALID, "sampler kind");
    bad = rq;
    bad.total_samples = spp - 1;
    ok &= expect(rp_sample(scene, &bad, rec, (uint64_t)npx * spp), RP_INVALID, "total sample budget");
    bad = rq;
    bad.threadgroup_width = 96;
    ok &= expect(rp_sample(scene, &bad, rec, (uint64_t)npx * spp), RP_INVALID, "threadgroup width");
    ok &= expect(rp_sample(scene, &rq, rec, (uint64_t)npx * spp - 1), RP_INVALID, "output capacity");

    ok &= expect(rp_sample(scene, &rq, rec, (uint64_t)npx * spp), RP_SUCCESS, "rp_sample");
    double rec_sum = 0.0;
    uint64_t index_xor = 0, failed = 0;
    for (size_t i = 0; i < (size_t)npx * spp; ++i) {
        rec_sum += (double)rec[i].xyz[1];
        index_xor ^= rec[i].sobol_index;
        failed += rec[i].status != RP_PATH_SUCCESS;
        if (rec[i].filter_weight != 1.0f || rec[i].splat_count != 0) ok = 0;
    }
    free(rec);

    rp_frame_config fc;
    memset(&fc, 0, sizeof fc);
    fc.mode = RP_MODE_WAVEFRONT;
    fc.spp = spp;
    fc.seed = 1;
    fc.randomization = RP_RANDOMIZE_FASTOWEN;
    float *film = calloc((size_t)npx * 3, sizeof *film);
    if (!film) return 1;
    rp_frame_stats st;
    memset(&st, 0, sizeof st);
    ok &= expect(rp_render_frame(scene, &fc, film, &st), RP_SUCCESS, "rp_render_frame");
    double film_sum = 0.0;
    for (size_t i = 0; i < (size_t)npx * 3; ++i) film_sum += (double)film[i];
    free(film);
    rp_scene_destroy(scene);

    printf("records %u xyz_y_sum %.17g sobol_xor %llu failed %llu film_sum %.17g paths %llu\n", npx * spp, rec_sum,
           (unsigned long long)index_xor, (unsigned long long)failed, film_sum, (unsigned long long)st.paths);
    return ok ? 0 : 1;
}
