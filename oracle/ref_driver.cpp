// C entry points over the reference's own portable host code, for tests.
//
// TEST INFRASTRUCTURE ONLY.  `make -C oracle ref` compiles this file together
// with the reference's sources where they lie (src/sampling.cpp,
// src/light_sampling.cpp, src/rgb_spectrum.cpp and the headers they include)
// into oracle/_ref/libref.so.  Nothing of the reference is copied here: this
// file only calls it.  The reference's symbols are built hidden and the
// linker drops every function these entry points do not reach, so the parts
// of those files that need the image decoders or the CMake-generated sRGB
// table are never linked (and never called).
//
// Covered (reference file:line of what each entry calls):
//   ref_sobol_*          src/sobol_shared.h:56-154 with the tables of src/sampling.cpp:14-39
//   ref_hero_*           src/hero_spectrum_shared.h:199-244 (renderer hero wavelengths)
//   ref_packet_visible   src/packet_spectrum_shared.h:27-58 (light power proxy wavelengths)
//   ref_light_*          src/light_sampling.cpp:17-27 (power proxy), 294-327 (patch
//                        rectangle / area), 329-381 (positional and normal cones),
//                        382-571 (alias table and light BVH)
#include "hero_spectrum_shared.h"
#include "light_sampling.hpp"
#include "packet_spectrum_shared.h"
#include "sampling.hpp"
#include "sobol_shared.h"

#include <array>
#include <cstdint>
#include <cstring>
#include <span>
#include <string>
#include <vector>

#define REF_API extern "C" __attribute__((visibility("default")))

REF_API uint64_t ref_sobol_index(uint32_t m, uint32_t x, uint32_t y, uint64_t sample) {
    return rypath::sobol::index(rypath::sobol_tables(), m, x, y, sample);
}
REF_API uint32_t ref_sobol_word(uint64_t index, uint32_t dimension, uint32_t mode, uint32_t seed_bits) {
    return rypath::sobol::word(rypath::sobol_tables(), index, dimension, mode, seed_bits);
}
REF_API uint32_t ref_dimension_hash(uint32_t dimension, uint32_t seed_bits) {
    return rypath::sobol::dimension_hash(dimension, seed_bits);
}
REF_API float ref_sobol_uniform(uint32_t word) { return rypath::sobol::uniform(word); }
REF_API float ref_camera_uniform(uint32_t word, uint32_t m, uint32_t pixel) {
    return rypath::sobol::camera_uniform(word, m, pixel);
}

static void copy_wavelengths(const rypath_packet_spectrum::Wavelengths &w, float *lambda, float *pdf) {
    for (int k = 0; k < 4; ++k) {
        lambda[k] = w.lambda[k];
        pdf[k] = w.pdf[k];
    }
}
REF_API void ref_hero_uniform(float u, float *lambda, float *pdf) {
    copy_wavelengths(rypath_hero::sample_uniform(u), lambda, pdf);
}
REF_API void ref_hero_visible(float u, float *lambda, float *pdf) {
    copy_wavelengths(rypath_hero::sample_visible(u), lambda, pdf);
}
REF_API void ref_packet_visible(float u, float *lambda, float *pdf) {
    copy_wavelengths(rypath_packet_spectrum::sample_visible(u), lambda, pdf);
}

REF_API double ref_light_spectrum_power(const float *table471) {
    std::array<float, 471> s{};
    std::memcpy(s.data(), table471, sizeof(float) * 471);
    return rypath::light_spectrum_power(s);
}

static std::array<std::array<double, 3>, 4> corners(const double *p) {
    std::array<std::array<double, 3>, 4> c{};
    for (int i = 0; i < 4; ++i)
        for (int a = 0; a < 3; ++a) c[i][a] = p[3 * i + a];
    return c;
}
// p = (p00, p10, p01, p11), 12 doubles
REF_API int ref_light_patch_rectangle(const double *p) { return rypath::light_patch_rectangle(corners(p)) ? 1 : 0; }
REF_API double ref_light_patch_area(const double *p) { return rypath::light_patch_area(corners(p)); }

// The normal cone of n normals (3n doubles): direction and cos_outer.
REF_API void ref_light_normal_cone(const double *normals, uint32_t n, float *direction, float *cos_outer) {
    std::vector<std::array<double, 3>> v(n);
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) v[i][a] = normals[3 * i + a];
    rypath::LightSamplingInput in;
    rypath::light_normal_cone(in, v);
    for (int a = 0; a < 3; ++a) direction[a] = in.direction[a];
    *cos_outer = in.cos_outer;
}
// The positional cone of a spot light whose local-from-world rotation is m
// (9 floats, rows): direction, cos_outer, cos_emission, two_sided.
REF_API void ref_light_positional_cone(const float *m, float outer, float emission, float *direction,
                                       float *cos_outer, float *cos_emission, int *two_sided) {
    std::array<float, 9> mm{};
    for (int i = 0; i < 9; ++i) mm[i] = m[i];
    rypath::LightSamplingInput in;
    rypath::light_positional_cone(in, mm, outer, emission);
    for (int a = 0; a < 3; ++a) direction[a] = in.direction[a];
    *cos_outer = in.cos_outer;
    *cos_emission = in.cos_emission;
    *two_sided = in.two_sided ? 1 : 0;
}

// build_light_sampling over n inputs given as planar arrays.  Outputs (the
// caller sizes node arrays for 2n - 1 nodes): per light the alias threshold,
// pmf, alias and the leaf node; per node its box, direction, normalised
// power, cones, child / parent / light links and flags.  Returns the node
// count, or -1 with the reference's diagnostic in `error`.
REF_API int ref_build_light_sampling(uint32_t n, const double *power, const double *bounds_power, const float *low,
                                     const float *high, const float *direction, const float *cos_outer,
                                     const float *cos_emission, const uint8_t *two_sided, const uint8_t *infinite,
                                     float *alias_threshold, float *alias_pmf, uint32_t *alias, uint32_t *leaf,
                                     float *node_low, float *node_high, float *node_direction, float *node_power,
                                     float *node_cos_outer, float *node_cos_emission, uint32_t *node_child,
                                     uint32_t *node_parent, uint32_t *node_light, uint32_t *node_flags, char *error,
                                     size_t error_size) {
    std::vector<rypath::LightSamplingInput> in(n);
    for (uint32_t i = 0; i < n; ++i) {
        auto &v = in[i];
        v.power = power[i];
        v.bounds_power = bounds_power[i];
        for (int a = 0; a < 3; ++a) {
            v.low[a] = low[3 * i + a];
            v.high[a] = high[3 * i + a];
            v.direction[a] = direction[3 * i + a];
        }
        v.cos_outer = cos_outer[i];
        v.cos_emission = cos_emission[i];
        v.two_sided = two_sided[i] != 0;
        v.infinite = infinite[i] != 0;
    }
    std::string diagnostic;
    auto tables = rypath::build_light_sampling(in, size_t(1) << 40, diagnostic);
    if (!tables) {
        if (error_size) {
            std::strncpy(error, diagnostic.c_str(), error_size - 1);
            error[error_size - 1] = 0;
        }
        return -1;
    }
    for (uint32_t i = 0; i < n; ++i) {
        alias_threshold[i] = tables->alias[i].threshold;
        alias_pmf[i] = tables->alias[i].pmf;
        alias[i] = tables->alias[i].alias;
        leaf[i] = tables->records[i].leaf;
    }
    for (size_t k = 0; k < tables->nodes.size(); ++k) {
        const auto &nd = tables->nodes[k];
        const float lo[3] = {nd.low.x, nd.low.y, nd.low.z}, hi[3] = {nd.high.x, nd.high.y, nd.high.z},
                    dir[3] = {nd.direction.x, nd.direction.y, nd.direction.z};
        for (int a = 0; a < 3; ++a) {
            node_low[3 * k + a] = lo[a];
            node_high[3 * k + a] = hi[a];
            node_direction[3 * k + a] = dir[a];
        }
        node_power[k] = nd.power;
        node_cos_outer[k] = nd.cos_outer;
        node_cos_emission[k] = nd.cos_emission;
        node_child[k] = nd.child;
        node_parent[k] = nd.parent;
        node_light[k] = nd.light;
        node_flags[k] = nd.flags;
    }
    return int(tables->nodes.size());
}
