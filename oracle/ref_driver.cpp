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
//   ref_rect_*, ref_tri_* src/light_geometry_sampling_shared.h:84-260 (bilinear
//                        warp, spherical rectangle and triangle samples and inverses),
//                        composed as the patch and triangle branches of
//                        shaders/light_geometry_sampling.h:114-209
//   ref_hero_sequence    src/hero_spectrum_shared.h:102-160 (scaled likelihoods,
//                        advance, balance)
//   ref_conductor_*      tests/conductor_reference.hpp:24-103 (double complex Fresnel,
//                        Smith GGX, Heitz VNDF; the reference's own test model)
//   ref_dielectric_*     tests/dielectric_reference.hpp:27-139 (double rough dielectric)
#include "hero_spectrum_shared.h"
#include "light_geometry_sampling_shared.h"
#include "light_sampling.hpp"
#include "packet_spectrum_shared.h"
#include "sampling.hpp"
#include "sobol_shared.h"
// the reference's double-precision test models of its two microfacet BSDFs
#include "tests/conductor_reference.hpp"
#include "tests/dielectric_reference.hpp"

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

namespace lg = rypath_light_geometry_math;
static lg::V3 v3(const float *p) { return lg::V3(p[0], p[1], p[2]); }

// The solid-angle density of the patch proposal of rectangle (p00, p10, p01,
// p11) (12 floats) at rectangle coordinates uv from `reference` with receiver
// normal (zero: none), as the reference's pdf() for a patch: -1 when the
// spherical rectangle is not used (not an affine rectangle, or 1e-4 sr or less).
REF_API float ref_rect_pdf(const float *reference, const float *corners, const float *normal, const float *uv) {
    const lg::V3 p[4] = {v3(corners), v3(corners + 3), v3(corners + 6), v3(corners + 9)};
    const lg::V3 r = v3(reference), n = v3(normal);
    const float area = lg::rectangle_solid_angle(p[0], p[1], p[2], p[3], r);
    if (!(lg::rectangle(p[0], p[1], p[2], p[3]) && area > 1e-4f)) return -1.0f;
    const auto frame = lg::rectangle_frame(r, p[0], p[1] - p[0], p[2] - p[0]);
    const lg::V2 square = lg::rectangle_inverse(frame, lg::V2(uv[0], uv[1]));
    return (lg::dot(n, n) > 0 ? lg::bilinear_pdf(square, lg::rectangle_cosines(p[0], p[1], p[2], p[3], r, n)) : 1.f) /
           area;
}
// The patch proposal's sample: rectangle coordinates uv and their density.
// Returns 0 (nothing written) when the spherical rectangle is not used.
REF_API int ref_rect_sample(const float *reference, const float *corners, const float *normal, float u0, float u1,
                            float *uv, float *pdf) {
    const lg::V3 p[4] = {v3(corners), v3(corners + 3), v3(corners + 6), v3(corners + 9)};
    const lg::V3 r = v3(reference), n = v3(normal);
    lg::V2 u(std::min(u0, lg::one_minus_epsilon), std::min(u1, lg::one_minus_epsilon));
    const float area = lg::rectangle_solid_angle(p[0], p[1], p[2], p[3], r);
    if (!(lg::rectangle(p[0], p[1], p[2], p[3]) && area > 1e-4f)) return 0;
    if (lg::dot(n, n) > 0) u = lg::bilinear_sample(u, lg::rectangle_cosines(p[0], p[1], p[2], p[3], r, n));
    const lg::V2 q = lg::rectangle_sample(lg::rectangle_frame(r, p[0], p[1] - p[0], p[2] - p[0]), u);
    uv[0] = q.x;
    uv[1] = q.y;
    *pdf = ref_rect_pdf(reference, corners, normal, uv);
    return 1;
}

// A sequence of hero events from the initial density: before step i the
// balance weight of densities pdf_i against other_i, then advance with
// (num_i, pdf_i).  Returns the number of steps whose update was valid.
REF_API uint32_t ref_hero_sequence(uint32_t steps, const float *beta0, const float *num, const float *pdf,
                                   const float *other, float *beta_out, float *weight_out) {
    rypath_hero::Density density = rypath_hero::initial_density();
    rypath_hero::Vector beta = {beta0[0], beta0[1], beta0[2], beta0[3]};
    for (uint32_t i = 0; i < steps; ++i) {
        const rypath_hero::Vector pv = {pdf[4 * i], pdf[4 * i + 1], pdf[4 * i + 2], pdf[4 * i + 3]};
        const rypath_hero::Vector ov = {other[4 * i], other[4 * i + 1], other[4 * i + 2], other[4 * i + 3]};
        const rypath_hero::Vector nv = {num[4 * i], num[4 * i + 1], num[4 * i + 2], num[4 * i + 3]};
        const rypath_hero::Scalar w = rypath_hero::balance(density, pv, ov);
        weight_out[i] = w.valid ? w.value : -1.0f;
        const rypath_hero::Update u = rypath_hero::advance(density, beta, nv, pv);
        if (!u.valid) return i;
        density = u.density;
        beta = u.beta;
        for (int k = 0; k < 4; ++k) beta_out[4 * i + k] = beta[k];
    }
    return steps;
}

// The reference's double models: evaluate (f, pdf) and sample (direction,
// weight f |cos| / pdf, pdf; returns whether the sample was accepted).
namespace cr = conductor_reference;
namespace dr = dielectric_reference;
REF_API double ref_conductor_fresnel(double c, double eta, double k) { return cr::fresnel(c, eta, k); }
REF_API double ref_dielectric_fresnel(double c, double eta) { return dr::fresnel(c, eta); }
REF_API void ref_conductor_eval(double eta, double k, double ax, double ay, int remap, const double *wo,
                                const double *wi, double *f, double *pdf) {
    const auto e = cr::evaluate(cr::prepare(eta, k, ax, ay, remap != 0), {wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]});
    *f = e.f;
    *pdf = e.pdf;
}
REF_API int ref_conductor_sample(double eta, double k, double ax, double ay, int remap, const double *wo, double u,
                                 double v, double *wi, double *weight, double *pdf) {
    const auto r = cr::sample(cr::prepare(eta, k, ax, ay, remap != 0), {wo[0], wo[1], wo[2]}, u, v);
    wi[0] = r.i.x, wi[1] = r.i.y, wi[2] = r.i.z;
    *weight = r.weight;
    *pdf = r.pdf;
    return r.accepted ? 1 : 0;
}
REF_API void ref_dielectric_eval(double eta, double ax, double ay, int remap, const double *wo, const double *wi,
                                 double *f, double *pdf) {
    const auto e = dr::evaluate(dr::prepare(eta, ax, ay, remap != 0), {wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]});
    *f = e.f;
    *pdf = e.pdf;
}
REF_API int ref_dielectric_sample(double eta, double ax, double ay, int remap, const double *wo, double uc, double u,
                                  double v, double *wi, double *weight, double *pdf) {
    const auto r = dr::sample(dr::prepare(eta, ax, ay, remap != 0), {wo[0], wo[1], wo[2]}, uc, u, v);
    wi[0] = r.i.x, wi[1] = r.i.y, wi[2] = r.i.z;
    *weight = r.weight;
    *pdf = r.pdf;
    return r.accepted ? 1 : 0;
}

// The triangle proposal of triangle (p0, p1, p2) (9 floats) from `reference`
// with receiver normal (zero: none), as the triangle branch of the
// reference's sample() composes it: barycentrics (b1, b2) of p1 and p2, and
// the density pdf() gives at the sampled point.  Returns 1 for the
// spherical triangle, 0 for area sampling (bary set, pdf untouched), -1 when
// the spherical sample fails.
REF_API float ref_tri_pdf(const float *reference, const float *corners, const float *normal, const float *q);
REF_API int ref_tri_sample(const float *reference, const float *corners, const float *normal, float u0, float u1,
                           float *bary, float *pdf) {
    const lg::V3 p0 = v3(corners), p1 = v3(corners + 3), p2 = v3(corners + 6), r = v3(reference), n = v3(normal);
    lg::V2 u(std::min(u0, lg::one_minus_epsilon), std::min(u1, lg::one_minus_epsilon));
    const float area = lg::triangle_solid_angle(p0, p1, p2, r);
    if (!lg::spherical_triangle(area)) {
        const lg::V2 uv = lg::triangle_uniform_uv(u);
        bary[0] = uv.x;
        bary[1] = uv.y;
        return 0;
    }
    if (lg::dot(n, n) > 0) u = lg::bilinear_sample(u, lg::triangle_cosines(p0, p1, p2, r, n));
    lg::V2 uv;
    if (!lg::triangle_sample(p0, p1, p2, r, u, uv)) return -1;
    bary[0] = uv.x;
    bary[1] = uv.y;
    const lg::V3 q = (1 - uv.x - uv.y) * p0 + uv.x * p1 + uv.y * p2;
    const float qf[3] = {q.x, q.y, q.z};
    *pdf = ref_tri_pdf(reference, corners, normal, qf);
    return 1;
}
// pdf() of the triangle branch at point q: -1 when the spherical triangle is
// not used.
REF_API float ref_tri_pdf(const float *reference, const float *corners, const float *normal, const float *q) {
    const lg::V3 p0 = v3(corners), p1 = v3(corners + 3), p2 = v3(corners + 6), r = v3(reference), n = v3(normal);
    const float area = lg::triangle_solid_angle(p0, p1, p2, r);
    if (!lg::spherical_triangle(area)) return -1.0f;
    const lg::V2 square = lg::triangle_inverse(p0, p1, p2, r, lg::unit(v3(q) - r));
    return (lg::dot(n, n) > 0 ? lg::bilinear_pdf(square, lg::triangle_cosines(p0, p1, p2, r, n)) : 1.f) / area;
}
