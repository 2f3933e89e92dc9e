// Scalar (one-wavelength) transport on the device: the reference's scalar
// estimator for diffuse surfaces (shaders/scalar_transport.h:446-899,
// docs/rendering.md:41-60).  One thread runs one complete path.
//
// Draw blocks (render.metal:154-177): block 0 holds the camera x/y and
// wavelength words; scattering b draws block 1 + 2b (continuation x/y,
// roulette word z) and block 2 + 2b (light selection and position).  The
// independent sampler is Philox4x32-10 with counter (pixel, sample low,
// sample high, block) and key (seed low, seed high); the Sobol sampler's
// block is dimensions 4 block .. 4 block + 3.  The demo model
// (RP_TRANSPORT_SCALAR_DEMO) lights only by the environment, sampled over the
// uniform sphere; the scalar model selects lights with the scene's light
// sampler.  Both weight with the power heuristic and survive with
// probability 3/4 from the third scattering on.
#pragma once
#include "rp_path.h"

namespace rpd {

struct ScalarResult {
    uint32_t status;
    float wavelength, radiance;
    uint32_t scatterings;
    uint32_t words[4];
    uint32_t ext_rays, shadow_rays;
};

// Power heuristic a^2 / (a^2 + b^2) on values scaled by their maximum.
RP_DEV float power_heuristic(float a, float b) {
    const float scale = omax(a, b);
    if (!(scale > 0) || !__builtin_isfinite(scale)) return __builtin_nanf("");
    a = a / scale;
    b = b / scale;
    return a * a / (a * a + b * b);
}
RP_DEV F3 sphere_dir(float u, float v) {
    const float z = 1.0f - 2.0f * u;
    const float r = __builtin_sqrtf(omax(0.0f, 1.0f - z * z));
    float sn, cs;
    sincos_turns(v, sn, cs);
    return {r * cs, r * sn, z};
}
RP_DEV F3 hemisphere_dir(F3 n, float u, float v) {
    const float r = __builtin_sqrtf(u);
    float sn, cs;
    sincos_turns(v, sn, cs);
    const F3 axis = fabsf(n.z) < 0.999f ? F3{0, 0, 1} : F3{0, 1, 0};
    const F3 t = normalize(cross3(axis, n));
    const F3 b = cross3(n, t);
    return normalize((r * cs) * t + (r * sn) * b + __builtin_sqrtf(1.0f - u) * n);
}
// Demo weights with p_light = 1/(4 pi) and p_bsdf = cos/pi.
RP_DEV float demo_light_weight(float c) {
    const float q = 4.0f * c;
    return 1.0f / (1.0f + q * q);
}
RP_DEV float demo_bsdf_weight(float c) {
    const float q = 4.0f * c;
    const float q2 = q * q;
    return q2 / (1.0f + q2);
}
// f cos / p_light times the power-heuristic weight (none for delta lights).
RP_DEV float scattering_direct(float f_cos, float bsdf_pdf, float light_pdf, bool delta) {
    if (!(f_cos >= 0) || !(bsdf_pdf >= 0) || !(light_pdf > 0) || !__builtin_isfinite(f_cos) ||
        !__builtin_isfinite(bsdf_pdf) || !__builtin_isfinite(light_pdf))
        return __builtin_nanf("");
    if (f_cos == 0) return 0;
    return (f_cos / light_pdf) * (delta ? 1.0f : power_heuristic(light_pdf, bsdf_pdf));
}

template <bool kInterfaces>
RP_DEV void trace_scalar(const Scene &S, const PathParams &P, Stack &st, uint32_t pixel, uint64_t sample,
                         bool independent, ScalarResult &out) {
    const bool demo = S.transport == RP_TRANSPORT_SCALAR_DEMO;
    const uint32_t W = S.camera.width, H = S.camera.height;
    const uint32_t px = pixel % W, py = pixel / W;
    const uint32_t seed_lo = uint32_t(P.seed), seed_hi = uint32_t(P.seed >> 32);
    out.status = RP_PATH_SUCCESS;
    out.radiance = 0.0f;
    out.scatterings = 0;
    out.ext_rays = out.shadow_rays = 0;
    uint64_t index = 0;
    uint4 w0;
    if (independent) {
        w0 = philox(make_uint4(pixel, uint32_t(sample), uint32_t(sample >> 32), 0u), make_uint2(seed_lo, seed_hi));
    } else {
        index = sobol_index(S, px, py, sample);
        w0 = sobol_block(S, index, 0, RP_RANDOMIZE_NONE, 0);
        w0.z = scramble(w0.z, 2, P.mode, seed_lo);
        w0.w = scramble(w0.w, 3, P.mode, seed_lo);
    }
    out.words[0] = w0.x, out.words[1] = w0.y, out.words[2] = w0.z, out.words[3] = w0.w;
    auto su = [&](uint32_t w) { return independent ? philox_u01(w) : sobol_uniform(w); };
    const float lambda = 360.0f + 470.0f * su(w0.z);
    out.wavelength = lambda;
    Wavelengths wl;
    for (int k = 0; k < 4; ++k) wl.lambda[k] = lambda, wl.pdf[k] = 1.0f / 470.0f;
    const float ux = independent ? su(w0.x) : camera_uniform(w0.x, S.m, px);
    const float uy = independent ? su(w0.y) : camera_uniform(w0.y, S.m, py);
    const float fx = (float(px) + ux) / float(W), fy = (float(py) + uy) / float(H);
    const rp_camera &cam = S.camera;
    F3 d = {cam.corner[0] + fx * cam.dx[0] + fy * cam.dy[0], cam.corner[1] + fx * cam.dx[1] + fy * cam.dy[1],
            cam.corner[2] + fx * cam.dx[2] + fy * cam.dy[2]};
    const float scale = omax(fabsf(d.x), omax(fabsf(d.y), fabsf(d.z)));
    d = normalize(d / scale);
    F3 o = {cam.origin[0], cam.origin[1], cam.origin[2]};
    const float E = S.has_environment ? spectrum_eval(S.environment, lambda, S.spectra) : 0.0f;
    float throughput = 1.0f, escape_weight = 1.0f, prev_pdf = 0.0f;
    F3 prev_p = o, prev_n = F3{0, 0, 0};
    uint32_t crossings = 0;
    float L = 0.0f;
    for (;;) {
        Hit h;
        traverse<false>(S, st, o, d, __builtin_inff(), h);
        ++out.ext_rays;
        if (h.prim == kEmpty) {
            if (!demo && out.scatterings != 0)
                escape_weight = power_heuristic(
                    prev_pdf, S.env_light == kEmpty ? 0.0f : light_pmf(S, prev_p, prev_n, S.env_light) * kInv4Pi);
            L = L + throughput * escape_weight * E;
            break;
        }
        const uint32_t bounce = out.scatterings;
        const SurfacePoint sp = surface_point(S, h);
        const uint32_t obj_id = __float_as_uint(S.tris[3 * h.prim + 1].w);
        const rp_object obj = S.objects[obj_id];
        const rp_material &mat = S.materials[obj.material];
        if (kInterfaces && mat.kind == RP_MATERIAL_INTERFACE) {
            if (crossings++ >= P.boundary) { out.status = RP_PATH_BOUNDARY_EXHAUSTION; break; }
            o = offset_origin(sp.p, sp.err, sp.n, d);
            continue;
        }
        const uint32_t light_id = __float_as_uint(S.tris[3 * h.prim + 2].w);
        if (obj.emission != RP_NONE && light_id != RP_NONE) {
            const rp_emission &em = S.emissions[obj.emission];
            const float Le = spectrum_eval(em.radiance, lambda, S.spectra);
            if (Le > 0.0f && (em.two_sided || dot(sp.n, -d) >= 0.0f)) {
                float weight = 1.0f;
                if (bounce != 0)
                    weight = power_heuristic(prev_pdf, area_light_pdf(S, light_id, prev_p, prev_n, sp.p, sp.n));
                if (!__builtin_isfinite(weight)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
                L = L + throughput * (weight * Le);
            }
        }
        if (!demo && bounce >= S.max_depth) break;
        if (mat.kind != RP_MATERIAL_DIFFUSE) { out.status = RP_PATH_INVALID_HIT; break; }
        const float rho = spectrum_eval(mat.reflectance, lambda, S.spectra);
        if (!__builtin_isfinite(rho) || !(rho >= 0.0f && rho <= 1.0f)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
        if (rho == 0.0f) break;
        if (bounce >= P.emergency) { out.status = RP_PATH_EMERGENCY_DEPTH; break; }
        const F3 n = dot(sp.n, d) > 0.0f ? -sp.n : sp.n;
        uint4 bw, lw;
        if (independent) {
            const uint4 c = make_uint4(pixel, uint32_t(sample), uint32_t(sample >> 32), 0u);
            const uint2 key = make_uint2(seed_lo, seed_hi);
            lw = philox(make_uint4(c.x, c.y, c.z, 2 + 2 * bounce), key);
            bw = philox(make_uint4(c.x, c.y, c.z, 1 + 2 * bounce), key);
        } else {
            uint4 rb, rl;
            sobol_raw2(S, index, 1 + 2 * bounce, 2 + 2 * bounce, rb, rl);
            bw = sobol_scramble4(rb, 1 + 2 * bounce, P.mode, seed_lo);
            lw = sobol_scramble4(rl, 2 + 2 * bounce, P.mode, seed_lo);
        }
        if (demo) {
            const F3 ldir = sphere_dir(su(lw.x), su(lw.y));
            const float lc = dot(n, ldir);
            if (E > 0.0f && lc > 0.0f) {
                float tr[4];
                ++out.shadow_rays;
                if (transmittance<kInterfaces>(S, st, offset_origin(sp.p, sp.err, sp.n, ldir), ldir, __builtin_inff(),
                                               RP_NONE, wl, tr, P.boundary))
                    L = L + throughput * rho * ((4.0f * lc) * demo_light_weight(lc)) * E * tr[0];
            }
        } else {
            const LightSample ls = sample_light(S, sp.p, sp.n, wl, su(lw.x), su(lw.y), su(lw.z));
            if (ls.ok && ls.pdf > 0 && ls.le[0] > 0) {
                const float cosv = dot(n, ls.wi);
                if (cosv > 0) {
                    const float f_cos = rho * cosv / kPi, bsdf_pdf = cosv / kPi;
                    const float factor = scattering_direct(f_cos, bsdf_pdf, ls.pdf, ls.delta);
                    if (!__builtin_isfinite(factor)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
                    ShadowReq sr;
                    shadow_setup(ls, offset_origin(sp.p, sp.err, sp.n, ls.wi), RP_NONE, sr);
                    float tr[4];
                    ++out.shadow_rays;
                    if (transmittance<kInterfaces>(S, st, sr.o, sr.d, sr.tmax, RP_NONE, wl, tr, P.boundary))
                        L = L + throughput * factor * ls.le[0] * tr[0];
                }
            }
        }
        const F3 wi = hemisphere_dir(n, su(bw.x), su(bw.y));
        throughput = throughput * rho;
        if (!__builtin_isfinite(throughput) || throughput < 0.0f) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
        ++out.scatterings;
        if (out.scatterings >= 3) {
            if (bw.z >= 0xc0000000u) break;
            throughput = throughput / 0.75f;
        }
        const float cosine = dot(n, wi);
        if (!(cosine > 0.0f)) { out.status = RP_PATH_INVALID_ARITHMETIC; break; }
        prev_pdf = cosine / kPi;
        escape_weight = demo ? demo_bsdf_weight(cosine) : power_heuristic(prev_pdf, 1.0f / (4.0f * kPi));
        prev_p = sp.p;
        prev_n = sp.n;
        o = offset_origin(sp.p, sp.err, sp.n, wi);
        d = wi;
        crossings = 0;
    }
    out.radiance = L;
}

// The film weight of a scalar path: 470 CIE(round lambda) L, or for the demo
// model CIE linearly interpolated at lambda times 470 / Y L (Y = the
// trapezoid integral of y-bar, 106.857).
RP_DEV void scalar_xyz(const Scene &S, const ScalarResult &r, float xyz[3]) {
    xyz[0] = xyz[1] = xyz[2] = 0.0f;
    if (r.status != RP_PATH_SUCCESS) return;
    if (S.transport == RP_TRANSPORT_SCALAR_DEMO) {
        const float off = r.wavelength - 360.0f;
        int lo = int(off);
        lo = lo < 0 ? 0 : (lo > 470 ? 470 : lo);
        const int hi = lo + 1 < 470 ? lo + 1 : 470;
        const float f = off - float(lo);
        const float scale = 470.0f / 106.856914916766900f;
        const float4 a = S.cie[lo], b = S.cie[hi];
        xyz[0] = ((a.x + (b.x - a.x) * f) * scale) * r.radiance;
        xyz[1] = ((a.y + (b.y - a.y) * f) * scale) * r.radiance;
        xyz[2] = ((a.z + (b.z - a.z) * f) * scale) * r.radiance;
        return;
    }
    int i = int(floorf(r.wavelength + 0.5f)) - 360;
    i = i < 0 ? 0 : (i > 470 ? 470 : i);
    const float4 c = S.cie[i];
    const float w = 470.0f * r.radiance;
    xyz[0] = c.x * w;
    xyz[1] = c.y * w;
    xyz[2] = c.z * w;
}

} // namespace rpd
