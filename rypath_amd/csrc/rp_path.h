// Hero-wavelength path events on the device.  One event consumes the
// closest hit of the current ray and produces (a) the next ray and (b) at
// most one shadow request (next-event estimation).  The megakernel loops
// trace -> event -> shadow in registers; the wavefront kernels run the same
// functions over queues with the state in HBM.  Event order, sample
// dimensions and arithmetic follow the reference's packet event
// (shaders/packet_pt_event_body.inc) with the hero helpers of
// shaders/hero_pt_helpers.h and src/hero_spectrum_shared.h.
#pragma once
#include "rp_device.h"

namespace rpd {

struct PathParams {
    uint64_t seed;
    uint32_t mode; // Sobol randomization
    uint32_t emergency, boundary;
    uint32_t roulette; // 1: throughput roulette (the reference's transport); 0: off
};

constexpr uint32_t kFlagPrevDelta = 1u, kFlagRoulette = 2u;

struct PathState {
    F3 o, d;
    float tmax;
    float beta[4], density[4], prev_density[4], prev_pdf[4], eta_scale[4], L[4];
    F3 prev_p;
    uint32_t prev_nc; // the previous vertex's light-sampling normal (octahedral code; kNoNormal in media)
    uint32_t pixel;
    uint64_t sample, index;
    uint32_t w2;           // wavelength Sobol word (dimension 2)
    uint32_t count, crossings, medium, mdraw, flags, status, primary;
};

struct ShadowReq {
    bool active, delta_or_nomis;
    F3 o, d;
    float tmax;
    uint32_t medium;
    float val[4];
    float w;
};

// Event hooks for the path-replay adjoint.  The forward renderer uses NoHook,
// whose empty members compile away; the adjoint receives every contribution
// added to the path radiance and the log-derivatives of every factor.
struct NoHook {
    static constexpr bool active = false;
    RP_DEV void add(const float *) {}
    RP_DEV void medium_factor(uint32_t, const Wavelengths &, float, const MediumCoeff &) {}
    RP_DEV void medium_pass(uint32_t, const Wavelengths &, float) {}
    RP_DEV void continuation(const rp_material &, uint32_t, const Lane *, F3, F3, const Wavelengths &) {}
    RP_DEV void nee(const rp_material &, uint32_t, const Lane *, F3, F3, const Wavelengths &) {}
    RP_DEV void nee_medium() {}
};

RP_DEV Wavelengths path_wavelengths(const Scene &S, uint32_t w2) { return sample_wavelengths(sobol_uniform(w2), S.hero_mode); }

// The initial state of path (pixel, sample) apart from its direction, Sobol
// index and wavelength word (path_begin computes those).
RP_DEV void path_init(const Scene &S, const PathParams &P, uint32_t pixel, uint64_t sample, PathState &ps) {
    const rp_camera &cam = S.camera;
    ps.pixel = pixel;
    ps.sample = sample;
    ps.o = {cam.origin[0], cam.origin[1], cam.origin[2]};
    ps.tmax = __builtin_inff();
    for (int k = 0; k < 4; ++k) {
        ps.beta[k] = 1.0f;
        ps.density[k] = 1.0f;
        ps.prev_density[k] = 1.0f;
        ps.prev_pdf[k] = 0.0f;
        ps.eta_scale[k] = 1.0f;
        ps.L[k] = 0.0f;
    }
    ps.prev_p = ps.o;
    ps.prev_nc = kNoNormal;
    ps.count = 1;
    ps.crossings = 0;
    ps.medium = S.camera.medium;
    ps.mdraw = 0;
    ps.flags = P.roulette ? kFlagRoulette : 0u;
    ps.status = RP_PATH_SUCCESS;
    ps.primary = RP_NONE;
}

// Camera sample: Sobol index, camera words, wavelengths, primary ray
// (shaders/render.metal:238-312 begin_sample, :458-477 primary_ray, box filter).
RP_DEV void path_begin(const Scene &S, const PathParams &P, uint32_t pixel, uint64_t sample, PathState &ps,
                       uint32_t &w0, uint32_t &w1) {
    const uint32_t W = S.camera.width, H = S.camera.height;
    const uint32_t px = pixel % W, py = pixel / W;
    uint32_t seed_bits = uint32_t(P.seed);
    path_init(S, P, pixel, sample, ps);
    ps.index = sobol_index(S, px, py, sample);
    // dimensions 0-2 from one 16-byte direction load per index bit (as
    // sobol_block does); the camera words are unscrambled, the wavelength
    // word takes the frame's randomization
    const uint4 w = sobol_block(S, ps.index, 0, RP_RANDOMIZE_NONE, 0);
    w0 = w.x;
    w1 = w.y;
    ps.w2 = scramble(w.z, 2, P.mode, seed_bits);
    float ux = camera_uniform(w0, S.m, px), uy = camera_uniform(w1, S.m, py);
    float fx = (float(px) + ux) / float(W), fy = (float(py) + uy) / float(H);
    const rp_camera &cam = S.camera;
    F3 d = {cam.corner[0] + fx * cam.dx[0] + fy * cam.dy[0], cam.corner[1] + fx * cam.dx[1] + fy * cam.dy[1],
            cam.corner[2] + fx * cam.dx[2] + fy * cam.dy[2]};
    float scale = omax(fabsf(d.x), omax(fabsf(d.y), fabsf(d.z)));
    ps.d = normalize(d / scale);
}

// True while the path still needs its current ray traced.
RP_DEV bool path_alive(const Scene &S, const PathState &ps) {
    if (ps.status != RP_PATH_SUCCESS) return false;
    if (ps.count >= S.max_depth + 2) return false;
    return ps.beta[0] > 0 || ps.beta[1] > 0 || ps.beta[2] > 0 || ps.beta[3] > 0;
}

RP_DEV void add4(float L[4], const float v[4]) {
    for (int k = 0; k < 4; ++k) L[k] = L[k] + v[k];
}

// Builds the shadow request of a light sample taken from point p.
RP_DEV void shadow_setup(const LightSample &ls, F3 from, uint32_t medium, ShadowReq &sr) {
    if (ls.infinite) {
        sr.d = ls.wi;
        sr.tmax = __builtin_inff();
    } else {
        F3 to = ls.delta ? ls.p : offset_origin(ls.p, ls.err, ls.n, from - ls.p);
        sr.d = to - from;
        sr.tmax = 1.0f - 1e-4f;
    }
    sr.o = from;
    sr.medium = medium;
    sr.active = true;
}

// One path event.  Returns true when the path continues with ps.o/ps.d.
template <class Hook>
RP_DEV bool path_event(const Scene &S, const PathParams &P, PathState &ps, const Hit &h, ShadowReq &sr, Hook &hk) {
    sr.active = false;
    const Wavelengths wl = path_wavelengths(S, ps.w2);
    const uint32_t seed_bits = uint32_t(P.seed);
    const uint32_t capacity = S.max_depth + 2;
    if (ps.count == 1 && ps.crossings == 0) ps.primary = h.id;
    float seg = h.prim == kEmpty ? __builtin_inff() : h.t;
    bool prev_delta = (ps.flags & kFlagPrevDelta) != 0;
    const bool rr_enabled = (ps.flags & kFlagRoulette) != 0;

    if (ps.medium != RP_NONE) {
        MediumCoeff mc = medium_coeff(S, ps.medium, wl);
        uint4 mw = philox(make_uint4(ps.pixel, uint32_t(ps.sample), uint32_t(ps.sample >> 32), 0x20000000u | ps.mdraw),
                          make_uint2(uint32_t(P.seed), uint32_t(P.seed >> 32) ^ 0x6d656431u));
        ++ps.mdraw;
        float uf = philox_u01(mw.x);
        float tf = mc.st[0] > 0 ? -log_f(1.0f - uf) / mc.st[0] : __builtin_inff();
        if (tf < seg) {
            float num[4], pdf[4];
            for (int k = 0; k < 4; ++k) {
                float tr = exp_f(-mc.st[k] * tf);
                num[k] = mc.ss[k] * tr;
                pdf[k] = mc.st[k] * tr;
            }
            if (!advance(ps.density, ps.beta, num, pdf)) { ps.status = RP_PATH_INVALID_ARITHMETIC; return false; }
            if (Hook::active) hk.medium_factor(ps.medium, wl, tf, mc);
            for (int k = 0; k < 4; ++k) ps.prev_pdf[k] = ps.prev_pdf[k] * pdf[k];
            F3 mp = ps.o + tf * ps.d;
            ++ps.count;
            ps.crossings = 0;
            if (ps.count == capacity) return false;
            if (ps.count - 2 >= P.emergency) { ps.status = RP_PATH_EMERGENCY_DEPTH; return false; }
            uint4 draw, lraw; // scattering block 4 + depth, light block 133 + depth
            sobol_raw2(S, ps.index, 4 + ps.count - 2, 133 + ps.count - 2, draw, lraw);
            const uint4 dr = sobol_scramble4(draw, 4 + ps.count - 2, P.mode, seed_bits);
            float direct_density[4], direct_beta[4];
            for (int k = 0; k < 4; ++k) direct_density[k] = ps.density[k];
            bool direct_allowed = hero_continue(ps.beta, ps.eta_scale, rr_enabled, ps.count - 1, sobol_uniform(dr.w));
            for (int k = 0; k < 4; ++k) direct_beta[k] = ps.beta[k];
            F3 wo = -ps.d;
            if (!direct_allowed) return false;
            float ppdf;
            F3 wi = hg_sample(wo, mc.g, sobol_uniform(dr.y), sobol_uniform(dr.z), ppdf);
            for (int k = 0; k < 4; ++k) ps.prev_density[k] = ps.density[k];
            float pnum[4] = {ppdf, ppdf, ppdf, ppdf}, ppdfs[4] = {ppdf, ppdf, ppdf, ppdf};
            if (!advance(ps.density, ps.beta, pnum, ppdfs)) { ps.status = RP_PATH_INVALID_ARITHMETIC; return false; }
            for (int k = 0; k < 4; ++k) ps.prev_pdf[k] = ppdf;
            ps.flags &= ~kFlagPrevDelta;
            ps.prev_p = mp;
            ps.prev_nc = kNoNormal;
            const uint4 lw = sobol_scramble4(lraw, 133 + ps.count - 2, P.mode, seed_bits);
            LightSample ls = sample_light(S, mp, F3{0, 0, 0}, wl, sobol_uniform(lw.x), sobol_uniform(lw.y),
                                          sobol_uniform(lw.z));
            if (ls.ok) {
                float ph = hg(dot(wo, ls.wi), mc.g);
                bool any = false;
                for (int k = 0; k < 4; ++k) {
                    sr.val[k] = ((direct_beta[k] * ph) * (ls.le[k] / ls.pdf));
                    any = any || sr.val[k] > 0;
                }
                if (any) {
                    shadow_setup(ls, mp, ps.medium, sr);
                    if (Hook::active) hk.nee_medium();
                    sr.delta_or_nomis = ls.delta;
                    sr.w = 1.0f;
                    if (!ls.delta) {
                        float lp[4] = {ls.pdf, ls.pdf, ls.pdf, ls.pdf};
                        float pp[4] = {ph, ph, ph, ph};
                        float a = mixture(direct_density, lp), b = mixture(direct_density, pp);
                        sr.w = a / (a + b);
                    }
                }
            }
            ps.o = mp;
            ps.d = wi;
            ps.tmax = __builtin_inff();
            return true;
        }
        if (__builtin_isfinite(seg)) {
            float num[4], pdf[4];
            for (int k = 0; k < 4; ++k) {
                float tr = exp_f(-mc.st[k] * seg);
                num[k] = tr;
                pdf[k] = tr;
            }
            if (!advance(ps.density, ps.beta, num, pdf)) { ps.status = RP_PATH_INVALID_ARITHMETIC; return false; }
            if (Hook::active) hk.medium_pass(ps.medium, wl, seg);
            for (int k = 0; k < 4; ++k) ps.prev_pdf[k] = ps.prev_pdf[k] * pdf[k];
        } else {
            return false;
        }
    }

    if (h.prim == kEmpty) {
        if (S.has_environment) {
            float val[4];
            for (int k = 0; k < 4; ++k) val[k] = ps.beta[k] * spectrum_eval(S.environment, wl.lambda[k], S.spectra);
            if (!(ps.count == 1 || prev_delta)) {
                float lp = S.env_pdf;
                float lpv[4] = {lp, lp, lp, lp};
                float a = mixture(ps.prev_density, ps.prev_pdf), b = mixture(ps.prev_density, lpv);
                float w = a / (a + b);
                for (int k = 0; k < 4; ++k) val[k] = val[k] * w;
            }
            add4(ps.L, val);
            if (Hook::active) hk.add(val);
        }
        return false;
    }
    SurfacePoint sp = surface_point(S, h);
    const uint32_t obj_id = __float_as_uint(S.tris[3 * h.prim + 1].w);
    const rp_object obj = S.objects[obj_id];
    const rp_material &mat = S.materials[obj.material];
    if (mat.kind == RP_MATERIAL_INTERFACE) {
        if (ps.crossings++ >= P.boundary) { ps.status = RP_PATH_BOUNDARY_EXHAUSTION; return false; }
        ps.medium = outgoing_medium(obj, ps.medium, sp.n, ps.d);
        ps.o = offset_origin(sp.p, sp.err, sp.n, ps.d);
        ps.tmax = __builtin_inff();
        return true;
    }
    F3 wo = -ps.d;
    ++ps.count;
    ps.crossings = 0;
    const uint32_t light_id = __float_as_uint(S.tris[3 * h.prim + 2].w);
    if (obj.emission != RP_NONE && light_id != RP_NONE) {
        const rp_emission &em = S.emissions[obj.emission];
        float cos_l = dot(sp.n, wo);
        if (em.two_sided || cos_l > 0) {
            float val[4];
            for (int k = 0; k < 4; ++k) val[k] = ps.beta[k] * spectrum_eval(em.radiance, wl.lambda[k], S.spectra);
            if (!(ps.count == 2 || prev_delta)) {
                float lp = area_light_pdf(S, light_id, ps.prev_p, oct_decode(ps.prev_nc), sp.p, sp.n);
                float lpv[4] = {lp, lp, lp, lp};
                float a = mixture(ps.prev_density, ps.prev_pdf), b = mixture(ps.prev_density, lpv);
                float w = a / (a + b);
                for (int k = 0; k < 4; ++k) val[k] = val[k] * w;
            }
            add4(ps.L, val);
            if (Hook::active) hk.add(val);
        }
    }
    if (ps.count == capacity) return false;
    if (ps.count - 2 >= P.emergency) { ps.status = RP_PATH_EMERGENCY_DEPTH; return false; }
    Frame fr = frame_from_z(sp.n);
    F3 wol = fr.local(wo);
    Lane lanes[4];
    for (int k = 0; k < 4; ++k) lanes[k] = prepare_lane(S, mat, wl.lambda[k]);
    bool connectible = !lanes[0].smooth || lanes[0].kind == RP_MATERIAL_DIFFUSE;
    uint4 draw, lraw; // scattering block 4 + depth, light block 133 + depth
    sobol_raw2(S, ps.index, 4 + ps.count - 2, 133 + ps.count - 2, draw, lraw);
    const uint4 dr = sobol_scramble4(draw, 4 + ps.count - 2, P.mode, seed_bits);
    float direct_density[4], direct_beta[4];
    for (int k = 0; k < 4; ++k) direct_density[k] = ps.density[k];
    bool direct_allowed = hero_continue(ps.beta, ps.eta_scale, rr_enabled, ps.count - 1, sobol_uniform(dr.w));
    for (int k = 0; k < 4; ++k) direct_beta[k] = ps.beta[k];
    bool cont = direct_allowed;
    if (cont) {
        float uc = sobol_uniform(dr.x), u0 = sobol_uniform(dr.y), u1 = sobol_uniform(dr.z);
        Sampled hs = sample_lane(lanes[0], wol, uc, u0, u1);
        if (!hs.ok) cont = false;
        else {
            float num[4], pdf[4], etas[4];
            if (hs.delta) {
                for (int k = 0; k < 4; ++k) {
                    if (k == 0) {
                        float m, p;
                        dirac_lane(lanes[0], wol, hs.wi, u0, u1, m, p);
                        num[0] = m;
                        pdf[0] = hs.pdf;
                    } else dirac_lane(lanes[k], wol, hs.wi, u0, u1, num[k], pdf[k]);
                    etas[k] = hs.transmission ? (wol.z > 0 ? lanes[k].eta : 1.0f / lanes[k].eta) : 1.0f;
                }
            } else {
                Eval ev[4];
                eval_hero(lanes, wol, hs.wi, ev);
                for (int k = 0; k < 4; ++k) {
                    num[k] = ev[k].f * fabsf(hs.wi.z);
                    pdf[k] = ev[k].pdf;
                    etas[k] = hs.transmission ? (wol.z > 0 ? lanes[k].eta : 1.0f / lanes[k].eta) : 1.0f;
                }
                pdf[0] = hs.pdf;
            }
            for (int k = 0; k < 4; ++k) ps.prev_density[k] = direct_density[k];
            if (!advance(ps.density, ps.beta, num, pdf)) {
                ps.status = RP_PATH_INVALID_ARITHMETIC;
                cont = false;
            } else {
                if (Hook::active && !hs.delta) hk.continuation(mat, obj.material, lanes, wol, hs.wi, wl);
                for (int k = 0; k < 4; ++k) ps.prev_pdf[k] = pdf[k];
                if (hs.transmission && rr_enabled)
                    for (int k = 0; k < 4; ++k) ps.eta_scale[k] = ps.eta_scale[k] * (etas[k] * etas[k]);
                if (hs.delta) ps.flags |= kFlagPrevDelta;
                else ps.flags &= ~kFlagPrevDelta;
                ps.prev_p = sp.p;
                ps.prev_nc = oct_encode(sp.n);
                F3 wi = normalize(fr.world(hs.wi));
                ps.medium = outgoing_medium(obj, ps.medium, sp.n, wi);
                ps.o = offset_origin(sp.p, sp.err, sp.n, wi);
                ps.d = wi;
                ps.tmax = __builtin_inff();
            }
        }
    }
    if (ps.status == RP_PATH_SUCCESS && direct_allowed && connectible) {
        const uint4 lw = sobol_scramble4(lraw, 133 + ps.count - 2, P.mode, seed_bits);
        LightSample ls = sample_light(S, sp.p, oct_decode(oct_encode(sp.n)), wl, sobol_uniform(lw.x),
                                      sobol_uniform(lw.y), sobol_uniform(lw.z));
        if (ls.ok) {
            F3 wil = fr.local(ls.wi);
            float cosv = fabsf(dot(sp.n, ls.wi));
            float bp[4];
            bool any = false;
            Eval ev[4];
            eval_hero(lanes, wol, wil, ev);
            for (int k = 0; k < 4; ++k) {
                bp[k] = ev[k].pdf;
                sr.val[k] = ((direct_beta[k] * ev[k].f) * (ls.le[k] / ls.pdf)) * cosv;
                any = any || sr.val[k] > 0;
            }
            if (any) {
                F3 from = offset_origin(sp.p, sp.err, sp.n, ls.wi);
                shadow_setup(ls, from, outgoing_medium(obj, ps.medium, sp.n, ls.wi), sr);
                if (Hook::active) hk.nee(mat, obj.material, lanes, wol, wil, wl);
                sr.delta_or_nomis = ls.delta;
                sr.w = 1.0f;
                if (!ls.delta) {
                    float lp[4] = {ls.pdf, ls.pdf, ls.pdf, ls.pdf};
                    float a = mixture(direct_density, lp), b = mixture(direct_density, bp);
                    sr.w = a / (a + b);
                }
            }
        }
    }
    return cont;
}

// Applies a resolved shadow request: L += (val * T) * w when visible.
RP_DEV void shadow_apply(PathState &ps, const ShadowReq &sr, bool visible, const float tr[4]) {
    if (!visible) return;
    float val[4];
    for (int k = 0; k < 4; ++k) val[k] = sr.val[k] * tr[k];
    if (!sr.delta_or_nomis)
        for (int k = 0; k < 4; ++k) val[k] = val[k] * sr.w;
    add4(ps.L, val);
}

struct PathResult {
    uint32_t status;
    float L[4];
    Wavelengths wl;
    uint64_t index;
    uint32_t primary, vertices, w0, w1;
    uint32_t ext_rays, shadow_rays;
};

// Megakernel form: the whole path in registers.
// Runs the path in ps from its current ray to its end: per bounce the
// closest hit, the event and its shadow ray, in the order the wavefront
// kernels apply them (the shadow contribution after the event).  Adds the
// extension and shadow rays traced to ext and shr.  kInterfaces = false when
// the scene has no null interfaces.
template <bool kInterfaces>
RP_DEV void run_path(const Scene &S, const PathParams &P, Stack &st, PathState &ps, const Wavelengths &wl,
                     uint32_t &ext, uint32_t &shr) {
    for (;;) {
        Hit h;
        traverse<false>(S, st, ps.o, ps.d, ps.tmax, h);
        ++ext;
        ShadowReq sr;
        NoHook nh;
        const bool cont = path_event(S, P, ps, h, sr, nh) && path_alive(S, ps);
        if (sr.active) {
            ++shr;
            float tr[4];
            const bool vis = transmittance<kInterfaces>(S, st, sr.o, sr.d, sr.tmax, sr.medium, wl, tr, P.boundary);
            shadow_apply(ps, sr, vis, tr);
        }
        if (!cont) break;
    }
}

RP_DEV void trace_path(const Scene &S, const PathParams &P, Stack &st, uint32_t pixel, uint64_t sample,
                       PathResult &out) {
    PathState ps;
    path_begin(S, P, pixel, sample, ps, out.w0, out.w1);
    out.ext_rays = out.shadow_rays = 0;
    const Wavelengths wl = path_wavelengths(S, ps.w2);
    if (path_alive(S, ps)) run_path<true>(S, P, st, ps, wl, out.ext_rays, out.shadow_rays);
    out.status = ps.status;
    for (int k = 0; k < 4; ++k) out.L[k] = ps.L[k];
    out.wl = wl;
    out.index = ps.index;
    out.primary = ps.primary;
    out.vertices = ps.count;
}

} // namespace rpd
